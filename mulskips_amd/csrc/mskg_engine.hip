// mskg_engine.hip -- MI355X KMC event engine: kernels and the C ABI of
// include/mskg.h.
//
// Kernels
//   k_rate_sweep   Get_Prob over every lattice point (HBM-bound stream of the
//                  packed words + neighbour gathers); the rate-update sweep.
//   k_replay       applies a logged event sequence with the reference's
//                  update rules (one wave; events are strictly ordered).
//   k_tile_scan    per-tile rate totals and active flags (start of a run).
//   k_pass         one colour of the synchronous-sublattice sweep: each
//                  workgroup (one 64-wide wave) runs rejection-free KMC inside
//                  one tile for the window tau, selecting events with a
//                  wavefront prefix scan over the tile's FP32 rates held in
//                  LDS, FP64 partial sums and Philox-4x32-10 draws.
//   k_sweep_end    FP64 reductions of tile rates -> next tau and the clock.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstdarg>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mskg.h"
#include "mskg_events.h"

using namespace mskg;

namespace {

constexpr int WAVE = 64;

// Device-resident run state (one per engine).
struct RunState {
  double tau;            // current window
  double clock;          // KMC time (FP64)
  double rtot;           // total rate estimate (isolated-atom rule)
  double rmax;           // largest tile rate of the last sweep
  double events_per_pass;
  unsigned long long events;
  unsigned long long counts[8];
  unsigned long long sweeps;
  int err;
  int err_site[3];
  int stop;
};

struct PassParams {
  Lattice L;
  RunState *st;
  double *tile_r;
  const int32_t *tiles;  // tiles of this colour
  int ntiles;
  int ntz;
  int cap;
  uint32_t key0, key1;
  int colour;
};

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
  return v;
}
__device__ inline double wave_incl_scan(double v, int lane) {
#pragma unroll
  for (int off = 1; off < WAVE; off <<= 1) {
    double t = __shfl_up(v, off, WAVE);
    if (lane >= off) v += t;
  }
  return v;
}

struct TileGeom {
  int tx, ty, tz, z0, nz, half;
  int64_t npt;
};
__device__ inline TileGeom tile_geom(int t, int T, int ntx, int nty, int lz) {
  TileGeom g;
  g.tx = t % ntx;
  g.ty = (t / ntx) % nty;
  g.tz = t / (ntx * nty);
  g.z0 = g.tz * T;
  g.nz = min(T, lz + 1 - g.z0);
  g.half = T / 2;
  g.npt = (int64_t)g.half * g.half * g.nz;
  return g;
}
__device__ inline P3 tile_point(const Box &b, const TileGeom &g, int T, int k) {
  int pp = g.half * g.half;
  int zl = k / pp, r = k - zl * pp;
  int j = r / g.half, i = r - j * g.half;
  P3 p;
  p.z = g.z0 + zl;
  int par = p.z & 1;
  p.x = g.tx * T + 2 * i + par;
  p.y = g.ty * T + 2 * j + par;
  p.i = pidx(b, p.x, p.y, p.z);
  return p;
}
__device__ inline int tile_local(const TileGeom &g, int T, const P3 &p) {
  if (p.i < 0) return -1;
  if (p.x / T != g.tx || p.y / T != g.ty || p.z < g.z0 || p.z >= g.z0 + g.nz) return -1;
  return (p.z - g.z0) * g.half * g.half + ((p.y - g.ty * T) >> 1) * g.half + ((p.x - g.tx * T) >> 1);
}

// ---- rate-update sweep ----------------------------------------------------
// Every point: an MC particle is an empty or coverage site with coordination
// 1..3 or an atom with coordination <= 3 (Get_Prob.f:54-331 is evaluated for
// it); everything else gets class 0.
__global__ void __launch_bounds__(256) k_rate_sweep(Lattice L, double rtot, int *err) {
  const int64_t n = L.box.npts;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
    const Site sp = L.site[p];
    int kind = s_kind(sp), coor = s_coor(sp);
    bool mc = (kind == KIND_EMPTY && coor >= 1 && coor <= 3) || (kind >= KIND_COV0 && coor <= 3) ||
              (kind_species(kind) && coor <= 3);
    int cls = 0;
    double r = 0.0;
    if (mc) {
      P3 pt;
      pt.i = p;
      pxyz(L.box, p, pt.x, pt.y, pt.z);
      int e = ERR_NONE;
      get_prob(L, rtot, pt, sp, coor, list_of(L, pt, sp), cls, r, e);
      if (e) atomicExch(err, e);
    }
    L.rate[p] = cls ? (float)r : 0.0f;
    L.cls[p] = (uint8_t)cls;
  }
}

// ---- event-log replay -------------------------------------------------------
__global__ void __launch_bounds__(WAVE) k_replay(Lattice L, const mskg_event *log, int64_t n, int clock_mode,
                                                 RunState *st) {
  if (threadIdx.x != 0) return;
  double clock = st->clock;
  for (int64_t k = 0; k < n; k++) {
    const mskg_event ev = log[k];
    P3 s = at(L, ev.site[0], ev.site[1], ev.site[2]);
    if (s.i < 0 || L.cls[s.i] == 0) {
      st->err = ERR_NOT_MC;
      st->err_site[0] = s.x; st->err_site[1] = s.y; st->err_site[2] = s.z;
      st->events = (unsigned long long)k;
      return;
    }
    if (clock_mode == 0) clock = clock + 1.0 / ev.rtot;
    Choice ch;
    ch.arm = ev.arm;
    ch.u_arm = 0.0;
    EvResult res;
    execute(L, ev.rtot, s, ev.itrans, ch, res);
    if (res.err) {
      st->err = res.err;
      st->err_site[0] = s.x; st->err_site[1] = s.y; st->err_site[2] = s.z;
      st->events = (unsigned long long)k;
      return;
    }
    int slot = ev.itrans <= 3 ? ev.itrans - 1 : (ev.itrans == 4 ? 3 : (ev.itrans == 5 ? 4 : ev.itrans - 1));
    if (slot >= 0 && slot < 8) st->counts[slot]++;
  }
  st->events += (unsigned long long)n;
  st->clock = clock;
}

// ---- synchronous sublattice -------------------------------------------------
__global__ void __launch_bounds__(256) k_tile_scan(Lattice L, int ntiles, int ntz, double *tile_r) {
  // one wave per tile
  int lane = threadIdx.x & (WAVE - 1);
  int t = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
  if (t >= ntiles) return;
  TileGeom g = tile_geom(t, L.tile, L.ntx, L.nty, L.box.lz);
  double acc = 0.0;
  int any = 0;
  for (int k = lane; k < g.npt; k += WAVE) {
    P3 p = tile_point(L.box, g, L.tile, k);
    uint8_t c = L.cls[p.i];
    if (c) {
      any = 1;
      if (c != CLS_ISOLATED) acc += (double)L.rate[p.i];
    }
  }
  acc = wave_sum(acc);
  any = __any(any);
  if (lane == 0) {
    tile_r[t] = acc;
    L.tile_active[t] = any ? 1 : 0;
  }
}

__global__ void __launch_bounds__(WAVE) k_pass(PassParams P) {
  extern __shared__ unsigned char smem[];
  const int lane = threadIdx.x;
  const Lattice &L = P.L;
  const int cap = P.cap;
  int32_t *lk = (int32_t *)smem;                  // tile-local point offsets
  float *lr = (float *)(lk + cap);                // rates
  uint8_t *lc = (uint8_t *)(lr + cap);            // classes
  __shared__ int s_n, s_sel, s_stop, s_nt;
  __shared__ int s_tk[5];
  __shared__ float s_tr[5];
  __shared__ uint8_t s_tc[5];

  if (blockIdx.x >= P.ntiles) return;
  const int t = P.tiles[blockIdx.x];
  RunState *st = P.st;
  const double tau = st->tau;
  if (!(tau > 0.0) || st->err) return;
  if (!L.tile_active[t]) {
    if (lane == 0) P.tile_r[t] = 0.0;
    return;
  }
  TileGeom g = tile_geom(t, L.tile, L.ntx, L.nty, L.box.lz);

  // 1. collect the tile's MC particles into LDS (ballot + prefix compaction)
  int n = 0;
  bool overflow = false;
  for (int k0 = 0; k0 < g.npt; k0 += WAVE) {
    int k = k0 + lane;
    uint8_t c = 0;
    float r = 0.0f;
    if (k < g.npt) {
      P3 p = tile_point(L.box, g, L.tile, k);
      c = L.cls[p.i];
      if (c) r = L.rate[p.i];
    }
    unsigned long long m = __ballot(c != 0);
    int pre = __popcll(m & ((1ull << lane) - 1ull));
    if (c) {
      int slot = n + pre;
      if (slot < cap) {
        lk[slot] = k;
        lr[slot] = r;
        lc[slot] = c;
      }
    }
    n += __popcll(m);
    if (n > cap) { overflow = true; break; }
  }
  if (overflow) {
    if (lane == 0) {
      st->err = ERR_CAPACITY;
      P3 p = tile_point(L.box, g, L.tile, 0);
      st->err_site[0] = p.x; st->err_site[1] = p.y; st->err_site[2] = p.z;
    }
    return;
  }
  if (n == 0) {
    if (lane == 0) {
      L.tile_active[t] = 0;
      P.tile_r[t] = 0.0;
    }
    return;
  }

  // 2. rejection-free KMC inside the tile for the window tau
  const double rtot = st->rtot;
  double tloc = 0.0;
  unsigned long long nev = 0, cnt[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  U4 ctr;
  ctr.x = (uint32_t)t;
  ctr.y = (uint32_t)st->sweeps;
  ctr.z = (uint32_t)(st->sweeps >> 32) ^ ((uint32_t)P.colour << 29);
  ctr.w = 0;
  int err = 0;
  // every wave ends its pass: the window tau bounds the expected number of
  // events, and this cap bounds the worst case
  for (int guard = 0; guard < (1 << 16); guard++) {
    double part = 0.0;
    for (int j = lane; j < n; j += WAVE) part += (double)lr[j];
    const double R = wave_sum(part);
    if (!(R > 0.0)) break;
    U4 d = philox4x32_10(ctr, P.key0, P.key1);
    ctr.w++;
    double u1 = u01_open(d.x, d.y);
    double dt = -log(u1) / R;
    if (tloc + dt > tau) break;
    tloc += dt;
    const double target = u01_open(d.z, d.w) * R;
    // select the first particle whose inclusive prefix exceeds target
    double acc = 0.0;
    int sel = -1, last_nz = -1;
    for (int base = 0; base < n; base += WAVE) {
      int j = base + lane;
      double v = j < n ? (double)lr[j] : 0.0;
      unsigned long long nzm = __ballot(v > 0.0);
      if (nzm) last_nz = base + 63 - __clzll((long long)nzm);
      double inc = wave_incl_scan(v, lane) + acc;
      unsigned long long hit = __ballot(j < n && v > 0.0 && inc > target);
      if (hit) {
        sel = base + __ffsll((long long)hit) - 1;
        break;
      }
      acc = __shfl(inc, WAVE - 1, WAVE);
    }
    if (sel < 0) sel = last_nz;  // rounding at the top end of the scan
    if (sel < 0) break;
    // execute on one lane
    if (lane == 0) {
      P3 s = tile_point(L.box, g, L.tile, lk[sel]);
      U4 d2 = philox4x32_10(ctr, P.key0, P.key1);
      Choice ch;
      ch.arm = -1;
      ch.u_arm = u01_open(d2.z, d2.w);
      int itrans = pick_itrans(L, s, lc[sel], u01_open(d2.x, d2.y));
      EvResult res;
      execute(L, rtot, s, itrans, ch, res);
      err = res.err;
      int slot = itrans <= 3 ? itrans - 1 : (itrans == 4 ? 3 : (itrans == 5 ? 4 : itrans - 1));
      if (slot >= 0 && slot < 8) cnt[slot]++;
      int nt = 0;
      for (int q = 0; q < res.ntouched; q++) {
        int kl = tile_local(g, L.tile, res.touched[q]);
        if (kl < 0) continue;
        s_tk[nt] = kl;
        s_tr[nt] = L.rate[res.touched[q].i];
        s_tc[nt] = L.cls[res.touched[q].i];
        nt++;
      }
      s_nt = nt;
      s_stop = err;
    }
    ctr.w++;
    nev++;
    __syncthreads();
    if (s_stop) {
      err = s_stop;
      break;
    }
    // refresh the on-chip list for touched points inside the tile
    const int nt = s_nt;
    for (int q = 0; q < nt; q++) {
      const int kl = s_tk[q];
      int found = -1;
      for (int base = 0; base < n; base += WAVE) {
        int j = base + lane;
        unsigned long long h = __ballot(j < n && lk[j] == kl);
        if (h) { found = base + __ffsll((long long)h) - 1; break; }
      }
      if (lane == 0) {
        if (found >= 0) {
          lr[found] = s_tr[q];
          lc[found] = s_tc[q];
          s_n = n;
        } else if (s_tc[q]) {
          if (n < cap) {
            lk[n] = kl;
            lr[n] = s_tr[q];
            lc[n] = s_tc[q];
            s_n = n + 1;
          } else {
            s_n = -1;
          }
        } else {
          s_n = n;
        }
      }
      __syncthreads();
      if (s_n < 0) { err = ERR_CAPACITY; break; }
      n = s_n;
    }
    if (err) break;
  }
  // 3. tile totals for the next window; counters
  double part = 0.0;
  int anyc = 0;
  for (int j = lane; j < n; j += WAVE) {
    if (lc[j]) anyc = 1;
    if (lc[j] != CLS_ISOLATED) part += (double)lr[j];
  }
  const double R = wave_sum(part);
  anyc = __any(anyc);
  if (lane == 0) {
    P.tile_r[t] = R;
    if (!anyc) L.tile_active[t] = 0;
    if (nev) {
      atomicAdd(&st->events, nev);
      for (int q = 0; q < 8; q++)
        if (cnt[q]) atomicAdd(&st->counts[q], cnt[q]);
    }
    if (err) {
      st->err = err;
      P3 p = tile_point(L.box, g, L.tile, 0);
      st->err_site[0] = p.x; st->err_site[1] = p.y; st->err_site[2] = p.z;
    }
  }
}

// After the eight colour passes: the clock advances by the window just used
// (operator splitting of the master equation over the colours), and the next
// window is events_per_pass / (largest tile rate).
__global__ void __launch_bounds__(1024) k_sweep_end(RunState *st, const double *tile_r, int ntiles) {
  __shared__ double smax[1024], ssum[1024];
  double m = 0.0, s = 0.0;
  for (int t = threadIdx.x; t < ntiles; t += blockDim.x) {
    double r = tile_r[t];
    m = fmax(m, r);
    s += r;
  }
  smax[threadIdx.x] = m;
  ssum[threadIdx.x] = s;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + off]);
      ssum[threadIdx.x] += ssum[threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (st->tau > 0.0 && !st->err) st->clock += st->tau;
    st->sweeps++;
    st->rmax = smax[0];
    if (ssum[0] > 0.0) st->rtot = ssum[0];
    st->tau = smax[0] > 0.0 ? st->events_per_pass / smax[0] : 0.0;
    if (!(smax[0] > 0.0)) st->stop = MSKG_STOP_NO_EVENTS;
  }
}

// ---- host side ---------------------------------------------------------------
#define HIPCHK(e, x)                                                                  \
  do {                                                                                \
    hipError_t _r = (x);                                                              \
    if (_r != hipSuccess) return fail_((e), MSKG_EDEVICE, "%s: %s", #x, hipGetErrorString(_r)); \
  } while (0)

}  // namespace

struct mskg_engine {
  mskg_config cfg;
  Box box;
  Geometry geo;
  Tables tab;
  int ncrystal, ncov, list_cov[3];
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // device memory
  Site *d_site = nullptr;
  float *d_rate = nullptr;
  uint8_t *d_cls = nullptr;
  uint8_t *d_tile_active = nullptr;
  double *d_tile_r = nullptr;
  int32_t *d_tiles = nullptr;  // concatenated per-colour tile lists
  Geometry *d_geo = nullptr;
  Tables *d_tab = nullptr;
  RunState *d_st = nullptr;
  int *d_err = nullptr;
  int tile = 0, ntx = 0, nty = 0, ntz = 0, ntiles = 0, cap = 1024;
  std::vector<int> colour_off, colour_n;
  double run_ms = 0, sweep_ms = 0;
  int64_t sweep_points = 0, sweeps_done = 0;
  bool loaded = false;
  std::string err;
};

static int fail_(mskg_engine *e, int code, const char *fmt, ...) {
  if (e) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    e->err = buf;
  }
  return code;
}

// cos/sin of DFLOAT(2*irot)*pi/3. with the REAL(4) pi of FIND_NN.f:27,
// evaluated with the host C library like the reference; the device uses
// these values rather than its own cos/sin.
static void build_geometry(Geometry &g) {
  const double pi = (double)3.14159265358979323f;
  for (int irot = 1; irot <= 2; irot++) {
    double ang = (double)(2 * irot) * pi / (double)3.0f;
    g.cth[irot - 1] = std::cos(ang);
    g.sth[irot - 1] = std::sin(ang);
  }
  g.arm_shift = (double)8.6602254037f;
}

static Lattice make_lattice(mskg_engine *e) {
  Lattice L;
  L.site = e->d_site;
  L.rate = e->d_rate;
  L.cls = e->d_cls;
  L.tile_active = e->d_tile_active;
  L.box = e->box;
  L.geo = e->d_geo;
  L.tab = e->d_tab;
  L.tile = e->tile;
  L.ntx = e->ntx;
  L.nty = e->nty;
  return L;
}

static int choose_tile(int lx, int ly, int want) {
  auto ok = [&](int t) {
    return t >= 16 && t % 2 == 0 && lx % t == 0 && ly % t == 0 && (lx / t) % 2 == 0 && (ly / t) % 2 == 0;
  };
  if (want) return ok(want) ? want : -1;
  const int pref[] = {24, 20, 18, 28, 30, 32, 36, 40, 42, 48, 16, 22, 26, 34, 38, 44, 46};
  for (int t : pref)
    if (ok(t)) return t;
  for (int t = 16; t <= std::min(lx, ly) / 2; t += 2)
    if (ok(t)) return t;
  return -1;
}

static const char *err_name(int c) {
  switch (c) {
    case ERR_COORD: return "coordination inconsistent with neighbour counts (Get_Prob.f:293-298)";
    case ERR_GEOM: return "neighbour displacement does not fit the site record";
    case ERR_NOT_MC: return "logged event on a site that is not an MC particle";
    case ERR_COVERAGE: return "coverage species not in ListCov";
    case ERR_CAPACITY: return "tile holds more MC particles than tile_capacity";
    case ERR_TEMPERATURE: return "zero temperature in laser-annealing mode";
    default: return "unknown";
  }
}

// Map one LattCoo word (defsystem.f:87-91) to a site kind.
static int word_kind(const mskg_engine *e, int32_t v, int *kind) {
  int cov = v & 255, occ = (v >> 23) & 1, sic = (v >> 27) & 7;
  if (cov == 113) { *kind = KIND_WALL; return 0; }
  if (cov != 0) {
    if (occ) return -1;
    for (int k = 0; k < e->ncov; k++)
      if (e->list_cov[k] == cov) { *kind = KIND_COV0 + k; return 0; }
    return -2;
  }
  if (occ) {
    if (sic < 1 || sic > 3) return -3;
    *kind = sic;
    return 0;
  }
  if (sic) return -4;
  *kind = KIND_EMPTY;
  return 0;
}

static int read_state(mskg_engine *e, RunState &st) {
  HIPCHK(e, hipMemcpyAsync(&st, e->d_st, sizeof st, hipMemcpyDeviceToHost, e->stream));
  HIPCHK(e, hipStreamSynchronize(e->stream));
  return MSKG_OK;
}

extern "C" {

const char *mskg_last_error(const mskg_engine *e) { return e ? e->err.c_str() : "null engine"; }

int mskg_create(const mskg_config *cfg, mskg_engine **out) {
  if (!cfg || !out) return MSKG_EINVAL;
  *out = nullptr;
  mskg_engine *e = new (std::nothrow) mskg_engine();
  if (!e) return MSKG_ENOMEM;
  *out = e;
  e->cfg = *cfg;
  if (cfg->len_x <= 0 || cfg->len_y <= 0 || cfg->len_z <= 0 || cfg->len_x % 12 || cfg->len_y % 12)
    return fail_(e, MSKG_EINVAL, "box %d x %d x %d: len_x and len_y must be positive multiples of 12", cfg->len_x,
                 cfg->len_y, cfg->len_z);
  if (cfg->len_z > 8190) return fail_(e, MSKG_EINVAL, "len_z must be <= 8190");
  if (cfg->n_crystal < 1 || cfg->n_crystal > 3 || cfg->n_cov < 0 || cfg->n_cov > 3)
    return fail_(e, MSKG_EINVAL, "n_crystal must be 1..3 and n_cov 0..3");
  if (!cfg->ptrans_e || !cfg->ptrans_d) return fail_(e, MSKG_EINVAL, "ptrans_e and ptrans_d are required");
  if (cfg->n_cov > 0 && (!cfg->ptrans_abs || !cfg->ptrans_des))
    return fail_(e, MSKG_EINVAL, "ptrans_abs and ptrans_des are required when n_cov > 0");
  e->box = make_box(cfg->len_x, cfg->len_y, cfg->len_z);
  e->ncrystal = cfg->n_crystal;
  e->ncov = cfg->n_cov;
  for (int i = 0; i < 3; i++) e->list_cov[i] = cfg->list_cov[i];
  memset(&e->tab, 0, sizeof e->tab);
  memcpy(e->tab.pe, cfg->ptrans_e, sizeof e->tab.pe);
  memcpy(e->tab.pd, cfg->ptrans_d, sizeof e->tab.pd);
  if (cfg->ptrans_abs) memcpy(e->tab.pabs, cfg->ptrans_abs, sizeof e->tab.pabs);
  if (cfg->ptrans_des) memcpy(e->tab.pdes, cfg->ptrans_des, sizeof e->tab.pdes);
  e->tab.ptrans_zig = cfg->ptrans_zig;
  e->tab.ncrystal = cfg->n_crystal;
  e->tab.ncov = cfg->n_cov;
  e->tab.la_mode = cfg->la_mode;
  e->tab.p0 = cfg->p0;
  e->tab.tm = cfg->tm;
  e->tab.sigma = cfg->sigma;
  e->tab.tflex = cfg->tflex;
  build_geometry(e->geo);
  e->tile = choose_tile(cfg->len_x, cfg->len_y, cfg->tile);
  if (e->tile < 0)
    return fail_(e, MSKG_EINVAL,
                 "no valid tile edge for box %d x %d (requested %d): need an even tile >= 16 with an even "
                 "number of tiles along x and y",
                 cfg->len_x, cfg->len_y, cfg->tile);
  e->cap = cfg->tile_capacity > 0 ? cfg->tile_capacity : 1024;
  e->ntx = cfg->len_x / e->tile;
  e->nty = cfg->len_y / e->tile;
  e->ntz = (cfg->len_z + 1 + e->tile - 1) / e->tile;
  e->ntiles = e->ntx * e->nty * e->ntz;

  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail_(e, MSKG_EDEVICE, "no HIP device available");
  HIPCHK(e, hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
  HIPCHK(e, hipEventCreate(&e->ev0));
  HIPCHK(e, hipEventCreate(&e->ev1));
  const int64_t n = e->box.npts;
  HIPCHK(e, hipMalloc(&e->d_site, n * sizeof(Site)));
  HIPCHK(e, hipMalloc(&e->d_rate, n * sizeof(float)));
  HIPCHK(e, hipMalloc(&e->d_cls, n * sizeof(uint8_t)));
  HIPCHK(e, hipMalloc(&e->d_tile_active, e->ntiles));
  HIPCHK(e, hipMalloc(&e->d_tile_r, e->ntiles * sizeof(double)));
  HIPCHK(e, hipMalloc(&e->d_tiles, e->ntiles * sizeof(int32_t)));
  HIPCHK(e, hipMalloc(&e->d_geo, sizeof(Geometry)));
  HIPCHK(e, hipMalloc(&e->d_tab, sizeof(Tables)));
  HIPCHK(e, hipMalloc(&e->d_st, sizeof(RunState)));
  HIPCHK(e, hipMalloc(&e->d_err, sizeof(int)));
  HIPCHK(e, hipMemcpyAsync(e->d_geo, &e->geo, sizeof(Geometry), hipMemcpyHostToDevice, e->stream));
  HIPCHK(e, hipMemcpyAsync(e->d_tab, &e->tab, sizeof(Tables), hipMemcpyHostToDevice, e->stream));
  HIPCHK(e, hipMemsetAsync(e->d_site, 0, n * sizeof(Site), e->stream));
  HIPCHK(e, hipMemsetAsync(e->d_rate, 0, n * sizeof(float), e->stream));
  HIPCHK(e, hipMemsetAsync(e->d_cls, 0, n, e->stream));
  HIPCHK(e, hipMemsetAsync(e->d_tile_active, 0, e->ntiles, e->stream));
  HIPCHK(e, hipMemsetAsync(e->d_tile_r, 0, e->ntiles * sizeof(double), e->stream));
  // tile lists per colour (colour = parity of the tile coordinates)
  std::vector<int32_t> tl;
  e->colour_off.assign(8, 0);
  e->colour_n.assign(8, 0);
  for (int c = 0; c < 8; c++) {
    e->colour_off[c] = (int)tl.size();
    for (int t = 0; t < e->ntiles; t++) {
      int tx = t % e->ntx, ty = (t / e->ntx) % e->nty, tz = t / (e->ntx * e->nty);
      if (((tx & 1) | ((ty & 1) << 1) | ((tz & 1) << 2)) == c) tl.push_back(t);
    }
    e->colour_n[c] = (int)tl.size() - e->colour_off[c];
  }
  HIPCHK(e, hipMemcpyAsync(e->d_tiles, tl.data(), tl.size() * sizeof(int32_t), hipMemcpyHostToDevice, e->stream));
  RunState st;
  memset(&st, 0, sizeof st);
  HIPCHK(e, hipMemcpyAsync(e->d_st, &st, sizeof st, hipMemcpyHostToDevice, e->stream));
  HIPCHK(e, hipStreamSynchronize(e->stream));
  return MSKG_OK;
}

void mskg_destroy(mskg_engine *e) {
  if (!e) return;
  if (e->stream) hipStreamSynchronize(e->stream);
  hipFree(e->d_site); hipFree(e->d_rate); hipFree(e->d_cls);
  hipFree(e->d_tile_active); hipFree(e->d_tile_r); hipFree(e->d_tiles);
  hipFree(e->d_geo); hipFree(e->d_tab); hipFree(e->d_st); hipFree(e->d_err);
  if (e->ev0) hipEventDestroy(e->ev0);
  if (e->ev1) hipEventDestroy(e->ev1);
  if (e->stream) hipStreamDestroy(e->stream);
  delete e;
}

int mskg_load_state(mskg_engine *e, const int32_t *latt, int32_t n_ad, const int32_t *ad_xyz, const int32_t *ad_nn,
                    const int32_t *ad_event, const double *ad_prob, int32_t n_bulk, const int32_t *bulk_xyz,
                    const int32_t *bulk_nn, int32_t n_void, const int32_t *void_xyz, const int32_t *void_nn,
                    double kmc_time) {
  if (!e || !latt) return MSKG_EINVAL;
  if (!e->d_site) return fail_(e, MSKG_EINVAL, "engine not created");
  const Box &b = e->box;
  std::vector<Site> site((size_t)b.npts);
  memset(site.data(), 0, site.size() * sizeof(Site));
  std::vector<float> rate((size_t)b.npts, 0.0f);
  std::vector<uint8_t> cls((size_t)b.npts, 0);
  for (int z = 0; z <= b.lz; z++)
    for (int y = 0; y < b.ly; y++)
      for (int x = (z ^ y) & 1 ? -1 : (z & 1); x < b.lx; x += 2) {
        if (x < 0) break;  // rows with y and z of different parity hold no lattice points
        int32_t v = latt[(int64_t)x + (int64_t)b.lx * ((int64_t)y + (int64_t)b.ly * z)];
        int kind = 0;
        if (word_kind(e, v, &kind))
          return fail_(e, MSKG_ESTATE, "LattCoo(%d,%d,%d)=0x%08x is not a representable site state", x, y, z, v);
        Site &s = site[(size_t)pidx(b, x, y, z)];
        s.kc = (uint8_t)(kind | (((v >> 24) & 7) << 3));
        s.temp = (uint16_t)((v >> 9) & 0x1FFF);
      }
  double rsum = 0.0;
  auto put = [&](int which, int32_t n, const int32_t *xyz, const int32_t *nn) -> int {
    for (int32_t i = 0; i < n; i++) {
      const int32_t *s = xyz + 3 * i;
      int64_t p = pidx_checked(b, s[0], s[1], s[2]);
      if (p < 0)
        return fail_(e, MSKG_ESTATE, "list %d record %d: site (%d,%d,%d) is not a lattice point of the box", which, i,
                     s[0], s[1], s[2]);
      NL l;
      for (int k = 0; k < 4; k++) {
        l.x[k] = nn[12 * i + 3 * k];
        l.y[k] = nn[12 * i + 3 * k + 1];
        l.z[k] = nn[12 * i + 3 * k + 2];
      }
      if (!set_list(b, site[(size_t)p], s[0], s[1], s[2], l))
        return fail_(e, MSKG_ESTATE, "list %d record %d at (%d,%d,%d): a NextNXYZ entry is too far away", which, i,
                     s[0], s[1], s[2]);
      if (which == 0) {
        int ev = ad_event ? ad_event[i] : 1;
        double pr = ad_prob ? ad_prob[i] : 0.0;
        int kind = s_kind(site[(size_t)p]), coor = s_coor(site[(size_t)p]);
        int c;
        if (ev == 4) c = (kind_species(kind) && coor == 0) ? CLS_ISOLATED : CLS_EVAP;
        else if (ev == 5) c = CLS_DESORB;
        else if (kind == KIND_EMPTY && coor == 1 && (pr == (double)1.e-8f || pr == (double)1.e-6f)) c = CLS_PLACEHOLDER;
        else c = CLS_ADSORB;
        if (!(pr > 0.0)) return fail_(e, MSKG_ESTATE, "MC record %d has a non-positive rate", i);
        cls[(size_t)p] = (uint8_t)c;
        rate[(size_t)p] = (float)pr;
        rsum += pr;
      }
    }
    return 0;
  };
  int rc;
  if ((rc = put(0, n_ad, ad_xyz, ad_nn))) return rc;
  if ((rc = put(1, n_bulk, bulk_xyz, bulk_nn))) return rc;
  if ((rc = put(2, n_void, void_xyz, void_nn))) return rc;
  HIPCHK(e, hipMemcpyAsync(e->d_site, site.data(), site.size() * sizeof(Site), hipMemcpyHostToDevice, e->stream));
  HIPCHK(e, hipMemcpyAsync(e->d_rate, rate.data(), rate.size() * 4, hipMemcpyHostToDevice, e->stream));
  HIPCHK(e, hipMemcpyAsync(e->d_cls, cls.data(), cls.size(), hipMemcpyHostToDevice, e->stream));
  RunState st;
  memset(&st, 0, sizeof st);
  st.clock = kmc_time;
  st.rtot = rsum;
  st.events_per_pass = e->cfg.events_per_pass > 0 ? e->cfg.events_per_pass : 1.0;
  HIPCHK(e, hipMemcpyAsync(e->d_st, &st, sizeof st, hipMemcpyHostToDevice, e->stream));
  HIPCHK(e, hipStreamSynchronize(e->stream));
  e->loaded = true;
  return MSKG_OK;
}

int mskg_set_temperature(mskg_engine *e, const int32_t *temp_k) {
  if (!e || !temp_k) return MSKG_EINVAL;
  if (!e->loaded) return fail_(e, MSKG_EINVAL, "no state loaded");
  const Box &b = e->box;
  std::vector<Site> site((size_t)b.npts);
  HIPCHK(e, hipMemcpyAsync(site.data(), e->d_site, site.size() * sizeof(Site), hipMemcpyDeviceToHost, e->stream));
  HIPCHK(e, hipStreamSynchronize(e->stream));
  for (int z = 0; z <= b.lz; z++)
    for (int y = 0; y < b.ly; y++)
      for (int x = 0; x < b.lx; x++)
        if (valid_parity(x, y, z)) {
          int32_t v = temp_k[(int64_t)x + (int64_t)b.lx * ((int64_t)y + (int64_t)b.ly * z)];
          if (v < 0 || v > 8191) return fail_(e, MSKG_EINVAL, "temperature %d K out of the 13-bit range", v);
          site[(size_t)pidx(b, x, y, z)].temp = (uint16_t)v;
        }
  HIPCHK(e, hipMemcpyAsync(e->d_site, site.data(), site.size() * sizeof(Site), hipMemcpyHostToDevice, e->stream));
  HIPCHK(e, hipStreamSynchronize(e->stream));
  return MSKG_OK;
}

int mskg_refresh_rates(mskg_engine *e) {
  if (!e || !e->loaded) return fail_(e, MSKG_EINVAL, "no state loaded");
  RunState st;
  int rc = read_state(e, st);
  if (rc) return rc;
  HIPCHK(e, hipMemsetAsync(e->d_err, 0, sizeof(int), e->stream));
  Lattice L = make_lattice(e);
  L.tile_active = nullptr;
  int blocks = (int)std::min<int64_t>((e->box.npts + 255) / 256, 256 * 64);
  HIPCHK(e, hipEventRecord(e->ev0, e->stream));
  hipLaunchKernelGGL(k_rate_sweep, dim3(blocks), dim3(256), 0, e->stream, L, st.rtot, e->d_err);
  HIPCHK(e, hipGetLastError());
  HIPCHK(e, hipEventRecord(e->ev1, e->stream));
  int err = 0;
  HIPCHK(e, hipMemcpyAsync(&err, e->d_err, sizeof(int), hipMemcpyDeviceToHost, e->stream));
  HIPCHK(e, hipStreamSynchronize(e->stream));
  float ms = 0;
  HIPCHK(e, hipEventElapsedTime(&ms, e->ev0, e->ev1));
  e->sweep_ms = ms;
  e->sweep_points = e->box.npts;
  if (err) return fail_(e, MSKG_ESTATE, "rate sweep: %s", err_name(err));
  return MSKG_OK;
}

int mskg_replay(mskg_engine *e, const mskg_event *log, int64_t n, int32_t clock_mode) {
  if (!e || !e->loaded || (n > 0 && !log)) return fail_(e, MSKG_EINVAL, "replay: no state or no log");
  if (n == 0) return MSKG_OK;
  mskg_event *d_log = nullptr;
  HIPCHK(e, hipMalloc(&d_log, n * sizeof(mskg_event)));
  hipError_t ce = hipMemcpyAsync(d_log, log, n * sizeof(mskg_event), hipMemcpyHostToDevice, e->stream);
  if (ce != hipSuccess) {
    hipFree(d_log);
    return fail_(e, MSKG_EDEVICE, "replay upload: %s", hipGetErrorString(ce));
  }
  Lattice L = make_lattice(e);
  L.tile_active = nullptr;
  hipLaunchKernelGGL(k_replay, dim3(1), dim3(WAVE), 0, e->stream, L, d_log, n, clock_mode, e->d_st);
  hipError_t le = hipGetLastError();
  RunState st;
  int rc = read_state(e, st);
  hipFree(d_log);
  if (le != hipSuccess) return fail_(e, MSKG_EDEVICE, "k_replay launch: %s", hipGetErrorString(le));
  if (rc) return rc;
  if (st.err)
    return fail_(e, MSKG_ESTATE, "replay stopped at event %llu, site (%d,%d,%d): %s", st.events, st.err_site[0],
                 st.err_site[1], st.err_site[2], err_name(st.err));
  return MSKG_OK;
}

int mskg_run(mskg_engine *e, int64_t max_events, double max_time, int64_t *events_done, double *kmc_time,
             int32_t *stop_reason) {
  if (!e || !e->loaded) return fail_(e, MSKG_EINVAL, "no state loaded");
  if (max_events < 0) return fail_(e, MSKG_EINVAL, "max_events < 0");
  Lattice L = make_lattice(e);
  RunState st;
  int rc = read_state(e, st);
  if (rc) return rc;
  if (st.err) return fail_(e, MSKG_ESTATE, "engine is in an error state: %s", err_name(st.err));
  const unsigned long long ev_start = st.events;
  // tile totals and active flags from the current rates, then the first window
  hipLaunchKernelGGL(k_tile_scan, dim3((e->ntiles + 3) / 4), dim3(256), 0, e->stream, L, e->ntiles, e->ntz,
                     e->d_tile_r);
  HIPCHK(e, hipGetLastError());
  {
    RunState s0 = st;
    s0.tau = 0.0;  // k_sweep_end with tau = 0 only sets the next window
    s0.stop = 0;
    s0.events_per_pass = e->cfg.events_per_pass > 0 ? e->cfg.events_per_pass : 1.0;
    HIPCHK(e, hipMemcpyAsync(e->d_st, &s0, sizeof s0, hipMemcpyHostToDevice, e->stream));
  }
  hipLaunchKernelGGL(k_sweep_end, dim3(1), dim3(1024), 0, e->stream, e->d_st, e->d_tile_r, e->ntiles);
  HIPCHK(e, hipGetLastError());
  const size_t shm = (size_t)e->cap * (4 + 4 + 1) + 16;
  int32_t stop = MSKG_STOP_LIMIT;
  HIPCHK(e, hipEventRecord(e->ev0, e->stream));
  int64_t sweeps = 0;
  int batch = 1;
  for (;;) {
    for (int sb = 0; sb < batch; sb++) {
      for (int c = 0; c < 8; c++) {
        if (e->colour_n[c] == 0) continue;
        PassParams P;
        P.L = L;
        P.st = e->d_st;
        P.tile_r = e->d_tile_r;
        P.tiles = e->d_tiles + e->colour_off[c];
        P.ntiles = e->colour_n[c];
        P.ntz = e->ntz;
        P.cap = e->cap;
        P.key0 = (uint32_t)e->cfg.seed;
        P.key1 = (uint32_t)(e->cfg.seed >> 32);
        P.colour = c;
        hipLaunchKernelGGL(k_pass, dim3(P.ntiles), dim3(WAVE), shm, e->stream, P);
      }
      hipLaunchKernelGGL(k_sweep_end, dim3(1), dim3(1024), 0, e->stream, e->d_st, e->d_tile_r, e->ntiles);
    }
    HIPCHK(e, hipGetLastError());
    sweeps += batch;
    rc = read_state(e, st);
    if (rc) return rc;
    if (st.err) break;
    int64_t done = (int64_t)(st.events - ev_start);
    if (st.stop == MSKG_STOP_NO_EVENTS) { stop = MSKG_STOP_NO_EVENTS; break; }
    if (done >= max_events || st.clock >= max_time) break;
    // size the next batch of sweeps from the observed events per sweep
    double per = (double)done / (double)sweeps;
    int64_t want = per > 0 ? (int64_t)((double)(max_events - done) / per * 0.9) : 16;
    batch = (int)std::max<int64_t>(1, std::min<int64_t>(want, 256));
  }
  HIPCHK(e, hipEventRecord(e->ev1, e->stream));
  HIPCHK(e, hipStreamSynchronize(e->stream));
  float ms = 0;
  HIPCHK(e, hipEventElapsedTime(&ms, e->ev0, e->ev1));
  e->run_ms = ms;
  e->sweeps_done = sweeps;
  if (events_done) *events_done = (int64_t)(st.events - ev_start);
  if (kmc_time) *kmc_time = st.clock;
  if (stop_reason) *stop_reason = stop;
  if (st.err)
    return fail_(e, MSKG_ESTATE, "run stopped in the tile at (%d,%d,%d): %s", st.err_site[0], st.err_site[1],
                 st.err_site[2], err_name(st.err));
  return MSKG_OK;
}

int mskg_store_state(mskg_engine *e, int32_t *latt, int32_t cap_ad, int32_t *n_ad, int32_t *ad_xyz, int32_t *ad_nn,
                     int32_t *ad_event, double *ad_prob, int32_t cap_bulk, int32_t *n_bulk, int32_t *bulk_xyz,
                     int32_t *bulk_nn, int32_t cap_void, int32_t *n_void, int32_t *void_xyz, int32_t *void_nn) {
  if (!e || !e->loaded) return fail_(e, MSKG_EINVAL, "no state loaded");
  const Box &b = e->box;
  std::vector<Site> site((size_t)b.npts);
  std::vector<float> rate((size_t)b.npts);
  std::vector<uint8_t> cls((size_t)b.npts);
  HIPCHK(e, hipMemcpyAsync(site.data(), e->d_site, site.size() * sizeof(Site), hipMemcpyDeviceToHost, e->stream));
  HIPCHK(e, hipMemcpyAsync(rate.data(), e->d_rate, rate.size() * 4, hipMemcpyDeviceToHost, e->stream));
  HIPCHK(e, hipMemcpyAsync(cls.data(), e->d_cls, cls.size(), hipMemcpyDeviceToHost, e->stream));
  HIPCHK(e, hipStreamSynchronize(e->stream));
  int32_t na = 0, nb = 0, nv = 0;
  bool over = false;
  for (int z = 0; z <= b.lz; z++)
    for (int y = 0; y < b.ly; y++)
      for (int x = 0; x < b.lx; x++) {
        if (!valid_parity(x, y, z)) continue;  // off-lattice grid points are left as the caller has them
        int64_t g = (int64_t)x + (int64_t)b.lx * ((int64_t)y + (int64_t)b.ly * z);
        size_t p = (size_t)pidx(b, x, y, z);
        const Site &s = site[p];
        int kind = s_kind(s), coor = s_coor(s);
        if (latt) {
          int32_t v = 0;
          if (kind == KIND_WALL) v = (1 << 23) | 113;
          else if (kind_species(kind)) v = (1 << 23) | (kind << 27);
          else if (kind >= KIND_COV0) v = e->list_cov[kind - KIND_COV0];
          v |= coor << 24;
          v |= (int32_t)s.temp << 9;
          latt[g] = v;
        }
        int32_t xyz[3] = {x, y, z};
        NL l = site_list(b, s, x, y, z);
        int32_t nn[12];
        for (int i = 0; i < 4; i++) {
          nn[3 * i] = l.x[i];
          nn[3 * i + 1] = l.y[i];
          nn[3 * i + 2] = l.z[i];
        }
        if (cls[p]) {
          if (na < cap_ad) {
            if (ad_xyz) memcpy(ad_xyz + 3 * na, xyz, 12);
            if (ad_nn) memcpy(ad_nn + 12 * na, nn, 48);
            int c = cls[p];
            if (ad_event) ad_event[na] = (c == CLS_EVAP || c == CLS_ISOLATED) ? 4 : c == CLS_DESORB ? 5 : 1;
            if (ad_prob) ad_prob[na] = (double)rate[p];
          } else {
            over = true;
          }
          na++;
        } else if (kind_species(kind)) {
          if (nb < cap_bulk) {
            if (bulk_xyz) memcpy(bulk_xyz + 3 * nb, xyz, 12);
            if (bulk_nn) memcpy(bulk_nn + 12 * nb, nn, 48);
          } else {
            over = true;
          }
          nb++;
        } else if ((kind == KIND_EMPTY || kind >= KIND_COV0) && coor == 4) {
          if (nv < cap_void) {
            if (void_xyz) memcpy(void_xyz + 3 * nv, xyz, 12);
            if (void_nn) memcpy(void_nn + 12 * nv, nn, 48);
          } else {
            over = true;
          }
          nv++;
        }
      }
  if (n_ad) *n_ad = na;
  if (n_bulk) *n_bulk = nb;
  if (n_void) *n_void = nv;
  if (over) return fail_(e, MSKG_ECAPACITY, "list capacity too small (need %d MC, %d bulk, %d void)", na, nb, nv);
  return MSKG_OK;
}

int mskg_site_rates(mskg_engine *e, int64_t n, const int32_t *xyz, int32_t *cls_out, float *rate_out) {
  if (!e || !e->loaded) return fail_(e, MSKG_EINVAL, "no state loaded");
  const Box &b = e->box;
  std::vector<float> rate((size_t)b.npts);
  std::vector<uint8_t> cls((size_t)b.npts);
  HIPCHK(e, hipMemcpyAsync(rate.data(), e->d_rate, rate.size() * 4, hipMemcpyDeviceToHost, e->stream));
  HIPCHK(e, hipMemcpyAsync(cls.data(), e->d_cls, cls.size(), hipMemcpyDeviceToHost, e->stream));
  HIPCHK(e, hipStreamSynchronize(e->stream));
  for (int64_t i = 0; i < n; i++) {
    const int32_t *s = xyz + 3 * i;
    int64_t p = pidx_checked(b, s[0], s[1], s[2]);
    if (p < 0) return fail_(e, MSKG_EINVAL, "site %lld (%d,%d,%d) is not a lattice point", (long long)i, s[0], s[1], s[2]);
    cls_out[i] = cls[(size_t)p];
    rate_out[i] = rate[(size_t)p];
  }
  return MSKG_OK;
}

int mskg_counters(mskg_engine *e, int64_t counts[8], double *kmc_time, int64_t *n_adatom, int64_t *n_atoms,
                  int64_t *n_voids) {
  if (!e || !e->loaded) return fail_(e, MSKG_EINVAL, "no state loaded");
  RunState st;
  int rc = read_state(e, st);
  if (rc) return rc;
  if (counts)
    for (int i = 0; i < 8; i++) counts[i] = (int64_t)st.counts[i];
  if (kmc_time) *kmc_time = st.clock;
  if (n_adatom || n_atoms || n_voids) {
    int32_t na = 0, nb = 0, nv = 0;
    rc = mskg_store_state(e, nullptr, 0, &na, nullptr, nullptr, nullptr, nullptr, 0, &nb, nullptr, nullptr, 0, &nv,
                          nullptr, nullptr);
    if (rc && rc != MSKG_ECAPACITY) return rc;
    e->err.clear();
    if (n_adatom) *n_adatom = na;
    if (n_atoms) *n_atoms = nb;
    if (n_voids) *n_voids = nv;
  }
  return MSKG_OK;
}

int mskg_last_timing(mskg_engine *e, double *run_ms, double *sweep_ms, int64_t *sweep_points, int64_t *sweeps_done) {
  if (!e) return MSKG_EINVAL;
  if (run_ms) *run_ms = e->run_ms;
  if (sweep_ms) *sweep_ms = e->sweep_ms;
  if (sweep_points) *sweep_points = e->sweep_points;
  if (sweeps_done) *sweeps_done = e->sweeps_done;
  return MSKG_OK;
}

int mskg_geometry_find_nn(const int32_t s1[3], const int32_t s2[3], const int32_t sc[3], int32_t len_x, int32_t len_y,
                          int32_t out[6]) {
  Geometry g;
  build_geometry(g);
  int a[3] = {s1[0], s1[1], s1[2]}, b[3] = {s2[0], s2[1], s2[2]}, c[3] = {sc[0], sc[1], sc[2]}, o[2][3];
  find_nn(g, len_x, len_y, a, b, c, o);
  for (int r = 0; r < 2; r++)
    for (int k = 0; k < 3; k++) out[3 * r + k] = o[r][k];
  return MSKG_OK;
}

int mskg_geometry_find_sn(const int32_t site[3], const int32_t sitec[3], const int32_t sn[9], int32_t len_x,
                          int32_t len_y, int32_t zig[9], int32_t arm[9]) {
  Geometry g;
  build_geometry(g);
  int s[3] = {site[0], site[1], site[2]}, c[3] = {sitec[0], sitec[1], sitec[2]}, n[3][3], zz[3][3], aa[3][3];
  for (int i = 0; i < 3; i++)
    for (int k = 0; k < 3; k++) n[i][k] = sn[3 * i + k];
  find_sn(g, len_x, len_y, s, c, n, zz, aa);
  for (int i = 0; i < 3; i++)
    for (int k = 0; k < 3; k++) {
      zig[3 * i + k] = zz[i][k];
      arm[3 * i + k] = aa[i][k];
    }
  return MSKG_OK;
}

}  // extern "C"
