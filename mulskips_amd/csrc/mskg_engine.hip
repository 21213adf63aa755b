// mskg_engine.hip -- MI355X KMC event engine: kernels and the C ABI of
// include/mskg.h.
//
// Kernels
//   k_rate_sweep   Get_Prob over every lattice point (HBM-bound stream of the
//                  packed words + neighbour gathers); the rate-update sweep.
//   k_replay       applies a logged event sequence with the reference's
//                  update rules (one wave; events are strictly ordered).
//   k_build_lists  per-tile candidate lists from the classes (after an upload,
//                  a rate sweep or a replay); k_tile_rates: their rate totals.
//   k_pass         one colour of the synchronous-sublattice sweep: each
//                  workgroup (one 64-wide wave) runs rejection-free KMC inside
//                  one tile for the window tau, selecting events with a
//                  wavefront prefix scan over the tile's FP32 rates held in
//                  LDS, FP64 partial sums and Philox-4x32-10 draws.
//   k_sweep        the eight colours of a sweep in one launch, each tile
//                  starting when its lower-colour neighbours are done (same
//                  result as eight k_pass launches, without their barriers).
//   k_rsum         fixed-order FP64 sum of the tile rate totals;
//   k_sweep_end    -> next tau and the clock.
//   k_serial       serial BKL over a batch of lattices (one wave each).
//   k_adjoint      log replay with per-class likelihood statistics.
//   k_build_slab, k_frame_count / k_frame_write: device builder, frame records.
// The event kernels are templates on the rate model (laser annealing,
// coverage species), chosen at launch from the tables.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstdarg>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mskg.h"
#include "mskg_events.h"

using namespace mskg;

namespace {

constexpr int WAVE = 64;

// Per-sweep accumulators, striped over slots so that thousands of tile
// waves do not serialise on one address; k_sweep_end folds them.
constexpr int NSLOT = 128;
// Device-side stop at the run's event or time limit (reported as MSKG_STOP_LIMIT).
constexpr int STOP_AT_LIMIT = 100;
// Events one tile may execute in one window (a bound on the wave's work).
constexpr int PASS_EVENT_CAP = 1 << 16;
// With only the bottom planes stored (stored_len_z), an event's site stays
// this many planes below the top stored plane (Lattice::ztop): an event
// reads at most 12 planes and writes at most 6 planes away from its site.
constexpr int TOP_MARGIN = 24;
struct Slot {
  unsigned long long rmax;  // bit pattern of the largest (non-negative) tile rate
  unsigned long long n, tiles, events, counts[8];
};
// The FP64 total rate is summed without atomics, in a fixed order: every
// tile keeps its total (Lattice::tile_r), NPART blocks each sum a fixed
// chunk of tiles, and k_sweep_end sums the NPART partials.  A rerun with
// the same seed therefore repeats the same windows.
constexpr int NPART = 2 * NSLOT;

// Device-resident run state (one per engine).
struct RunState {
  double tau;            // current window
  double clock;          // KMC time (FP64)
  double rtot;           // total rate estimate (isolated-atom rule)
  double rmax;           // largest tile rate of the last sweep
  double events_per_pass;
  int tau_policy;        // 0: tau = k / max tile rate; 1: tau = k * active tiles / total rate
  unsigned long long events;
  unsigned long long counts[8];
  unsigned long long sweeps;      // sweep counter (RNG stream position)
  unsigned long long sweeps_run;  // sweeps that advanced the clock
  long long n_mc;                 // MC particles at the end of the last sweep
  unsigned long long log_n;
  double max_time;                  // the run's limits: the sweep that reaches one stops the run
  unsigned long long max_events;
  int err;
  int err_site[3];
  int stop;
  // the laser-annealing build's crystal counts (la_variant): CountCrystal,
  // InitCountCrystal, the CountLiquid the rates read (refreshed per event in
  // a replay, per sweep in the parallel engine) and a sweep's changes
  int la_variant;
  long long cc[3], cinit[3], cl[3];
  unsigned long long dcc[3];
  unsigned long long desorbed[3];  // desorptions by coverage slot (CountCov)
  int final_sweep;                 // the window was shortened to end at max_time
  unsigned long long guard_exits;  // tile windows cut short by the per-window event cap
  // the sweep's working sums, last: the host reads and writes only the
  // fields before them (RS_HEAD bytes)
  Slot slot[NSLOT];
  double rpart[NPART];  // partial sums of the tile totals (k_rsum)
};
constexpr size_t RS_HEAD = offsetof(RunState, slot);

struct PassParams {
  Lattice L;
  RunState *st;
  mskg_event *log;          // optional event log (valid serialisation of the run)
  unsigned long long log_cap;
  int ntz;
  int cap;
  uint32_t key0, key1;
  int colour;
  int rsum_chunk;           // tiles per block of the fixed-order tile sum (k_rsum)
  // dependency-scheduled sweep: per-tile completion stamps, and sched[0]
  // the sweep's stamp, [1] queue head, [2] abort flag, [3] / [4] the first /
  // last tile layer holding candidates when the sweep starts
  unsigned int *done;
  unsigned int *sched;
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

// The value of v in another lane of the row (DPP, both halves of the
// double moved by the same control); lanes the control gives no source, and
// rows outside ROW_MASK, get 0.0.  Every lane of the wave must be active.
template <int CTRL, int ROW_MASK>
__device__ inline double dpp_take(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, ROW_MASK, 0xf, false);
  const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, ROW_MASK, 0xf, false);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
// Inclusive prefix sum over the wave's lanes in registers: prefix sums
// within each row of 16 lanes (row_shr 1, 2, 4, 8), then row 0's total into
// row 1 and row 2's into row 3 (row_bcast:15), then rows 0-1's total into
// rows 2 and 3 (row_bcast:31).  Every lane of the wave must be active.
__device__ inline double wave_incl_scan_dpp(double v) {
  v += dpp_take<0x111, 0xf>(v);  // row_shr:1
  v += dpp_take<0x112, 0xf>(v);  // row_shr:2
  v += dpp_take<0x114, 0xf>(v);  // row_shr:4
  v += dpp_take<0x118, 0xf>(v);  // row_shr:8
  v += dpp_take<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
  v += dpp_take<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
  return v;
}
// v of lane l (l the same in every lane).
__device__ inline double read_lane(double v, int l) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__device__ inline void slot_add(RunState *st, int b, double r, unsigned long long n, unsigned long long nev,
                                const unsigned long long *cnt) {
  Slot &sl = st->slot[b & (NSLOT - 1)];
  if (n) {
    const unsigned long long rb = (unsigned long long)__double_as_longlong(r);
    if (rb > sl.rmax) atomicMax(&sl.rmax, rb);
    atomicAdd(&sl.n, n);
    atomicAdd(&sl.tiles, 1ull);
  }
  if (nev) {
    atomicAdd(&sl.events, nev);
    if (cnt)
      for (int q = 0; q < 8; q++)
        if (cnt[q]) atomicAdd(&sl.counts[q], cnt[q]);
  }
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
// The point of a packed tile offset (tile_pack).
__device__ inline P3 tile_point(const Box &b, const TileGeom &g, int T, int k) {
  P3 p;
  p.z = g.z0 + (k >> 14);
  const int par = p.z & 1;
  p.x = g.tx * T + 2 * (k & 127) + par;
  p.y = g.ty * T + 2 * ((k >> 7) & 127) + par;
  p.i = pidx(b, p.x, p.y, p.z);
  return p;
}
// The packed offset of the k-th point of a tile in (z, y, x) order.
__device__ inline int tile_nth(const TileGeom &g, int k) {
  const int pp = g.half * g.half;
  const int zl = k / pp, r = k - zl * pp;
  const int j = r / g.half, i = r - j * g.half;
  return (zl << 14) | (j << 7) | i;
}
// The packed offset of p in tile g, -1 when p is outside the tile.
__device__ inline int tile_local(const TileGeom &g, int T, const P3 &p) {
  if (p.i < 0) return -1;
  const int xl = p.x - g.tx * T, yl = p.y - g.ty * T, zl = p.z - g.z0;
  if ((unsigned)xl >= (unsigned)T || (unsigned)yl >= (unsigned)T || (unsigned)zl >= (unsigned)g.nz) return -1;
  return tile_pack(zl, yl, xl);
}

// ---- rate-update sweep ----------------------------------------------------
// Every point: an MC particle is an empty or coverage site with coordination
// 1..3 or an atom with coordination <= 3 (Get_Prob.f:54-331 is evaluated for
// it); everything else gets class 0.
// With atoms_only, only atoms of coordination 1..3 are re-evaluated and every
// other point keeps its rate (used after the device slab builder, whose
// atoms need the second-neighbour factor).
//
// Each thread takes four consecutive points: four 16-byte record loads in
// flight, one 16-byte store of the rates and one 4-byte store of the
// classes.  Nearly every point is not an MC particle (a few per cent of the
// stored points are), so the Get_Prob branch is rare and its registers
// (spilled at four waves per SIMD) cost little.
template <bool LA, bool COV>
__device__ inline void sweep_point(const Lattice &L, double rtot, int64_t p, const Site &sp, int atoms_only,
                                   int *err, float &r_out, uint8_t &c_out, bool &keep) {
  const int kind = s_kind(sp), coor = s_coor(sp);
  keep = atoms_only && !(kind_species(kind) && coor >= 1 && coor <= 3);
  const bool mc = (kind == KIND_EMPTY && coor >= 1 && coor <= 3) || (kind >= KIND_COV0 && coor <= 3) ||
                  (kind_species(kind) && coor <= 3);
  int cls = 0, key = -1;
  double r = 0.0;
  if (mc && !keep) {
    P3 pt;
    pt.i = p;
    pxyz(L.box, p, pt.x, pt.y, pt.z);
    int e = ERR_NONE;
    get_prob<LA, COV>(L, rtot, pt, sp, coor, list_of(L, pt, sp), cls, r, e, nullptr, &key);
    if (e) atomicExch(err, e);
  }
  r_out = cls ? (float)r : 0.0f;
  c_out = (uint8_t)cls;
  if (L.key && !keep) L.key[p] = cls ? key : -1;
}

// LA, COV: the rate model's laser-annealing and coverage branches are compiled in.
// Four consecutive points from p0, or the points p0..n-1 at the end.
template <bool LA, bool COV>
__device__ inline void sweep_group(const Lattice &L, double rtot, int *err, int atoms_only, int64_t p0, int64_t n) {
  if (p0 + 4 <= n) {
    Site sp[4];
    for (int k = 0; k < 4; k++) sp[k] = L.site[p0 + k];
    float r[4];
    uint8_t c[4];
    bool keep[4];
    for (int k = 0; k < 4; k++) sweep_point<LA, COV>(L, rtot, p0 + k, sp[k], atoms_only, err, r[k], c[k], keep[k]);
    if (!(keep[0] || keep[1] || keep[2] || keep[3])) {
      *(float4 *)(L.rate + p0) = make_float4(r[0], r[1], r[2], r[3]);
      *(uchar4 *)(L.cls + p0) = make_uchar4(c[0], c[1], c[2], c[3]);
    } else {
      for (int k = 0; k < 4; k++)
        if (!keep[k]) {
          L.rate[p0 + k] = r[k];
          L.cls[p0 + k] = c[k];
        }
    }
  } else {
    for (int64_t q = p0; q < n; q++) {
      float r;
      uint8_t c;
      bool keep;
      sweep_point<LA, COV>(L, rtot, q, L.site[q], atoms_only, err, r, c, keep);
      if (!keep) {
        L.rate[q] = r;
        L.cls[q] = c;
      }
    }
  }
}

// LA, COV: the rate model's laser-annealing and coverage branches are compiled in.
template <bool LA, bool COV>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8)))
k_rate_sweep(Lattice L, double rtot, int *err, int atoms_only) {
  const int64_t n = L.box.npts;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x * 4;
  for (int64_t p0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4; p0 < n; p0 += stride) {
    if (p0 + 4 <= n) {
      Site sp[4];
      for (int k = 0; k < 4; k++) sp[k] = L.site[p0 + k];
      float r[4];
      uint8_t c[4];
      bool keep[4];
      for (int k = 0; k < 4; k++) sweep_point<LA, COV>(L, rtot, p0 + k, sp[k], atoms_only, err, r[k], c[k], keep[k]);
      if (!(keep[0] || keep[1] || keep[2] || keep[3])) {
        *(float4 *)(L.rate + p0) = make_float4(r[0], r[1], r[2], r[3]);
        *(uchar4 *)(L.cls + p0) = make_uchar4(c[0], c[1], c[2], c[3]);
      } else {
        for (int k = 0; k < 4; k++)
          if (!keep[k]) {
            L.rate[p0 + k] = r[k];
            L.cls[p0 + k] = c[k];
          }
      }
    } else {
      for (int64_t q = p0; q < n; q++) {
        float r;
        uint8_t c;
        bool keep;
        sweep_point<LA, COV>(L, rtot, q, L.site[q], atoms_only, err, r, c, keep);
        if (!keep) {
          L.rate[q] = r;
          L.cls[q] = c;
        }
      }
    }
  }
}

// The same over the points [first, end) of the stored order (first a
// multiple of 4), in groups of four counted in 32 bits (the largest lattice
// holds fewer than 2^31 groups).  A kernel of its own: the whole-lattice
// sweep keeps its registers without the range.
template <bool LA, bool COV>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8)))
k_rate_sweep_range(Lattice L, double rtot, int *err, int64_t first, int64_t end) {
  const int g_first = (int)(first >> 2), g_end = (int)((end + 3) >> 2);
  const int gstride = gridDim.x * blockDim.x;
  for (int g = g_first + blockIdx.x * blockDim.x + threadIdx.x; g < g_end; g += gstride)
    sweep_group<LA, COV>(L, rtot, err, 0, (int64_t)g * 4, end);
}

// ---- event-log replay -------------------------------------------------------
template <bool LA, bool COV>
__global__ void __launch_bounds__(WAVE) k_replay(Lattice L, const mskg_event *log, int64_t n, int clock_mode,
                                                 RunState *st) {
  __shared__ EvScratch sc;
  __shared__ int s_bad;
  const int lane = threadIdx.x;
  double clock = st->clock;
  for (int64_t k = 0; k < n; k++) {
    const mskg_event ev = log[k];
    const P3 s = at(L, ev.site[0], ev.site[1], ev.site[2]);
    if (lane == 0) s_bad = (s.i < 0 || L.cls[s.i] == 0) ? 1 : 0;
    __syncthreads();
    if (s_bad) {
      if (lane == 0) {
        st->err = ERR_NOT_MC;
        st->err_site[0] = s.x; st->err_site[1] = s.y; st->err_site[2] = s.z;
        st->events = (unsigned long long)k;
      }
      return;
    }
    if (clock_mode == 0) clock = clock + 1.0 / ev.rtot;
    Choice ch;
    ch.arm = ev.arm;
    ch.u_arm = 0.0;
    EvResult res;
    execute<LA, COV>(L, ev.rtot, s, ev.itrans, ch, res, sc, lane);
    if (lane == 0) s_bad = res.err;
    __syncthreads();
    if (s_bad) {
      if (lane == 0) {
        st->err = res.err;
        st->err_site[0] = s.x; st->err_site[1] = s.y; st->err_site[2] = s.z;
        st->events = (unsigned long long)k;
      }
      return;
    }
    if (lane == 0) {
      int slot = ev.itrans <= 3 ? ev.itrans - 1 : (ev.itrans == 4 ? 3 : (ev.itrans == 5 ? 4 : ev.itrans - 1));
      if (slot >= 0 && slot < 8) st->counts[slot]++;
      if (COV && ev.itrans == 5 && res.species) st->desorbed[res.species - 1]++;
      if (LA && L.cl && res.species && ev.itrans <= 4) {  // CountCrystal, then the next iteration's CountLiquid
        st->cc[res.species - 1] += ev.itrans <= 3 ? 1 : -1;
        for (int q = 0; q < 3; q++) st->cl[q] = st->cinit[q] - st->cc[q];
      }
    }
    __syncthreads();  // the next event's rates read the counts
  }
  if (lane == 0) {
    st->events += (unsigned long long)n;
    st->clock = clock;
  }
}

// ---- synchronous sublattice -------------------------------------------------
// Build every tile's candidate list from the classes (after a state upload,
// a rate sweep or a replay).  One wave per tile.
__global__ void __launch_bounds__(256) k_build_lists(Lattice L, int ntiles) {
  int lane = threadIdx.x & (WAVE - 1);
  int t = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
  if (t >= ntiles) return;
  TileGeom g = tile_geom(t, L.tile, L.ntx, L.nty, L.box.lz);
  int n = 0;
  for (int k0 = 0; k0 < g.npt; k0 += WAVE) {
    int k = k0 + lane;
    bool on = false;
    int kp = 0;
    if (k < g.npt) {
      kp = tile_nth(g, k);
      P3 p = tile_point(L.box, g, L.tile, kp);
      on = L.cls[p.i] != 0;
      L.inlist[p.i] = on ? 1 : 0;
    }
    unsigned long long m = __ballot(on);
    int slot = n + __popcll(m & ((1ull << lane) - 1ull));
    if (on) {
      if (slot < L.capg) L.tl_idx[(int64_t)t * L.capg + slot] = kp;
      else *L.list_err = ERR_TILE_LIST;
    }
    n += __popcll(m);
  }
  if (lane == 0) {
    L.tl_n[t] = n;
    L.tl_own[t] = min(n, L.capg);
    if (n) widen_window(L, tile_colour(g.tx, g.ty, g.tz), g.tz);
  }
}

// Window of tile layers from lists that are already built.
__global__ void __launch_bounds__(256) k_list_window(Lattice L, int ntiles) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntiles) return;
  if (L.tl_n[t] > 0) {
    const int tx = t % L.ntx, ty = (t / L.ntx) % L.nty, tz = t / (L.ntx * L.nty);
    widen_window(L, tile_colour(tx, ty, tz), tz);
  }
}

// Rate total (without isolated atoms) and population of every tile's list.
__global__ void __launch_bounds__(256) k_tile_rates(Lattice L, int ntiles, RunState *st) {
  int lane = threadIdx.x & (WAVE - 1);
  int t = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
  if (t >= ntiles) return;
  TileGeom g = tile_geom(t, L.tile, L.ntx, L.nty, L.box.lz);
  const int m = min(L.tl_n[t], L.capg);
  double acc = 0.0;
  int cnt = 0;
  for (int j = lane; j < m; j += WAVE) {
    P3 p = tile_point(L.box, g, L.tile, L.tl_idx[(int64_t)t * L.capg + j]);
    uint8_t c = L.cls[p.i];
    if (c) cnt++;
    if (c && c != CLS_ISOLATED) acc += (double)L.rate[p.i];
  }
  acc = wave_sum(acc);
  for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, WAVE);
  if (lane == 0) {
    L.tile_r[t] = cnt ? acc : 0.0;
    if (cnt) slot_add(st, t, acc, (unsigned long long)cnt, 0, nullptr);
  }
}

// Sorts lk/lr/lc[lo, hi) by offset (the offsets are distinct): a bitonic
// network across the lanes for up to 64 entries, odd-even transposition
// rounds beyond (rare: appends from other tiles between two passes).
__device__ inline void sort_by_offset(int32_t *lk, float *lr, uint8_t *lc, int lo, int hi, int lane) {
  const int m = hi - lo;
  if (m <= 1) return;
  if (m <= WAVE) {
    int key = lane < m ? lk[lo + lane] : 0x7fffffff;
    int idx = lane;
    for (int k = 2; k <= WAVE; k <<= 1)
      for (int j = k >> 1; j > 0; j >>= 1) {
        const int ok = __shfl_xor(key, j, WAVE), oi = __shfl_xor(idx, j, WAVE);
        const bool keep_min = ((lane & k) == 0) == ((lane & j) == 0);
        if (keep_min ? ok < key : ok > key) {
          key = ok;
          idx = oi;
        }
      }
    float r = 0.0f;
    uint8_t c = 0;
    if (lane < m) {
      r = lr[lo + idx];
      c = lc[lo + idx];
    }
    __syncthreads();
    if (lane < m) {
      lk[lo + lane] = key;
      lr[lo + lane] = r;
      lc[lo + lane] = c;
    }
    __syncthreads();
    return;
  }
  for (int round = 0; round < m; round++) {
    for (int i = lo + 2 * lane + (round & 1); i + 1 < hi; i += 2 * WAVE)
      if (lk[i] > lk[i + 1]) {
        const int tk = lk[i];
        lk[i] = lk[i + 1];
        lk[i + 1] = tk;
        const float tr = lr[i];
        lr[i] = lr[i + 1];
        lr[i + 1] = tr;
        const uint8_t tc = lc[i];
        lc[i] = lc[i + 1];
        lc[i + 1] = tc;
      }
    __syncthreads();
  }
}

// The LDS of one tile's pass besides its candidate list (which is dynamic).
struct PassLds {
  unsigned long long cnt[8];  // events by type (lane 0)
  long long dcc[3];           // crystal count changes (laser-annealing build's form)
  unsigned long long des[3];  // desorptions by coverage slot
  EvScratch sc;
};

// One tile's share of a colour pass: its candidate list into LDS, rejection-
// free KMC inside the tile for the window tau, the list back.  Returns the
// error that ends the wave's work (0: none).
template <bool LA, bool COV>
__device__ inline int pass_tile(const PassParams &P, Lattice &L, RunState *st, const double tau, const int t,
                                const int colour, PassLds &S, int32_t *lk, float *lr, uint8_t *lc, const int lane) {
  const int cap = P.cap;
  // The tile's first loads leave together, in one round trip: its list
  // length and own count, the first WAVE list entries (read whatever the
  // length; a lane uses its entry only below it, and the index stays in the
  // tile's list storage), and the run state the window needs.
  const int n_in = L.tl_n[t];
  const int own = L.tl_own[t];
  const int k_first = L.tl_idx[(int64_t)t * L.capg + min(lane, L.capg - 1)];
  const double rtot = st->rtot;
  const unsigned long long sweeps = st->sweeps;
  if (n_in == 0) return 0;
  L.cur_tile = t;
  TileGeom g = tile_geom(t, L.tile, L.ntx, L.nty, L.box.lz);

  // 1. the tile's candidate list -> LDS, dropping entries no longer active;
  // the entries other tiles appended since the tile's last pass (an order
  // that depends on timing) are then sorted by offset
  const int m_in = min(n_in, L.capg);
  int n = 0, nh = 0;
  for (int j0 = 0; j0 < m_in; j0 += WAVE) {
    int j = j0 + lane;
    uint8_t c = 0;
    float r = 0.0f;
    int k = 0;
    if (j < m_in) {
      k = j0 == 0 ? k_first : L.tl_idx[(int64_t)t * L.capg + j];
      P3 p = tile_point(L.box, g, L.tile, k);
      // class and rate in one round trip (the rate is used only with a class)
      c = L.cls[p.i];
      r = L.rate[p.i];
      if (!c) L.inlist[p.i] = 0;
    }
    unsigned long long m = __ballot(c != 0);
    int slot = n + __popcll(m & ((1ull << lane) - 1ull));
    if (c && slot < cap) {
      lk[slot] = k;
      lr[slot] = r;
      lc[slot] = c;
    }
    n += __popcll(m);
    nh += __popcll(__ballot(c != 0 && j < own));
  }
  if (n > cap) {
    if (lane == 0) {
      st->err = ERR_CAPACITY;
      P3 p = tile_point(L.box, g, L.tile, 0);
      st->err_site[0] = p.x; st->err_site[1] = p.y; st->err_site[2] = p.z;
    }
    return ERR_CAPACITY;
  }
  __syncthreads();
  sort_by_offset(lk, lr, lc, nh, n, lane);

  // 2. rejection-free KMC inside the tile for the window tau
  double tloc = 0.0;
  unsigned long long nev = 0;
  if (lane < 8) S.cnt[lane] = 0;
  if (lane < 3) S.dcc[lane] = 0;
  if (COV && lane < 3) S.des[lane] = 0;
  U4 ctr;
  ctr.x = (uint32_t)t;
  ctr.y = (uint32_t)sweeps;
  ctr.z = (uint32_t)(sweeps >> 32) ^ ((uint32_t)colour << 29);
  ctr.w = 0;
  int err = 0;
  // every wave ends its pass: the window tau bounds the expected number of
  // events, and this cap bounds the worst case (a window cut short by it is
  // counted in RunState::guard_exits and reported by the run)
  int guard = 0;
  for (; guard < PASS_EVENT_CAP; guard++) {
    // lane-major order: lane l owns entries l, l + 64, ...; one wave scan of
    // the lanes' partial sums gives the total and the lane that holds the pick
    double part = 0.0;
    for (int j = lane; j < n; j += WAVE) part += (double)lr[j];
    const double inc = wave_incl_scan_dpp(part);
    const double R = read_lane(inc, WAVE - 1);
    if (!(R > 0.0)) break;
    // one Philox block per event: waiting time, selection, event type and
    // stacking choice, 32 random bits each (the reference's ran3 has fewer)
    const U4 d = philox4x32_10(ctr, P.key0, P.key1);
    ctr.w++;
    const double dt = -log(u01_32(d.x)) / R;  // FP64 waiting time
    if (tloc + dt > tau) break;
    tloc += dt;
    const double target = u01_32(d.y) * R;
    // the first lane whose inclusive prefix exceeds target (the last lane
    // with a rate when rounding leaves target at the top end), then the
    // first of its entries whose running prefix exceeds target
    const unsigned long long hit = __ballot(part > 0.0 && inc > target);
    const unsigned long long nzl = __ballot(part > 0.0);
    if (!nzl) break;
    const int owner = hit ? __ffsll((long long)hit) - 1 : 63 - __clzll((long long)nzl);
    int sel = -1;
    if (lane == owner) {
      double acc = inc - part;
      int last = -1;
      for (int j = lane; j < n; j += WAVE) {
        const double v = (double)lr[j];
        if (v > 0.0) {
          last = j;
          acc += v;
          if (acc > target) {
            sel = j;
            break;
          }
        }
      }
      if (sel < 0) sel = last;
    }
    sel = __builtin_amdgcn_readlane(sel, owner);
    if (sel < 0) break;
    {
      const P3 s = tile_point(L.box, g, L.tile, lk[sel]);
      Choice ch;
      ch.arm = -1;
      ch.u_arm = u01_32(d.w);
      const int itrans = la_swap<LA>(L, pick_itrans<LA, COV>(L, s, lc[sel], u01_32(d.z)));
      const int coor0 = P.log ? s_coor(L.site[s.i]) : 0;
      __syncthreads();  // every lane has read the site before it changes
      EvResult res;
      execute<LA, COV>(L, rtot, s, itrans, ch, res, S.sc, lane);
      __syncthreads();  // neighbour updates by lanes 1..3 are visible to lane 0
      err = __builtin_amdgcn_readfirstlane(res.err);  // res is complete on lane 0
      if (lane == 0) {
        if (P.log) {
          const unsigned long long li = atomicAdd(&st->log_n, 1ull);
          if (li < P.log_cap) {
            mskg_event ev;
            ev.iter = (int64_t)li + 1;
            ev.site[0] = s.x; ev.site[1] = s.y; ev.site[2] = s.z;
            ev.itrans = itrans;
            ev.arm = res.arm;
            ev.coor = coor0;
            ev.rtot = rtot;
            ev.time = st->clock + tloc;
            P.log[li] = ev;
          }
        }
        int slot = itrans <= 3 ? itrans - 1 : (itrans == 4 ? 3 : (itrans == 5 ? 4 : itrans - 1));
        if (slot >= 0 && slot < 8) S.cnt[slot]++;
        if (LA && L.cl && res.species && itrans <= 4) S.dcc[res.species - 1] += itrans <= 3 ? 1 : -1;
        if (COV && itrans == 5 && res.species) S.des[res.species - 1]++;
      }
    }
    nev++;
    if (err) break;
    // the touched points' new rates, one lane each, as execute() left them
    // on chip (S.sc holds them until the next event); a point that got no
    // new rate keeps its list entry as it is
    const int nt = __builtin_amdgcn_readfirstlane(S.sc.ntouched);
    int kl = -1;  // tile-local offset of touched point `lane` (-1: none / outside the tile)
    float tr = 0.0f;
    uint8_t tc = 0;
    if (lane < nt) {
      const P3 tp = S.sc.touched[lane];
      const int ts = S.sc.tset[lane];
      kl = ts ? tile_local(g, L.tile, tp) : -1;
      if (kl >= 0) {
        if (ts == 1) {
          tr = S.sc.trate[lane];
          tc = S.sc.tcls[lane];
        } else {
          tr = L.rate[tp.i];
          tc = L.cls[tp.i];
        }
      }
    }
    // refresh the on-chip list: the slot of each touched point, one ballot
    // of the list against its offset (list entries are distinct, so each
    // touched point has at most one slot; a point touched twice finds the
    // same slot), and whether an earlier touched point has the same offset
    int found = -1;
    bool dup = false;
    const int k_lane = lane < n ? lk[lane] : -1;  // the list's first WAVE entries, one a lane
    for (int q = 0; q < nt; q++) {
      const int kq = __builtin_amdgcn_readlane(kl, q);
      if (q < lane && kq == kl) dup = true;
      if (kq < 0) continue;
      unsigned long long m = __ballot(k_lane == kq);
      int f = m ? __ffsll((long long)m) - 1 : -1;
      for (int j0 = WAVE; f < 0 && j0 < n; j0 += WAVE) {
        const int kj = j0 + lane < n ? lk[j0 + lane] : -1;
        m = __ballot(kj == kq);
        if (m) f = j0 + __ffsll((long long)m) - 1;
      }
      if (lane == q) found = f;
    }
    // ... updates in place, and appends the new MC particles (first
    // occurrence of each) after the list in touched order
    bool app = false;
    if (kl >= 0) {
      if (found >= 0) {
        lr[found] = tr;
        lc[found] = tc;
      } else if (tc) {
        app = !dup;
      }
    }
    const unsigned long long am = __ballot(app);
    if (app) {
      const int slot = n + __popcll(am & ((1ull << lane) - 1ull));
      if (slot < cap) {
        lk[slot] = kl;
        lr[slot] = tr;
        lc[slot] = tc;
      }
    }
    if (n + __popcll(am) > cap) {  // the list keeps its first cap entries
      err = ERR_CAPACITY;
      break;
    }
    n += __popcll(am);
    __syncthreads();
  }
  if (guard >= PASS_EVENT_CAP && lane == 0) atomicAdd(&st->guard_exits, 1ull);
  // 3. write the list back (active entries only), tile totals, counters
  double part = 0.0;
  int nout = 0;
  for (int j0 = 0; j0 < n; j0 += WAVE) {
    int j = j0 + lane;
    bool keep = j < n && lc[j] != 0;
    if (j < n) L.inlist[tile_point(L.box, g, L.tile, lk[j]).i] = keep ? 1 : 0;
    if (keep && lc[j] != CLS_ISOLATED) part += (double)lr[j];
    unsigned long long m = __ballot(keep);
    int slot = nout + __popcll(m & ((1ull << lane) - 1ull));
    if (keep) L.tl_idx[(int64_t)t * L.capg + slot] = lk[j];
    nout += __popcll(m);
  }
  const double R = wave_sum(part);
  if (lane == 0) {
    L.tl_n[t] = nout;
    L.tl_own[t] = nout;
    L.tile_r[t] = R;
    slot_add(st, t, R, (unsigned long long)nout, nev, S.cnt);
    if (LA && L.cl)
      for (int k = 0; k < 3; k++)
        if (S.dcc[k]) atomicAdd(&st->dcc[k], (unsigned long long)S.dcc[k]);
    if (COV)
      for (int k = 0; k < 3; k++)
        if (S.des[k]) atomicAdd(&st->desorbed[k], S.des[k]);
    if (nout) {
      if (g.tz < L.win[4 * colour + 2]) atomic_min_i32(&L.win[4 * colour + 2], g.tz);
      if (g.tz > L.win[4 * colour + 3]) atomic_max_i32(&L.win[4 * colour + 3], g.tz);
    }
    if (err) {
      st->err = err;
      P3 p = tile_point(L.box, g, L.tile, 0);
      st->err_site[0] = p.x; st->err_site[1] = p.y; st->err_site[2] = p.z;
    }
  }
  err = __builtin_amdgcn_readfirstlane(err);  // the same in every lane
  if (err) return err;
  __syncthreads();
  return 0;
}

// The tiles of `colour` in the colour's window of tile layers that hold
// candidates (win[4c], win[4c+1], fixed during a sweep): their number, and
// zfirst, the lowest of those layers with the colour's layer parity.
__device__ inline int colour_tiles(const Lattice &L, int ntz, int colour, int &zfirst) {
  const int c2 = (colour >> 2) & 1;
  const int zlo = max(L.win[4 * colour], 0), zhi = min(L.win[4 * colour + 1], ntz - 1);
  zfirst = 0;
  if (zlo > zhi) return 0;  // no tile of this colour holds candidates
  zfirst = zlo + ((c2 - zlo) & 1);
  const int nlayers = zfirst <= zhi ? (zhi - zfirst) / 2 + 1 : 0;
  return nlayers * (L.ntx / 2) * (L.nty / 2);
}
// The wi-th of those tiles: layer by layer, y then x.
__device__ inline int colour_tile(const Lattice &L, int colour, int zfirst, int wi) {
  const int c0 = colour & 1, c1 = (colour >> 1) & 1;
  const int hx = L.ntx / 2, per_layer = hx * (L.nty / 2);
  const int layer = wi / per_layer, rr = wi - layer * per_layer;
  return ((zfirst + 2 * layer) * L.nty + 2 * (rr / hx) + c1) * L.ntx + 2 * (rr % hx) + c0;
}

// One colour pass: the tiles of `colour` in the window of tile layers that
// hold candidates, each workgroup (one wave) striding over them.  Three
// waves per SIMD (168 VGPRs): 3072 tiles in flight on 256 CUs, against the
// ~4096 tiles of a colour pass at the 512^3 workload.
template <bool LA, bool COV>
__global__ void __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(3, 3))) k_pass(PassParams P) {
  extern __shared__ unsigned char smem[];
  __shared__ PassLds S;
  const int lane = threadIdx.x;
  Lattice L = P.L;
  int32_t *lk = (int32_t *)smem;        // tile-local point offsets
  float *lr = (float *)(lk + P.cap);    // rates
  uint8_t *lc = (uint8_t *)(lr + P.cap);  // classes
  RunState *st = P.st;
  const double tau = st->tau;
  if (!(tau > 0.0) || st->err || st->stop) return;
  int zfirst;
  const int nwin = colour_tiles(L, P.ntz, P.colour, zfirst);
  for (int wi = blockIdx.x; wi < nwin; wi += gridDim.x) {
    const int t = colour_tile(L, P.colour, zfirst, wi);
    if (pass_tile<LA, COV>(P, L, st, tau, t, P.colour, S, lk, lr, lc, lane)) return;
  }
}

// ---- a sweep scheduled by tile dependencies ------------------------------
// The eight colour passes of a sweep in one launch without a barrier
// between colours: a tile's pass starts as soon as its neighbour tiles of
// the lower colours have finished theirs in this sweep (tiles of the higher
// colours finished theirs in the previous launch).  A tile's events read
// and write only its own tile and the 24-unit shell of its 26 neighbour
// tiles, so every tile sees exactly the state it sees with a barrier after
// each colour, and the sweep's result -- lattice, lists, totals, random
// streams -- is the same bit for bit.  Tiles that do not neighbour each
// other (same colour, or farther apart) run concurrently as before.
//
// Work distribution: a queue of the sweep's tiles in colour order; every
// wave takes the next entry with one atomic, waits for the entry's lower-
// colour neighbours, runs the tile and publishes its completion.  A wave
// only waits on entries of lower colours, which were taken earlier by waves
// that are running, so the queue always drains.  The completion flag of a
// tile is the sweep's stamp (sched[0], advanced by k_sweep_end), stored
// with agent-scope release after the tile's writes; a waiting wave loads it
// at agent scope and takes an agent-scope acquire before it reads the
// neighbours' records, so the tiles' records travel between XCDs.

// The tile layers of colours with layer parity `odd` that a sweep's queue
// holds, for candidates in layers [zlo, zhi] when the sweep starts: the
// even colours (0-3) run first, so their layers outside the window stay
// empty until they have run; the odd colours (4-7) run after the even
// layers next to them, whose events can enlist points one layer out.
__device__ inline void queue_layers(int zlo, int zhi, int odd, int ntz, int &first, int &count) {
  const int lo = max(odd ? zlo - 1 : zlo, 0), hi = min(odd ? zhi + 1 : zhi, ntz - 1);
  first = lo + ((odd - lo) & 1);
  count = first <= hi ? (hi - first) / 2 + 1 : 0;
}

// A wave waits at most this long (wall-clock ticks of 10 ns) for a
// neighbour tile; the run then stops with ERR_SCHED instead of hanging.
constexpr long long SCHED_WAIT_TICKS = 200000000ll;

template <bool LA, bool COV>
__global__ void __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(3, 3))) k_sweep(PassParams P) {
  extern __shared__ unsigned char smem[];
  __shared__ PassLds S;
  const int lane = threadIdx.x;
  Lattice L = P.L;
  int32_t *lk = (int32_t *)smem;
  float *lr = (float *)(lk + P.cap);
  uint8_t *lc = (uint8_t *)(lr + P.cap);
  RunState *st = P.st;
  const double tau = st->tau;
  if (!(tau > 0.0) || st->err || st->stop) return;
  // the window as the sweep starts (L.win's first words widen during it)
  const int zlo = (int)P.sched[3], zhi = (int)P.sched[4];
  if (zlo > zhi) return;
  int first[2], count[2];
  queue_layers(zlo, zhi, 0, P.ntz, first[0], count[0]);
  queue_layers(zlo, zhi, 1, P.ntz, first[1], count[1]);
  const int hx = L.ntx / 2, per_layer = hx * (L.nty / 2);
  const int per_even = count[0] * per_layer, per_odd = count[1] * per_layer;
  const int total = 4 * (per_even + per_odd);
  const unsigned int stamp = P.sched[0];
  unsigned int *const head = P.sched + 1;
  // neighbour k of 26 (lanes 0..25): offsets -1..1 along each axis
  const int kk = lane + (lane >= 13 ? 1 : 0);
  const int ndx = kk % 3 - 1, ndy = (kk / 3) % 3 - 1, ndz = kk / 9 - 1;
  const int nflip = (ndx & 1) | ((ndy & 1) << 1) | ((ndz & 1) << 2);
  // Every decision of this loop is wave-uniform by construction (scalar
  // values from readfirstlane / ballot) and no memory access sits under a
  // lane-dependent branch: the compiler may otherwise restructure the loop
  // so that lane 0, which would own such accesses, runs a different
  // iteration than the other lanes.
  for (;;) {
    // the next entry: every lane adds 1 (one atomic of 64 for the wave), so
    // the head counts entries in units of 64
    const unsigned int took = __hip_atomic_fetch_add(head, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int q = __builtin_amdgcn_readfirstlane((int)(took >> 6));
    if (q >= total) break;
    // entry -> colour, tile
    int colour, wi;
    if (q < 4 * per_even) {
      colour = q / per_even;
      wi = q - colour * per_even;
    } else {
      const int r = q - 4 * per_even;
      colour = 4 + r / per_odd;
      wi = r - (colour - 4) * per_odd;
    }
    const int odd = colour >> 2;
    const int layer = wi / per_layer, rr = wi - layer * per_layer;
    const int tz = first[odd] + 2 * layer, ty = 2 * (rr / hx) + ((colour >> 1) & 1), tx = 2 * (rr % hx) + (colour & 1);
    const int t = (tz * L.nty + ty) * L.ntx + tx;
    // the neighbours of lower colours that this sweep's queue holds
    bool dep = false;
    int n = 0;
    if (lane < 26) {
      const int nc = colour ^ nflip, nz = tz + ndz;
      if (nc < colour && nz >= 0 && nz < P.ntz) {
        const int nodd = nc >> 2;
        const int rel = nz - first[nodd];
        if (rel >= 0 && rel < 2 * count[nodd]) {
          int nx = tx + ndx, ny = ty + ndy;
          nx += nx < 0 ? L.ntx : (nx >= L.ntx ? -L.ntx : 0);
          ny += ny < 0 ? L.nty : (ny >= L.nty ? -L.nty : 0);
          n = (nz * L.nty + ny) * L.ntx + nx;
          dep = true;
        }
      }
    }
    const unsigned int *const flag = P.done + (dep ? n : t);  // a lane without a neighbour reads its own tile's
    bool abort = false;
    if (__ballot(dep) != 0ull) {
      const long long t0 = wall_clock64();
      for (;;) {
        const unsigned int f = __hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (__ballot(dep && f != stamp) == 0ull) break;
        const unsigned int ab = __hip_atomic_load(&P.sched[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const bool late = wall_clock64() - t0 > SCHED_WAIT_TICKS;
        if (__builtin_amdgcn_readfirstlane((int)ab) != 0 || late) {
          if (late) {  // every lane stores the same values
            __hip_atomic_store(&P.sched[2], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            atomicCAS(&st->err, 0, ERR_SCHED);
          }
          abort = true;
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    const int err_now = __builtin_amdgcn_readfirstlane(__hip_atomic_load(&st->err, __ATOMIC_RELAXED,
                                                                         __HIP_MEMORY_SCOPE_AGENT));
    const int n_in = __builtin_amdgcn_readfirstlane(L.tl_n[t]);
    if (!abort && !err_now && n_in > 0) {
      pass_tile<LA, COV>(P, L, st, tau, t, colour, S, lk, lr, lc, lane);
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");  // the tile's writes before its flag
    }
    // every lane stores the same stamp
    __hip_atomic_store(&P.done[t], stamp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// The fixed-order sum of the tile totals: block b sums its chunk of tiles
// (each thread a fixed stride, then a tree in LDS) into rpart[b].
__device__ inline void rsum_block(const double *tile_r, int ntiles, int chunk, RunState *st, int b, double *s) {
  const int lo = b * chunk, hi = min(lo + chunk, ntiles);
  double acc = 0.0;
  for (int t = lo + (int)threadIdx.x; t < hi; t += 256) acc += tile_r[t];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) st->rpart[b] = s[0];
}
__global__ void __launch_bounds__(256) k_rsum(const double *tile_r, int ntiles, int chunk, RunState *st) {
  __shared__ double s[256];
  rsum_block(tile_r, ntiles, chunk, st, blockIdx.x, s);
}

// After the eight colour passes: the clock advances by the window just used
// (operator splitting of the master equation over the colours), the next
// window is events_per_pass / (largest tile rate), and the reference's
// early stop applies when at most 20 MC particles remain (main.f:307-310).
struct SweepEndLds {
  double srsum[NSLOT];
  unsigned long long srmax[NSLOT], sn[NSLOT], stl[NSLOT], sev[NSLOT], scnt[8][NSLOT];
};
__device__ inline void sweep_end(RunState *st, int *win, const int *list_err, unsigned int *sched,
                                 SweepEndLds &E) {
  double *srsum = E.srsum;
  unsigned long long *srmax = E.srmax, *sn = E.sn, *stl = E.stl, *sev = E.sev;
  unsigned long long(*scnt)[NSLOT] = E.scnt;
  const int i = threadIdx.x;
  Slot &sl = st->slot[i];
  srsum[i] = st->rpart[i] + st->rpart[i + NSLOT];
  srmax[i] = sl.rmax; sn[i] = sl.n; stl[i] = sl.tiles; sev[i] = sl.events;
  for (int q = 0; q < 8; q++) scnt[q][i] = sl.counts[q];
  sl.rmax = 0ull; sl.n = 0ull; sl.tiles = 0ull; sl.events = 0ull;
  for (int q = 0; q < 8; q++) sl.counts[q] = 0ull;
  __syncthreads();
  for (int off = NSLOT / 2; off > 0; off >>= 1) {
    if (i < off) {
      srsum[i] += srsum[i + off];
      srmax[i] = srmax[i] > srmax[i + off] ? srmax[i] : srmax[i + off];
      sn[i] += sn[i + off];
      stl[i] += stl[i + off];
      sev[i] += sev[i + off];
      for (int q = 0; q < 8; q++) scnt[q][i] += scnt[q][i + off];
    }
    __syncthreads();
  }
  if (i != 0) return;
  const double rmax = __longlong_as_double((long long)srmax[0]);
  const double rsum = srsum[0];
  const long long nmc = (long long)sn[0];
  st->events += sev[0];
  for (int q = 0; q < 8; q++) st->counts[q] += scnt[q][0];
  if (st->la_variant)  // CountLiquid for the next sweep (main.f:637 of the laser-annealing build)
    for (int k = 0; k < 3; k++) {
      st->cc[k] += (long long)st->dcc[k];
      st->dcc[k] = 0ull;
      st->cl[k] = st->cinit[k] - st->cc[k];
    }
  // the passes of the next sweep visit the layers that held candidates; a
  // sweep whose passes were skipped (the run had stopped) collected none,
  // and the windows stay those of the last sweep that ran
  const bool passes_ran = !st->stop;
  int zlo = 0x7fffffff, zhi = -1;
  for (int c = 0; c < 8; c++) {
    if (passes_ran) {
      win[4 * c] = win[4 * c + 2];
      win[4 * c + 1] = win[4 * c + 3];
      win[4 * c + 2] = 0x7fffffff;
      win[4 * c + 3] = -1;
    }
    zlo = min(zlo, win[4 * c]);
    zhi = max(zhi, win[4 * c + 1]);
  }
  // the dependency-scheduled sweep: a new stamp, an empty queue, the window
  sched[0]++;
  sched[1] = 0u;
  sched[2] = 0u;
  sched[3] = (unsigned int)zlo;
  sched[4] = (unsigned int)zhi;
  if (*list_err && !st->err) st->err = *list_err;
  if (st->tau > 0.0 && !st->err && !st->stop) {
    // a window shortened to end at the time limit lands on it exactly
    st->clock = st->final_sweep ? st->max_time : st->clock + st->tau;
    st->sweeps_run++;
  }
  st->final_sweep = 0;
  st->sweeps++;
  st->rmax = rmax;
  st->n_mc = nmc;
  if (rsum > 0.0) st->rtot = rsum;
  if (st->tau_policy == 1)
    st->tau = rsum > 0.0 ? st->events_per_pass * (double)stl[0] / rsum : 0.0;
  else
    st->tau = rmax > 0.0 ? st->events_per_pass / rmax : 0.0;
  if (!st->stop) {
    if (nmc <= 20) st->stop = MSKG_STOP_FEW_MC;
    else if (!(rmax > 0.0)) st->stop = MSKG_STOP_NO_EVENTS;
    else if (st->clock >= st->max_time || st->events >= st->max_events) st->stop = STOP_AT_LIMIT;
  }
  // the last window of a run to a time limit ends at that limit: every tile
  // stops at max_time, as the serial loop stops at the first event past it
  if (!st->stop && st->tau > 0.0 && st->clock + st->tau >= st->max_time) {
    st->tau = st->max_time - st->clock;
    st->final_sweep = 1;
  }
}
__global__ void __launch_bounds__(NSLOT) k_sweep_end(RunState *st, int *win, const int *list_err,
                                                     unsigned int *sched) {
  __shared__ SweepEndLds E;
  sweep_end(st, win, list_err, sched, E);
}

// ---- the batched ensemble: one sweep over many lattices ------------------------
// Every replica keeps its own lattice, lists, windows, run state and random
// stream; a colour pass is one launch over that colour's tiles of all the
// replicas still running.  Before each colour pass a plan lays that
// colour's tiles of the replicas end to end -- made just before the pass,
// as k_pass reads its window when it starts, because the passes of the
// earlier colours widen the windows of the later ones when they enlist
// points in a new tile layer:
//   plan[c * (n + 1) + r]      c = 0..7, r = 0..n: tiles of colour c in the
//                              windows of replicas 0..r-1 (replicas that do
//                              not run hold none)
//   plan[8 * (n + 1) + c]      queue head of colour c, in units of 64
//   plan[8 * (n + 1) + 8]      replicas running the coming sweep
//   plan[8 * (n + 1) + 9 + r]  replica r runs the coming sweep
// The plan of colour 0 (start = 1) comes after the sweep's end and also
// decides which replicas run the next sweep.  Persistent waves take the
// entries of a colour's queue one at a time, so the tiles of large and
// small replicas share the GPU and the slowest tiles of a pass overlap
// other tiles' work instead of idling the rest.
constexpr int ENS_MAX = 1024;
static inline int64_t ens_plan_words(int n) { return 8 * (int64_t)(n + 1) + 9 + n; }

__global__ void __launch_bounds__(ENS_MAX) k_ens_plan(const PassParams *__restrict__ reps, int n, int *plan,
                                                      int colour, int start) {
  __shared__ int sc[ENS_MAX];
  __shared__ int nrun;
  const int r = threadIdx.x;
  int cnt = 0, run = 0;
  if (r == 0) nrun = 0;
  if (r < n) {
    const PassParams &P = reps[r];
    const RunState *st = P.st;
    if (start) {
      run = st->tau > 0.0 && !st->err && !st->stop;
      plan[8 * (n + 1) + 9 + r] = run;
    } else {
      run = plan[8 * (n + 1) + 9 + r];
    }
    // k_pass's condition, checked as its colour pass starts
    if (run && !st->err) {
      int zf;
      cnt = colour_tiles(P.L, P.ntz, colour, zf);
    }
  }
  __syncthreads();
  if (start && run) atomicAdd(&nrun, 1);
  // inclusive scan over the replicas, in replica order
  sc[r] = cnt;
  __syncthreads();
  for (int off = 1; off < ENS_MAX; off <<= 1) {
    const int v = r >= off ? sc[r - off] : 0;
    __syncthreads();
    sc[r] += v;
    __syncthreads();
  }
  if (r < n) plan[colour * (n + 1) + r + 1] = sc[r];
  if (r == 0) {
    plan[colour * (n + 1)] = 0;
    plan[8 * (n + 1) + colour] = 0;
    if (start) plan[8 * (n + 1) + 8] = nrun;
  }
}

// A pointer read from device memory is a generic (flat) pointer to the
// compiler, and every access through it a flat instruction, which waits on
// both the vector-memory and the LDS counters and takes 64-bit addresses in
// vector registers.  Every pointer of a replica's parameters is to global
// memory: a round trip through the global address space states that.  In
// k_pass_ens the lattice is laundered where a wave changes replica but not
// its first value, so the loop's merge of the two stays generic and most of
// the tile's accesses remain flat: laundering the first value as well makes
// them global but adds spills, and measured 2% slower on the 64-replica
// SiGe ensemble at tile 16 (660.7 against 646.5 ms).
template <typename T>
__device__ inline T *as_global(T *p) {
  // through an integer: a direct cast there and back folds away
  __attribute__((address_space(1))) T *g = (__attribute__((address_space(1))) T *)(uintptr_t)p;
  return (T *)g;
}
__device__ inline void to_global(Lattice &L) {
  L.site = as_global(L.site); L.rate = as_global(L.rate); L.cls = as_global(L.cls);
  L.inlist = as_global(L.inlist); L.tl_idx = as_global(L.tl_idx); L.tl_n = as_global(L.tl_n);
  L.tl_own = as_global(L.tl_own); L.tile_r = as_global(L.tile_r); L.list_err = as_global(L.list_err);
  L.win = as_global(L.win); L.geo = as_global(L.geo); L.tab = as_global(L.tab); L.key = as_global(L.key);
  L.cl = as_global(L.cl);
}

// The replica whose tiles hold entry q of colour pc's queue: the last r with
// pc[r] <= q (pc is the colour's prefix, pc[n] > q), 64 prefixes at a time.
__device__ inline int ens_replica(const int *pc, int n, int q, int lane) {
  const int j1 = 64 * lane;
  const int v1 = pc[min(j1, n)];  // every lane loads (an index clamped to the prefix)
  const int b = __popcll(__ballot(j1 <= n && v1 <= q)) - 1;  // pc[0] = 0 <= q: b >= 0
  const int j2 = 64 * b + lane;
  const int v2 = pc[min(j2, n)];
  return 64 * b + __popcll(__ballot(j2 <= n && v2 <= q)) - 1;
}

template <bool LA, bool COV>
__global__ void __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(3, 3)))
k_pass_ens(const PassParams *__restrict__ reps, int n, int *plan, int colour) {
  extern __shared__ unsigned char smem[];
  __shared__ PassLds S;
  const int lane = threadIdx.x;
  const int *pc = plan + colour * (n + 1);
  const int total = __builtin_amdgcn_readfirstlane(pc[n]);
  unsigned int *const head = (unsigned int *)(plan + 8 * (n + 1) + colour);
  // the replica of the last entry: its entries [lo, hi) of the queue, its
  // lattice and run state, the window's first layer and the window tau
  int r = -1, lo = 0, hi = 0, zfirst = 0, failed = 0;
  double tau = 0.0;
  Lattice L = reps[0].L;
  RunState *st = nullptr;
  // Every decision of this loop is wave-uniform (readfirstlane / ballot)
  // and no memory access sits under a lane-dependent branch.  A wave's
  // first entry is its own (blockIdx.x), so that the waves do not all meet
  // at the queue head when the pass starts (same-address atomics are
  // served one after another), and an empty pass takes no atomic at all;
  // the entries from gridDim.x on are taken from the queue.
  int q = blockIdx.x;
  for (;;) {
    if (q >= total) break;
    if (q < lo || q >= hi) {
      r = __builtin_amdgcn_readfirstlane(ens_replica(pc, n, q, lane));
      lo = __builtin_amdgcn_readfirstlane(pc[r]);
      hi = __builtin_amdgcn_readfirstlane(pc[r + 1]);
      L = reps[r].L;
      to_global(L);
      st = as_global(reps[r].st);
      // a replica that has failed in this sweep runs no further tiles here
      // (k_pass checks at the start of each colour pass)
      failed = __builtin_amdgcn_readfirstlane(__hip_atomic_load(&st->err, __ATOMIC_RELAXED,
                                                                __HIP_MEMORY_SCOPE_AGENT));
      tau = st->tau;
      colour_tiles(L, reps[r].ntz, colour, zfirst);
    }
    if (!failed) {
      const PassParams &P = reps[r];
      const int t = colour_tile(L, colour, zfirst, q - lo);
      int32_t *lk = (int32_t *)smem;
      float *lr = (float *)(lk + P.cap);
      uint8_t *lc = (uint8_t *)(lr + P.cap);
      // a wave stops its tiles of a replica whose tile failed (as k_pass's wave returns)
      failed = __builtin_amdgcn_readfirstlane(pass_tile<LA, COV>(P, L, st, tau, t, colour, S, lk, lr, lc, lane));
    }
    // the next entry: every lane adds 1 (one atomic of 64 for the wave)
    const unsigned int took = __hip_atomic_fetch_add(head, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    q = (int)gridDim.x + __builtin_amdgcn_readfirstlane((int)(took >> 6));
  }
}

// Tile sums and the sweep's end for every replica that ran the sweep (the
// same per-replica work as k_rsum and k_sweep_end, so the replica's window,
// clock and stop are those it gets alone).
__global__ void __launch_bounds__(256) k_rsum_ens(const PassParams *__restrict__ reps, int n, const int *plan) {
  __shared__ double s[256];
  const int r = blockIdx.y;
  if (!plan[8 * (n + 1) + 9 + r]) return;
  const PassParams &P = reps[r];
  rsum_block(P.L.tile_r, P.L.ntx * P.L.nty * P.ntz, P.rsum_chunk, P.st, blockIdx.x, s);
}
__global__ void __launch_bounds__(NSLOT) k_sweep_end_ens(const PassParams *__restrict__ reps, int n,
                                                         const int *plan) {
  __shared__ SweepEndLds E;
  const int r = blockIdx.x;
  if (!plan[8 * (n + 1) + 9 + r]) return;
  const PassParams &P = reps[r];
  sweep_end(P.st, P.L.win, P.L.list_err, P.sched, E);
}

// ---- device-side initial state ----------------------------------------------
// Diamond sites of the reference grid: x, y, z multiples of 3 whose thirds
// have equal parity; (x+y+z)/3 = 0 (mod 4) is fcc sublattice A (bonds
// JumpDia(:,1:4)), = 3 (mod 4) sublattice B (JumpDia(:,5:8))
// (initial/SetSi.f:36-45, modules/defsystem.f:112-113).
__device__ inline int dia_type(int x, int y, int z) {
  if (x % 3 || y % 3 || z % 3) return 0;
  int x1 = x / 3, y1 = y / 3, z1 = z / 3;
  if ((x1 & 1) != (y1 & 1) || (x1 & 1) != (z1 & 1)) return 0;
  int m = (x1 + y1 + z1) & 3;
  return m == 0 ? 1 : (m == 3 ? 2 : 0);
}
__constant__ int c_jump_dia[8][3] = {{3, 3, 3}, {3, -3, -3}, {-3, -3, 3}, {-3, 3, -3},
                                     {3, 3, -3}, {3, -3, 3}, {-3, -3, -3}, {-3, 3, 3}};

// The flat (001) slab of SetSi.f:25-220 (sic=0: Si on both sublattices,
// wall layers z=0..5 and LenZ-6..LenZ-1) or SetSiC3C.f:25-198 (sic=1: Si on
// A, C on B, no walls): bulk below len-3, the top atomic layer as MC
// particles with PTransE(s,2,0,...), the empty sites above with
// coordination 2 and their deposition rate.  Every point is written.
__global__ void __launch_bounds__(256) k_build_slab(Lattice L, int len, int sic, double fraction, uint32_t key0,
                                                   uint32_t key1, int top_lz) {
  const Tables &t = *L.tab;
  const int64_t n = L.box.npts;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
    int x, y, z;
    pxyz(L.box, p, x, y, z);
    Site s;
    for (int a = 0; a < 4; a++) s.d[a][0] = s.d[a][1] = s.d[a][2] = 0;
    s.kc = 0; s.valid = 0; s.temp = 0;
    int cls = CLS_NONE;
    double r = 0.0;
    // top_lz: LenZ of the whole box (the top wall planes lie above the stored ones when only the bottom planes are stored)
    const bool wall = sic != 1 && ((z >= 0 && z <= 5) || (z >= top_lz - 6 && z <= top_lz - 1));
    const int ty = dia_type(x, y, z);
    if (wall) {
      s.kc = KIND_WALL;
    } else if (ty && z <= len) {
      int sp = (sic == 1 && ty == 2) ? 2 : 1;
      if (sic == 2) {  // random substitutional alloy: species 2 with probability `fraction`
        U4 c;
        c.x = (uint32_t)p; c.y = (uint32_t)(p >> 32); c.z = 0x5e6e0001u; c.w = 0;
        const U4 d = philox4x32_10(c, key0, key1);
        if (u01_open(d.x, d.y) < fraction) sp = 2;
      }
      NL l;
      for (int i = 0; i < 4; i++) {
        const int *j = c_jump_dia[(ty == 1 ? 0 : 4) + i];
        l.x[i] = x + j[0]; l.y[i] = y + j[1]; l.z[i] = z + j[2];
        if (l.x[i] >= L.box.lx) l.x[i] -= L.box.lx;
        if (l.x[i] < 0) l.x[i] += L.box.lx;
        if (l.y[i] >= L.box.ly) l.y[i] -= L.box.ly;
        if (l.y[i] < 0) l.y[i] += L.box.ly;
      }
      set_list(L.box, s, x, y, z, l);
      if (z <= len - 4) {
        s.kc = (uint8_t)(sp | (4 << 3));
      } else if (z <= len - 1) {
        s.kc = (uint8_t)(sp | (2 << 3));
        if (z >= t.frozen_below_z) {
          cls = CLS_EVAP;
          r = PEv(t, sp, 2, 0, 0, 0, 0, 0);
        }
      } else {
        s.kc = (uint8_t)(KIND_EMPTY | (2 << 3));
        cls = CLS_ADSORB;
        r = pd_of(t, 1, 2, 0) + pd_of(t, 2, 2, 0) + pd_of(t, 3, 2, 0) + t.pabs[0 + 6] + t.pabs[1 + 6] + t.pabs[2 + 6];
      }
    }
    L.site[p] = s;
    L.rate[p] = cls ? (float)r : 0.0f;
    L.cls[p] = (uint8_t)cls;
  }
}


// ---- the shaped initial states (main.f:224-294) ----------------------------------
// SetNC / SetNCSp / SetInvPOff / SetSiFINFET / SetCAD all place atoms and
// walls first, then run one coordination pass (an atom with no occupied
// neighbour is removed) and one probability pass that fills the lists.  On
// the device: k_shape_place writes every stored point's kind from the
// builder's geometry, k_shape_isolated removes the isolated atoms,
// k_shape_records gives every site its coordination and neighbour list,
// and the rate sweep evaluates the MC particles.  An isolated atom's
// neighbours are all empty, so no removal changes a coordination another
// removal reads, and the result does not depend on the order of the
// reference's sequential pass.
enum : int { SHAPE_BOX = 0, SHAPE_SPHERE = 1, SHAPE_INVP = 2, SHAPE_FINFET = 3, SHAPE_CAD = 4 };
struct ShapeSpec {
  int shape;
  int l1, l2, l3, l4;   // Len1..Len4 as the builder takes them
  int cx, cy, cz;       // SetNC / SetNCSp centre (LenX/2, LenY/2, LenZ/2)
  int lz;               // LenZ
  int xa, xb;           // SetSiFINFET: the fin lies in xa <= x < xb
  int zfix;             // SetInvPOff: atoms at z <= zfix are bulk of coordination 4 (-1: none)
  int walls;            // an empty site's coordination leaves walls out
  const int8_t *ids;    // SetCAD geometry ids, [z][y][x] for z < LenZ
  double fraction;      // SetCAD: species 2 with this probability (Philox)
  uint32_t key0, key1;
};

// The builder's content of grid point (x, y, z): KIND_WALL, a species, or
// KIND_EMPTY.  Walls lie on every grid point of their region, atoms on the
// diamond sites of theirs (SetNC.f, SetNCSp.f, SetInvPOff.f:21-286,
// SetSiFINFET.f:25-312, SetCAD.f:84-190).
__device__ inline int shape_content(const ShapeSpec &S, int x, int y, int z, int lx, int ly, int64_t p) {
  const int t = dia_type(x, y, z);
  switch (S.shape) {
    case SHAPE_BOX:
    case SHAPE_SPHERE: {
      const int l2 = S.shape == SHAPE_BOX ? S.l2 : S.l1, l3 = S.shape == SHAPE_BOX ? S.l3 : S.l1;
      if (!t || x < S.cx - S.l1 / 2 || x > S.cx + S.l1 / 2 - 1 || y < S.cy - l2 / 2 || y > S.cy + l2 / 2 - 1 ||
          z < S.cz - l3 / 2 || z > S.cz + l3 / 2 - 1)
        return KIND_EMPTY;
      if (S.shape == SHAPE_SPHERE) {
        const int d2 = (x - S.cx) * (x - S.cx) + (y - S.cy) * (y - S.cy) + (z - S.cz) * (z - S.cz);
        if (!(d2 <= S.l1 * S.l1 / 4)) return KIND_EMPTY;
      }
      return t == 1 ? 1 : 2;
    }
    case SHAPE_INVP: {  // l1: Len (pyramid size), l2: Len3 (substrate top)
      if (!t) return KIND_EMPTY;
      if (z <= 20) return t == 1 ? 1 : 2;
      if (z > S.l2 - 1) return KIND_EMPTY;
      // the reference's single-precision test: integer sums minus
      // 1.10*(z-Len3) in REAL(4) against REAL(4) Len
      const int lenh = lx / 2 - 1;
      const float dz = 1.10f * (float)(z - S.l2), fl = (float)S.l1;
      const bool hole = ((float)(abs(x - lenh) + abs(y - lenh)) - dz <= fl) ||
                        ((float)(abs(x) + abs(y)) - dz <= fl) || ((float)(abs(x - lx) + abs(y)) - dz <= fl) ||
                        ((float)(abs(x) + abs(y - ly)) - dz <= fl) ||
                        ((float)(abs(x - lx) + abs(y - ly)) - dz <= fl);
      return hole ? KIND_EMPTY : (t == 1 ? 1 : 2);
    }
    case SHAPE_FINFET: {  // l1: Len (Si top), l2: wall bottom, l3: wall top
      const bool wall = z <= 5 || (z >= S.lz - 6 && z <= S.lz - 1) || (z >= S.l3 && z <= S.lz - 7) ||
                        (z >= S.l2 && z <= S.lz - 7 && (x < S.xa || x >= S.xb));
      if (wall) return KIND_WALL;
      return (t && z >= 6 && z <= S.l1 - 1) ? 1 : KIND_EMPTY;
    }
    default: {  // SHAPE_CAD
      if (z <= 3 || (z >= S.lz - 4 && z <= S.lz - 1)) return KIND_WALL;
      if (z > S.lz - 5) return KIND_EMPTY;
      const int id = S.ids[(int64_t)x + (int64_t)lx * ((int64_t)y + (int64_t)ly * z)];
      if (id == 10) return KIND_WALL;
      if (id != 1 || !t) return KIND_EMPTY;
      if (S.fraction > 0.0) {  // the reference's RAND() < alloyfraction, from the engine's stream
        U4 c;
        c.x = (uint32_t)p; c.y = (uint32_t)(p >> 32); c.z = 0x5e6e0002u; c.w = 0;
        const U4 d = philox4x32_10(c, S.key0, S.key1);
        if (u01_open(d.x, d.y) < S.fraction) return 2;
      }
      return 1;
    }
  }
}

__global__ void __launch_bounds__(256) k_shape_place(Lattice L, ShapeSpec S) {
  const int64_t n = L.box.npts;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
    int x, y, z;
    pxyz(L.box, p, x, y, z);
    Site s;
    for (int a = 0; a < 4; a++) s.d[a][0] = s.d[a][1] = s.d[a][2] = 0;
    s.valid = 0;
    s.temp = L.site[p].temp;  // a temperature set before the build stays (SetCAD.f:62-65)
    s.kc = (uint8_t)shape_content(S, x, y, z, L.box.lx, L.box.ly, p);
    L.site[p] = s;
    L.rate[p] = 0.0f;
    L.cls[p] = 0;
    L.inlist[p] = 0;
  }
}

// The four neighbour sites of diamond site (x, y, z) of type t, x and y
// wrapped once (the builders' JumpDia lists).
__device__ inline NL dia_list(const Box &b, int t, int x, int y, int z) {
  NL l;
  for (int i = 0; i < 4; i++) {
    const int *j = c_jump_dia[(t == 1 ? 0 : 4) + i];
    l.x[i] = x + j[0]; l.y[i] = y + j[1]; l.z[i] = z + j[2];
    if (l.x[i] >= b.lx) l.x[i] -= b.lx;
    if (l.x[i] < 0) l.x[i] += b.lx;
    if (l.y[i] >= b.ly) l.y[i] -= b.ly;
    if (l.y[i] < 0) l.y[i] += b.ly;
  }
  return l;
}
__device__ inline int kind_at(const Lattice &L, int x, int y, int z) {
  const int64_t q = pidx_checked(L.box, x, y, z);
  return q < 0 ? KIND_EMPTY : s_kind(L.site[q]);
}

// An atom none of whose four neighbours is occupied (an atom or a wall) is
// removed (SetCAD.f:229-246, SetNC.f's same pass); atoms of a fixed bulk
// region are left alone.
__global__ void __launch_bounds__(256) k_shape_isolated(Lattice L, ShapeSpec S) {
  const int64_t n = L.box.npts;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
    const Site sp = L.site[p];
    if (!kind_species(s_kind(sp))) continue;
    int x, y, z;
    pxyz(L.box, p, x, y, z);
    if (z <= S.zfix) continue;
    const int t = dia_type(x, y, z);
    if (!t) continue;
    const NL l = dia_list(L.box, t, x, y, z);
    int occ = 0;
    for (int i = 0; i < 4; i++) occ += kind_occupied(kind_at(L, l.x[i], l.y[i], l.z[i]));
    if (!occ) {
      Site e = sp;
      e.kc = KIND_EMPTY;  // LattCoo = 0, the temperature kept
      L.site[p] = e;
    }
  }
}

// Coordination and neighbour list of every diamond site (SetCAD.f:194-416):
// an atom counts its occupied neighbours (walls included), an empty site
// its atoms (with walls = 1; every occupied neighbour otherwise); an empty
// site of coordination 1 keeps only that neighbour, in entry 1.  Atoms and
// empty sites of coordination >= 1 get their list; walls and the empty
// sites of coordination 0 keep a bare record.
__global__ void __launch_bounds__(256) k_shape_records(Lattice L, ShapeSpec S, int *err) {
  const int64_t n = L.box.npts;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
    Site s = L.site[p];
    const int kind = s_kind(s);
    if (kind == KIND_WALL) continue;
    int x, y, z;
    pxyz(L.box, p, x, y, z);
    const int t = dia_type(x, y, z);
    if (!t) continue;
    NL l = dia_list(L.box, t, x, y, z);
    int coor = 0, last = -1;
    for (int i = 0; i < 4; i++) {
      const int k = kind_at(L, l.x[i], l.y[i], l.z[i]);
      if (kind_species(kind) ? kind_occupied(k) : (S.walls ? kind_species(k) : kind_occupied(k))) {
        coor++;
        last = i;
      }
    }
    if (kind_species(kind) && z <= S.zfix) coor = 4;
    if (!kind_species(kind) && coor == 0) continue;
    if (!kind_species(kind) && coor == 1) {
      // the NBuff compaction of SetCAD.f:316-327: entry 1 the one atom, the rest unfixed
      const int lx0 = l.x[last], ly0 = l.y[last], lz0 = l.z[last];
      for (int i = 0; i < 4; i++) l.x[i] = l.y[i] = l.z[i] = 0;
      l.x[0] = lx0; l.y[0] = ly0; l.z[0] = lz0;
    }
    s_set_coor(s, coor);
    if (!set_list(L.box, s, x, y, z, l)) atomicExch(err, ERR_GEOM);
    L.site[p] = s;
  }
}

// The placeholder deposition rate of an initial state is Get_Prob_Ini's
// REAL(4) 1.e-6 (Get_Prob_Ini.f:104), not Get_Prob's 1.e-8.
__global__ void __launch_bounds__(256) k_initial_placeholders(Lattice L) {
  const int64_t n = L.box.npts;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x)
    if (L.cls[p] == CLS_PLACEHOLDER) L.rate[p] = (float)(double)1.e-6f;
}

// Walls on the grid points off the stored lattice, one bit per grid point
// (what the phase maps report): thread w sets word w.
__global__ void __launch_bounds__(256) k_shape_offwall(Lattice L, ShapeSpec S, uint32_t *bits, int64_t words) {
  const int lx = L.box.lx, ly = L.box.ly;
  for (int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; w < words; w += (int64_t)gridDim.x * blockDim.x) {
    uint32_t v = 0;
    for (int b = 0; b < 32; b++) {
      const int64_t g = w * 32 + b;
      const int x = (int)(g % lx), y = (int)((g / lx) % ly), z = (int)(g / ((int64_t)lx * ly));
      if (z > L.box.lz || valid_parity(x, y, z)) continue;
      if (shape_content(S, x, y, z, lx, ly, 0) == KIND_WALL) v |= 1u << b;
    }
    bits[w] = v;
  }
}

// ---- serial BKL over a batch of independent lattices --------------------------
// One workgroup (one wave) per replica runs the reference's serial algorithm:
// every event picks one MC particle of the whole lattice with probability
// rate/R (R the FP64 total) and advances the clock by -ln(u)/R.  The rates
// are organised as a three-level sum: the FP32 rates of each tile's
// candidate list, FP64 tile sums, and FP64 sums of 64 consecutive tiles; an
// event rescans only the lists of the tiles it touched.
struct SerialDesc {
  Lattice L;
  double *tsum;   // [ntiles]
  double *bsum;   // [nblk]
  RunState *st;
  mskg_event *log;
  unsigned long long log_cap;
  long long max_events;
  double max_time;
  uint32_t key0, key1;
  int ntiles, nblk, replica;
};

// Rate sum of one tile's list (every class); entries whose class fell to 0
// leave the list.  The compaction is in place: an entry moves only to a
// position at or before the one it was read from in the same step.
__device__ double tile_sum_compact(const Lattice &L, int t, int lane) {
  const TileGeom g = tile_geom(t, L.tile, L.ntx, L.nty, L.box.lz);
  const int m = min(L.tl_n[t], L.capg);
  int32_t *idx = L.tl_idx + (int64_t)t * L.capg;
  double part = 0.0;
  int nout = 0;
  for (int j0 = 0; j0 < m; j0 += WAVE) {
    const int j = j0 + lane;
    int k = 0;
    bool keep = false;
    if (j < m) {
      k = idx[j];
      const P3 p = tile_point(L.box, g, L.tile, k);
      const uint8_t c = L.cls[p.i];
      keep = c != 0;
      if (keep) part += (double)L.rate[p.i];
      else L.inlist[p.i] = 0;
    }
    const unsigned long long msk = __ballot(keep);
    const int slot = nout + __popcll(msk & ((1ull << lane) - 1ull));
    if (keep) idx[slot] = k;
    nout += __popcll(msk);
  }
  part = wave_sum(part);
  if (lane == 0) L.tl_n[t] = nout;
  return part;
}

// Index of the first element of v[0..n) (FP64, read from memory) whose
// inclusive prefix exceeds target; *below receives the prefix before it.
// Rounding at the top end falls back to the last positive element.
__device__ int scan_pick(const double *v, int n, double target, int lane, double *below) {
  double acc = 0.0;
  int last_nz = -1;
  double acc_last = 0.0;
  for (int base = 0; base < n; base += WAVE) {
    const int j = base + lane;
    const double x = j < n ? v[j] : 0.0;
    const double inc = wave_incl_scan(x, lane) + acc;
    const unsigned long long nzm = __ballot(x > 0.0);
    if (nzm) {
      const int hi = 63 - __clzll((long long)nzm);
      last_nz = base + hi;
      acc_last = __shfl(inc - x, hi, WAVE);
    }
    const unsigned long long hit = __ballot(j < n && x > 0.0 && inc > target);
    if (hit) {
      const int l = __ffsll((long long)hit) - 1;
      *below = __shfl(inc - x, l, WAVE);
      return base + l;
    }
    acc = __shfl(inc, WAVE - 1, WAVE);
  }
  *below = acc_last;
  return last_nz;
}

// Replicas of up to SERIAL_LDS_TILES tiles keep their tile and block sums on
// chip; larger ones use the engine's global arrays.
constexpr int SERIAL_LDS_TILES = 4096;
template <bool LA, bool COV>
__global__ void __launch_bounds__(WAVE) k_serial(const SerialDesc *descs) {
  __shared__ EvScratch sc;
  __shared__ int s_tiles[MAX_TOUCHED], s_ntl, s_err;
  __shared__ double s_tsum[SERIAL_LDS_TILES], s_bsum[SERIAL_LDS_TILES / WAVE];
  __shared__ unsigned long long s_cnt[8];
  const int lane = threadIdx.x;
  const SerialDesc D = descs[blockIdx.x];
  Lattice L = D.L;
  L.cur_tile = -1;
  RunState *st = D.st;
  if (st->err) return;
  const bool on_chip = D.ntiles <= SERIAL_LDS_TILES;
  double *tsum = on_chip ? s_tsum : D.tsum;
  double *bsum = on_chip ? s_bsum : D.bsum;
  if (lane < 8) s_cnt[lane] = 0;
  // initial sums
  for (int t = 0; t < D.ntiles; t++) {
    const double ts = tile_sum_compact(L, t, lane);
    if (lane == 0) tsum[t] = ts;
  }
  __syncthreads();
  for (int b = 0; b < D.nblk; b++) {
    const int t = b * WAVE + lane;
    const double v = wave_sum(t < D.ntiles ? tsum[t] : 0.0);
    if (lane == 0) bsum[b] = v;
  }
  __syncthreads();
  double clock = st->clock;
  unsigned long long events = st->events, nev = 0;
  int stop = MSKG_STOP_LIMIT, err = 0;
  U4 ctr;
  ctr.x = 0x5e71a1u;  // the stream depends on the engine's seed only, not on its place in the batch
  while ((long long)nev < D.max_events && clock < D.max_time) {
    double part = 0.0;
    for (int b = lane; b < D.nblk; b += WAVE) part += bsum[b];
    const double R = wave_sum(part);
    if (!(R > 0.0)) { stop = MSKG_STOP_NO_EVENTS; break; }
    ctr.y = (uint32_t)events;
    ctr.z = (uint32_t)(events >> 32);
    ctr.w = 0;
    // one Philox block per event: waiting time, selection, event type and
    // stacking choice, 32 random bits each
    const U4 d = philox4x32_10(ctr, D.key0, D.key1);
    const double dt = -log(u01_32(d.x)) / R;
    double target = u01_32(d.y) * R, below = 0.0;
    // block -> tile -> list entry
    const int b = scan_pick(bsum, D.nblk, target, lane, &below);
    if (b < 0) { stop = MSKG_STOP_NO_EVENTS; break; }
    target -= below;
    const int t0 = b * WAVE;
    const int tl = scan_pick(tsum + t0, min(WAVE, D.ntiles - t0), target, lane, &below);
    if (tl < 0) { err = ERR_SUMS; break; }
    const int t = t0 + tl;
    target -= below;
    const TileGeom g = tile_geom(t, L.tile, L.ntx, L.nty, L.box.lz);
    const int m = min(L.tl_n[t], L.capg);
    const int32_t *idx = L.tl_idx + (int64_t)t * L.capg;
    double acc = 0.0;
    int sel = -1, last_nz = -1;
    for (int base = 0; base < m; base += WAVE) {
      const int j = base + lane;
      double v = 0.0;
      if (j < m) {
        const P3 p = tile_point(L.box, g, L.tile, idx[j]);
        if (L.cls[p.i]) v = (double)L.rate[p.i];
      }
      const unsigned long long nzm = __ballot(v > 0.0);
      if (nzm) last_nz = base + 63 - __clzll((long long)nzm);
      const double inc = wave_incl_scan(v, lane) + acc;
      const unsigned long long hit = __ballot(j < m && v > 0.0 && inc > target);
      if (hit) { sel = base + __ffsll((long long)hit) - 1; break; }
      acc = __shfl(inc, WAVE - 1, WAVE);
    }
    if (sel < 0) sel = last_nz;
    if (sel < 0) { err = ERR_SUMS; break; }
    const P3 s = tile_point(L.box, g, L.tile, idx[sel]);
    Choice ch;
    ch.arm = -1;
    ch.u_arm = u01_32(d.w);
    const int itrans = pick_itrans<LA, COV>(L, s, L.cls[s.i], u01_32(d.z));
    const int coor0 = s_coor(L.site[s.i]);
    __syncthreads();
    EvResult res;
    execute<LA, COV>(L, R, s, itrans, ch, res, sc, lane);
    __syncthreads();
    clock += dt;
    if (lane == 0) {
      s_err = res.err;
      if (D.log) {
        const unsigned long long li = st->log_n + nev;
        if (li < D.log_cap) {
          mskg_event ev;
          ev.iter = (int64_t)(events + 1);
          ev.site[0] = s.x; ev.site[1] = s.y; ev.site[2] = s.z;
          ev.itrans = itrans;
          ev.arm = res.arm;
          ev.coor = coor0;
          ev.rtot = R;
          ev.time = clock;
          D.log[li] = ev;
        }
      }
      const int slot = itrans <= 3 ? itrans - 1 : (itrans == 4 ? 3 : (itrans == 5 ? 4 : itrans - 1));
      if (slot >= 0 && slot < 8) s_cnt[slot]++;
      if (COV && itrans == 5 && res.species) st->desorbed[res.species - 1]++;
      int nt = 0;
      for (int q = 0; q < sc.ntouched; q++) {
        if (sc.touched[q].i < 0) continue;
        const int tq = tile_of(L, sc.touched[q]);
        bool dup = false;
        for (int r = 0; r < nt; r++) dup |= s_tiles[r] == tq;
        if (!dup) s_tiles[nt++] = tq;
      }
      s_ntl = nt;
    }
    __syncthreads();
    events++;
    nev++;
    if (s_err) { err = s_err; break; }
    const int ntl = s_ntl;
    for (int q = 0; q < ntl; q++) {
      const double ts = tile_sum_compact(L, s_tiles[q], lane);
      if (lane == 0) tsum[s_tiles[q]] = ts;
    }
    __syncthreads();
    for (int q = 0; q < ntl; q++) {
      const int bq = s_tiles[q] / WAVE;
      bool dup = false;
      for (int r = 0; r < q; r++) dup |= s_tiles[r] / WAVE == bq;
      if (dup) continue;
      const int tt = bq * WAVE + lane;
      const double v = wave_sum(tt < D.ntiles ? tsum[tt] : 0.0);
      if (lane == 0) bsum[bq] = v;
    }
    __syncthreads();
  }
  if (lane == 0) {
    st->clock = clock;
    st->events = events;
    st->log_n += nev;
    st->stop = stop;
    for (int q = 0; q < 8; q++) st->counts[q] += s_cnt[q];
    if (err) st->err = err;
  }
}

// ---- adjoint: replay with per-class accounting ---------------------------------
// For a recorded trajectory {(event k, dt_k)} the log-likelihood is
//   sum_k [ log r(event k) - R_k dt_k ]  with R_k = sum_c n_c(k) r_c,
// where n_c(k) is the number of points whose rate comes from class c before
// event k.  Replaying the log, this kernel accumulates S_c (selections of
// class c) and A_c = sum_k n_c(k) dt_k; the likelihood and its gradient with
// respect to any parameter of the rate tables follow on the host as
// sum_c S_c log r_c - A_c r_c.  A_c is accumulated lazily per class in LDS
// (A_c += n_c (t - tau_c) whenever n_c changes) on a clock that restarts
// every ADJ_EPOCH events, so that the interval lengths keep full precision.
constexpr int ADJ_EPOCH = 1024;
constexpr int ADJ_NS = KEY_N + 9;  // + deposition of species s at coordination c
struct AdjDesc {
  Lattice L;                 // L.key set
  int32_t *key_seen;         // key each point was last accounted under
  float *rate_seen;          // its rate then (isolated atoms)
  const mskg_event *log;
  const double *dt;
  long long n;
  const int *cnt0;           // [KEY_N + 1] initial class populations, then the unsupported points
  const double *iso0;        // initial sum of isolated-atom rates
  const int *sweep_err;      // the rate sweep's error flag (the engine's)
  double *S;                 // [ADJ_NS]
  double *A;                 // [KEY_N]
  double *scal;              // [4]: sum dt, integral of isolated rates, sum of their log rates, events
  RunState *st;
  int32_t *hist;             // may be null: [n / every][KEY_N] class populations after every `every` events
  long long every;
  double *acc;               // [2][KEY_N]: each class's population integral, then the time of its last change
};

__global__ void __launch_bounds__(WAVE) k_adjoint(const AdjDesc *descs) {
  extern __shared__ unsigned char smem[];
  int *cnt = (int *)smem;                 // [KEY_N] class populations
  __shared__ EvScratch sc;
  __shared__ int s_bad;
  __shared__ int s_o[MAX_TOUCHED], s_nw[MAX_TOUCHED];
  __shared__ float s_ro[MAX_TOUCHED], s_rn[MAX_TOUCHED];
  const int lane = threadIdx.x;
  const AdjDesc D = descs[blockIdx.x];
  Lattice L = D.L;
  L.cur_tile = -1;
  // the per-class integrals and change times in the segment's own device
  // memory (LDS holds the populations only, so that more waves fit a CU)
  double *const A = D.acc, *const tau = D.acc + KEY_N;
  for (int c = lane; c < KEY_N; c += WAVE) {
    A[c] = 0.0;
    tau[c] = 0.0;
    cnt[c] = D.cnt0[c];
  }
  for (int c = lane; c < ADJ_NS; c += WAVE) D.S[c] = 0.0;
  __syncthreads();
  double t = 0.0, ttot = 0.0, r_iso = *D.iso0, tau_iso = 0.0, iso_int = 0.0, iso_log = 0.0;
  int err = 0;
  long long k = 0;
  // the log is read one event ahead: its loads overlap the event before
  mskg_event ev_next;
  double dt_next = 0.0;
  if (D.n > 0) {
    ev_next = D.log[0];
    dt_next = D.dt[0];
  }
  for (; k < D.n; k++) {
    const mskg_event ev = ev_next;
    const double dtk = dt_next;
    if (k + 1 < D.n) {
      ev_next = D.log[k + 1];
      dt_next = D.dt[k + 1];
    }
    t += dtk;
    ttot += dtk;
    const P3 s = at(L, ev.site[0], ev.site[1], ev.site[2]);
    if (lane == 0) {
      int bad = 0;
      const int ks = s.i >= 0 ? D.key_seen[s.i] : -1;
      if (s.i < 0 || L.cls[s.i] == 0) bad = ERR_NOT_MC;
      else if (ks < 0) bad = ERR_COVERAGE;  // a rate class the likelihood does not cover
      else if (ks >= KEY_ADS && ks < KEY_PH) {
        if (ev.itrans < 1 || ev.itrans > 3) bad = ERR_NOT_MC;
        else atomicAdd(&D.S[KEY_N + (ev.itrans - 1) * 3 + (ks - KEY_ADS)], 1.0);
      } else if (ks == KEY_ISO) {
        iso_log += log((double)D.rate_seen[s.i]);
        atomicAdd(&D.S[KEY_ISO], 1.0);
      } else {
        atomicAdd(&D.S[ks], 1.0);  // no return value: the wave does not wait for it
      }
      s_bad = bad;
    }
    __syncthreads();
    if (s_bad) { err = s_bad; break; }
    Choice ch;
    ch.arm = ev.arm;
    ch.u_arm = 0.0;
    EvResult res;
    execute<false, false>(L, ev.rtot, s, ev.itrans, ch, res, sc, lane);  // the likelihood excludes LA and coverage tables
    __syncthreads();
    // the touched points' keys and rates, before and after, one lane each
    // (a point listed twice is taken once); every class change of the event
    // happens at time t, so the accounting below does not depend on order
    const int nt = __builtin_amdgcn_readfirstlane(sc.ntouched);
    int o = -3, nw = -1;  // o = -3: a point an earlier lane holds (or none)
    float ro = 0.0f, rn = 0.0f;
    if (lane < nt) {
      const int64_t p = sc.touched[lane].i;
      bool first = p >= 0;
      for (int q = 0; q < lane; q++)
        if (sc.touched[q].i == p) first = false;
      if (first) {
        o = D.key_seen[p];
        nw = L.key[p];
        ro = D.rate_seen[p];
        rn = L.rate[p];
        D.key_seen[p] = nw;
        D.rate_seen[p] = rn;
      }
      s_o[lane] = o;
      s_nw[lane] = nw;
      s_ro[lane] = ro;
      s_rn[lane] = rn;
    }
    int bad = __builtin_amdgcn_readfirstlane(res.err);  // res is complete on lane 0
    if (!bad && __ballot(o != -3 && nw == KEY_UNSUPPORTED)) bad = ERR_COVERAGE;
    if (!bad) {
      // Every class change of the event happens at time t: a class's
      // integral takes its population up to t once (a second change at t
      // adds population x 0), so the earliest lane holding a class, as old
      // or new class in touched order, updates it with the population before
      // the event, and the populations then change by atomics.
      const bool ch = o != -3 && o != nw;
      const int co = ch && o >= 0 ? o : -1, cn = ch && nw >= 0 ? nw : -1;
      bool do_o = co >= 0, do_n = cn >= 0;
      for (int q = 0; q < nt; q++) {
        const int qo = __builtin_amdgcn_readlane(co, q), qn = __builtin_amdgcn_readlane(cn, q);
        if (q < lane) {
          if (qo == co || qn == co) do_o = false;
          if (qo == cn || qn == cn) do_n = false;
        }
      }
      if (do_o) { A[co] += cnt[co] * (t - tau[co]); tau[co] = t; }
      if (do_n) { A[cn] += cnt[cn] * (t - tau[cn]); tau[cn] = t; }
      __syncthreads();  // every first update has read its population
      if (co >= 0) atomicAdd(&cnt[co], -1);
      if (cn >= 0) atomicAdd(&cnt[cn], 1);
      // the isolated-atom rate sum changes in touched order (lane 0; rare)
      const bool iso_ch = o != -3 && (o == KEY_ISO ? (double)ro : 0.0) != (nw == KEY_ISO ? (double)rn : 0.0);
      if (__ballot(iso_ch) && lane == 0)
        for (int q = 0; q < nt; q++) {
          const int oq = s_o[q], nq = s_nw[q];
          if (oq == -3) continue;
          const double roq = oq == KEY_ISO ? (double)s_ro[q] : 0.0;
          const double rnq = nq == KEY_ISO ? (double)s_rn[q] : 0.0;
          if (roq != rnq) {
            iso_int += r_iso * (t - tau_iso);
            tau_iso = t;
            r_iso += rnq - roq;
          }
        }
    }
    __syncthreads();
    if (bad) { err = bad; k++; break; }
    if (D.hist && (k + 1) % D.every == 0) {  // checkpoint: the class populations after event k
      int32_t *row = D.hist + ((k + 1) / D.every - 1) * KEY_N;
      for (int c = lane; c < KEY_N; c += WAVE) row[c] = cnt[c];
    }
    if ((k + 1) % ADJ_EPOCH == 0) {  // restart the accounting clock
      for (int c = lane; c < KEY_N; c += WAVE) {
        A[c] += cnt[c] * (t - tau[c]);
        tau[c] = 0.0;
      }
      iso_int += r_iso * (t - tau_iso);
      tau_iso = 0.0;
      t = 0.0;
      __syncthreads();
    }
  }
  for (int c = lane; c < KEY_N; c += WAVE) D.A[c] = A[c] + cnt[c] * (t - tau[c]);
  if (lane == 0) {
    iso_int += r_iso * (t - tau_iso);
    D.scal[0] = ttot;
    D.scal[1] = iso_int;
    D.scal[2] = iso_log;
    D.scal[3] = (double)k;
    if (err) D.st->err = err;
  }
}

// Before the replay: each segment's rate-sweep error and unsupported-point
// count (out[2i], out[2i + 1]); after it: each replica's run error (out[i]).
__global__ void __launch_bounds__(256) k_adj_check(const AdjDesc *descs, int n, int *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[2 * i] = descs[i].sweep_err[0];
  out[2 * i + 1] = descs[i].cnt0[KEY_N];
}
__global__ void __launch_bounds__(256) k_adj_errs(const AdjDesc *descs, int n, int *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = descs[i].st->err;
}

// Initial class populations and isolated-rate sum of one lattice.
__global__ void __launch_bounds__(256) k_key_hist(Lattice L, int *cnt, double *iso) {
  const int64_t n = L.box.npts;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
    const int k = L.key[p];
    if (k == KEY_ISO) atomicAdd(iso, (double)L.rate[p]);
    if (k >= 0) atomicAdd(&cnt[k], 1);
    else if (k == KEY_UNSUPPORTED) atomicAdd(&cnt[KEY_N], 1);
  }
}


// ---- frame compaction (the xyz output of inpout/WriteMolMolXYZFile.f) --------
// Ordered record lists in the reference's grid order (z, then y, then x; the
// stored-point order restricted to one z plane is the same), z <= LenZ-1.
// FRAME_MAIN, the main and _d frames:
//   list 0 the points the main frame lists: coverage species and atoms of
//     coordination <= 3 (WriteMolMolXYZFile.f:131-166);
//   list 1 vacancies (empty, coordination 4);
//   list 2 non-epitaxial bulk atoms: coordination 4 off the cubic diamond
//     sites (the _d frame, :77-121);
//   and two header counts over z = 1..LenZ-1 (:93-110): atoms of
//   coordination <= 3, covered points.
// FRAME_DEFECTS, the SiC defect frames:
//   list 0 vacancies with the species sum of their four listed neighbours
//     (the _v frame's classification, :222-260);
//   list 1 the _w frame (:264-398): bulk atoms with z = 1..LenZ-1 and three
//     first neighbours of the other species (Si around C, C around Si), read
//     while the isolated wrong atoms -- bulk atoms with no such neighbour --
//     are flipped to the other species; code = the species as flipped.
//     k_wrong_mark marks those flips (bit 7 of the record's `valid` byte,
//     which the neighbour lists leave unused) and clears the marks after.
// Each block owns a contiguous chunk of points; k_frame_count counts,
// the host scans, k_frame_write writes in order with a block-wide prefix
// scan.
constexpr int FRAME_THREADS = 256;
enum : int { FRAME_MAIN = 0, FRAME_DEFECTS = 1 };
template <int MODE> struct FrameShape;
template <> struct FrameShape<FRAME_MAIN> { static constexpr int NLIST = 3, NCNT = 5; };
template <> struct FrameShape<FRAME_DEFECTS> { static constexpr int NLIST = 2, NCNT = 2; };
constexpr uint8_t WRONG_FLIP = 0x80;

__device__ inline int frame_cat(const Lattice &L, int64_t p, const Site &sp, int *code, bool hdr[2]) {
  int x, y, z;
  pxyz(L.box, p, x, y, z);
  hdr[0] = hdr[1] = false;
  const int kind = s_kind(sp), coor = s_coor(sp);
  if (kind == KIND_WALL || z > L.box.lz - 1) return -1;
  const bool inner = z >= 1;
  if (kind_species(kind)) {
    if (coor <= 3) {
      *code = kind;
      hdr[0] = inner;
      return 0;
    }
    if (!dia_type(x, y, z)) {
      *code = kind;
      return 2;
    }
    return -1;
  }
  if (kind_cov_slot(kind)) {
    *code = -kind_cov_slot(kind);  // coverage slot, negated; the host maps it to ListCov
    hdr[1] = inner;
    return coor == 4 ? 1 : 0;
  }
  return coor == 4 ? 1 : -1;  // an empty point with coordination 4 is a vacancy
}

// Species of a record as the _w analysis sees it: the wrong-site flip swaps
// Si (1) and C (2).
__device__ inline int flipped_species(const Site &s) {
  const int k = s_kind(s);
  return (s.valid & WRONG_FLIP) ? 3 - k : k;
}
// NCoorCorr (WriteMolMolXYZFile.f:284-296) of a bulk atom: its listed
// neighbours carrying the other species' bit (the C bit around Si, the Si
// bit around C), with the flips marked or not.
__device__ inline int wrong_corr(const Lattice &L, const P3 &pt, const Site &sp, bool flips) {
  const int k = flips ? flipped_species(sp) : s_kind(sp);
  const NL l = list_of(L, pt, sp);
  int corr = 0;
  for (int i = 0; i < 4; i++) {
    const Site q = rd(L, entry(L, l, i).i);
    int kq = s_kind(q);
    if (!kind_species(kq)) continue;
    if (flips) kq = flipped_species(q);
    corr += (k == 1) ? ((kq >> 1) & 1) : (k == 2 ? (kq & 1) : 0);
  }
  return corr;
}

// Membership of point p in the lists of a frame, the code per list and the
// header counts.
template <int MODE>
__device__ inline void frame_point(const Lattice &L, int64_t p, const Site &sp, int sic, bool in[],
                                   int code[], unsigned long long cnt[]) {
  if (MODE == FRAME_MAIN) {
    int c = 0;
    bool hdr[2];
    const int cat = frame_cat(L, p, sp, &c, hdr);
    in[0] = cat == 0 || (cat == 1 && c < 0);  // a covered vacancy is also listed in the main frame
    in[1] = cat == 1;
    in[2] = cat == 2;
    code[0] = code[2] = c;
    code[1] = 0;
    for (int k = 0; k < 3; k++) cnt[k] += in[k];
    cnt[3] += hdr[0];
    cnt[4] += hdr[1];
  } else {
    in[0] = in[1] = false;
    code[0] = code[1] = 0;
    const int kind = s_kind(sp), coor = s_coor(sp);
    if (kind == KIND_WALL || coor != 4) return;
    P3 pt;
    pt.i = p;
    pxyz(L.box, p, pt.x, pt.y, pt.z);
    if (pt.z > L.box.lz - 1) return;
    if (kind == KIND_EMPTY || kind >= KIND_COV0) {  // a vacancy: the species sum of its neighbours
      const NL l = list_of(L, pt, sp);
      int sum = 0;
      for (int i = 0; i < 4; i++) {
        const int kq = s_kind(rd(L, entry(L, l, i).i));
        sum += kind_species(kq) ? kq : 0;
      }
      in[0] = true;
      code[0] = sum;
    } else if (sic && kind_species(kind) && pt.z >= 1 && wrong_corr(L, pt, sp, true) == 3) {
      in[1] = true;
      code[1] = flipped_species(sp);
    }
    cnt[0] += in[0];
    cnt[1] += in[1];
  }
}

// The isolated wrong atoms of the SiC analysis (WriteMolMolXYZFile.f:269-299).
// The reference visits the bulk atoms in grid order and flips each one with
// no neighbour of the other species on the spot, so a flip is seen by the
// atoms visited after it.  A flip only raises its neighbours' counts (its
// neighbours share its species until it flips), so an atom flips iff it has
// no neighbour of the other species to begin with (a candidate) and no
// candidate neighbour before it in grid order flips: the candidates are few,
// k_wrong_cand lists them with their neighbours, the host resolves that
// order, and k_wrong_set marks (or unmarks) the flips.
__global__ void __launch_bounds__(256) k_wrong_cand(Lattice L, int64_t cap, unsigned long long *n, int4 *out) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < L.box.npts;
       p += (int64_t)gridDim.x * blockDim.x) {
    const Site sp = L.site[p];
    const int kind = s_kind(sp);
    if (!(kind == 1 || kind == 2) || s_coor(sp) != 4) continue;
    P3 pt;
    pt.i = p;
    pxyz(L.box, p, pt.x, pt.y, pt.z);
    if (pt.z < 1 || pt.z > L.box.lz - 1 || wrong_corr(L, pt, sp, false) != 0) continue;
    const unsigned long long i = atomicAdd(n, 1ull);
    if ((int64_t)i >= cap) continue;  // counted; the host grows the list and asks again
    const NL l = list_of(L, pt, sp);
    out[5 * i] = make_int4(pt.x, pt.y, pt.z, 0);
    for (int k = 0; k < 4; k++) out[5 * i + 1 + k] = make_int4(l.x[k], l.y[k], l.z[k], 0);
  }
}
__global__ void __launch_bounds__(256) k_wrong_set(Lattice L, const int4 *pos, int64_t n, int set) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t p = pidx(L.box, pos[i].x, pos[i].y, pos[i].z);
    const uint8_t v = L.site[p].valid;
    L.site[p].valid = (uint8_t)(set ? (v | WRONG_FLIP) : (v & ~WRONG_FLIP));
  }
}

template <int MODE>
__global__ void __launch_bounds__(FRAME_THREADS) k_frame_count(Lattice L, int64_t chunk, int sic,
                                                               unsigned long long *counts) {
  constexpr int NL_ = FrameShape<MODE>::NLIST, NC = FrameShape<MODE>::NCNT;
  __shared__ unsigned long long sc[NC][FRAME_THREADS / WAVE];
  const int64_t lo = (int64_t)blockIdx.x * chunk, hi = min(lo + chunk, L.box.npts);
  unsigned long long c[NC];
  for (int k = 0; k < NC; k++) c[k] = 0;
  for (int64_t p = lo + threadIdx.x; p < hi; p += FRAME_THREADS) {
    bool in[NL_];
    int code[NL_];
    frame_point<MODE>(L, p, L.site[p], sic, in, code, c);
  }
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
  for (int k = 0; k < NC; k++) {
    unsigned long long v = c[k];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
    if (lane == 0) sc[k][w] = v;
  }
  __syncthreads();
  if (threadIdx.x < NC) {
    unsigned long long v = 0;
    for (int q = 0; q < FRAME_THREADS / WAVE; q++) v += sc[threadIdx.x][q];
    counts[(int64_t)blockIdx.x * NC + threadIdx.x] = v;
  }
}

template <int MODE>
__global__ void __launch_bounds__(FRAME_THREADS) k_frame_write(Lattice L, int64_t chunk, int sic,
                                                               const unsigned long long *offs, int4 *out0,
                                                               int4 *out1, int4 *out2, int64_t cap0, int64_t cap1,
                                                               int64_t cap2) {
  constexpr int NL_ = FrameShape<MODE>::NLIST, NC = FrameShape<MODE>::NCNT;
  __shared__ unsigned int wsum[NL_][FRAME_THREADS / WAVE];
  const int64_t lo = (int64_t)blockIdx.x * chunk, hi = min(lo + chunk, L.box.npts);
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
  unsigned long long base[NL_];
  for (int k = 0; k < NL_; k++) base[k] = offs[(int64_t)blockIdx.x * NL_ + k];
  int4 *const outs[3] = {out0, out1, out2};
  const unsigned long long caps[3] = {(unsigned long long)cap0, (unsigned long long)cap1, (unsigned long long)cap2};
  for (int64_t p0 = lo; p0 < hi; p0 += FRAME_THREADS) {
    const int64_t p = p0 + threadIdx.x;
    bool in[NL_];
    int code[NL_];
    unsigned long long cdummy[NC];
    for (int k = 0; k < NL_; k++) in[k] = false;
    if (p < hi) frame_point<MODE>(L, p, L.site[p], sic, in, code, cdummy);
    int x = 0, y = 0, z = 0;
    bool any = false;
    for (int k = 0; k < NL_; k++) any |= in[k];
    if (any) pxyz(L.box, p, x, y, z);
    unsigned int rank[NL_];
    for (int k = 0; k < NL_; k++) {
      const unsigned long long m = __ballot(in[k]);
      rank[k] = __popcll(m & ((1ull << lane) - 1ull));
      if (lane == 0) wsum[k][w] = __popcll(m);
    }
    __syncthreads();
    for (int k = 0; k < NL_; k++) {
      unsigned int before = 0, total = 0;
      for (int q = 0; q < FRAME_THREADS / WAVE; q++) {
        if (q < w) before += wsum[k][q];
        total += wsum[k][q];
      }
      const unsigned long long at = base[k] + before + rank[k];
      if (in[k] && at < caps[k]) outs[k][at] = make_int4(x, y, z, code[k]);
      base[k] += total;
    }
    __syncthreads();
  }
}

// The block offsets of a frame from the block counts, on the device: one
// block scans the counts of every list in order and writes the offsets and
// the totals (NCNT of them).
template <int MODE>
__global__ void __launch_bounds__(256) k_frame_scan(const unsigned long long *counts, int nblk,
                                                    unsigned long long *offs, int64_t *totals) {
  constexpr int NL_ = FrameShape<MODE>::NLIST, NC = FrameShape<MODE>::NCNT;
  __shared__ unsigned long long part[256];
  for (int k = 0; k < NC; k++) {
    // each thread sums a contiguous run of blocks, then the runs are scanned
    const int per = (nblk + 255) / 256;
    const int lo = min((int)threadIdx.x * per, nblk), hi = min(lo + per, nblk);
    unsigned long long s = 0;
    for (int b = lo; b < hi; b++) s += counts[(int64_t)b * NC + k];
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long acc = 0;
      for (int q = 0; q < 256; q++) {
        const unsigned long long v = part[q];
        part[q] = acc;
        acc += v;
      }
      totals[k] = (int64_t)acc;
    }
    __syncthreads();
    if (k < NL_) {
      unsigned long long acc = part[threadIdx.x];
      for (int b = lo; b < hi; b++) {
        offs[(int64_t)b * NL_ + k] = acc;
        acc += counts[(int64_t)b * NC + k];
      }
    }
    __syncthreads();
  }
}

// Scratch of the allocation-free frame entry point (mskg_frame_device): the
// block counts (5 per block) and offsets (3 per block) of one frame.  Every
// GPU the library runs on holds its own copy until the process ends; host
// code reaches it only through the address hipGetSymbolAddress gives for the
// device (noalloc_scratch), never through the array's name.
constexpr int NOALLOC_FRAME_BLOCKS = 16384;
constexpr int64_t NOALLOC_MAX_CHUNK = 1ll << 30;  // lattice points per block
static __device__ __align__(256) unsigned long long scratch[NOALLOC_FRAME_BLOCKS * 8];

// ---- census: atoms by species and covered points by coverage slot ---------
// Every deposition adds one atom of its species and every evaporation
// removes one (absorption / desorption likewise for coverage), so the
// evaporations and desorptions by species since a state was loaded follow
// from the depositions and absorptions counted by the events and the
// census then and now, with no per-event work.
__global__ void __launch_bounds__(256) k_census(Lattice L, unsigned long long *out) {
  __shared__ unsigned long long s[6];
  if (threadIdx.x < 6) s[threadIdx.x] = 0ull;
  __syncthreads();
  unsigned long long c[6] = {0ull, 0ull, 0ull, 0ull, 0ull, 0ull};
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < L.box.npts;
       p += (int64_t)gridDim.x * blockDim.x) {
    const int k = s_kind(L.site[p]);
    if (kind_species(k)) c[k - 1]++;
    else if (kind_cov_slot(k)) c[2 + kind_cov_slot(k)]++;
  }
  for (int q = 0; q < 6; q++) {
    unsigned long long v = c[q];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
    if ((threadIdx.x & (WAVE - 1)) == 0 && v) atomicAdd(&s[q], v);
  }
  __syncthreads();
  if (threadIdx.x < 6 && s[threadIdx.x]) atomicAdd(&out[threadIdx.x], s[threadIdx.x]);
}

// ---- laser-annealing coupling: temperature map in, phase maps out ----------
// The laser-annealing build exchanges whole-grid int32 maps with its FEM
// driver: the temperature map b and the phase maps c (coordination) and d
// (species), DIMENSION(LenZ, LenX*LenY) (mulskips-source-LaserAnnealing/
// modules/defsystem.f:88-91), or the temperature file's planes
// (mulskips-source-LaserAnnealing/initial/SetT.f:56-66).  Both are read and
// written here on the device, so a cycle moves 4 B per grid point over PCIe
// instead of the lattice.
enum : int { FIELD_LATTCOO = 2 };  // x fastest over the LenZ+1 planes of LattCoo
MSKG_HD int64_t field_index(const Box &b, int layout, int x, int y, int z) {
  if (layout == MSKG_FIELD_SOCKET) return z + (int64_t)b.lz * (x + (int64_t)b.lx * y);
  return x + (int64_t)b.lx * (y + (int64_t)b.ly * z);
}

// Range check of a strict temperature map at the stored points (the 13 bits
// of LattCoo's T field), before anything is written.
__global__ void __launch_bounds__(256) k_check_field(Lattice L, const int32_t *f, int layout, int *err) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < L.box.npts;
       p += (int64_t)gridDim.x * blockDim.x) {
    int x, y, z;
    pxyz(L.box, p, x, y, z);
    const int32_t v = f[field_index(L.box, layout, x, y, z)];
    if (v < 0 || v > 8191) atomicExch(err, 1);
  }
}

// SetT.f:38-48 (the laser-annealing build): every point of the planes
// 0..LenZ-1 takes the low LenT = 13 bits of its map value (MVBITS), except
// where the value is 10, the FEM side's wall marker, which leaves the point
// as it is.  strict: the map
// covers LattCoo's LenZ+1 planes and every value is taken as it is (range
// checked before).  One thread per stored point, records read and written
// whole.
__global__ void __launch_bounds__(256) k_set_temperature(Lattice L, const int32_t *f, int layout, int strict) {
  const int64_t n = L.box.npts;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
    int x, y, z;
    pxyz(L.box, p, x, y, z);
    if (!strict && z >= L.box.lz) continue;
    const int32_t v = f[field_index(L.box, layout, x, y, z)];
    if (!strict && v == 10) continue;
    Site s = L.site[p];
    s.temp = (uint16_t)(v & 8191);
    L.site[p] = s;
  }
}

// main.f:796-818 of the laser-annealing build: for the points of planes
// 0..LenZ-1, c = coordination and d = species of an occupied point, 10 and
// 10 on a wall (index 113), 0 and 0 elsewhere (grid points off the diamond
// super-lattice are never occupied).  Rows of the map (a z column of the
// socket layout, an x row of the plane layout) are strided over the grid;
// writes are in map order.
__global__ void __launch_bounds__(256) k_phases(Lattice L, int layout, const uint32_t *offwall, int wall_planes,
                                               int32_t *c, int32_t *d) {
  const Box &b = L.box;
  const int64_t rows = layout == MSKG_FIELD_SOCKET ? (int64_t)b.lx * b.ly : (int64_t)b.ly * b.lz;
  const int len = layout == MSKG_FIELD_SOCKET ? b.lz : b.lx;
  for (int64_t r = blockIdx.x; r < rows; r += gridDim.x) {
    int x = 0, y = 0, z = 0;
    if (layout == MSKG_FIELD_SOCKET) {
      x = (int)(r % b.lx);
      y = (int)(r / b.lx);
    } else {
      y = (int)(r % b.ly);
      z = (int)(r / b.ly);
    }
    for (int k = threadIdx.x; k < len; k += blockDim.x) {
      if (layout == MSKG_FIELD_SOCKET) z = k;
      else x = k;
      int32_t cv = 0, dv = 0;
      if (valid_parity(x, y, z)) {
        const Site s = L.site[pidx(b, x, y, z)];
        const int kind = s_kind(s);
        if (kind == KIND_WALL) {
          cv = dv = 10;
        } else if (kind_species(kind)) {
          cv = s_coor(s);
          dv = kind;
        }
      } else {
        const int64_t g = x + (int64_t)b.lx * (y + (int64_t)b.ly * z);
        const bool wall = (offwall && ((offwall[g >> 5] >> (g & 31)) & 1)) ||
                          (wall_planes && (z <= 5 || (z >= b.lz - 6 && z <= b.lz - 1)));
        if (wall) cv = dv = 10;
      }
      const int64_t o = r * len + k;
      if (c) c[o] = cv;
      if (d) d[o] = dv;
    }
  }
}

// The kernel instance for a rate model (all four share one signature).
template <class F>
F pick_variant(bool la, bool cov, F tt, F tf, F ft, F ff) {
  return la ? (cov ? tt : tf) : (cov ? ft : ff);
}
#define MSKG_VARIANT(K, la, cov) pick_variant((la) != 0, (cov) != 0, K<true, true>, K<true, false>, K<false, true>, K<false, false>)

// ---- host side ---------------------------------------------------------------
#define HIPCHK(e, x)                                                                  \
  do {                                                                                \
    hipError_t _r = (x);                                                              \
    if (_r != hipSuccess) return fail_((e), MSKG_EDEVICE, "%s: %s", #x, hipGetErrorString(_r)); \
  } while (0)

}  // namespace

struct mskg_engine {
  mskg_config cfg;
  bool cov_model = true;  // the kernels with the coverage-species branches
  Box box;               // the stored planes (all of them unless cfg.stored_len_z)
  int logical_lz = 0;    // the reference's LenZ
  bool windowed = false; // only the bottom planes are stored
  Geometry geo;
  Tables tab;
  int ncrystal, ncov, list_cov[3];
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // device memory
  Site *d_site = nullptr;
  float *d_rate = nullptr;
  uint8_t *d_cls = nullptr;
  uint8_t *d_inlist = nullptr;
  int32_t *d_tl_idx = nullptr;
  int32_t *d_tl_n = nullptr;
  int32_t *d_tl_own = nullptr;
  double *d_tile_r = nullptr;
  int32_t *d_tile_n = nullptr;
  int32_t *d_field[2] = {nullptr, nullptr};  // grid-sized maps of the laser-annealing coupling (on first use)
  // walls on grid points off the diamond super-lattice (the reference's
  // builders fill whole wall regions), which only the phase maps report:
  // one bit per grid point from mskg_load_state (null: none), or the wall
  // planes of the device builders
  uint32_t *d_offwall = nullptr;
  bool offwall_planes = false;
  // InitSt 'ST' reference tables (one allocation: temperatures, then the tables)
  void *d_st_tables = nullptr;
  // census (atoms by species, covered points by slot) of the state as loaded
  int64_t census0[6] = {0, 0, 0, 0, 0, 0};
  int *d_list_err = nullptr;
  int *d_win = nullptr;
  // per-tile completion stamps of the dependency-scheduled sweep, then its
  // stamp / queue / window words (PassParams::done, ::sched)
  unsigned int *d_done = nullptr;
  int schedule = MSKG_SCHEDULE_COLOURS;
  Geometry *d_geo = nullptr;
  Tables *d_tab = nullptr;
  RunState *d_st = nullptr;
  int *d_err = nullptr;
  int device = 0;          // the HIP device the engine lives on
  int tile = 0, ntx = 0, nty = 0, ntz = 0, ntiles = 0, cap = 1024, capg = 512;
  int rsum_chunk = 1;      // tiles per block of the fixed-order tile sum
  bool lists_dirty = true;
  int pass_grid = 4096;  // workgroups per colour pass (they stride over the window)
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  mskg_event *d_log = nullptr;
  int64_t log_cap = 0;
  // serial mode: tile and block rate sums
  double *d_tsum = nullptr, *d_bsum = nullptr;
  // likelihood accounting: class key of every point, and as last accounted
  int32_t *d_key = nullptr, *d_key_seen = nullptr;
  float *d_rate_seen = nullptr;
  bool track = false;       // keep every point's class key current through runs
  bool keys_valid = false;  // d_key matches the rates (a keyed sweep or a tracked run since)
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
  L.inlist = e->d_inlist;
  L.tl_idx = e->d_tl_idx;
  L.tl_n = e->d_tl_n;
  L.tl_own = e->d_tl_own;
  L.tile_r = e->d_tile_r;
  L.list_err = e->d_list_err;
  L.win = e->d_win;
  L.capg = e->capg;
  L.cur_tile = -1;
  L.box = e->box;
  L.geo = e->d_geo;
  L.tab = e->d_tab;
  L.tile = e->tile;
  L.inv_tile = 1.0f / (float)e->tile;
  L.ntx = e->ntx;
  L.nty = e->nty;
  L.key = e->track ? e->d_key : nullptr;
  L.cl = e->tab.la_variant ? (const long long *)((const char *)e->d_st + offsetof(RunState, cl)) : nullptr;
  L.ztop = e->windowed ? e->box.lz - TOP_MARGIN : 0x7fffffff;
  return L;
}

// The captured sweep holds the k_pass variant (InitSt 'ST' tables on or off)
// and what make_lattice gave at capture (CountLiquid, class keys) and the
// event log: a change to any of these needs a new capture at the next run.
static void drop_graph(mskg_engine *e) {
  if (e->graph_exec) { hipGraphExecDestroy(e->graph_exec); e->graph_exec = nullptr; }
  if (e->graph) { hipGraphDestroy(e->graph); e->graph = nullptr; }
}

// Same-colour tiles are one tile apart, so two concurrent events are at
// least tile+1 grid units apart along some axis.  An event writes records
// and rates at most 3 units from its site per axis (6 with the
// second-neighbour term, which refreshes the second neighbours) and reads at
// most 6 (12) units away, so a tile edge of 16 (24) keeps what one event
// writes out of what any other reads, with room for the rare neighbour
// entries that are not bond vectors.
static int choose_tile(int lx, int ly, int want, int min_tile) {
  auto ok = [&](int t) {
    return t >= min_tile && t <= 254 && t % 2 == 0 && lx % t == 0 && ly % t == 0 && (lx / t) % 2 == 0 &&
           (ly / t) % 2 == 0;
  };
  if (want) return ok(want) ? want : -1;
  const int pref[] = {24, 20, 18, 28, 30, 32, 36, 40, 42, 48, 16, 22, 26, 34, 38, 44, 46};
  for (int t : pref)
    if (ok(t)) return t;
  for (int t = min_tile; t <= std::min(lx, ly) / 2; t += 2)
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
    case ERR_TILE_LIST: return "a tile's candidate list overflowed tile_capacity";
    case ERR_SUMS: return "serial selection found no event inside a positive rate sum";
    case ERR_LIQUID: return "a species' CountLiquid is negative: its solidification rate would be negative";
    case ERR_ST_RANGE: return "a site's temperature lies outside the InitSt 'ST' reference temperatures";
    case ERR_SCHED: return "a tile waited more than 2 s for its neighbour tiles in a dependency-scheduled sweep";
    case ERR_TOP:
      return "an event site within 24 planes of the top of the stored planes (stored_len_z): store more planes";
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

static int la_variant_off(mskg_engine *e);
static int census(mskg_engine *e, int64_t out[6]);
// The run state's fields before the sweep sums (slot, rpart: left as they
// are in st).
static int read_state(mskg_engine *e, RunState &st) {
  HIPCHK(e, hipMemcpyAsync(&st, e->d_st, RS_HEAD, hipMemcpyDeviceToHost, e->stream));
  HIPCHK(e, hipStreamSynchronize(e->stream));
  return MSKG_OK;
}
// ... and back (the device's sweep sums kept).
static int write_state_head(mskg_engine *e, const RunState &st) {
  HIPCHK(e, hipMemcpyAsync(e->d_st, &st, RS_HEAD, hipMemcpyHostToDevice, e->stream));
  return MSKG_OK;
}

// The ordered record lists of one frame mode (k_frame_count / k_frame_write):
// hosts[k] receives list k (caps[k] records), counts the NCNT totals.  A call
// with too small capacities only sizes the lists (MSKG_ECAPACITY when it
// brought buffers).
template <int MODE>
static int frame_lists(mskg_engine *e, int sic, int32_t *const hosts[], const int64_t caps[], int64_t counts[]) {
  constexpr int NL_ = FrameShape<MODE>::NLIST, NC = FrameShape<MODE>::NCNT;
  const int64_t n = e->box.npts;
  const int nblk = (int)std::min<int64_t>(std::max<int64_t>(1, (n + 65535) / 65536), 65536);
  const int64_t chunk = (n + nblk - 1) / nblk;
  Lattice L = make_lattice(e);
  unsigned long long *d_counts = nullptr, *d_offs = nullptr;
  int4 *d_out[3] = {nullptr, nullptr, nullptr};
  auto cleanup = [&]() {
    hipFree(d_counts);
    hipFree(d_offs);
    for (int k = 0; k < 3; k++) hipFree(d_out[k]);
  };
  if (hipMalloc(&d_counts, (size_t)nblk * NC * sizeof(unsigned long long)) != hipSuccess ||
      hipMalloc(&d_offs, (size_t)nblk * NL_ * sizeof(unsigned long long)) != hipSuccess) {
    cleanup();
    return fail_(e, MSKG_ENOMEM, "frame: no device memory for the block counts");
  }
  hipLaunchKernelGGL(k_frame_count<MODE>, dim3(nblk), dim3(FRAME_THREADS), 0, e->stream, L, chunk, sic, d_counts);
  std::vector<unsigned long long> cnt((size_t)nblk * NC), offs((size_t)nblk * NL_);
  hipError_t ce = hipGetLastError();
  if (ce == hipSuccess)
    ce = hipMemcpyAsync(cnt.data(), d_counts, cnt.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                        e->stream);
  if (ce == hipSuccess) ce = hipStreamSynchronize(e->stream);
  if (ce != hipSuccess) {
    cleanup();
    return fail_(e, MSKG_EDEVICE, "frame count: %s", hipGetErrorString(ce));
  }
  unsigned long long tot[NC];
  for (int k = 0; k < NC; k++) tot[k] = 0;
  for (int b = 0; b < nblk; b++)
    for (int k = 0; k < NC; k++) {
      if (k < NL_) offs[(size_t)b * NL_ + k] = tot[k];
      tot[k] += cnt[(size_t)b * NC + k];
    }
  for (int k = 0; k < NC; k++) counts[k] = (int64_t)tot[k];
  bool fits = true, brought = false;
  for (int k = 0; k < NL_; k++) {
    fits &= (int64_t)tot[k] <= caps[k];
    brought |= hosts[k] && caps[k];
  }
  if (!fits) {
    cleanup();
    if (brought) return fail_(e, MSKG_ECAPACITY, "frame: the records do not fit the buffers given");
    return MSKG_OK;  // sizing call
  }
  for (int k = 0; k < NL_; k++)
    if (hipMalloc(&d_out[k], std::max<unsigned long long>(1, tot[k]) * sizeof(int4)) != hipSuccess) {
      cleanup();
      return fail_(e, MSKG_ENOMEM, "frame: no device memory for the records");
    }
  ce = hipMemcpyAsync(d_offs, offs.data(), offs.size() * sizeof(unsigned long long), hipMemcpyHostToDevice,
                      e->stream);
  if (ce == hipSuccess) {
    int64_t capv[3] = {0, 0, 0};
    for (int k = 0; k < NL_; k++) capv[k] = (int64_t)tot[k];
    hipLaunchKernelGGL(k_frame_write<MODE>, dim3(nblk), dim3(FRAME_THREADS), 0, e->stream, L, chunk, sic, d_offs,
                       d_out[0], d_out[1], d_out[2], capv[0], capv[1], capv[2]);
    ce = hipGetLastError();
  }
  for (int k = 0; k < NL_ && ce == hipSuccess; k++)
    if (tot[k] && hosts[k])
      ce = hipMemcpyAsync(hosts[k], d_out[k], tot[k] * sizeof(int4), hipMemcpyDeviceToHost, e->stream);
  if (ce == hipSuccess) ce = hipStreamSynchronize(e->stream);
  cleanup();
  if (ce != hipSuccess) return fail_(e, MSKG_EDEVICE, "frame write: %s", hipGetErrorString(ce));
  return MSKG_OK;
}

// The flips of the SiC wrong-site analysis in the reference's grid order
// (see k_wrong_cand): the positions of the atoms that flip.
static int wrong_flips(mskg_engine *e, std::vector<int4> &flips) {
  flips.clear();
  Lattice L = make_lattice(e);
  const int blocks = (int)std::min<int64_t>((e->box.npts + 255) / 256, 256 * 64);
  int64_t cap = 1 << 16;
  for (;;) {
    unsigned long long *d_n = nullptr;
    int4 *d_c = nullptr;
    if (hipMalloc(&d_n, sizeof(unsigned long long)) != hipSuccess ||
        hipMalloc(&d_c, (size_t)cap * 5 * sizeof(int4)) != hipSuccess) {
      hipFree(d_n);
      return fail_(e, MSKG_ENOMEM, "frame: no device memory for the wrong-site candidates");
    }
    unsigned long long n = 0;
    std::vector<int4> c;
    hipError_t ce = hipMemsetAsync(d_n, 0, sizeof n, e->stream);
    if (ce == hipSuccess) {
      hipLaunchKernelGGL(k_wrong_cand, dim3(blocks), dim3(256), 0, e->stream, L, cap, d_n, d_c);
      ce = hipGetLastError();
    }
    if (ce == hipSuccess) ce = hipMemcpyAsync(&n, d_n, sizeof n, hipMemcpyDeviceToHost, e->stream);
    if (ce == hipSuccess) ce = hipStreamSynchronize(e->stream);
    if (ce == hipSuccess && (int64_t)n <= cap) {
      c.resize((size_t)n * 5);
      if (n) ce = hipMemcpy(c.data(), d_c, c.size() * sizeof(int4), hipMemcpyDeviceToHost);
    }
    hipFree(d_n);
    hipFree(d_c);
    if (ce != hipSuccess) return fail_(e, MSKG_EDEVICE, "frame wrong-site candidates: %s", hipGetErrorString(ce));
    if ((int64_t)n > cap) {
      cap = (int64_t)n;
      continue;
    }
    // grid order (z, then y, then x); a candidate flips unless a candidate
    // neighbour before it did
    const Box &b = e->box;
    auto key = [&](const int4 &q) { return ((int64_t)q.z * b.ly + q.y) * b.lx + q.x; };
    std::vector<size_t> order(n);
    for (size_t i = 0; i < n; i++) order[i] = i;
    std::sort(order.begin(), order.end(), [&](size_t a, size_t bb) { return key(c[5 * a]) < key(c[5 * bb]); });
    std::vector<int64_t> flipped_keys;
    for (size_t oi = 0; oi < n; oi++) {
      const size_t i = order[oi];
      bool flip = true;
      for (int k = 0; k < 4 && flip; k++) {
        const int64_t q = key(c[5 * i + 1 + k]);
        flip = !std::binary_search(flipped_keys.begin(), flipped_keys.end(), q);
      }
      if (flip) {
        flipped_keys.push_back(key(c[5 * i]));  // keys arrive in increasing order
        flips.push_back(c[5 * i]);
      }
    }
    return MSKG_OK;
  }
}

static int set_wrong_marks(mskg_engine *e, const std::vector<int4> &flips, int set) {
  if (flips.empty()) return MSKG_OK;
  int4 *d = nullptr;
  HIPCHK(e, hipMalloc(&d, flips.size() * sizeof(int4)));
  hipError_t ce = hipMemcpyAsync(d, flips.data(), flips.size() * sizeof(int4), hipMemcpyHostToDevice, e->stream);
  if (ce == hipSuccess) {
    const int blocks = (int)std::min<size_t>((flips.size() + 255) / 256, 4096);
    hipLaunchKernelGGL(k_wrong_set, dim3(blocks), dim3(256), 0, e->stream, make_lattice(e), d,
                       (int64_t)flips.size(), set);
    ce = hipGetLastError();
  }
  if (ce == hipSuccess) ce = hipStreamSynchronize(e->stream);
  hipFree(d);
  if (ce != hipSuccess) return fail_(e, MSKG_EDEVICE, "frame wrong-site marks: %s", hipGetErrorString(ce));
  return MSKG_OK;
}

// The device address of `scratch` on each device, taken once per device
// with that device current (mskg_create).
constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static unsigned long long *g_scratch[MAX_DEVICES] = {};

static int noalloc_scratch_init(mskg_engine *e) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (e->device < 0 || e->device >= MAX_DEVICES)
    return fail_(e, MSKG_EDEVICE, "device %d beyond the %d the library tracks", e->device, MAX_DEVICES);
  if (g_scratch[e->device]) return MSKG_OK;
  void *ptr = nullptr;
  const hipError_t rc = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (rc != hipSuccess || !ptr)
    return fail_(e, MSKG_EDEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(rc));
  g_scratch[e->device] = (unsigned long long *)ptr;
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
  if (cfg->stored_len_z < 0 || cfg->stored_len_z > cfg->len_z + 1 ||
      (cfg->stored_len_z > 0 && cfg->stored_len_z < 2 * TOP_MARGIN))
    return fail_(e, MSKG_EINVAL, "stored_len_z %d: 0 (all planes) or %d..%d", cfg->stored_len_z, 2 * TOP_MARGIN,
                 cfg->len_z + 1);
  if (cfg->n_crystal < 1 || cfg->n_crystal > 3 || cfg->n_cov < 0 || cfg->n_cov > 3)
    return fail_(e, MSKG_EINVAL, "n_crystal must be 1..3 and n_cov 0..3");
  if (!cfg->ptrans_e || !cfg->ptrans_d) return fail_(e, MSKG_EINVAL, "ptrans_e and ptrans_d are required");
  if (cfg->n_cov > 0 && (!cfg->ptrans_abs || !cfg->ptrans_des))
    return fail_(e, MSKG_EINVAL, "ptrans_abs and ptrans_des are required when n_cov > 0");
  e->logical_lz = cfg->len_z;
  e->windowed = cfg->stored_len_z > 0 && cfg->stored_len_z < cfg->len_z + 1;
  e->box = make_box(cfg->len_x, cfg->len_y, e->windowed ? cfg->stored_len_z - 1 : cfg->len_z);
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
  e->cov_model = cfg->n_cov > 0;  // coverage species, or absorption / desorption rates in the tables
  for (int i = 0; i < 12; i++) e->cov_model |= e->tab.pabs[i] != 0.0 || e->tab.pdes[i] != 0.0;
  e->tab.p0 = cfg->p0;
  e->tab.tm = cfg->tm;
  e->tab.sigma = cfg->sigma;
  e->tab.tflex = cfg->tflex;
  e->tab.frozen_below_z = cfg->frozen_below_z;
  e->tab.use_f2 = cfg->second_factor != nullptr;
  if (cfg->second_factor)
    for (int i = 0; i < 13; i++) {
      e->tab.f2[i] = cfg->second_factor[i];
      if (!(e->tab.f2[i] > 0.0) || !std::isfinite(e->tab.f2[i]))
        return fail_(e, MSKG_EINVAL, "second_factor[%d] must be positive and finite", i);
    }
  build_geometry(e->geo);
  const int min_tile = e->tab.use_f2 ? 24 : 16;
  e->tile = choose_tile(cfg->len_x, cfg->len_y, cfg->tile, min_tile);
  if (e->tile < 0)
    return fail_(e, MSKG_EINVAL,
                 "no valid tile edge for box %d x %d (requested %d): need an even tile >= %d with an even "
                 "number of tiles along x and y",
                 cfg->len_x, cfg->len_y, cfg->tile, min_tile);
  // the pass kernel holds a tile's list in LDS at 9 B per entry: at most
  // 6144 entries (55 KB) keep one wave's LDS under 64 KB
  if (cfg->tile_capacity < 0 || cfg->tile_capacity > 6144)
    return fail_(e, MSKG_EINVAL, "tile_capacity %d: 0 (1024) or 1..6144", cfg->tile_capacity);
  e->cap = cfg->tile_capacity > 0 ? cfg->tile_capacity : 1024;
  e->capg = e->cap;
  e->ntx = cfg->len_x / e->tile;
  e->nty = cfg->len_y / e->tile;
  e->ntz = (e->box.lz + 1 + e->tile - 1) / e->tile;
  e->ntiles = e->ntx * e->nty * e->ntz;
  // the tile sums are taken in chunks of the full box's tiles, so that a
  // box storing only its bottom planes sums the same values in the same
  // order (the tiles it does not store hold no rates)
  e->rsum_chunk = (e->ntx * e->nty * ((cfg->len_z + 1 + e->tile - 1) / e->tile) + NPART - 1) / NPART;

  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail_(e, MSKG_EDEVICE, "no HIP device available");
  HIPCHK(e, hipGetDevice(&e->device));
  if (int rc = noalloc_scratch_init(e)) return rc;
  HIPCHK(e, hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
  HIPCHK(e, hipEventCreate(&e->ev0));
  HIPCHK(e, hipEventCreate(&e->ev1));
  const int64_t n = e->box.npts;
  HIPCHK(e, hipMalloc(&e->d_site, n * sizeof(Site)));
  HIPCHK(e, hipMalloc(&e->d_rate, n * sizeof(float)));
  HIPCHK(e, hipMalloc(&e->d_cls, n * sizeof(uint8_t)));
  HIPCHK(e, hipMalloc(&e->d_inlist, n));
  HIPCHK(e, hipMalloc(&e->d_tl_idx, (size_t)e->ntiles * e->capg * sizeof(int32_t)));
  HIPCHK(e, hipMalloc(&e->d_tl_n, e->ntiles * sizeof(int32_t)));
  HIPCHK(e, hipMalloc(&e->d_tl_own, e->ntiles * sizeof(int32_t)));
  HIPCHK(e, hipMalloc(&e->d_tile_r, e->ntiles * sizeof(double)));
  HIPCHK(e, hipMalloc(&e->d_list_err, sizeof(int)));
  HIPCHK(e, hipMalloc(&e->d_win, 32 * sizeof(int)));
  HIPCHK(e, hipMalloc(&e->d_done, ((size_t)e->ntiles + 8) * sizeof(unsigned int)));
  HIPCHK(e, hipMalloc(&e->d_geo, sizeof(Geometry)));
  HIPCHK(e, hipMalloc(&e->d_tab, sizeof(Tables)));
  HIPCHK(e, hipMalloc(&e->d_st, sizeof(RunState)));
  HIPCHK(e, hipMalloc(&e->d_err, sizeof(int)));
  HIPCHK(e, hipMemcpyAsync(e->d_geo, &e->geo, sizeof(Geometry), hipMemcpyHostToDevice, e->stream));
  HIPCHK(e, hipMemcpyAsync(e->d_tab, &e->tab, sizeof(Tables), hipMemcpyHostToDevice, e->stream));
  HIPCHK(e, hipMemsetAsync(e->d_site, 0, n * sizeof(Site), e->stream));
  HIPCHK(e, hipMemsetAsync(e->d_rate, 0, n * sizeof(float), e->stream));
  HIPCHK(e, hipMemsetAsync(e->d_cls, 0, n, e->stream));
  HIPCHK(e, hipMemsetAsync(e->d_inlist, 0, n, e->stream));
  HIPCHK(e, hipMemsetAsync(e->d_tl_n, 0, e->ntiles * sizeof(int32_t), e->stream));
  HIPCHK(e, hipMemsetAsync(e->d_tl_own, 0, e->ntiles * sizeof(int32_t), e->stream));
  HIPCHK(e, hipMemsetAsync(e->d_tile_r, 0, e->ntiles * sizeof(double), e->stream));
  HIPCHK(e, hipMemsetAsync(e->d_list_err, 0, sizeof(int), e->stream));
  {  // empty windows: lists built before any run widen them
    int w0[32];
    for (int c = 0; c < 8; c++) {
      w0[4 * c] = w0[4 * c + 2] = 0x7fffffff;
      w0[4 * c + 1] = w0[4 * c + 3] = -1;
    }
    HIPCHK(e, hipMemcpyAsync(e->d_win, w0, sizeof w0, hipMemcpyHostToDevice, e->stream));
  }
  // completion stamps: 0 is no sweep's (the first k_sweep_end makes it 1)
  HIPCHK(e, hipMemsetAsync(e->d_done, 0, ((size_t)e->ntiles + 8) * sizeof(unsigned int), e->stream));
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
  if (e->graph_exec) hipGraphExecDestroy(e->graph_exec);
  if (e->graph) hipGraphDestroy(e->graph);
  hipFree(e->d_inlist); hipFree(e->d_tl_idx); hipFree(e->d_tl_n); hipFree(e->d_tl_own); hipFree(e->d_tile_r); hipFree(e->d_list_err);
  hipFree(e->d_field[0]); hipFree(e->d_field[1]); hipFree(e->d_offwall); hipFree(e->d_st_tables); hipFree(e->d_win);
  hipFree(e->d_done);
  hipFree(e->d_geo); hipFree(e->d_tab); hipFree(e->d_st); hipFree(e->d_err); hipFree(e->d_log);
  hipFree(e->d_tsum); hipFree(e->d_bsum); hipFree(e->d_key); hipFree(e->d_key_seen); hipFree(e->d_rate_seen);
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
  // planes above the stored ones: vacuum or walls only, which no event reaches
  for (int z = b.lz + 1; z <= e->logical_lz; z++)
    for (int y = 0; y < b.ly; y++)
      for (int x = 0; x < b.lx; x++) {
        const int32_t v = latt[(int64_t)x + (int64_t)b.lx * ((int64_t)y + (int64_t)b.ly * z)];
        if (v != 0 && (v & 255) != 113)
          return fail_(e, MSKG_ESTATE, "LattCoo(%d,%d,%d)=0x%08x lies above the %d stored planes", x, y, z, v,
                       b.lz + 1);
      }
  // walls the caller's LattCoo holds off the lattice (index 113)
  std::vector<uint32_t> offwall;
  for (int z = 0; z <= b.lz; z++)
    for (int y = 0; y < b.ly; y++)
      for (int x = 0; x < b.lx; x++) {
        if (valid_parity(x, y, z)) continue;
        const int64_t g = (int64_t)x + (int64_t)b.lx * ((int64_t)y + (int64_t)b.ly * z);
        if ((latt[g] & 255) != 113) continue;
        if (offwall.empty()) offwall.assign((size_t)((b.lx * (int64_t)b.ly * (b.lz + 1) + 31) / 32), 0u);
        offwall[(size_t)(g >> 5)] |= 1u << (g & 31);
      }
  hipFree(e->d_offwall);
  e->d_offwall = nullptr;
  e->offwall_planes = false;
  if (!offwall.empty()) {
    HIPCHK(e, hipMalloc(&e->d_offwall, offwall.size() * sizeof(uint32_t)));
    HIPCHK(e, hipMemcpy(e->d_offwall, offwall.data(), offwall.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
  }
  double rsum = 0.0;
  auto put = [&](int which, int32_t n, const int32_t *xyz, const int32_t *nn) -> int {
    for (int32_t i = 0; i < n; i++) {
      const int32_t *s = xyz + 3 * i;
      int64_t p = pidx_checked(b, s[0], s[1], s[2]);
      if (p < 0)
        return fail_(e, MSKG_ESTATE, "list %d record %d: site (%d,%d,%d) is not a lattice point of the box", which, i,
                     s[0], s[1], s[2]);
      if (which == 0 && e->windowed && s[2] > b.lz - TOP_MARGIN)
        return fail_(e, MSKG_ESTATE, "MC record %d at (%d,%d,%d) lies within %d planes of the stored top (%d)", i,
                     s[0], s[1], s[2], TOP_MARGIN, b.lz);
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
  if (int rc_la = la_variant_off(e)) return rc_la;  // a new state: the counts are the caller's to give again
  if (int rc_c = census(e, e->census0)) return rc_c;
  e->loaded = true;
  e->lists_dirty = true;
  e->keys_valid = false;
  return MSKG_OK;
}

// A grid-sized device map of the coupling (LenX * LenY * (LenZ+1) int32).
static int32_t *field_buffer(mskg_engine *e, int k) {
  if (!e->d_field[k]) {
    const Box &b = e->box;
    if (hipMalloc(&e->d_field[k], (size_t)b.lx * b.ly * (b.lz + 1) * sizeof(int32_t)) != hipSuccess) {
      e->d_field[k] = nullptr;
      fail_(e, MSKG_ENOMEM, "no device memory for a %d x %d x %d map", b.lx, b.ly, b.lz + 1);
    }
  }
  return e->d_field[k];
}

static int set_temperature_map(mskg_engine *e, const int32_t *map, int layout, int on_device) {
  const Box &b = e->box;
  const bool strict = layout == FIELD_LATTCOO;
  const int64_t n = (int64_t)b.lx * b.ly * (strict ? b.lz + 1 : b.lz);
  const int32_t *f = map;
  if (!on_device) {
    int32_t *buf = field_buffer(e, 0);
    if (!buf) return MSKG_ENOMEM;
    HIPCHK(e, hipMemcpyAsync(buf, map, n * sizeof(int32_t), hipMemcpyHostToDevice, e->stream));
    f = buf;
  }
  const int blocks = (int)std::min<int64_t>((std::max<int64_t>(n, b.npts) + 255) / 256, 256 * 64);
  if (strict) {
    int bad = 0;
    HIPCHK(e, hipMemsetAsync(e->d_err, 0, sizeof(int), e->stream));
    hipLaunchKernelGGL(k_check_field, dim3(blocks), dim3(256), 0, e->stream, make_lattice(e), f, layout, e->d_err);
    HIPCHK(e, hipGetLastError());
    HIPCHK(e, hipMemcpyAsync(&bad, e->d_err, sizeof(int), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    if (bad) return fail_(e, MSKG_EINVAL, "a temperature is out of the 13-bit range 0..8191 K");
  }
  hipLaunchKernelGGL(k_set_temperature, dim3(blocks), dim3(256), 0, e->stream, make_lattice(e), f, layout,
                     strict ? 1 : 0);
  HIPCHK(e, hipGetLastError());
  HIPCHK(e, hipStreamSynchronize(e->stream));
  return MSKG_OK;
}

// The laser-annealing maps and forms cover the whole grid: they need every plane stored.
static int whole_grid(mskg_engine *e, const char *what) {
  if (e->windowed)
    return fail_(e, MSKG_EINVAL, "%s needs every grid plane stored (stored_len_z %d of %d)", what, e->box.lz + 1,
                 e->logical_lz + 1);
  return MSKG_OK;
}

int mskg_set_temperature(mskg_engine *e, const int32_t *temp_k) {
  if (!e || !temp_k) return MSKG_EINVAL;
  if (!e->loaded) return fail_(e, MSKG_EINVAL, "no state loaded");
  if (int rc = whole_grid(e, "the temperature map")) return rc;
  return set_temperature_map(e, temp_k, FIELD_LATTCOO, 0);
}

int mskg_set_temperature_field(mskg_engine *e, const int32_t *field, int32_t layout, int32_t on_device) {
  if (!e || !field) return MSKG_EINVAL;
  if (!e->loaded) return fail_(e, MSKG_EINVAL, "no state loaded");
  if (int rc = whole_grid(e, "the temperature map")) return rc;
  if (layout != MSKG_FIELD_SOCKET && layout != MSKG_FIELD_PLANES)
    return fail_(e, MSKG_EINVAL, "field layout %d: 0 (socket) or 1 (planes)", layout);
  return set_temperature_map(e, field, layout, on_device);
}

int mskg_phases(mskg_engine *e, int32_t *coor, int32_t *species, int32_t layout, int32_t on_device) {
  if (!e) return MSKG_EINVAL;
  if (!e->loaded) return fail_(e, MSKG_EINVAL, "no state loaded");
  if (int rc = whole_grid(e, "the phase maps")) return rc;
  if (layout != MSKG_FIELD_SOCKET && layout != MSKG_FIELD_PLANES)
    return fail_(e, MSKG_EINVAL, "phase map layout %d: 0 (socket) or 1 (planes)", layout);
  if (!coor && !species) return MSKG_OK;
  const Box &b = e->box;
  const int64_t n = (int64_t)b.lx * b.ly * b.lz;
  int32_t *out[2] = {coor, species};
  if (!on_device)
    for (int k = 0; k < 2; k++)
      if (out[k] && !(out[k] = field_buffer(e, k))) return MSKG_ENOMEM;
  const int64_t rows = layout == MSKG_FIELD_SOCKET ? (int64_t)b.lx * b.ly : (int64_t)b.ly * b.lz;
  const int blocks = (int)std::min<int64_t>(rows, 256 * 64);
  hipLaunchKernelGGL(k_phases, dim3(blocks), dim3(256), 0, e->stream, make_lattice(e), (int)layout, e->d_offwall,
                     e->offwall_planes ? 1 : 0, out[0], out[1]);
  HIPCHK(e, hipGetLastError());
  if (!on_device) {
    if (coor) HIPCHK(e, hipMemcpyAsync(coor, out[0], n * sizeof(int32_t), hipMemcpyDeviceToHost, e->stream));
    if (species) HIPCHK(e, hipMemcpyAsync(species, out[1], n * sizeof(int32_t), hipMemcpyDeviceToHost, e->stream));
  }
  HIPCHK(e, hipStreamSynchronize(e->stream));
  return MSKG_OK;
}

// Back to the main build's laser-annealing form (a new state was loaded).
static int la_variant_off(mskg_engine *e) {
  if (!e->tab.la_variant) return MSKG_OK;
  e->tab.la_variant = 0;
  drop_graph(e);
  HIPCHK(e, hipMemcpyAsync(e->d_tab, &e->tab, sizeof(Tables), hipMemcpyHostToDevice, e->stream));
  return MSKG_OK;
}

int mskg_set_la_variant(mskg_engine *e, const double p0[3], const double tm[3], const double sigma[3],
                        const double tflex[3], const int64_t init_count[3], const int64_t count[3]) {
  if (!e || !p0 || !tm || !sigma || !tflex || !init_count || !count) return MSKG_EINVAL;
  if (!e->loaded) return fail_(e, MSKG_EINVAL, "no state loaded");
  if (!e->tab.la_mode) return fail_(e, MSKG_EINVAL, "the laser-annealing build's form needs la_mode");
  if (int rc = whole_grid(e, "the laser-annealing build's form")) return rc;
  for (int k = 0; k < 3; k++)
    if (k < e->ncrystal && !(sigma[k] != 0.0))
      return fail_(e, MSKG_EINVAL, "sigma of species %d is zero", k + 1);
  e->tab.la_variant = 1;
  drop_graph(e);
  e->tab.la_swap = e->ncrystal == 2 && e->cfg.list_crystal[0] == 14 && e->cfg.list_crystal[1] == 32;
  for (int k = 0; k < 3; k++) {
    e->tab.la_p0[k] = p0[k];
    e->tab.la_tm[k] = tm[k];
    e->tab.la_sigma[k] = k < e->ncrystal ? sigma[k] : 1.0;
    e->tab.la_tflex[k] = tflex[k];
  }
  HIPCHK(e, hipMemcpyAsync(e->d_tab, &e->tab, sizeof(Tables), hipMemcpyHostToDevice, e->stream));
  RunState st;
  int rc = read_state(e, st);
  if (rc) return rc;
  st.la_variant = 1;
  for (int k = 0; k < 3; k++) {
    st.cinit[k] = init_count[k];
    st.cc[k] = count[k];
    st.dcc[k] = 0ull;
    st.cl[k] = st.cinit[k] - st.cc[k];
  }
  if (int rc2 = write_state_head(e, st)) return rc2;
  HIPCHK(e, hipStreamSynchronize(e->stream));
  e->lists_dirty = true;
  return MSKG_OK;
}

int mskg_set_st_tables(mskg_engine *e, int32_t ntemp, const int32_t *temps, const double *pe, const double *pd,
                       const double *pabs, const double *pdes) {
  if (!e || ntemp < 0 || (ntemp > 0 && (!temps || !pe || !pd || !pabs || !pdes))) return MSKG_EINVAL;
  if (!e->d_tab) return fail_(e, MSKG_EINVAL, "engine not created");
  if (ntemp > 0 && e->tab.la_mode) return fail_(e, MSKG_EINVAL, "InitSt 'ST' tables and la_mode exclude each other");
  for (int k = 0; k < ntemp; k++) {
    if (temps[k] <= 0 || temps[k] > 8191) return fail_(e, MSKG_EINVAL, "reference temperature %d K out of range", temps[k]);
    if (k && temps[k] <= temps[k - 1])
      return fail_(e, MSKG_EINVAL, "reference temperatures must increase (InterpProbFromField.f:39 assumes it)");
  }
  HIPCHK(e, hipStreamSynchronize(e->stream));
  drop_graph(e);
  hipFree(e->d_st_tables);
  e->d_st_tables = nullptr;
  e->tab.st_n = 0;
  e->tab.st_temp = nullptr;
  e->tab.st_pe = e->tab.st_pd = e->tab.st_pabs = e->tab.st_pdes = nullptr;
  if (ntemp > 0) {
    const size_t n = (size_t)ntemp, off_pe = 256 * ((n * sizeof(int32_t) + 255) / 256);
    const size_t npe = n * 3 * 4096, npd = n * 9, nab = n * 12;
    const size_t bytes = off_pe + (npe + npd + 2 * nab) * sizeof(double);
    HIPCHK(e, hipMalloc(&e->d_st_tables, bytes));
    char *base = (char *)e->d_st_tables;
    double *dpe = (double *)(base + off_pe), *dpd = dpe + npe, *dab = dpd + npd, *dde = dab + nab;
    HIPCHK(e, hipMemcpyAsync(base, temps, n * sizeof(int32_t), hipMemcpyHostToDevice, e->stream));
    HIPCHK(e, hipMemcpyAsync(dpe, pe, npe * sizeof(double), hipMemcpyHostToDevice, e->stream));
    HIPCHK(e, hipMemcpyAsync(dpd, pd, npd * sizeof(double), hipMemcpyHostToDevice, e->stream));
    HIPCHK(e, hipMemcpyAsync(dab, pabs, nab * sizeof(double), hipMemcpyHostToDevice, e->stream));
    HIPCHK(e, hipMemcpyAsync(dde, pdes, nab * sizeof(double), hipMemcpyHostToDevice, e->stream));
    e->tab.st_n = ntemp;
    e->tab.st_temp = (const int *)base;
    e->tab.st_pe = dpe;
    e->tab.st_pd = dpd;
    e->tab.st_pabs = dab;
    e->tab.st_pdes = dde;
  }
  HIPCHK(e, hipMemcpyAsync(e->d_tab, &e->tab, sizeof(Tables), hipMemcpyHostToDevice, e->stream));
  HIPCHK(e, hipStreamSynchronize(e->stream));
  e->lists_dirty = true;
  return MSKG_OK;
}

int mskg_la_counts(mskg_engine *e, int64_t count[3], int64_t liquid[3]) {
  if (!e || !e->loaded) return fail_(e, MSKG_EINVAL, "no state loaded");
  RunState st;
  int rc = read_state(e, st);
  if (rc) return rc;
  for (int k = 0; k < 3; k++) {
    if (count) count[k] = st.cc[k] + (long long)st.dcc[k];
    if (liquid) liquid[k] = st.cinit[k] - (st.cc[k] + (long long)st.dcc[k]);
  }
  return MSKG_OK;
}

// The rate sweep over the stored points [first, end).
// A sweep of the whole lattice recomputes the classes from the records, so
// the tile lists are rebuilt at the next run; a sweep of the layers that
// hold the MC particles of engine-kept lists changes rates only (a point's
// class is zero exactly when its record makes it no MC particle, which a
// sweep does not change), so the lists stay.
static int refresh_range(mskg_engine *e, int64_t first, int64_t end, bool lists_stay) {
  RunState st;
  int rc = read_state(e, st);
  if (rc) return rc;
  HIPCHK(e, hipMemsetAsync(e->d_err, 0, sizeof(int), e->stream));
  Lattice L = make_lattice(e);
  L.inlist = nullptr;
  if (!lists_stay) e->lists_dirty = true;
  HIPCHK(e, hipEventRecord(e->ev0, e->stream));
  const bool la = e->tab.la_mode || e->tab.st_n;
  if (first == 0 && end == e->box.npts) {
    const int blocks = (int)std::min<int64_t>((e->box.npts + 255) / 256, 256 * 64);
    hipLaunchKernelGGL(MSKG_VARIANT(k_rate_sweep, la, e->cov_model), dim3(blocks), dim3(256), 0, e->stream, L, st.rtot,
                       e->d_err, 0);
    HIPCHK(e, hipGetLastError());
  } else if (end > first) {
    // one group of four points per thread up to 65536 blocks: the range of
    // the active layers (~57 M points at 512^3) runs in a single round of
    // groups (0.671 of HBM peak, against 0.647-0.650 at 16384 or 4096
    // blocks with several groups per thread)
    const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>((end - first + 1023) / 1024, 65536));
    hipLaunchKernelGGL(MSKG_VARIANT(k_rate_sweep_range, la, e->cov_model), dim3(blocks), dim3(256), 0, e->stream, L,
                       st.rtot, e->d_err, first, end);
    HIPCHK(e, hipGetLastError());
  }
  HIPCHK(e, hipEventRecord(e->ev1, e->stream));
  int err = 0;
  HIPCHK(e, hipMemcpyAsync(&err, e->d_err, sizeof(int), hipMemcpyDeviceToHost, e->stream));
  HIPCHK(e, hipStreamSynchronize(e->stream));
  float ms = 0;
  HIPCHK(e, hipEventElapsedTime(&ms, e->ev0, e->ev1));
  e->sweep_ms = ms;
  e->sweep_points = end - first;
  if (err) return fail_(e, MSKG_ESTATE, "rate sweep: %s", err_name(err));
  return MSKG_OK;
}

int mskg_refresh_rates(mskg_engine *e) {
  if (!e || !e->loaded) return fail_(e, MSKG_EINVAL, "no state loaded");
  const int rc = refresh_range(e, 0, e->box.npts, false);
  if (!rc && e->track) e->keys_valid = true;
  return rc;
}

int mskg_refresh_active(mskg_engine *e) {
  if (!e || !e->loaded) return fail_(e, MSKG_EINVAL, "no state loaded");
  // the tile layers that hold MC particles: every tile layer that held a
  // list entry since the lists were built is in the window of some colour
  // (windows only widen between sweeps, and a sweep keeps the layers its
  // tiles and their appends touched).  With lists that are not current (an
  // upload, a replay, a whole-lattice sweep since) every point is swept.
  if (e->lists_dirty) return mskg_refresh_rates(e);
  int zlo = 0, zhi = e->ntz - 1;
  {
    int w[32];
    HIPCHK(e, hipMemcpyAsync(w, e->d_win, sizeof w, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    zlo = 0x7fffffff;
    zhi = -1;
    for (int c = 0; c < 8; c++) {
      zlo = std::min(zlo, w[4 * c]);
      zhi = std::max(zhi, w[4 * c + 1]);
    }
    if (zlo > zhi) zlo = 0, zhi = -1;  // no MC particle
    zlo = std::max(zlo, 0);
    zhi = std::min(zhi, e->ntz - 1);
  }
  const Box &b = e->box;
  const int64_t z0 = (int64_t)zlo * e->tile, z1 = zhi < zlo ? z0 : std::min<int64_t>((int64_t)(zhi + 1) * e->tile, b.lz + 1);
  return refresh_range(e, z0 * b.plane, z1 * b.plane, true);
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
  L.inlist = nullptr;
  e->lists_dirty = true;
  hipLaunchKernelGGL(MSKG_VARIANT(k_replay, e->tab.la_mode || e->tab.st_n, e->cov_model), dim3(1), dim3(WAVE), 0, e->stream, L, d_log, n, clock_mode, e->d_st);
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

// The start of a synchronous-sublattice run, on the engine's stream: tile
// lists current, the run's limits in the run state, the tile totals of the
// current rates and the first window (k_sweep_end with tau = 0).  st: the
// run state before it.
static int run_begin(mskg_engine *e, int64_t max_events, double max_time, RunState &st) {
  if (!e || !e->loaded) return fail_(e, MSKG_EINVAL, "no state loaded");
  if (max_events < 0) return fail_(e, MSKG_EINVAL, "max_events < 0");
  Lattice L = make_lattice(e);
  int rc = read_state(e, st);
  if (rc) return rc;
  if (st.err) return fail_(e, MSKG_ESTATE, "engine is in an error state: %s", err_name(st.err));
  const int tblocks = (e->ntiles + 3) / 4;
  {
    static int w0[32];
    for (int c = 0; c < 8; c++) {
      w0[4 * c] = w0[4 * c + 2] = 0x7fffffff;
      w0[4 * c + 1] = w0[4 * c + 3] = -1;
    }
    HIPCHK(e, hipMemcpyAsync(e->d_win, w0, sizeof w0, hipMemcpyHostToDevice, e->stream));
  }
  if (e->lists_dirty) {
    HIPCHK(e, hipMemsetAsync(e->d_list_err, 0, sizeof(int), e->stream));
    hipLaunchKernelGGL(k_build_lists, dim3(tblocks), dim3(256), 0, e->stream, L, e->ntiles);
    HIPCHK(e, hipGetLastError());
    e->lists_dirty = false;
  } else {
    hipLaunchKernelGGL(k_list_window, dim3(tblocks), dim3(256), 0, e->stream, L, e->ntiles);
    HIPCHK(e, hipGetLastError());
  }
  {
    RunState s0 = st;
    s0.tau = 0.0;  // k_sweep_end with tau = 0 only sets the next window
    s0.stop = 0;
    s0.final_sweep = 0;
    s0.tau_policy = e->cfg.tau_policy;
    s0.events_per_pass = e->cfg.events_per_pass > 0 ? e->cfg.events_per_pass : 1.0;
    s0.max_time = max_time;
    s0.max_events = st.events + (unsigned long long)max_events;
    if (int rc2 = write_state_head(e, s0)) return rc2;
    HIPCHK(e, hipMemsetAsync((char *)e->d_st + RS_HEAD, 0, sizeof s0.slot, e->stream));
  }
  // tile totals of the current rates, then the first window
  hipLaunchKernelGGL(k_tile_rates, dim3(tblocks), dim3(256), 0, e->stream, L, e->ntiles, e->d_st);
  HIPCHK(e, hipGetLastError());
  hipLaunchKernelGGL(k_rsum, dim3(NPART), dim3(256), 0, e->stream, e->d_tile_r, e->ntiles, e->rsum_chunk,
                       e->d_st);
  hipLaunchKernelGGL(k_sweep_end, dim3(1), dim3(NSLOT), 0, e->stream, e->d_st, e->d_win, e->d_list_err,
                       e->d_done + e->ntiles);
  HIPCHK(e, hipGetLastError());
  return MSKG_OK;
}

// The pass kernels' parameters of an engine (colour 0).
static PassParams pass_params(mskg_engine *e) {
  PassParams P;
  P.L = make_lattice(e);
  P.st = e->d_st;
  P.log = e->d_log;
  P.log_cap = (unsigned long long)e->log_cap;
  P.ntz = e->ntz;
  P.cap = e->cap;
  P.key0 = (uint32_t)e->cfg.seed;
  P.key1 = (uint32_t)(e->cfg.seed >> 32);
  P.colour = 0;
  P.rsum_chunk = e->rsum_chunk;
  P.done = e->d_done;
  P.sched = e->d_done + e->ntiles;
  return P;
}

int mskg_run(mskg_engine *e, int64_t max_events, double max_time, int64_t *events_done, double *kmc_time,
             int32_t *stop_reason) {
  RunState st;
  int rc = run_begin(e, max_events, max_time, st);
  if (rc) return rc;
  const unsigned long long ev_start = st.events;
  const size_t shm = (size_t)e->cap * (4 + 4 + 1) + 16;
  // one sweep = eight colour passes (or one dependency-scheduled launch) +
  // the window/clock update, captured once as a graph and replayed
  if (!e->graph_exec) {
    PassParams P = pass_params(e);
    const bool la = e->tab.la_mode || e->tab.st_n;
    int grid = std::max(1, std::min(e->ntiles / 8 + 1, e->pass_grid));
    if (e->schedule == MSKG_SCHEDULE_TILES) {
      // every wave the device holds at once: the waves take tiles from the
      // sweep's queue until it is empty
      int per_cu = 0, cus = 0;
      HIPCHK(e, hipOccupancyMaxActiveBlocksPerMultiprocessor(
                    &per_cu, reinterpret_cast<const void *>(MSKG_VARIANT(k_sweep, la, e->cov_model)), WAVE, shm));
      HIPCHK(e, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, e->device));
      grid = std::max(1, per_cu * cus);
    }
    HIPCHK(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    if (e->schedule == MSKG_SCHEDULE_TILES) {
      hipLaunchKernelGGL(MSKG_VARIANT(k_sweep, la, e->cov_model), dim3(grid), dim3(WAVE), shm, e->stream, P);
    } else {
      for (int c = 0; c < 8; c++) {
        P.colour = c;
        hipLaunchKernelGGL(MSKG_VARIANT(k_pass, la, e->cov_model), dim3(grid), dim3(WAVE), shm, e->stream, P);
      }
    }
    hipLaunchKernelGGL(k_rsum, dim3(NPART), dim3(256), 0, e->stream, e->d_tile_r, e->ntiles, e->rsum_chunk,
                       e->d_st);
    hipLaunchKernelGGL(k_sweep_end, dim3(1), dim3(NSLOT), 0, e->stream, e->d_st, e->d_win, e->d_list_err,
                       e->d_done + e->ntiles);
    HIPCHK(e, hipStreamEndCapture(e->stream, &e->graph));
    HIPCHK(e, hipGraphInstantiate(&e->graph_exec, e->graph, nullptr, nullptr, 0));
  }
  int32_t stop = MSKG_STOP_LIMIT;
  HIPCHK(e, hipEventRecord(e->ev0, e->stream));
  int64_t sweeps = 0;
  int batch = 1;
  for (;;) {
    for (int sb = 0; sb < batch; sb++) HIPCHK(e, hipGraphLaunch(e->graph_exec, e->stream));
    sweeps += batch;
    rc = read_state(e, st);
    if (rc) return rc;
    if (st.err) break;
    int64_t done = (int64_t)(st.events - ev_start);
    if (st.stop) { stop = st.stop == STOP_AT_LIMIT ? MSKG_STOP_LIMIT : st.stop; break; }
    if (done >= max_events || st.clock >= max_time) break;
    // size the next batch of sweeps from the observed events per sweep, at
    // most doubling the sweeps run so far: sweeps queued past the stop
    // still cost a launch each
    double per = (double)done / (double)sweeps;
    int64_t want = per > 0 ? (int64_t)((double)(max_events - done) / per * 0.9) : 16;
    batch = (int)std::max<int64_t>(1, std::min<int64_t>(std::min<int64_t>(want, 512), sweeps));
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
        // the reference's list membership: an atom of coordination <= 3 is an MC
        // particle even when it has no rate (a frozen-substrate atom)
        if (cls[p] || (kind_species(kind) && coor <= 3)) {
          if (na < cap_ad) {
            if (ad_xyz) memcpy(ad_xyz + 3 * na, xyz, 12);
            if (ad_nn) memcpy(ad_nn + 12 * na, nn, 48);
            int c = cls[p];
            if (ad_event) ad_event[na] = (c == CLS_EVAP || c == CLS_ISOLATED || !c) ? 4 : c == CLS_DESORB ? 5 : 1;
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

static int census(mskg_engine *e, int64_t out[6]) {
  unsigned long long *d = nullptr;
  if (hipMalloc(&d, 6 * sizeof(unsigned long long)) != hipSuccess)
    return fail_(e, MSKG_ENOMEM, "census: no device memory");
  unsigned long long h[6];
  hipError_t ce = hipMemsetAsync(d, 0, 6 * sizeof(unsigned long long), e->stream);
  if (ce == hipSuccess) {
    const int blocks = (int)std::min<int64_t>((e->box.npts + 255) / 256, 256 * 64);
    hipLaunchKernelGGL(k_census, dim3(blocks), dim3(256), 0, e->stream, make_lattice(e), d);
    ce = hipGetLastError();
  }
  if (ce == hipSuccess) ce = hipMemcpyAsync(h, d, sizeof h, hipMemcpyDeviceToHost, e->stream);
  if (ce == hipSuccess) ce = hipStreamSynchronize(e->stream);
  hipFree(d);
  if (ce != hipSuccess) return fail_(e, MSKG_EDEVICE, "census: %s", hipGetErrorString(ce));
  for (int q = 0; q < 6; q++) out[q] = (int64_t)h[q];
  return MSKG_OK;
}

int mskg_species_counts(mskg_engine *e, int64_t evaporated[3], int64_t desorbed[3]) {
  if (!e || !e->loaded) return fail_(e, MSKG_EINVAL, "no state loaded");
  RunState st;
  int rc = read_state(e, st);
  if (rc) return rc;
  int64_t now[6];
  if ((rc = census(e, now))) return rc;
  for (int k = 0; k < 3; k++) {
    if (evaporated) evaporated[k] = (int64_t)st.counts[k] - (now[k] - e->census0[k]);
    if (desorbed) desorbed[k] = (int64_t)st.desorbed[k];
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

int mskg_set_schedule(mskg_engine *e, int32_t schedule) {
  if (!e) return MSKG_EINVAL;
  if (schedule != MSKG_SCHEDULE_TILES && schedule != MSKG_SCHEDULE_COLOURS)
    return fail_(e, MSKG_EINVAL, "schedule %d: MSKG_SCHEDULE_TILES (1) or MSKG_SCHEDULE_COLOURS (0)", schedule);
  if (schedule != e->schedule) {
    HIPCHK(e, hipStreamSynchronize(e->stream));
    drop_graph(e);
    e->schedule = schedule;
  }
  return MSKG_OK;
}

int mskg_run_diagnostics(mskg_engine *e, int64_t *guard_exits, int64_t *sweeps, double *tau, int64_t *n_mc) {
  if (!e || !e->loaded) return fail_(e, MSKG_EINVAL, "no state loaded");
  RunState st;
  int rc = read_state(e, st);
  if (rc) return rc;
  if (guard_exits) *guard_exits = (int64_t)st.guard_exits;
  if (sweeps) *sweeps = (int64_t)st.sweeps_run;
  if (tau) *tau = st.tau;
  if (n_mc) *n_mc = (int64_t)st.n_mc;
  return MSKG_OK;
}

static int build_slab(mskg_engine *e, int32_t mode, int32_t len, double fraction);

int mskg_build_slab(mskg_engine *e, int32_t mode, int32_t len) {
  if (!e || !e->d_site) return MSKG_EINVAL;
  if (mode != 0 && mode != 1) return fail_(e, MSKG_EINVAL, "mode must be 0 (SetSi) or 1 (SetSiC3C)");
  return build_slab(e, mode, len, 0.0);
}

int mskg_build_alloy_slab(mskg_engine *e, int32_t len, double fraction) {
  if (!e || !e->d_site) return MSKG_EINVAL;
  if (!(fraction >= 0.0 && fraction <= 1.0)) return fail_(e, MSKG_EINVAL, "alloy fraction must be in [0, 1]");
  if (e->ncrystal < 2) return fail_(e, MSKG_EINVAL, "an alloy needs n_crystal >= 2");
  return build_slab(e, 2, len, fraction);
}

static int build_slab(mskg_engine *e, int32_t mode, int32_t len, double fraction) {
  if (len < 6 || len > e->logical_lz - (mode != 1 ? 7 : 1) || len % 3)
    return fail_(e, MSKG_EINVAL, "slab height %d does not fit the box height %d", len, e->logical_lz);
  if (e->windowed && len + 6 > e->box.lz - TOP_MARGIN)
    return fail_(e, MSKG_EINVAL, "slab height %d: the stored planes (stored_len_z %d) need %d more above it", len,
                 e->box.lz + 1, TOP_MARGIN + 6);
  Lattice L = make_lattice(e);
  int blocks = (int)std::min<int64_t>((e->box.npts + 255) / 256, 256 * 64);
  hipLaunchKernelGGL(k_build_slab, dim3(blocks), dim3(256), 0, e->stream, L, len, mode, fraction,
                     (uint32_t)e->cfg.seed, (uint32_t)(e->cfg.seed >> 32), e->logical_lz);
  HIPCHK(e, hipGetLastError());
  hipFree(e->d_offwall);  // the SetSi builders fill their wall planes at every grid point
  e->d_offwall = nullptr;
  e->offwall_planes = mode != 1;
  if (e->tab.use_f2 || mode == 2) {  // the surface atoms' rates from their neighbours (species, second shell)
    HIPCHK(e, hipMemsetAsync(e->d_err, 0, sizeof(int), e->stream));
    hipLaunchKernelGGL(MSKG_VARIANT(k_rate_sweep, e->tab.la_mode || e->tab.st_n, e->cov_model), dim3(blocks), dim3(256), 0, e->stream, L, 0.0, e->d_err, 1);
    HIPCHK(e, hipGetLastError());
  }
  // the total rate of the initial population seeds the isolated-atom rule
  RunState st;
  memset(&st, 0, sizeof st);
  st.events_per_pass = e->cfg.events_per_pass > 0 ? e->cfg.events_per_pass : 1.0;
  const double nsurf = (double)e->box.lx * e->box.ly / 36.0;  // 2 atoms per (6x6) surface cell per layer
  const Tables &t = e->tab;
  st.rtot = nsurf * (mode ? 0.5 * (t.pe[0 + 3 * 2] + t.pe[1 + 3 * 2]) : t.pe[0 + 3 * 2]) +
            nsurf * (t.pd[0 + 3] + t.pd[1 + 3] + t.pd[2 + 3]);
  HIPCHK(e, hipMemcpyAsync(e->d_st, &st, sizeof st, hipMemcpyHostToDevice, e->stream));
  HIPCHK(e, hipStreamSynchronize(e->stream));
  if (int rc_la = la_variant_off(e)) return rc_la;  // a new state: the counts are the caller's to give again
  if (int rc_c = census(e, e->census0)) return rc_c;
  e->loaded = true;
  e->lists_dirty = true;
  e->keys_valid = false;
  return MSKG_OK;
}

// A shaped initial state built on the device (k_shape_place ->
// k_shape_isolated -> k_shape_records -> rate sweep), and the off-lattice
// wall bits of the builders that fill wall regions.
static int build_shape(mskg_engine *e, ShapeSpec S, bool walls_off_lattice) {
  if (e->windowed)
    return fail_(e, MSKG_EINVAL, "the shaped builders fill the whole box: every plane must be stored");
  S.lz = e->logical_lz;
  S.key0 = (uint32_t)e->cfg.seed;
  S.key1 = (uint32_t)(e->cfg.seed >> 32);
  S.cx = e->box.lx / 2;
  S.cy = e->box.ly / 2;
  S.cz = e->logical_lz / 2;
  Lattice L = make_lattice(e);
  L.key = nullptr;
  const int blocks = (int)std::min<int64_t>((e->box.npts + 255) / 256, 256 * 64);
  HIPCHK(e, hipMemsetAsync(e->d_err, 0, sizeof(int), e->stream));
  hipLaunchKernelGGL(k_shape_place, dim3(blocks), dim3(256), 0, e->stream, L, S);
  hipLaunchKernelGGL(k_shape_isolated, dim3(blocks), dim3(256), 0, e->stream, L, S);
  hipLaunchKernelGGL(k_shape_records, dim3(blocks), dim3(256), 0, e->stream, L, S, e->d_err);
  hipLaunchKernelGGL(MSKG_VARIANT(k_rate_sweep, e->tab.la_mode || e->tab.st_n, e->cov_model), dim3(blocks),
                     dim3(256), 0, e->stream, L, 0.0, e->d_err, 0);
  hipLaunchKernelGGL(k_initial_placeholders, dim3(blocks), dim3(256), 0, e->stream, L);
  HIPCHK(e, hipGetLastError());
  hipFree(e->d_offwall);
  e->d_offwall = nullptr;
  e->offwall_planes = false;
  if (walls_off_lattice) {
    const int64_t words = (e->box.lx * (int64_t)e->box.ly * (e->box.lz + 1) + 31) / 32;
    HIPCHK(e, hipMalloc(&e->d_offwall, words * sizeof(uint32_t)));
    const int wb = (int)std::min<int64_t>((words + 255) / 256, 256 * 64);
    hipLaunchKernelGGL(k_shape_offwall, dim3(wb), dim3(256), 0, e->stream, L, S, e->d_offwall, words);
    HIPCHK(e, hipGetLastError());
  }
  int err = 0;
  HIPCHK(e, hipMemcpyAsync(&err, e->d_err, sizeof(int), hipMemcpyDeviceToHost, e->stream));
  RunState st;
  memset(&st, 0, sizeof st);
  st.events_per_pass = e->cfg.events_per_pass > 0 ? e->cfg.events_per_pass : 1.0;
  HIPCHK(e, hipMemcpyAsync(e->d_st, &st, sizeof st, hipMemcpyHostToDevice, e->stream));
  HIPCHK(e, hipStreamSynchronize(e->stream));
  e->loaded = true;  // the state is the builder's, whatever follows
  e->lists_dirty = true;
  e->keys_valid = false;
  if (err) {
    e->loaded = false;
    return fail_(e, MSKG_ESTATE, "shaped builder: %s", err_name(err));
  }
  if (int rc_la = la_variant_off(e)) return rc_la;
  if (int rc_c = census(e, e->census0)) return rc_c;
  return MSKG_OK;
}

int mskg_build_shape(mskg_engine *e, const char *init_st, const int32_t lens[4]) {
  if (!e || !e->d_site || !init_st || !lens) return MSKG_EINVAL;
  const std::string is(init_st);
  if (is == "F" || is == "FS") return mskg_build_slab(e, is == "F" ? 1 : 0, lens[0] - lens[0] % 24);
  ShapeSpec S;
  memset(&S, 0, sizeof S);
  S.zfix = -1;
  S.l1 = lens[0]; S.l2 = lens[1]; S.l3 = lens[2]; S.l4 = lens[3];
  if (is == "C" || is == "S") {  // SetNC.f:25-235, SetNCSp.f:25-237
    S.shape = is == "C" ? SHAPE_BOX : SHAPE_SPHERE;
    if (S.l1 <= 0 || (is == "C" && (S.l2 <= 0 || S.l3 <= 0)))
      return fail_(e, MSKG_EINVAL, "InitSt %s: the shape's lengths must be positive", init_st);
    return build_shape(e, S, false);
  }
  if (is == "I") {  // SetInvPOff.f:21-286: Len1 the pyramid, Len2 the substrate top
    S.shape = SHAPE_INVP;
    S.zfix = 20;
    if (S.l2 <= 21 || S.l2 + 10 > e->logical_lz)
      return fail_(e, MSKG_EINVAL, "InitSt I: Len2 %d must lie in 22..LenZ-10", S.l2);
    return build_shape(e, S, false);
  }
  if (is == "FF") {  // SetSiFINFET.f:25-312 with main.f:281-285's reductions
    S.shape = SHAPE_FINFET;
    S.l1 -= S.l1 % 24; S.l2 -= S.l2 % 24; S.l3 -= S.l3 % 24; S.l4 -= S.l4 % 24;
    if (S.l1 < 6 || S.l3 + 10 > e->logical_lz)
      return fail_(e, MSKG_EINVAL, "InitSt FF: lengths %d %d %d %d do not fit the box", S.l1, S.l2, S.l3, S.l4);
    S.xa = (int)(0.5f * (float)(e->box.lx - S.l4));
    S.xb = (int)(0.5f * (float)(e->box.lx + S.l4));
    S.walls = 1;
    return build_shape(e, S, true);
  }
  return fail_(e, MSKG_EINVAL, "InitSt '%s': the device builds F, FS, C, S, I, FF (and IN with mskg_build_cad)",
               init_st);
}

int mskg_build_cad(mskg_engine *e, const int8_t *ids, int32_t on_device, double alloy_fraction) {
  if (!e || !e->d_site || !ids) return MSKG_EINVAL;
  if (!(alloy_fraction >= 0.0 && alloy_fraction <= 1.0))
    return fail_(e, MSKG_EINVAL, "alloy fraction must be in [0, 1]");
  if (alloy_fraction > 0.0 && e->ncrystal < 2) return fail_(e, MSKG_EINVAL, "an alloy needs n_crystal >= 2");
  if (e->logical_lz < 10) return fail_(e, MSKG_EINVAL, "SetCAD needs LenZ >= 10");
  if (e->windowed)
    return fail_(e, MSKG_EINVAL, "the shaped builders fill the whole box: every plane must be stored");
  ShapeSpec S;
  memset(&S, 0, sizeof S);
  S.shape = SHAPE_CAD;
  S.zfix = -1;
  S.walls = 1;
  S.fraction = alloy_fraction;
  int8_t *d_ids = nullptr;
  const size_t nb = (size_t)e->box.lx * e->box.ly * e->logical_lz;
  if (on_device) {
    S.ids = ids;
  } else {
    HIPCHK(e, hipMalloc(&d_ids, nb));
    hipError_t ce = hipMemcpyAsync(d_ids, ids, nb, hipMemcpyHostToDevice, e->stream);
    if (ce != hipSuccess) {
      hipFree(d_ids);
      return fail_(e, MSKG_EDEVICE, "geometry upload: %s", hipGetErrorString(ce));
    }
    S.ids = d_ids;
  }
  const int rc = build_shape(e, S, true);
  hipStreamSynchronize(e->stream);
  hipFree(d_ids);
  return rc;
}

int mskg_track_classes(mskg_engine *e, int32_t on) {
  if (!e || !e->d_site) return MSKG_EINVAL;
  if (!on) {
    if (e->track) drop_graph(e);
    e->track = false;
    e->keys_valid = false;
    return MSKG_OK;
  }
  if (e->tab.ncov || e->tab.la_mode)
    return fail_(e, MSKG_EINVAL, "class keys cover models without coverage species and without LA rates");
  const int64_t np = e->box.npts;
  if (!e->d_key) {
    HIPCHK(e, hipMalloc(&e->d_key, np * sizeof(int32_t)));
    HIPCHK(e, hipMalloc(&e->d_key_seen, np * sizeof(int32_t)));
    HIPCHK(e, hipMalloc(&e->d_rate_seen, np * sizeof(float)));
  }
  if (!e->track) drop_graph(e);
  e->track = true;
  e->keys_valid = false;
  if (!e->loaded) return MSKG_OK;
  return mskg_refresh_rates(e);  // a keyed sweep: rates and keys from the state
}

int mskg_copy_state(mskg_engine *dst, const mskg_engine *src) {
  if (!dst || !src || !src->loaded) return MSKG_EINVAL;
  if (dst == src) return MSKG_OK;
  if (dst->box.lx != src->box.lx || dst->box.ly != src->box.ly || dst->box.lz != src->box.lz ||
      dst->logical_lz != src->logical_lz ||
      dst->tile != src->tile)
    return fail_(dst, MSKG_EINVAL, "copy_state: engines differ in box or tile");
  const int64_t np = dst->box.npts;
  HIPCHK(dst, hipStreamSynchronize(src->stream));
  HIPCHK(dst, hipMemcpyAsync(dst->d_site, src->d_site, np * sizeof(Site), hipMemcpyDeviceToDevice, dst->stream));
  HIPCHK(dst, hipMemcpyAsync(dst->d_rate, src->d_rate, np * sizeof(float), hipMemcpyDeviceToDevice, dst->stream));
  HIPCHK(dst, hipMemcpyAsync(dst->d_cls, src->d_cls, np, hipMemcpyDeviceToDevice, dst->stream));
  dst->keys_valid = false;
  hipFree(dst->d_offwall);
  dst->d_offwall = nullptr;
  dst->offwall_planes = src->offwall_planes;
  if (src->d_offwall) {
    const size_t words = (size_t)((src->box.lx * (int64_t)src->box.ly * (src->box.lz + 1) + 31) / 32);
    HIPCHK(dst, hipMalloc(&dst->d_offwall, words * sizeof(uint32_t)));
    HIPCHK(dst, hipMemcpyAsync(dst->d_offwall, src->d_offwall, words * sizeof(uint32_t), hipMemcpyDeviceToDevice,
                               dst->stream));
  }
  if (dst->track && src->track && src->keys_valid) {
    HIPCHK(dst, hipMemcpyAsync(dst->d_key, src->d_key, np * sizeof(int32_t), hipMemcpyDeviceToDevice, dst->stream));
    dst->keys_valid = true;
  }
  RunState st;
  HIPCHK(dst, hipMemcpyAsync(&st, src->d_st, sizeof st, hipMemcpyDeviceToHost, dst->stream));
  HIPCHK(dst, hipStreamSynchronize(dst->stream));
  RunState s0;
  memset(&s0, 0, sizeof s0);
  s0.clock = st.clock;
  s0.rtot = st.rtot;
  s0.events_per_pass = dst->cfg.events_per_pass > 0 ? dst->cfg.events_per_pass : 1.0;
  s0.la_variant = dst->tab.la_variant;
  for (int k = 0; k < 3; k++) {
    s0.cc[k] = st.cc[k] + (long long)st.dcc[k];
    s0.cinit[k] = st.cinit[k];
    s0.cl[k] = s0.cinit[k] - s0.cc[k];
  }
  HIPCHK(dst, hipMemcpyAsync(dst->d_st, &s0, sizeof s0, hipMemcpyHostToDevice, dst->stream));
  HIPCHK(dst, hipStreamSynchronize(dst->stream));
  if (int rc = census(dst, dst->census0)) return rc;  // the event counters restart here
  dst->loaded = true;
  dst->lists_dirty = true;
  return MSKG_OK;
}

int mskg_set_event_log(mskg_engine *e, int64_t capacity) {
  if (!e || capacity < 0) return MSKG_EINVAL;
  if (e->d_log) { HIPCHK(e, hipFree(e->d_log)); e->d_log = nullptr; }
  e->log_cap = 0;
  if (capacity > 0) {
    HIPCHK(e, hipMalloc(&e->d_log, capacity * sizeof(mskg_event)));
    e->log_cap = capacity;
  }
  drop_graph(e);
  if (e->loaded) {
    RunState st;
    int rc = read_state(e, st);
    if (rc) return rc;
    st.log_n = 0;
    if (int rc2 = write_state_head(e, st)) return rc2;
    HIPCHK(e, hipStreamSynchronize(e->stream));
  }
  return MSKG_OK;
}

int mskg_get_event_log(mskg_engine *e, mskg_event *out, int64_t cap, int64_t *n) {
  if (!e || !n) return MSKG_EINVAL;
  RunState st;
  int rc = read_state(e, st);
  if (rc) return rc;
  *n = (int64_t)st.log_n;
  if (st.log_n > (unsigned long long)e->log_cap)
    return fail_(e, MSKG_ECAPACITY, "event log overflowed: %lld events, capacity %lld", (long long)st.log_n,
                 (long long)e->log_cap);
  int64_t m = std::min<int64_t>(*n, cap);
  if (m > 0 && out) {
    HIPCHK(e, hipMemcpyAsync(out, e->d_log, m * sizeof(mskg_event), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
  }
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

int mskg_run_serial(mskg_engine *const *engines, int32_t n, const int64_t *max_events, const double *max_time,
                    int64_t *events_done, double *kmc_time, int32_t *stop_reason) {
  if (!engines || n <= 0 || !max_events || !max_time) return MSKG_EINVAL;
  mskg_engine *e0 = engines[0];
  if (!e0) return MSKG_EINVAL;
  std::vector<SerialDesc> descs(n);
  std::vector<unsigned long long> ev0(n), log0(n);
  bool any_la = false, any_cov = false;
  for (int i = 0; i < n; i++) {
    mskg_engine *e = engines[i];
    if (!e || !e->loaded) return fail_(e0, MSKG_EINVAL, "serial run: replica %d has no state loaded", i);
    any_la |= e->tab.la_mode != 0 || e->tab.st_n > 0;
    if (e->tab.la_variant)
      return fail_(e0, MSKG_EINVAL, "serial run: replica %d uses the laser-annealing build's rate form", i);
    any_cov |= e->cov_model;
    if (max_events[i] < 0) return fail_(e, MSKG_EINVAL, "max_events < 0");
    for (int j = 0; j < i; j++)
      if (engines[j] == e) return fail_(e0, MSKG_EINVAL, "serial run: replica %d appears twice", i);
    HIPCHK(e, hipStreamSynchronize(e->stream));
    const int nblk = (e->ntiles + WAVE - 1) / WAVE;
    if (!e->d_tsum) {
      HIPCHK(e, hipMalloc(&e->d_tsum, e->ntiles * sizeof(double)));
      HIPCHK(e, hipMalloc(&e->d_bsum, nblk * sizeof(double)));
    }
    Lattice L = make_lattice(e);
    if (e->lists_dirty) {
      HIPCHK(e, hipMemsetAsync(e->d_list_err, 0, sizeof(int), e->stream));
      hipLaunchKernelGGL(k_build_lists, dim3((e->ntiles + 3) / 4), dim3(256), 0, e->stream, L, e->ntiles);
      HIPCHK(e, hipGetLastError());
      e->lists_dirty = false;
    }
    RunState st;
    int rc = read_state(e, st);
    if (rc) return rc;
    if (st.err) return fail_(e, MSKG_ESTATE, "engine is in an error state: %s", err_name(st.err));
    int lerr = 0;
    HIPCHK(e, hipMemcpy(&lerr, e->d_list_err, sizeof(int), hipMemcpyDeviceToHost));
    if (lerr) return fail_(e, MSKG_ECAPACITY, "a tile holds more MC particles than tile_capacity");
    ev0[i] = st.events;
    log0[i] = st.log_n;
    SerialDesc &D = descs[i];
    D.L = L;
    D.tsum = e->d_tsum;
    D.bsum = e->d_bsum;
    D.st = e->d_st;
    D.log = e->d_log;
    D.log_cap = (unsigned long long)e->log_cap;
    D.max_events = max_events[i];
    D.max_time = max_time[i];
    D.key0 = (uint32_t)e->cfg.seed;
    D.key1 = (uint32_t)(e->cfg.seed >> 32) ^ 0x5e1a1u;  // a stream apart from the sublattice passes
    D.ntiles = e->ntiles;
    D.nblk = nblk;
    D.replica = i;
  }
  SerialDesc *d_descs = nullptr;
  HIPCHK(e0, hipMalloc(&d_descs, n * sizeof(SerialDesc)));
  hipError_t ce = hipMemcpyAsync(d_descs, descs.data(), n * sizeof(SerialDesc), hipMemcpyHostToDevice, e0->stream);
  if (ce == hipSuccess) {
    HIPCHK(e0, hipEventRecord(e0->ev0, e0->stream));
    hipLaunchKernelGGL(MSKG_VARIANT(k_serial, any_la, any_cov), dim3(n), dim3(WAVE), 0, e0->stream, d_descs);
    ce = hipGetLastError();
    if (ce == hipSuccess) ce = hipEventRecord(e0->ev1, e0->stream);
    if (ce == hipSuccess) ce = hipStreamSynchronize(e0->stream);
  }
  hipFree(d_descs);
  if (ce != hipSuccess) return fail_(e0, MSKG_EDEVICE, "k_serial: %s", hipGetErrorString(ce));
  float ms = 0;
  HIPCHK(e0, hipEventElapsedTime(&ms, e0->ev0, e0->ev1));
  int rc_all = MSKG_OK;
  for (int i = 0; i < n; i++) {
    mskg_engine *e = engines[i];
    e->run_ms = ms;
    e->lists_dirty = false;
    RunState st;
    int rc = read_state(e, st);
    if (rc) return rc;
    if (events_done) events_done[i] = (int64_t)(st.events - ev0[i]);
    if (kmc_time) kmc_time[i] = st.clock;
    if (stop_reason) stop_reason[i] = st.stop;
    if (st.err && rc_all == MSKG_OK)
      rc_all = fail_(e, MSKG_ESTATE, "serial run of replica %d stopped: %s", i, err_name(st.err));
    if (e->d_log && st.log_n > (unsigned long long)e->log_cap && rc_all == MSKG_OK)
      rc_all = fail_(e, MSKG_ECAPACITY, "event log of replica %d overflowed", i);
    (void)log0;
  }
  return rc_all;
}

int mskg_run_ensemble(mskg_engine *const *engines, int32_t n, const int64_t *max_events, const double *max_time,
                      int64_t *events_done, double *kmc_time, int32_t *stop_reason) {
  if (!engines || n <= 0 || !max_events || !max_time) return MSKG_EINVAL;
  mskg_engine *e0 = engines[0];
  if (!e0) return MSKG_EINVAL;
  if (n > ENS_MAX) return fail_(e0, MSKG_EINVAL, "ensemble of %d replicas: at most %d", n, ENS_MAX);
  const bool la = e0->tab.la_mode || e0->tab.st_n;
  const bool cov = e0->cov_model;
  int cap = 0;
  int64_t queue_bound = 0;
  for (int i = 0; i < n; i++) {
    mskg_engine *e = engines[i];
    if (!e) return fail_(e0, MSKG_EINVAL, "ensemble: replica %d is null", i);
    for (int j = 0; j < i; j++)
      if (engines[j] == e) return fail_(e0, MSKG_EINVAL, "ensemble: replica %d appears twice", i);
    if (e->device != e0->device) return fail_(e0, MSKG_EINVAL, "ensemble: replica %d lives on another device", i);
    if ((bool)(e->tab.la_mode || e->tab.st_n) != la || e->cov_model != cov)
      return fail_(e0, MSKG_EINVAL, "ensemble: replica %d needs another rate-model kernel than replica 0", i);
    cap = std::max(cap, e->cap);
    queue_bound += (int64_t)(e->ntx / 2) * (e->nty / 2) * ((e->ntz + 1) / 2);  // a colour's tiles at most
  }
  // a colour's queue head counts entries in units of 64 in 32 bits
  if (queue_bound > (1ll << 25)) return fail_(e0, MSKG_EINVAL, "ensemble: too many tiles per colour pass");
  std::vector<PassParams> reps(n);
  std::vector<unsigned long long> ev0(n);
  for (int i = 0; i < n; i++) {
    mskg_engine *e = engines[i];
    RunState st;
    if (int rc = run_begin(e, max_events[i], max_time[i], st)) {
      if (e != e0) fail_(e0, rc, "ensemble replica %d: %s", i, e->err.c_str());
      return rc;
    }
    ev0[i] = st.events;
    reps[i] = pass_params(e);
  }
  for (int i = 0; i < n; i++) HIPCHK(engines[i], hipStreamSynchronize(engines[i]->stream));
  hipStream_t sm = e0->stream;
  PassParams *d_reps = nullptr;
  int *d_plan = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  const int64_t pw = ens_plan_words(n);
  const size_t shm = (size_t)cap * (4 + 4 + 1) + 16;
  int nrun = 0;
  int64_t sweeps = 0;
  float ms = 0;
  auto kpass = MSKG_VARIANT(k_pass_ens, la, cov);
  hipError_t ce = hipMalloc(&d_reps, n * sizeof(PassParams));
  if (ce == hipSuccess) ce = hipMalloc(&d_plan, pw * sizeof(int));
  if (ce == hipSuccess) ce = hipMemcpyAsync(d_reps, reps.data(), n * sizeof(PassParams), hipMemcpyHostToDevice, sm);
  int grid = 0;
  if (ce == hipSuccess) {
    // persistent waves: as many as the device holds at once
    int per_cu = 0, cus = 0;
    ce = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(kpass), WAVE, shm);
    if (ce == hipSuccess) ce = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, e0->device);
    grid = std::max(1, per_cu * cus);
  }
  if (ce == hipSuccess) {
    hipLaunchKernelGGL(k_ens_plan, dim3(1), dim3(ENS_MAX), 0, sm, d_reps, n, d_plan, 0, 1);
    ce = hipGetLastError();
  }
  // one sweep of every running replica: the eight colour passes, the tile
  // sums and sweep ends, the next sweep's plan
  if (ce == hipSuccess) ce = hipStreamBeginCapture(sm, hipStreamCaptureModeThreadLocal);
  if (ce == hipSuccess) {
    for (int c = 0; c < 8; c++) {
      if (c) hipLaunchKernelGGL(k_ens_plan, dim3(1), dim3(ENS_MAX), 0, sm, d_reps, n, d_plan, c, 0);
      hipLaunchKernelGGL(kpass, dim3(grid), dim3(WAVE), shm, sm, d_reps, n, d_plan, c);
    }
    hipLaunchKernelGGL(k_rsum_ens, dim3(NPART, n), dim3(256), 0, sm, d_reps, n, d_plan);
    hipLaunchKernelGGL(k_sweep_end_ens, dim3(n), dim3(NSLOT), 0, sm, d_reps, n, d_plan);
    hipLaunchKernelGGL(k_ens_plan, dim3(1), dim3(ENS_MAX), 0, sm, d_reps, n, d_plan, 0, 1);
    ce = hipStreamEndCapture(sm, &graph);
  }
  if (ce == hipSuccess) ce = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (ce == hipSuccess) ce = hipEventRecord(e0->ev0, sm);
  int batch = 1;
  while (ce == hipSuccess) {
    for (int b = 0; b < batch && ce == hipSuccess; b++) ce = hipGraphLaunch(exec, sm);
    if (ce != hipSuccess) break;
    sweeps += batch;
    ce = hipMemcpyAsync(&nrun, d_plan + 8 * (n + 1) + 8, sizeof(int), hipMemcpyDeviceToHost, sm);
    if (ce == hipSuccess) ce = hipStreamSynchronize(sm);
    if (ce != hipSuccess || nrun == 0) break;
    // a sweep queued after every replica has stopped still costs its
    // launches: batches grow with the sweeps run, up to 32
    batch = (int)std::max<int64_t>(1, std::min<int64_t>(sweeps / 4, 32));
  }
  if (ce == hipSuccess) ce = hipEventRecord(e0->ev1, sm);
  // whatever happened, nothing queued may still use the plan or the
  // parameters when they are freed
  const hipError_t cs = hipStreamSynchronize(sm);
  if (ce == hipSuccess) ce = cs;
  if (ce == hipSuccess) ce = hipEventElapsedTime(&ms, e0->ev0, e0->ev1);
  if (exec) hipGraphExecDestroy(exec);
  if (graph) hipGraphDestroy(graph);
  hipFree(d_reps);
  hipFree(d_plan);
  if (ce != hipSuccess) return fail_(e0, MSKG_EDEVICE, "ensemble run: %s", hipGetErrorString(ce));
  int rc_all = MSKG_OK;
  for (int i = 0; i < n; i++) {
    mskg_engine *e = engines[i];
    e->run_ms = ms;
    e->sweeps_done = sweeps;
    RunState st;
    if (int rc = read_state(e, st)) return rc;
    if (events_done) events_done[i] = (int64_t)(st.events - ev0[i]);
    if (kmc_time) kmc_time[i] = st.clock;
    if (stop_reason) stop_reason[i] = st.stop == STOP_AT_LIMIT || !st.stop ? MSKG_STOP_LIMIT : st.stop;
    if (st.err && rc_all == MSKG_OK) {
      rc_all = fail_(e, MSKG_ESTATE, "run stopped in the tile at (%d,%d,%d): %s", st.err_site[0], st.err_site[1],
                     st.err_site[2], err_name(st.err));
      if (e != e0) fail_(e0, rc_all, "ensemble replica %d: %s", i, e->err.c_str());
    }
  }
  return rc_all;
}

int mskg_adjoint(mskg_engine *const *engines, int32_t n, const mskg_event *const *logs, const double *const *dts,
                 const int64_t *n_events, double *S, double *A, double *scalars) {
  return mskg_adjoint_checkpointed(engines, n, logs, dts, n_events, S, A, scalars, 0, nullptr);
}

int mskg_adjoint_checkpointed(mskg_engine *const *engines, int32_t n, const mskg_event *const *logs,
                              const double *const *dts, const int64_t *n_events, double *S, double *A, double *scalars,
                              int64_t every, int32_t *const *hist) {
  if (!engines || n <= 0 || !logs || !dts || !n_events || !S || !A || !scalars) return MSKG_EINVAL;
  if (every < 0 || (every > 0 && !hist) || (every == 0 && hist)) return fail_(engines[0], MSKG_EINVAL, "checkpoints: every %lld", (long long)every);
  mskg_engine *e0 = engines[0];
  if (!e0) return MSKG_EINVAL;
  // every replica's state is checked and its key arrays exist before anything is queued
  std::vector<char> fresh(n);
  std::vector<double> rtot(n);
  std::vector<int64_t> rows(n);
  for (int i = 0; i < n; i++) {
    mskg_engine *e = engines[i];
    if (!e || !e->loaded) return fail_(e0, MSKG_EINVAL, "adjoint: replica %d has no state loaded", i);
    for (int j = 0; j < i; j++)
      if (engines[j] == e) return fail_(e0, MSKG_EINVAL, "adjoint: replica %d appears twice", i);
    if (e->tab.ncov || e->tab.la_mode)
      return fail_(e, MSKG_EINVAL, "the likelihood covers models without coverage species and without LA rates");
    if (n_events[i] < 0 || (n_events[i] > 0 && (!logs[i] || !dts[i]))) return fail_(e, MSKG_EINVAL, "bad log");
    const int64_t np = e->box.npts;
    if (!e->d_key) {
      if (hipMalloc(&e->d_key, np * sizeof(int32_t)) != hipSuccess ||
          hipMalloc(&e->d_key_seen, np * sizeof(int32_t)) != hipSuccess ||
          hipMalloc(&e->d_rate_seen, np * sizeof(float)) != hipSuccess)
        return fail_(e, MSKG_ENOMEM, "adjoint: no device memory for the class keys");
    }
    RunState st;
    if (int rc = read_state(e, st)) return rc;  // also: nothing of e's own stream is still running
    if (st.err) return fail_(e, MSKG_ESTATE, "engine is in an error state: %s", err_name(st.err));
    // the trajectory starts from the engine's rates: a tracked engine's keys
    // are current; otherwise a sweep refreshes the rates and assigns the keys
    fresh[i] = !(e->track && e->keys_valid);
    rtot[i] = st.rtot;
    rows[i] = every > 0 ? n_events[i] / every : 0;
  }
  // One device block for every segment, in regions of one kind each so that
  // the results come back in one copy per kind: the logs and waiting times,
  // the initial class counts (+ the unsupported points) and isolated-rate
  // sums, S, A and the scalars in the caller's layouts, the checkpoints, the
  // check words, the descriptors, the replay's per-class integrals.
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  std::vector<size_t> log_off(n), dt_off(n), hist_off(n);
  size_t lb = 0, db = 0, hb = 0;
  for (int i = 0; i < n; i++) {
    const size_t ne = (size_t)std::max<int64_t>(1, n_events[i]);
    log_off[i] = lb; lb += up(ne * sizeof(mskg_event));
    dt_off[i] = db; db += up(ne * sizeof(double));
    hist_off[i] = hb; hb += (size_t)rows[i] * KEY_N * sizeof(int32_t);
  }
  const size_t cnt_b = up((size_t)n * (KEY_N + 1) * sizeof(int)), iso_b = up((size_t)n * sizeof(double)),
               s_b = up((size_t)n * ADJ_NS * sizeof(double)), a_b = up((size_t)n * KEY_N * sizeof(double)),
               sc_b = up((size_t)n * 4 * sizeof(double)), h_b = up(std::max<size_t>(hb, 1)),
               chk_b = up((size_t)n * 2 * sizeof(int)), d_b = up((size_t)n * sizeof(AdjDesc)),
               acc_b = up((size_t)n * 2 * KEY_N * sizeof(double));
  const size_t o_dt = lb, o_cnt = o_dt + db, o_iso = o_cnt + cnt_b, o_s = o_iso + iso_b, o_a = o_s + s_b,
               o_sc = o_a + a_b, o_h = o_sc + sc_b, o_chk = o_h + h_b, o_d = o_chk + chk_b, o_acc = o_d + d_b,
               total = o_acc + acc_b;
  char *arena = nullptr;
  if (hipMalloc(&arena, total) != hipSuccess) return fail_(e0, MSKG_ENOMEM, "adjoint: no device memory for the logs");
  int *const d_cnt = (int *)(arena + o_cnt);
  double *const d_iso = (double *)(arena + o_iso);
  int *const d_chk = (int *)(arena + o_chk);
  AdjDesc *const d_descs = (AdjDesc *)(arena + o_d);
  std::vector<AdjDesc> descs(n);
  for (int i = 0; i < n; i++) {
    mskg_engine *e = engines[i];
    Lattice L = make_lattice(e);
    L.inlist = nullptr;
    L.key = e->d_key;
    AdjDesc &D = descs[i];
    D.L = L;
    D.key_seen = e->d_key_seen;
    D.rate_seen = e->d_rate_seen;
    D.log = (const mskg_event *)(arena + log_off[i]);
    D.dt = (const double *)(arena + o_dt + dt_off[i]);
    D.n = n_events[i];
    D.cnt0 = d_cnt + (size_t)i * (KEY_N + 1);
    D.iso0 = d_iso + i;
    D.sweep_err = e->d_err;
    D.S = (double *)(arena + o_s) + (size_t)i * ADJ_NS;
    D.A = (double *)(arena + o_a) + (size_t)i * KEY_N;
    D.scal = (double *)(arena + o_sc) + (size_t)i * 4;
    D.st = e->d_st;
    D.hist = rows[i] > 0 ? (int32_t *)(arena + o_h + hist_off[i]) : nullptr;
    D.every = every > 0 ? every : 1;
    D.acc = (double *)(arena + o_acc) + (size_t)i * 2 * KEY_N;
  }
  hipStream_t sm = e0->stream;
  auto done = [&](int rc) {
    hipStreamSynchronize(sm);  // nothing queued may still use the block when it is freed
    hipFree(arena);
    return rc;
  };
  // the copies from host memory first, while the stream is still idle (a
  // copy from pageable memory returns only when done)
  hipError_t ce = hipMemcpyAsync(d_descs, descs.data(), n * sizeof(AdjDesc), hipMemcpyHostToDevice, sm);
  for (int i = 0; i < n && ce == hipSuccess; i++)
    if (n_events[i] > 0) {
      ce = hipMemcpyAsync(arena + log_off[i], logs[i], n_events[i] * sizeof(mskg_event), hipMemcpyHostToDevice, sm);
      if (ce == hipSuccess)
        ce = hipMemcpyAsync(arena + o_dt + dt_off[i], dts[i], n_events[i] * sizeof(double), hipMemcpyHostToDevice, sm);
    }
  if (ce == hipSuccess) ce = hipMemsetAsync(d_cnt, 0, cnt_b, sm);
  if (ce == hipSuccess) ce = hipMemsetAsync(d_iso, 0, iso_b, sm);
  // each segment's initial classes: the keys (a rate sweep if they are not
  // current), their populations, and the keys and rates as first seen
  for (int i = 0; i < n && ce == hipSuccess; i++) {
    mskg_engine *e = engines[i];
    const int64_t np = e->box.npts;
    const int blocks = (int)std::min<int64_t>((np + 255) / 256, 256 * 64);
    ce = hipMemsetAsync(e->d_err, 0, sizeof(int), sm);
    if (ce == hipSuccess && fresh[i]) {
      hipLaunchKernelGGL(MSKG_VARIANT(k_rate_sweep, e->tab.la_mode || e->tab.st_n, e->cov_model), dim3(blocks),
                         dim3(256), 0, sm, descs[i].L, rtot[i], e->d_err, 0);
      ce = hipGetLastError();
    }
    if (ce == hipSuccess) {
      hipLaunchKernelGGL(k_key_hist, dim3(blocks), dim3(256), 0, sm, descs[i].L, d_cnt + (size_t)i * (KEY_N + 1),
                         d_iso + i);
      ce = hipGetLastError();
    }
    if (ce == hipSuccess) ce = hipMemcpyAsync(e->d_key_seen, e->d_key, np * sizeof(int32_t), hipMemcpyDeviceToDevice, sm);
    if (ce == hipSuccess) ce = hipMemcpyAsync(e->d_rate_seen, e->d_rate, np * sizeof(float), hipMemcpyDeviceToDevice, sm);
    e->lists_dirty = true;
  }
  const int cblocks = (n + 255) / 256;
  std::vector<int> chk(2 * (size_t)n);
  if (ce == hipSuccess) {
    hipLaunchKernelGGL(k_adj_check, dim3(cblocks), dim3(256), 0, sm, d_descs, n, d_chk);
    ce = hipGetLastError();
  }
  if (ce == hipSuccess) ce = hipMemcpyAsync(chk.data(), d_chk, 2 * (size_t)n * sizeof(int), hipMemcpyDeviceToHost, sm);
  if (ce == hipSuccess) ce = hipStreamSynchronize(sm);
  if (ce != hipSuccess) return done(fail_(e0, MSKG_EDEVICE, "adjoint setup: %s", hipGetErrorString(ce)));
  for (int i = 0; i < n; i++) {
    if (chk[2 * i]) return done(fail_(engines[i], MSKG_ESTATE, "rate sweep: %s", err_name(chk[2 * i])));
    if (chk[2 * i + 1])
      return done(fail_(engines[i], MSKG_EINVAL, "%d points have rates the likelihood does not cover", chk[2 * i + 1]));
  }
  const size_t shm = (size_t)KEY_N * sizeof(int);
  ce = hipEventRecord(e0->ev0, sm);
  if (ce == hipSuccess) {
    hipLaunchKernelGGL(k_adjoint, dim3(n), dim3(WAVE), shm, sm, d_descs);
    ce = hipGetLastError();
  }
  if (ce == hipSuccess) ce = hipEventRecord(e0->ev1, sm);
  if (ce == hipSuccess) {
    hipLaunchKernelGGL(k_adj_errs, dim3(cblocks), dim3(256), 0, sm, d_descs, n, d_chk);
    ce = hipGetLastError();
  }
  // the results, one copy per kind (the checkpoints through a host buffer)
  std::vector<int32_t> hall(hb / sizeof(int32_t));
  if (ce == hipSuccess) ce = hipMemcpyAsync(S, arena + o_s, (size_t)n * ADJ_NS * sizeof(double), hipMemcpyDeviceToHost, sm);
  if (ce == hipSuccess) ce = hipMemcpyAsync(A, arena + o_a, (size_t)n * KEY_N * sizeof(double), hipMemcpyDeviceToHost, sm);
  if (ce == hipSuccess) ce = hipMemcpyAsync(scalars, arena + o_sc, (size_t)n * 4 * sizeof(double), hipMemcpyDeviceToHost, sm);
  if (ce == hipSuccess) ce = hipMemcpyAsync(chk.data(), d_chk, (size_t)n * sizeof(int), hipMemcpyDeviceToHost, sm);
  if (ce == hipSuccess && hb) ce = hipMemcpyAsync(hall.data(), arena + o_h, hb, hipMemcpyDeviceToHost, sm);
  if (ce == hipSuccess) ce = hipStreamSynchronize(sm);
  if (ce != hipSuccess) return done(fail_(e0, MSKG_EDEVICE, "k_adjoint: %s", hipGetErrorString(ce)));
  float ms = 0;
  hipEventElapsedTime(&ms, e0->ev0, e0->ev1);
  int rc_all = MSKG_OK;
  for (int i = 0; i < n; i++) {
    mskg_engine *e = engines[i];
    e->run_ms = ms;
    if (rows[i] > 0)
      memcpy(hist[i], hall.data() + hist_off[i] / sizeof(int32_t), (size_t)rows[i] * KEY_N * sizeof(int32_t));
    if (chk[i] && rc_all == MSKG_OK)
      rc_all = fail_(e, MSKG_ESTATE, "adjoint replay of replica %d stopped after %lld events: %s", i,
                     (long long)scalars[(size_t)i * 4 + 3], err_name(chk[i]));
  }
  return done(rc_all);
}

int mskg_frame(mskg_engine *e, int32_t *main_rec, int64_t cap_main, int32_t *void_rec, int64_t cap_void,
               int32_t *noepi_rec, int64_t cap_noepi, int64_t counts[5]) {
  if (!e || !e->loaded || !counts) return fail_(e, MSKG_EINVAL, "frame: no state loaded");
  int32_t *const hosts[3] = {main_rec, void_rec, noepi_rec};
  const int64_t caps[3] = {cap_main, cap_void, cap_noepi};
  return frame_lists<FRAME_MAIN>(e, 0, hosts, caps, counts);
}

int mskg_frame_device(mskg_engine *e, int32_t *main_rec, int64_t cap_main, int32_t *void_rec, int64_t cap_void,
                      int32_t *noepi_rec, int64_t cap_noepi, int64_t *counts) {
  if (!e || !e->loaded || !counts) return fail_(e, MSKG_EINVAL, "frame_device: no state loaded or no counts");
  if (cap_main < 0 || cap_void < 0 || cap_noepi < 0) return fail_(e, MSKG_EINVAL, "frame_device: negative capacity");
  const int64_t n = e->box.npts;
  const int nblk = (int)std::min<int64_t>(std::max<int64_t>(1, (n + 65535) / 65536), NOALLOC_FRAME_BLOCKS);
  const int64_t chunk = (n + nblk - 1) / nblk;
  if (chunk > NOALLOC_MAX_CHUNK)
    return fail_(e, MSKG_EINVAL, "frame_device: %lld lattice points exceed the scratch's %d blocks of %lld points",
                 (long long)n, NOALLOC_FRAME_BLOCKS, (long long)NOALLOC_MAX_CHUNK);
  unsigned long long *sc = g_scratch[e->device];
  if (!sc) return fail_(e, MSKG_EDEVICE, "frame_device: the device's scratch is not set up");
  unsigned long long *d_counts = sc, *d_offs = sc + (size_t)nblk * FrameShape<FRAME_MAIN>::NCNT;
  Lattice L = make_lattice(e);
  hipLaunchKernelGGL(k_frame_count<FRAME_MAIN>, dim3(nblk), dim3(FRAME_THREADS), 0, e->stream, L, chunk, 0,
                     d_counts);
  hipLaunchKernelGGL(k_frame_scan<FRAME_MAIN>, dim3(1), dim3(256), 0, e->stream, d_counts, nblk, d_offs, counts);
  hipLaunchKernelGGL(k_frame_write<FRAME_MAIN>, dim3(nblk), dim3(FRAME_THREADS), 0, e->stream, L, chunk, 0, d_offs,
                     (int4 *)main_rec, (int4 *)void_rec, (int4 *)noepi_rec, cap_main, cap_void, cap_noepi);
  hipError_t ce = hipGetLastError();
  if (ce == hipSuccess) ce = hipStreamSynchronize(e->stream);
  if (ce != hipSuccess) return fail_(e, MSKG_EDEVICE, "frame_device: %s", hipGetErrorString(ce));
  return MSKG_OK;
}

int mskg_frame_defects(mskg_engine *e, int32_t *vac_rec, int64_t cap_vac, int32_t *wrong_rec, int64_t cap_wrong,
                       int64_t counts[2]) {
  if (!e || !e->loaded || !counts) return fail_(e, MSKG_EINVAL, "frame: no state loaded");
  // the wrong-occupation analysis runs for two crystal species one of which is C
  const int sic = e->ncrystal == 2 &&
                  (e->cfg.list_crystal[0] == 6 || e->cfg.list_crystal[1] == 6 || e->cfg.list_crystal[2] == 6);
  std::vector<int4> flips;
  if (sic) {
    if (int rc = wrong_flips(e, flips)) return rc;
    if (int rc = set_wrong_marks(e, flips, 1)) return rc;
  }
  int32_t *const hosts[2] = {vac_rec, wrong_rec};
  const int64_t caps[2] = {cap_vac, cap_wrong};
  const int rc = frame_lists<FRAME_DEFECTS>(e, sic, hosts, caps, counts);
  if (sic) {
    const std::string msg = e->err;
    if (int rc2 = set_wrong_marks(e, flips, 0)) return rc2;  // the lattice is left as it was
    e->err = msg;
  }
  return rc;
}

}  // extern "C"
