// mskg_events.h -- per-site rate evaluation and event execution on the
// device-resident lattice.  Each routine follows the reference routine named
// in its comment (paths relative to the reference repository root) and keeps
// its update order, its absolute-coordinate comparisons and its rule for
// which sites get a fresh rate, so that a replayed event log reproduces the
// reference state bit for bit.
//
// The reference keeps three particle lists and a sum tree; here a site's
// neighbour list lives in its own record and its rate/class beside it, so
// AddMC/EraseMC/AddBulk/AddVoid become "set rate/class" or "clear rate".
#pragma once

#include "mskg_lattice.h"

namespace mskg {

enum : int {
  ERR_NONE = 0,
  ERR_COORD = 1,      // N_Si+N_C+N_B+N_wall != Coor (Get_Prob.f:293-298)
  ERR_GEOM = 2,       // a neighbour displacement that does not fit the record
  ERR_NOT_MC = 3,     // replayed event on a site that is not an MC particle
  ERR_COVERAGE = 4,   // coverage index not in ListCov
  ERR_CAPACITY = 5,   // tile candidate list overflow
  ERR_TEMPERATURE = 6, // zero temperature in LA mode
  ERR_TILE_LIST = 7,   // a tile's candidate list overflowed
  ERR_SUMS = 8,        // serial selection: the rate sums disagree with the lists
  ERR_LIQUID = 9,      // negative CountLiquid (laser-annealing build's form)
  ERR_ST_RANGE = 10,   // a temperature outside the interpolation tables (InitSt 'ST')
  ERR_SCHED = 11,      // a tile waited too long for its neighbour tiles (dependency-scheduled sweep)
  ERR_TOP = 12         // an event within 24 planes of the top of the stored planes (stored_len_z)
};

// Per-tile candidate lists (synchronous-sublattice mode): every MC particle
// of a tile is listed once (inlist[p] = 1) as its tile-local offset; entries
// whose class has dropped to 0 are removed when the tile is next processed.
struct Lattice {
  Site *site;
  float *rate;
  uint8_t *cls;
  uint8_t *inlist;        // may be null (no list maintenance)
  int32_t *tl_idx;        // [ntiles][capg] tile-local offsets
  int32_t *tl_n;          // [ntiles] entries used
  int32_t *tl_own;        // [ntiles] entries the tile itself wrote (the rest were appended by others)
  double *tile_r;         // [ntiles] FP64 rate total of the tile's list (isolated atoms excluded)
  int *list_err;          // set on overflow
  int *win;               // [8][4] per colour: tile-layer window lo, hi (read by that colour's
                          // pass), next lo, next hi (collected for the next sweep)
  Box box;
  const Geometry *geo;
  const Tables *tab;
  int tile, ntx, nty, capg;
  int cur_tile;           // the tile whose pass is running: its list lives on chip
  float inv_tile;         // 1 / tile
  int32_t *key;           // may be null: rate-class key of every point (likelihood accounting)
  const long long *cl;    // may be null: CountLiquid(1..3) the laser-annealing build's rates read
  int ztop;               // highest z of an event's site: an event reads at most 12 and writes at
                          // most 6 planes from its site, so with only the bottom planes stored an
                          // event above this would reach past them (INT_MAX: every plane stored)
};

// Rate-class keys for the trajectory likelihood (models without coverage
// species).  A key names the table entry a point's rate comes from:
//   evaporation  (species-1) + 3*(n_Si+n_wall + 4*(n_C + 4*n_B)) + 192*n2
//                (n2 = occupied second-neighbour positions, 0 without that term)
//   adsorption site of coordination c       KEY_ADS + c - 1
//   placeholder deposition (1e-8)           KEY_PH
//   isolated atom (1e5*R, per point)        KEY_ISO
// and -1 for no rate.
enum : int { KEY_EVAP_N = 192 * 13, KEY_ADS = KEY_EVAP_N, KEY_PH = KEY_ADS + 3, KEY_ISO = KEY_PH + 1,
             KEY_N = KEY_ISO + 1, KEY_UNSUPPORTED = -2 };

struct P3 {
  int x, y, z;
  int64_t i;  // -1 outside the box
};

MSKG_HD P3 at(const Lattice &L, int x, int y, int z) {
  P3 p;
  p.x = x; p.y = y; p.z = z;
  p.i = pidx_checked(L.box, x, y, z);
  return p;
}
MSKG_HD P3 entry(const Lattice &L, const NL &l, int i) { return at(L, l.x[i], l.y[i], l.z[i]); }
MSKG_HD Site rd(const Lattice &L, int64_t i) {
  if (i < 0) {
    Site z;
    for (int a = 0; a < 4; a++) z.d[a][0] = z.d[a][1] = z.d[a][2] = 0;
    z.kc = 0; z.valid = 0; z.temp = 0;
    return z;
  }
  return L.site[i];
}
MSKG_HD void wr(const Lattice &L, int64_t i, const Site &s) {
  if (i >= 0) L.site[i] = s;
}
MSKG_HD NL list_of(const Lattice &L, const P3 &p, const Site &s) { return site_list(L.box, s, p.x, p.y, p.z); }
MSKG_HD bool same(const NL &l, int j, const P3 &a) { return l.x[j] == a.x && l.y[j] == a.y && l.z[j] == a.z; }
// The reference's Recipr loops: does list l contain the position a?
MSKG_HD bool recip(const NL &l, const P3 &a) {
  for (int j = 0; j < 4; j++)
    if (same(l, j, a)) return true;
  return false;
}

// x / tile for 0 <= x < 2^20: the FP32 quotient is off by at most one.
MSKG_HD int tdiv(const Lattice &L, int x) {
  int q = (int)((float)x * L.inv_tile);
  const int r = x - q * L.tile;
  if (r < 0) q--;
  else if (r >= L.tile) q++;
  return q;
}
MSKG_HD int tile_of(const Lattice &L, const P3 &p) {
  return (tdiv(L, p.z) * L.nty + tdiv(L, p.y)) * L.ntx + tdiv(L, p.x);
}
// A point's offset inside its tile, packed as (z << 14) | ((y / 2) << 7) |
// (x / 2) in tile-local coordinates (tile edge <= 254), so that decoding
// needs no division; offsets sort like the points' (z, y, x) order.
MSKG_HD int tile_pack(int zl, int yl, int xl) { return (zl << 14) | ((yl >> 1) << 7) | (xl >> 1); }
MSKG_HD int tile_offset(const Lattice &L, const P3 &p) {
  const int tx = tdiv(L, p.x), ty = tdiv(L, p.y), tz = tdiv(L, p.z);
  return tile_pack(p.z - tz * L.tile, p.y - ty * L.tile, p.x - tx * L.tile);
}
MSKG_HD int atomic_add_i32(int32_t *a, int v) {
#ifdef __HIP_DEVICE_COMPILE__
  return atomicAdd(a, v);
#else
  int o = *a;
  *a += v;
  return o;
#endif
}
MSKG_HD void atomic_min_i32(int *a, int v) {
#ifdef __HIP_DEVICE_COMPILE__
  atomicMin(a, v);
#else
  if (v < *a) *a = v;
#endif
}
MSKG_HD void atomic_max_i32(int *a, int v) {
#ifdef __HIP_DEVICE_COMPILE__
  atomicMax(a, v);
#else
  if (v > *a) *a = v;
#endif
}
// Does the record sq of point q list position p?  The reference's Recipr
// loops (list_of(q) contains p), evaluated on the stored displacements: a
// valid entry holds p - q wrapped as set_list wraps it (no displacement is
// near half a box edge), an unused entry stands for the origin.
MSKG_HD bool lists_point(const Box &b, const P3 &q, const Site &sq, const P3 &p) {
  int wx = p.x - q.x, wy = p.y - q.y;
  const int wz = p.z - q.z;
  if (wx > b.lx / 2) wx -= b.lx;
  if (wx < -b.lx / 2) wx += b.lx;
  if (wy > b.ly / 2) wy -= b.ly;
  if (wy < -b.ly / 2) wy += b.ly;
  const bool origin = p.x == 0 && p.y == 0 && p.z == 0;
  bool hit = false;
  for (int j = 0; j < 4; j++) {
    if ((sq.valid >> j) & 1) hit |= sq.d[j][0] == wx && sq.d[j][1] == wy && sq.d[j][2] == wz;
    else hit |= origin;
  }
  return hit;
}
MSKG_HD int tile_colour(int tx, int ty, int tz) { return (tx & 1) | ((ty & 1) << 1) | ((tz & 1) << 2); }
// A tile layer that holds candidates joins the window of layers its colour's
// pass visits.  A colour's window is only widened by passes of the other
// colours (same-colour tiles never touch each other), so it is constant
// while its own pass runs.
MSKG_HD void widen_window(const Lattice &L, int colour, int tz) {
  if (!L.win) return;
  int *w = L.win + 4 * colour;
  // read first: the window rarely changes, and atomics on one word serialise
  if (tz < w[0]) atomic_min_i32(&w[0], tz);
  if (tz > w[1]) atomic_max_i32(&w[1], tz);
  if (tz < w[2]) atomic_min_i32(&w[2], tz);
  if (tz > w[3]) atomic_max_i32(&w[3], tz);
}
// A site that becomes an MC particle joins its tile's list (once).
MSKG_HD void enlist(const Lattice &L, const P3 &p) {
  if (!L.inlist) return;
  const int tx = tdiv(L, p.x), ty = tdiv(L, p.y), tz = tdiv(L, p.z);
  const int t = (tz * L.nty + ty) * L.ntx + tx;
  if (t == L.cur_tile) return;  // the running tile keeps its list on chip and flags it on write-back
  if (L.inlist[p.i]) return;
  L.inlist[p.i] = 1;
  widen_window(L, tile_colour(tx, ty, tz), tz);
  const int slot = atomic_add_i32(&L.tl_n[t], 1);
  if (slot < L.capg) L.tl_idx[(int64_t)t * L.capg + slot] = tile_pack(p.z - tz * L.tile, p.y - ty * L.tile, p.x - tx * L.tile);
  else *L.list_err = ERR_TILE_LIST;
}
MSKG_HD void set_rate(const Lattice &L, const P3 &p, int cls, double rate, int key = -1) {
  if (p.i < 0) return;
  L.rate[p.i] = cls ? (float)rate : 0.0f;
  L.cls[p.i] = (uint8_t)cls;
  if (L.key) L.key[p.i] = cls ? key : -1;
  if (cls) enlist(L, p);
}

// ---- rates ---------------------------------------------------------------
MSKG_HD double PEv(const Tables &t, int sp, int a, int b, int c, int d, int e, int f) {
  return t.pe[(sp - 1) + 3 * (a + 4 * (b + 4 * (c + 4 * (d + 4 * (e + 4 * f)))))];
}
// The reference's kb (Get_Prob.f:70).
#define MSKG_KB 0.00008617330350
// LA = false compiles the laser-annealing branches out (tables with la_mode 0).
template <bool LA = true>
MSKG_HD double pd_of(const Tables &t, int sp, int coor, int temperature, const long long *cl = nullptr) {
  double e = t.pd[(sp - 1) + 3 * (coor - 1)];
  if (!LA || !t.la_mode) return e;
  if (t.la_variant && cl) {
    // LaserAnnealing Get_Prob.f:148-151, evaluated left to right: 0.5d0 *
    // P0(k) * FLOAT(CountLiquid(k)) / SUM(CountLiquid) * EXP(..) * (1 + ERF(..)),
    // FLOAT the REAL(4) conversion.  With no liquid left (the reference
    // divides by zero) solidification stops.
    const long long sum = cl[0] + cl[1] + cl[2];
    if (sum <= 0 || sp > t.ncrystal) return 0.0;  // PD(k,:) is set for k <= NCrystal only
    return 0.5 * t.la_p0[sp - 1] * (double)(float)(int)cl[sp - 1] / (double)sum *
           exp(-e / (MSKG_KB * t.la_tm[sp - 1])) * (1.0 + erf((temperature - t.la_tflex[sp - 1]) / t.la_sigma[sp - 1])) *
           (e != 0.0 ? 1.0 : 0.0);
  }
  return 0.5 * t.p0 * exp(-e / (MSKG_KB * t.tm)) * (1.0 + erf((temperature - t.tflex) / t.sigma)) *
         (e != 0.0 ? 1.0 : 0.0);
}
template <bool LA = true>
MSKG_HD double pe_of(const Tables &t, int sp, int a, int b, int c, int d, int e, int f, int temperature) {
  double v = PEv(t, sp, a, b, c, d, e, f);
  if (!LA || !t.la_mode) return v;
  const double p0 = t.la_variant ? t.la_p0[sp - 1] : t.p0;  // LaserAnnealing Get_Prob.f:144-146
  return p0 * exp(-v / (MSKG_KB * temperature)) * (v != 0.0 ? 1.0 : 0.0);
}
// The laser-annealing build's SiGe rule (its main.f:676-681): a deposition of
// a species whose liquid is used up (CountCrystal >= InitCountCrystal, i.e.
// CountLiquid <= 0) deposits the other species.
template <bool LA = true>
MSKG_HD int la_swap(const Lattice &L, int itrans) {
  if (LA && L.cl && L.tab->la_mode && L.tab->la_variant && L.tab->la_swap && (itrans == 1 || itrans == 2) &&
      L.cl[itrans - 1] <= 0)
    return itrans % 2 + 1;
  return itrans;
}

// InterpProbFromField.f:28-54: the tables at temperature T are reference
// table iu when T equals its temperature (the first such), else the blend
// of tables il = iu - 1 and iu, iu the first of the smallest positive
// (reference temperature - T), with integer weights wl = |T_iu - T| on il
// and wu = |T_il - T| on iu.  False when T lies outside the references (the
// reference indexes out of its arrays there).
struct STW {
  int il, iu, wl, wu, exact;
};
MSKG_HD bool st_weights(const Tables &t, int temperature, STW &w) {
  int best = -1;
  w.exact = 0;
  for (int k = 0; k < t.st_n; k++) {
    const int diff = t.st_temp[k] - temperature;
    if (diff == 0) {
      w.exact = 1;
      w.il = w.iu = k;
      w.wl = w.wu = 0;
      return true;
    }
    if (diff > 0 && (best < 0 || diff < t.st_temp[best] - temperature)) best = k;
  }
  if (best <= 0) return false;
  w.iu = best;
  w.il = best - 1;
  const int dl = t.st_temp[w.iu] - temperature, du = t.st_temp[w.il] - temperature;
  w.wl = dl < 0 ? -dl : dl;
  w.wu = du < 0 ? -du : du;
  return true;
}
// One interpolated table entry: (wl * ref(il) + wu * ref(iu)) / (wl + wu),
// the integer weights converted as Fortran does.
MSKG_HD double st_entry(const double *ref, int stride, int idx, const STW &w) {
  if (w.exact) return ref[(int64_t)w.iu * stride + idx];
  return ((double)w.wl * ref[(int64_t)w.il * stride + idx] + (double)w.wu * ref[(int64_t)w.iu * stride + idx]) /
         (double)(w.wl + w.wu);
}
MSKG_HD int pe_flat(int sp, int a, int b, int c, int d, int e, int f) {
  return (sp - 1) + 3 * (a + 4 * (b + 4 * (c + 4 * (d + 4 * (e + 4 * f)))));
}

// Get_Prob, Get_Prob.f:54-331, evaluated for point p whose record is sp,
// with the coordination and neighbour list passed explicitly (Evaporation.f:121
// passes a stale list copy).  rtot is the total rate R used by the
// isolated-atom rule.  No random number is drawn: the deposition species /
// absorbed coverage is drawn when the event executes (class CLS_ADSORB).
// Records that are newer than HBM (updated by this event on another lane):
// reads of these points take the given record instead.  known_i /
// known_coor: a point whose coordination the caller already holds (the
// event's site during the neighbour updates), read by the coordination-1
// rule of an empty site without a gather.
struct Ovr {
  int n;
  int64_t idx[5];
  Site rec[5];
};
MSKG_HD Site rd_o(const Lattice &L, int64_t i, const Ovr *o) {
  if (o && i >= 0)
    for (int k = 0; k < o->n; k++)
      if (o->idx[k] == i) return o->rec[k];
  return rd(L, i);
}

// The 12 second-neighbour positions of a diamond super-lattice site (the
// fcc shell: (+-6,+-6,0) and permutations), for the optional
// second-neighbour energy term; no reference counterpart.
MSKG_HD P3 second_pos(const Lattice &L, const P3 &p, int k) {
  const int a = (k & 2) ? -6 : 6, b = (k & 1) ? -6 : 6, g = k >> 2;
  int dx = g == 2 ? 0 : a, dy = g == 0 ? b : (g == 1 ? 0 : a), dz = g == 0 ? 0 : b;
  int x = p.x + dx, y = p.y + dy;
  if (x >= L.box.lx) x -= L.box.lx;
  if (x < 0) x += L.box.lx;
  if (y >= L.box.ly) y -= L.box.ly;
  if (y < 0) y += L.box.ly;
  return at(L, x, y, p.z + dz);
}
// Index step from a point to its second-neighbour position k when neither
// wraps nor leaves the box: (dz * plane) + (dy / 2) * hx + dx / 2 with the
// vectors of second_pos.
MSKG_HD int64_t second_step(const Box &b, int k) {
  const int a = (k & 2) ? -6 : 6, c = (k & 1) ? -6 : 6, g = k >> 2;
  const int dx = g == 2 ? 0 : a, dy = g == 0 ? c : (g == 1 ? 0 : a), dz = g == 0 ? 0 : c;
  return (int64_t)dz * b.plane + (int64_t)(dy / 2) * b.hx + dx / 2;
}
// COV = false compiles the coverage-species branches out (tables without
// coverage species and without absorption or desorption rates).  STEP: the
// second-shell gather of interior points takes fixed index steps (fewer
// instructions, more registers: for the event kernels, not the sweep).
template <bool LA = true, bool COV = true, bool STEP = false>
MSKG_HD void get_prob(const Lattice &L, double rtot, const P3 &p, const Site &sp, int coor, const NL &nn, int &cls,
                      double &rate, int &err, const Ovr *ovr = nullptr, int *key = nullptr, int64_t known_i = -1,
                      int known_coor = 0) {
  const Tables &t = *L.tab;
  int kdummy;
  if (!key) key = &kdummy;
  *key = -1;
  int kind = s_kind(sp);
  int temperature = 0;
  STW w;
  const bool st = LA && t.st_n > 0;  // Get_Prob.f:77-90: tables at the site's temperature
  if (LA && (t.la_mode || st)) {
    temperature = (int)sp.temp;
    if (temperature == 0) { err = ERR_TEMPERATURE; cls = CLS_NONE; rate = 0.0; return; }
    if (st && !st_weights(t, temperature, w)) { err = ERR_ST_RANGE; cls = CLS_NONE; rate = 0.0; return; }
  }
  if (!COV && kind >= KIND_COV0) { err = ERR_COVERAGE; cls = CLS_NONE; rate = 0.0; return; }
  int slot = COV ? kind_cov_slot(kind) : 0;
  if (slot) {  // desorption, :93-123
    if (slot > t.ncov) { err = ERR_COVERAGE; cls = CLS_NONE; rate = 0.0; return; }
    if (coor <= 3) {
      cls = CLS_DESORB;
      rate = st ? st_entry(t.st_pdes, 12, (slot - 1) + 3 * coor, w) : t.pdes[(slot - 1) + 3 * coor];
      *key = KEY_UNSUPPORTED;
    }
    else { cls = CLS_NONE; rate = 0.0; }
    return;
  }
  if (kind == KIND_EMPTY) {  // :155-214
    if (LA && t.la_mode && t.la_variant && L.cl && (L.cl[0] < 0 || L.cl[1] < 0 || L.cl[2] < 0)) {
      // a negative solidification rate, which a sum tree cannot hold (the
      // reference's selection goes wrong there, or UpdateTree.f:37-41 stops
      // on a rate that cancels to zero): the engine stops with an error
      err = ERR_LIQUID; cls = CLS_NONE; rate = 0.0; return;
    }
    if (coor >= 1 && coor <= 3) {
      bool ok = true;
      if (coor == 1) {
        const int64_t i0 = entry(L, nn, 0).i;
        ok = (i0 >= 0 && i0 == known_i ? known_coor : s_coor(rd_o(L, i0, ovr))) >= 2;
      }
      if (ok) {
        double pdep, pabs;
        if (st) {
          pdep = st_entry(t.st_pd, 9, 3 * (coor - 1), w) + st_entry(t.st_pd, 9, 1 + 3 * (coor - 1), w) +
                 st_entry(t.st_pd, 9, 2 + 3 * (coor - 1), w);
          pabs = COV ? st_entry(t.st_pabs, 12, 3 * coor, w) + st_entry(t.st_pabs, 12, 1 + 3 * coor, w) +
                           st_entry(t.st_pabs, 12, 2 + 3 * coor, w)
                     : 0.0;
        } else {
          pdep = pd_of<LA>(t, 1, coor, temperature, L.cl) + pd_of<LA>(t, 2, coor, temperature, L.cl) +
                 pd_of<LA>(t, 3, coor, temperature, L.cl);
          pabs = COV ? t.pabs[0 + 3 * coor] + t.pabs[1 + 3 * coor] + t.pabs[2 + 3 * coor] : 0.0;
        }
        cls = CLS_ADSORB;
        rate = pdep + pabs;
        *key = (t.ncov || st) ? KEY_UNSUPPORTED : KEY_ADS + coor - 1;
      } else {
        cls = CLS_PLACEHOLDER;
        rate = (double)1.e-8f;  // REAL(4) literal, Get_Prob.f:183
        *key = KEY_PH;
      }
    } else {
      cls = CLS_NONE;  // Coor 0 leaves the reference's outputs unassigned; Coor 4 is a vacancy
      rate = 0.0;
    }
    return;
  }
  // evaporation, :215-324
  if (p.z < t.frozen_below_z) {  // frozen substrate (engine option, no reference counterpart)
    cls = CLS_NONE;
    rate = 0.0;
    return;
  }
  int n_si = 0, n_c = 0, n_b = 0, n_wall = 0, covnn[3] = {0, 0, 0};
  P3 qs[4];
  Site sqs[4];
  uint8_t kc2[12];
  for (int i = 0; i < 4; i++) qs[i] = entry(L, nn, i);
  for (int i = 0; i < 4; i++) sqs[i] = rd_o(L, qs[i].i, ovr);  // four independent gathers in flight
  if (t.use_f2) {  // ... and the twelve second-shell kinds with them
    // at least 6 units from every face the shell neither wraps nor leaves
    // the box, and its indices are p's plus fixed steps
    const bool inner = STEP && p.x >= 6 && p.x < L.box.lx - 6 && p.y >= 6 && p.y < L.box.ly - 6 && p.z >= 6 &&
                       p.z <= L.box.lz - 6;
    for (int k = 0; k < 12; k++) {
      const int64_t i2 = inner ? p.i + second_step(L.box, k) : second_pos(L, p, k).i;
      kc2[k] = i2 < 0 ? (uint8_t)0 : (ovr ? rd_o(L, i2, ovr).kc : L.site[i2].kc);
    }
  }
  for (int i = 0; i < 4; i++) {
    const P3 &q = qs[i];
    const Site &sq = sqs[i];
    int kq = s_kind(sq);
    if (kind_occupied(kq)) {
      if (kq != KIND_WALL) {
        if (lists_point(L.box, q, sq, p)) {
          if (kq == 1) n_si++;
          if (kq == 2) n_c++;
          if (kq == 3) n_b++;
        }
      } else {
        n_wall++;
      }
    } else if (COV && kind_cov_slot(kq)) {
      if (lists_point(L.box, q, sq, p)) {
        int s = kind_cov_slot(kq);
        if (s > t.ncov) { err = ERR_COVERAGE; cls = CLS_NONE; rate = 0.0; return; }
        covnn[s - 1]++;
      }
    }
  }
  int tot = n_si + n_c + n_b + n_wall;
  if (tot != coor) { err = ERR_COORD; cls = CLS_NONE; rate = 0.0; return; }
  if (tot > 3) {
    cls = CLS_NONE;
    rate = 0.0;
  } else if (tot == 0) {
    cls = CLS_ISOLATED;
    rate = (double)100000.0f * rtot;
    *key = KEY_ISO;
  } else {
    int c1 = t.ncov >= 1 ? covnn[0] : 0, c2 = t.ncov >= 2 ? covnn[1] : 0, c3 = t.ncov >= 3 ? covnn[2] : 0;
    cls = CLS_EVAP;
    rate = st ? st_entry(t.st_pe, 3 * 4096, pe_flat(kind, n_si + n_wall, n_c, n_b, c1, c2, c3), w)
              : pe_of<LA>(t, kind, n_si + n_wall, n_c, n_b, c1, c2, c3, temperature);
    int n2 = 0;
    if (t.use_f2) {
      for (int k = 0; k < 12; k++) {
        const int kq = kc2[k] & 7;
        if (kind_occupied(kq) && kq != KIND_WALL) n2++;
      }
      rate *= t.f2[n2];
    }
    *key = (t.ncov || t.la_mode || st) ? KEY_UNSUPPORTED
                                       : (kind - 1) + 3 * (n_si + n_wall + 4 * (n_c + 4 * n_b)) + 192 * n2;
  }
}

// Where a stochastic choice of the reference comes from: the log (replay)
// or the engine's RNG.
struct Choice {
  int arm;       // 0/1 forced arm-chair choice, -1 draw
  double u_arm;  // uniform used when the arm/zig coordination sums tie
};

constexpr int MAX_TOUCHED = 17;
struct EvResult {
  int arm;        // stacking choice made (-1 none)
  int err;
  int species;    // deposited or evaporated species (1..3), desorbed coverage slot (1..3), 0 otherwise
};

MSKG_HD bool put_list(const Lattice &L, Site &s, const P3 &p, const NL &l, EvResult &res) {
  if (!set_list(L.box, s, p.x, p.y, p.z, l)) { res.err = ERR_GEOM; return false; }
  return true;
}

// ---- event execution -------------------------------------------------------
// The whole workgroup (one 64-wide wave) executes an event together: lane 0
// updates the site itself, lanes 0..3 update its four neighbours at once.
// The reference visits the neighbours one after the other, but each
// neighbour's update writes only that neighbour's record and reads only
// records the other neighbour updates do not write (two neighbours of a
// site are never neighbours of each other), so the concurrent updates give
// the sequential result.  When two list entries resolve to the same point
// the neighbours are visited in order by lane 0 instead.
struct EvScratch {
  NL nn;            // the site's neighbour list as the update loop uses it
  P3 pts[4];
  int coor_old;
  int inc[4];       // per-neighbour change of the site's own coordination
  int err[4];
  int unfixed[4];   // evaporation: neighbour i was unfixed (Coor 2 -> 1)
  int found[4];     //   ... and found its reciprocal partner (NBuff) at found_pos[i]
  int found_pos[4][3];
  int seq;          // visit the neighbours in order on lane 0
  int old_kind;     // evaporation: the species that left the site
  int nvisit;       // neighbours the event visits (desorption at coordination 1: one)
  int abort;
  P3 sec[12];       // second-neighbour refresh: the point refreshed by lane k (i = -1: none)
  int err2[12];
  // every point whose rate may have changed: the site, its neighbours and
  // (second-neighbour term) the evaporating atoms on its second-neighbour
  // positions; complete after execute() returns
  P3 touched[MAX_TOUCHED];
  int ntouched;
  // the new rate and class of touched point q as stored in HBM, when tset[q]
  // is 1; 0: the event gave it no new rate; 2: read it from HBM (a list with
  // repeated entries, whose points take several rates in turn)
  float trate[MAX_TOUCHED];
  uint8_t tcls[MAX_TOUCHED];
  uint8_t tset[MAX_TOUCHED];
  float secrate[12];  // the second-shell refresh of lane k + 1
  uint8_t seccls[12];
};

// The value set_rate stores, kept on chip for touched point q.
__device__ inline void note_rate(EvScratch &sc, int q, int cls, double rate) {
  sc.trate[q] = cls ? (float)rate : 0.0f;
  sc.tcls[q] = (uint8_t)cls;
  sc.tset[q] = 1;
}

#define MSKG_SYNC() __syncthreads()

__device__ inline void prefetch4(const Lattice &L, const NL &nn, P3 q[4], Site sq[4]) {
  for (int i = 0; i < 4; i++) q[i] = entry(L, nn, i);
  for (int i = 0; i < 4; i++) sq[i] = rd(L, q[i].i);
}

__device__ inline void resolve_neighbours(const Lattice &L, EvScratch &sc, int n) {
  int seq = 0;
  for (int i = 0; i < 4; i++) sc.pts[i] = entry(L, sc.nn, i);
  for (int i = 0; i < n; i++)
    for (int j = i + 1; j < n; j++)
      if (sc.pts[i].i >= 0 && sc.pts[i].i == sc.pts[j].i) seq = 1;
  sc.seq = seq;
  for (int i = 0; i < 4; i++) {
    sc.inc[i] = 0;
    sc.err[i] = 0;
    sc.unfixed[i] = 0;
    sc.found[i] = 0;
  }
}

__device__ inline void touched_list(const P3 &s, EvScratch &sc, int n) {
  sc.ntouched = 0;
  sc.touched[sc.ntouched++] = s;
  for (int i = 0; i < n; i++) sc.touched[sc.ntouched++] = sc.pts[i];
  for (int q = 0; q <= n; q++) sc.tset[q] = 0;
}

// Second-neighbour term: after a change of occupancy at s, the evaporating
// atom (coordination 1..3) on second-neighbour position k of s, if any,
// gets a fresh rate from the final state of the event.  It runs beside the
// site's own refresh: its evaluation reads the site only for its kind (and,
// for the rare non-bond list entry, its list), which are final by then.
struct FinalRefresh {
  bool active;
  P3 p;
  Site sp;
  int coor;
  NL nn;
};
__device__ inline void second_setup(const Lattice &L, const P3 &s, int k, EvScratch &sc, FinalRefresh &f) {
  const P3 q = second_pos(L, s, k);
  sc.sec[k].i = -1;
  if (q.i < 0) return;
  const Site sq = rd(L, q.i);
  const int kq = s_kind(sq), cq = s_coor(sq);
  if (!kind_occupied(kq) || kq == KIND_WALL || cq < 1 || cq > 3) return;
  sc.sec[k] = q;
  f.active = true;
  f.p = q;
  f.sp = sq;
  f.coor = cq;
  f.nn = list_of(L, q, sq);
}

// A rate evaluation the neighbour phase asks for: the point, its record
// (already written back unless list_after) and the coordination and list
// the evaluation uses.  list_after: the record is written after the
// evaluation, with the list stored only when the point gets a rate (AddMC).
struct RefreshReq {
  bool need, list_after;
  P3 p;
  Site sp;
  int coor;
  NL nn;
};

// deposition.f:78-235, the update of neighbour i of the new atom at s.
__device__ void dep_prepare(const Lattice &L, const P3 &s, int i, EvScratch &sc, int &err, RefreshReq &q) {
  const Geometry &g = *L.geo;
  const int lx = L.box.lx, ly = L.box.ly;
  const P3 p = sc.pts[i];
  Site sp = rd(L, p.i);
  int kind = s_kind(sp), coornn = s_coor(sp);
  NL nnn = list_of(L, p, sp);
  bool need = false, list_after = false;
  if (!kind_occupied(kind)) {
    if (coornn == 0) {
      coornn = 1;
      s_set_coor(sp, coornn);
      for (int j = 0; j < 4; j++) nnn.x[j] = nnn.y[j] = nnn.z[j] = 0;
      nnn.x[0] = s.x; nnn.y[0] = s.y; nnn.z[0] = s.z;
      need = list_after = true;
    } else if (coornn == 1) {
      int dx = nnn.x[0] - s.x, dy = nnn.y[0] - s.y, dz = nnn.z[0] - s.z;
      dx = dx < 0 ? -dx : dx;
      dx = dx < lx - dx ? dx : lx - dx;
      dy = dy < 0 ? -dy : dy;
      dy = dy < ly - dy ? dy : ly - dy;
      if (dx * dx + dy * dy + dz * dz == 72) {
        coornn = 2;
        nnn.x[1] = s.x; nnn.y[1] = s.y; nnn.z[1] = s.z;
        int a1[3] = {nnn.x[0], nnn.y[0], nnn.z[0]}, a2[3] = {s.x, s.y, s.z}, pc[3] = {p.x, p.y, p.z}, ns[2][3];
        find_nn(g, lx, ly, a1, a2, pc, ns);
        P3 qq[2];
        Site sq[2];
        for (int j = 2; j < 4; j++) {
          nnn.x[j] = ns[j - 2][0]; nnn.y[j] = ns[j - 2][1]; nnn.z[j] = ns[j - 2][2];
          qq[j - 2] = entry(L, nnn, j);
        }
        sq[0] = rd(L, qq[0].i);
        sq[1] = rd(L, qq[1].i);
        for (int j = 0; j < 2; j++) {
          int kq = s_kind(sq[j]);
          if (kind_occupied(kq) && s_coor(sq[j]) <= 3 && kq != KIND_WALL) {
            if (lists_point(L.box, qq[j], sq[j], p)) coornn++;
          }
        }
        s_set_coor(sp, coornn);
        if (!set_list(L.box, sp, p.x, p.y, p.z, nnn)) { err = ERR_GEOM; return; }
        wr(L, p.i, sp);
        need = true;
      }
    } else if (coornn == 2 || coornn == 3) {
      if (recip(nnn, s)) {
        coornn++;
        s_set_coor(sp, coornn);
        wr(L, p.i, sp);
        need = true;
      }
    }
  } else if (kind != KIND_WALL) {
    if (coornn <= 3 && recip(nnn, s)) {
      coornn++;
      s_set_coor(sp, coornn);
      wr(L, p.i, sp);
      need = true;
      if (sc.coor_old == 1 && i >= 1) sc.inc[i] = 1;
    }
  } else {
    sc.inc[i] = 1;  // wall neighbour
  }
  q.need = need;
  q.list_after = list_after;
  q.p = p;
  q.sp = sp;
  q.coor = coornn;
  q.nn = nnn;
}

// The neighbours lane `lane` visits: neighbour `lane` (lanes 0..n-1), or all
// n in order on lane 0 when two list entries are the same point.
__device__ inline void visit_range(const EvScratch &sc, int lane, int n, int &i0, int &i1) {
  i0 = lane;
  i1 = lane < n ? lane + 1 : 0;
  if (sc.seq) {
    i0 = 0;
    i1 = lane == 0 ? n : 0;
  }
}

// Deposition, mcevent/deposition.f:24-74: the site takes the species and,
// at coordination 1, its stacking (FIND_SN + IF_Arm_Choice).  Lane 0.
__device__ void dep_site(const Lattice &L, const P3 &s, int species, const Choice &ch, EvResult &res,
                         EvScratch &sc) {
  const Geometry &g = *L.geo;
  Site ss = rd(L, s.i);
  int coor = s_coor(ss);
  sc.coor_old = coor;
  s_set_kind(ss, species);
  wr(L, s.i, ss);
  res.arm = -1;
  NL nn_site = list_of(L, s, ss);
  if (coor == 1) {
    P3 c = entry(L, nn_site, 0);
    Site scr = rd(L, c.i);
    NL lc = list_of(L, c, scr);
    int sn[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, j = 0;
    for (int i = 0; i < 4; i++)
      if (!same(lc, i, s)) {
        if (j < 3) { sn[j][0] = lc.x[i]; sn[j][1] = lc.y[i]; sn[j][2] = lc.z[i]; }
        j++;
      }
    int site[3] = {s.x, s.y, s.z}, cc[3] = {c.x, c.y, c.z}, zig[3][3], arm[3][3];
    find_sn(g, L.box.lx, L.box.ly, site, cc, sn, zig, arm);
    Site sa[3], sz[3];
    for (int k = 0; k < 3; k++) {
      sa[k] = rd(L, at(L, arm[k][0], arm[k][1], arm[k][2]).i);
      sz[k] = rd(L, at(L, zig[k][0], zig[k][1], zig[k][2]).i);
    }
    int carm = 0, czig = 0;  // IF_Arm_Choice.f:35-54
    for (int k = 0; k < 3; k++) {
      carm += s_coor(sa[k]);
      czig += s_coor(sz[k]);
    }
    int is_arm;
    if (ch.arm >= 0) is_arm = ch.arm;
    else if (czig > carm) is_arm = 0;
    else if (czig == carm) is_arm = (ch.u_arm <= L.tab->ptrans_zig) ? 0 : 1;
    else is_arm = 1;
    res.arm = is_arm;
    for (int k = 0; k < 3; k++) {
      nn_site.x[k + 1] = is_arm ? arm[k][0] : zig[k][0];
      nn_site.y[k + 1] = is_arm ? arm[k][1] : zig[k][1];
      nn_site.z[k + 1] = is_arm ? arm[k][2] : zig[k][2];
    }
    if (!set_list(L.box, ss, s.x, s.y, s.z, nn_site)) sc.abort = ERR_GEOM;
    else wr(L, s.i, ss);
  }
  sc.nn = nn_site;
  resolve_neighbours(L, sc, 4);
  touched_list(s, sc, 4);
}

// The scan of Evaporation.f:82-128 / 225-259 that unfixes a list when a site
// drops to coordination 1: the last occupied reciprocal neighbour (NBuff).
__device__ bool nbuff_of(const Lattice &L, const P3 &p, const NL &old, int pos[3]) {
  P3 q[4];
  Site sq[4];
  prefetch4(L, old, q, sq);
  bool f = false;
  for (int j = 0; j < 4; j++) {
    int kq = s_kind(sq[j]);
    if (kind_occupied(kq) && s_coor(sq[j]) <= 3 && kq != KIND_WALL) {
      if (lists_point(L.box, q[j], sq[j], p)) {
        pos[0] = old.x[j]; pos[1] = old.y[j]; pos[2] = old.z[j];
        f = true;
      }
    }
  }
  return f;
}
__device__ inline bool one_entry_list(const Lattice &L, Site &sp, const P3 &p, const int pos[3]) {
  NL nl;
  for (int j = 0; j < 4; j++) nl.x[j] = nl.y[j] = nl.z[j] = 0;
  nl.x[0] = pos[0]; nl.y[0] = pos[1]; nl.z[0] = pos[2];
  return set_list(L.box, sp, p.x, p.y, p.z, nl);
}

// Evaporation.f:36-208, the update of neighbour i of the atom leaving s.
// NBuff is resolved afterwards (evap_nbuff_one) because the reference
// carries it from one neighbour to the next.  A neighbour that keeps a rate
// drops one coordination and asks for one rate evaluation.
__device__ void evap_prepare(const Lattice &L, const P3 &s, int i, EvScratch &sc, int &err, RefreshReq &q) {
  const P3 p = sc.pts[i];
  Site sp = rd(L, p.i);
  const int kind = s_kind(sp), coornn = s_coor(sp);
  const NL nnn = list_of(L, p, sp);
  if (kind == KIND_WALL) {
    sc.inc[i] = -1;  // wall neighbour
    return;
  }
  const bool occ = kind_occupied(kind);
  if (!(occ || (coornn >= 1 && coornn <= 4)) || !recip(nnn, s)) return;
  if (!occ && coornn == 1) {  // EraseMC + LattCoo=0, keeping the temperature bits
    Site z = rd(L, -1);
    z.temp = sp.temp;
    wr(L, p.i, z);
    set_rate(L, p, CLS_NONE, 0.0);
    note_rate(sc, 1 + i, CLS_NONE, 0.0);
    return;
  }
  if (!occ && coornn == 2) {  // unfixed: the list keeps the reciprocal partner (NBuff)
    int pos[3] = {0, 0, 0};
    sc.unfixed[i] = 1;
    sc.found[i] = nbuff_of(L, p, nnn, pos) ? 1 : 0;
    for (int k = 0; k < 3; k++) sc.found_pos[i][k] = pos[k];
    if (!one_entry_list(L, sp, p, pos)) { err = ERR_GEOM; return; }
  }
  // coordination 4 -> 3 is EraseBulk/EraseVoid + AddMC; the rate of an
  // unfixed site comes from its old list copy (Evaporation.f:121)
  s_set_coor(sp, coornn - 1);
  wr(L, p.i, sp);
  q.need = true;
  q.p = p;
  q.sp = sp;
  q.coor = coornn - 1;
  q.nn = nnn;
}

// The NBuff the reference carries across the neighbours: a neighbour that
// found no reciprocal partner gets the last one found before it (or the
// initial (0,0,0)).  Applied neighbour by neighbour, in order; `carried`
// is the value carried out of the loop into the site's own unfix.
__device__ void evap_nbuff_one(const Lattice &L, EvScratch &sc, int i, int carried[3], int &err) {
  if (!sc.unfixed[i]) return;
  if (sc.found[i]) {
    for (int k = 0; k < 3; k++) carried[k] = sc.found_pos[i][k];
  } else if (carried[0] || carried[1] || carried[2]) {
    const P3 p = sc.pts[i];
    Site sp = rd(L, p.i);
    if (!one_entry_list(L, sp, p, carried)) { err = ERR_GEOM; return; }
    wr(L, p.i, sp);
  }
}

// Evaporation, mcevent/Evaporation.f:25-35: the site empties.  Lane 0.
__device__ inline void evap_site(const Lattice &L, const P3 &s, EvScratch &sc) {
  Site ss = rd(L, s.i);
  sc.coor_old = s_coor(ss);
  sc.old_kind = s_kind(ss);
  s_set_kind(ss, KIND_EMPTY);
  wr(L, s.i, ss);
  sc.nn = list_of(L, s, ss);
  resolve_neighbours(L, sc, 4);
  touched_list(s, sc, 4);
}

// absorption.f:25-78 / desorption.f:25-117: the site takes (absorption,
// slot = ListCov index 1..3) or loses (slot 0) its coverage species; at
// coordination 1 desorption visits only the first neighbour.  Lane 0.
__device__ inline void cov_site(const Lattice &L, const P3 &s, int slot, EvScratch &sc) {
  Site ss = rd(L, s.i);
  sc.coor_old = s_coor(ss);
  sc.old_kind = s_kind(ss);
  s_set_kind(ss, slot ? KIND_COV0 + slot - 1 : KIND_EMPTY);
  wr(L, s.i, ss);
  sc.nn = list_of(L, s, ss);
  sc.nvisit = (!slot && sc.coor_old == 1) ? 1 : 4;
  resolve_neighbours(L, sc, sc.nvisit);
  touched_list(s, sc, sc.nvisit);
}

// Their neighbour i: an atom bonded back to s gets a fresh rate
// (absorption only for coordination <= 3, absorption.f:45-60;
// desorption.f:45-105).
__device__ inline void cov_prepare(const Lattice &L, const P3 &s, int i, bool absorb, EvScratch &sc, RefreshReq &q) {
  const P3 p = sc.pts[i];
  const Site sp = rd(L, p.i);
  const int k = s_kind(sp);
  if (!kind_occupied(k) || k == KIND_WALL || (absorb && s_coor(sp) > 3)) return;
  const NL nnn = list_of(L, p, sp);
  if (!recip(nnn, s)) return;
  q.need = true;
  q.p = p;
  q.sp = sp;
  q.coor = s_coor(sp);
  q.nn = nnn;
}

// The reference event an MC particle of class `cls` executes.  For an
// adsorption site this draws the species / coverage as Get_Prob.f:160-176
// and IndexEvent_Abs (Get_Prob.f:22-52) do, with the draw taken at
// execution instead of at rate-update time (same distribution: the class
// is refreshed whenever the coordination changes).  Returns itrans in the
// reference's numbering (1..3 deposit, 4 evaporate, 5 desorb, 6.. absorb).
template <bool LA = true, bool COV = true>
MSKG_HD int pick_itrans(const Lattice &L, const P3 &s, int cls, double u) {
  const Tables &t = *L.tab;
  if (cls == CLS_PLACEHOLDER) return 1;
  if (cls == CLS_EVAP || cls == CLS_ISOLATED) return 4;
  if (cls == CLS_DESORB) return 5;
  Site ss = rd(L, s.i);
  int coor = s_coor(ss);
  int temperature = (LA && t.la_mode) ? (int)ss.temp : 0;
  STW w;
  // InitSt 'ST': the tables at the site's temperature (its rate was evaluated
  // with them, so the temperature lies inside the references)
  const bool st = LA && t.st_n > 0 && st_weights(t, (int)ss.temp, w);
  auto pabs_of = [&](int i) { return st ? st_entry(t.st_pabs, 12, i + 3 * coor, w) : t.pabs[i + 3 * coor]; };
  double p1, p2, p3;
  if (st) {
    p1 = st_entry(t.st_pd, 9, 3 * (coor - 1), w);
    p2 = st_entry(t.st_pd, 9, 1 + 3 * (coor - 1), w);
    p3 = st_entry(t.st_pd, 9, 2 + 3 * (coor - 1), w);
  } else {
    p1 = pd_of<LA>(t, 1, coor, temperature, L.cl);
    p2 = pd_of<LA>(t, 2, coor, temperature, L.cl);
    p3 = pd_of<LA>(t, 3, coor, temperature, L.cl);
  }
  double pdep = p1 + p2 + p3;
  double pabs = COV ? pabs_of(0) + pabs_of(1) + pabs_of(2) : 0.0;
  double pcurr = u * (pdep + pabs);
  if (!COV || pcurr <= pdep) {
    if (pcurr <= p1) return 1;
    if (pcurr > p1 + p2) return 3;
    return 2;
  }
  if (t.ncov == 1) return 6;
  int r = 6;
  double lo = pdep;
  for (int i = 1; i <= t.ncov; i++) {
    double hi = lo + pabs_of(i - 1);
    if (pcurr > lo && pcurr <= hi) r = 5 + i;
    lo = hi;
  }
  return r;
}

// Execute one event of type itrans at site s.  Called by every lane of the
// wave with the same arguments; res is complete on lane 0.  Three phases:
// lane 0 updates the site; lanes 0..3 update its neighbours (one rate
// evaluation each); lane 0 gives the site its final rate while, with the
// second-neighbour term, lanes 1..12 refresh the atoms of its second shell.
// Each phase has one rate-evaluation call site for every event type, which
// keeps the inlined kernel small.
template <bool LA = true, bool COV = true>
__device__ void execute(const Lattice &L, double rtot, const P3 &s, int itrans, const Choice &ch, EvResult &res,
                        EvScratch &sc, int lane) {
  res.err = ERR_NONE;
  res.arm = -1;
  res.species = 0;
  if (s.z > L.ztop) {  // before anything changes: the run stops with the state as it was
    res.err = ERR_TOP;
    return;
  }
  const bool dep = itrans <= 3, evap = itrans == 4, absorb = itrans >= 6;
  if (lane < 12) {
    sc.sec[lane].i = -1;
    sc.err2[lane] = 0;
  }
  if (lane == 0) {
    sc.ntouched = 0;
    sc.abort = 0;
    sc.nvisit = 4;
    if (dep) dep_site(L, s, itrans, ch, res, sc);
    else if (evap) evap_site(L, s, sc);
    else if (COV) cov_site(L, s, absorb ? itrans - 5 : 0, sc);
    res.species = dep ? itrans : (evap && kind_species(sc.old_kind) ? sc.old_kind : 0);
    if (COV && itrans == 5) res.species = kind_cov_slot(sc.old_kind);  // the coverage desorbed
  }
  MSKG_SYNC();
  if (sc.abort) {
    if (lane == 0) res.err = sc.abort;
    return;
  }
  // phase 2: the neighbours
  int nbuff[3] = {0, 0, 0};  // Evaporation.f's NBuff, carried on lane 0
  {
    int i0, i1;
    visit_range(sc, lane, sc.nvisit, i0, i1);
#pragma unroll 1
    for (int i = i0; i < i1; i++) {
      RefreshReq q;
      q.need = q.list_after = false;
      if (dep) dep_prepare(L, s, i, sc, sc.err[i], q);
      else if (evap) evap_prepare(L, s, i, sc, sc.err[i], q);
      else if (COV) cov_prepare(L, s, i, absorb, sc, q);
      if (q.need) {
        int cls = 0, key = -1;
        double r = 0.0;
        // the site's coordination is final only in phase 3: here it is still coor_old
        get_prob<LA, COV, true>(L, rtot, q.p, q.sp, q.coor, q.nn, cls, r, sc.err[i], nullptr, &key, s.i,
                                sc.coor_old);
        bool ok = true;
        if (q.list_after) {
          if (cls && !set_list(L.box, q.sp, q.p.x, q.p.y, q.p.z, q.nn)) {
            sc.err[i] = ERR_GEOM;
            ok = false;
          } else {
            wr(L, q.p.i, q.sp);
          }
        }
        if (ok) {
          set_rate(L, q.p, cls, r, key);
          note_rate(sc, 1 + i, cls, r);
        }
      }
      if (evap && sc.seq) evap_nbuff_one(L, sc, i, nbuff, sc.err[i]);  // in order, as each neighbour is visited
    }
  }
  MSKG_SYNC();
  // phase 3: the site's final rate and the second shell
  FinalRefresh f;
  f.active = false;
  Site ss;
  bool unfix = false;
  if (lane == 0) {
    if (sc.seq)  // a point listed twice took several rates in turn
      for (int q = 1; q < sc.ntouched; q++) sc.tset[q] = 2;
    int coor = sc.coor_old;
    for (int i = 0; i < 4; i++) {
      if (dep || evap) coor += sc.inc[i];
      if (sc.err[i]) res.err = sc.err[i];
    }
    if (evap && !sc.seq)
      for (int i = 0; i < 4; i++) evap_nbuff_one(L, sc, i, nbuff, res.err);
    ss = rd(L, s.i);
    if (coor != sc.coor_old) s_set_coor(ss, coor);
    if (evap && coor == 0) {  // EraseMC zeroes the record
      for (int a = 0; a < 4; a++) ss.d[a][0] = ss.d[a][1] = ss.d[a][2] = 0;
      ss.valid = 0;
      wr(L, s.i, ss);
      set_rate(L, s, CLS_NONE, 0.0);
      note_rate(sc, 0, CLS_NONE, 0.0);
    } else {
      if (coor != sc.coor_old || evap) wr(L, s.i, ss);
      f.active = true;
      f.p = s;
      f.sp = ss;
      f.coor = coor;
      f.nn = sc.nn;
      unfix = evap && coor == 1;  // the site keeps one reciprocal bond (NBuff)
    }
  } else if (lane <= 12 && L.tab->use_f2 && (dep || evap)) {
    second_setup(L, s, lane - 1, sc, f);
  }
  int cls = 0, key = -1, e = ERR_NONE;
  double r = 0.0;
  if (f.active) get_prob<LA, COV, true>(L, rtot, f.p, f.sp, f.coor, f.nn, cls, r, e, nullptr, &key);
  if (unfix) {
    int pos[3];
    if (nbuff_of(L, s, sc.nn, pos))
      for (int k = 0; k < 3; k++) nbuff[k] = pos[k];
    if (!one_entry_list(L, ss, s, nbuff)) e = ERR_GEOM;
    else wr(L, s.i, ss);
  }
  if (f.active) {
    set_rate(L, f.p, cls, r, key);
    if (lane == 0) {
      note_rate(sc, 0, cls, r);
    } else {
      sc.secrate[lane - 1] = cls ? (float)r : 0.0f;
      sc.seccls[lane - 1] = (uint8_t)cls;
    }
    if (e) {
      if (lane == 0) res.err = e;
      else sc.err2[lane - 1] = e;
    }
  }
  if (L.tab->use_f2 && (dep || evap)) {
    MSKG_SYNC();  // the second-shell refreshes are done
    if (lane == 0)
      for (int k = 0; k < 12; k++) {
        if (sc.err2[k]) res.err = sc.err2[k];
        if (sc.sec[k].i >= 0) {
          sc.trate[sc.ntouched] = sc.secrate[k];
          sc.tcls[sc.ntouched] = sc.seccls[k];
          sc.tset[sc.ntouched] = 1;
          sc.touched[sc.ntouched++] = sc.sec[k];
        }
      }
  }
}

}  // namespace mskg
