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
  ERR_TEMPERATURE = 6 // zero temperature in LA mode
};

struct Lattice {
  Site *site;
  float *rate;
  uint8_t *cls;
  uint8_t *tile_active;   // may be null
  Box box;
  const Geometry *geo;
  const Tables *tab;
  int tile, ntx, nty;     // tile edge in grid units and tile counts (for tile_active)
};

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

MSKG_HD void mark_tile(const Lattice &L, const P3 &p) {
  if (!L.tile_active || p.i < 0) return;
  int tx = p.x / L.tile, ty = p.y / L.tile, tz = p.z / L.tile;
  L.tile_active[((int64_t)tz * L.nty + ty) * L.ntx + tx] = 1;
}
MSKG_HD void set_rate(const Lattice &L, const P3 &p, int cls, double rate) {
  if (p.i < 0) return;
  L.rate[p.i] = cls ? (float)rate : 0.0f;
  L.cls[p.i] = (uint8_t)cls;
  if (cls) mark_tile(L, p);
}

// ---- rates ---------------------------------------------------------------
MSKG_HD double PEv(const Tables &t, int sp, int a, int b, int c, int d, int e, int f) {
  return t.pe[(sp - 1) + 3 * (a + 4 * (b + 4 * (c + 4 * (d + 4 * (e + 4 * f)))))];
}
// The reference's kb (Get_Prob.f:70).
#define MSKG_KB 0.00008617330350
MSKG_HD double pd_of(const Tables &t, int sp, int coor, int temperature) {
  double e = t.pd[(sp - 1) + 3 * (coor - 1)];
  if (!t.la_mode) return e;
  return 0.5 * t.p0 * exp(-e / (MSKG_KB * t.tm)) * (1.0 + erf((temperature - t.tflex) / t.sigma)) *
         (e != 0.0 ? 1.0 : 0.0);
}
MSKG_HD double pe_of(const Tables &t, int sp, int a, int b, int c, int d, int e, int f, int temperature) {
  double v = PEv(t, sp, a, b, c, d, e, f);
  if (!t.la_mode) return v;
  return t.p0 * exp(-v / (MSKG_KB * temperature)) * (v != 0.0 ? 1.0 : 0.0);
}

// Get_Prob, Get_Prob.f:54-331, evaluated for point p whose record is sp,
// with the coordination and neighbour list passed explicitly (Evaporation.f:121
// passes a stale list copy).  rtot is the total rate R used by the
// isolated-atom rule.  No random number is drawn: the deposition species /
// absorbed coverage is drawn when the event executes (class CLS_ADSORB).
MSKG_HD void get_prob(const Lattice &L, double rtot, const P3 &p, const Site &sp, int coor, const NL &nn, int &cls,
                      double &rate, int &err) {
  const Tables &t = *L.tab;
  int kind = s_kind(sp);
  int temperature = 0;
  if (t.la_mode) {
    temperature = (int)sp.temp;
    if (temperature == 0) { err = ERR_TEMPERATURE; cls = CLS_NONE; rate = 0.0; return; }
  }
  int slot = kind_cov_slot(kind);
  if (slot) {  // desorption, :93-123
    if (slot > t.ncov) { err = ERR_COVERAGE; cls = CLS_NONE; rate = 0.0; return; }
    if (coor <= 3) { cls = CLS_DESORB; rate = t.pdes[(slot - 1) + 3 * coor]; }
    else { cls = CLS_NONE; rate = 0.0; }
    return;
  }
  if (kind == KIND_EMPTY) {  // :155-214
    if (coor >= 1 && coor <= 3) {
      bool ok = true;
      if (coor == 1) ok = s_coor(rd(L, entry(L, nn, 0).i)) >= 2;
      if (ok) {
        double pdep = pd_of(t, 1, coor, temperature) + pd_of(t, 2, coor, temperature) + pd_of(t, 3, coor, temperature);
        double pabs = t.pabs[0 + 3 * coor] + t.pabs[1 + 3 * coor] + t.pabs[2 + 3 * coor];
        cls = CLS_ADSORB;
        rate = pdep + pabs;
      } else {
        cls = CLS_PLACEHOLDER;
        rate = (double)1.e-8f;  // REAL(4) literal, Get_Prob.f:183
      }
    } else {
      cls = CLS_NONE;  // Coor 0 leaves the reference's outputs unassigned; Coor 4 is a vacancy
      rate = 0.0;
    }
    return;
  }
  // evaporation, :215-324
  int n_si = 0, n_c = 0, n_b = 0, n_wall = 0, covnn[3] = {0, 0, 0};
  for (int i = 0; i < 4; i++) {
    P3 q = entry(L, nn, i);
    Site sq = rd(L, q.i);
    int kq = s_kind(sq);
    if (kind_occupied(kq)) {
      if (kq != KIND_WALL) {
        if (recip(list_of(L, q, sq), p)) {
          if (kq == 1) n_si++;
          if (kq == 2) n_c++;
          if (kq == 3) n_b++;
        }
      } else {
        n_wall++;
      }
    } else if (kind_cov_slot(kq)) {
      if (recip(list_of(L, q, sq), p)) {
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
  } else {
    int c1 = t.ncov >= 1 ? covnn[0] : 0, c2 = t.ncov >= 2 ? covnn[1] : 0, c3 = t.ncov >= 3 ? covnn[2] : 0;
    cls = CLS_EVAP;
    rate = pe_of(t, kind, n_si + n_wall, n_c, n_b, c1, c2, c3, temperature);
  }
}

// Where a stochastic choice of the reference comes from: the log (replay)
// or the engine's RNG.
struct Choice {
  int arm;       // 0/1 forced arm-chair choice, -1 draw
  double u_arm;  // uniform used when the arm/zig coordination sums tie
};

struct EvResult {
  P3 touched[5];  // the site and its neighbours: every point whose rate may have changed
  int ntouched;
  int arm;        // stacking choice made (-1 none)
  int err;
};

// Get_Prob followed by UpdateTree / AddMC / EraseMC+Add{Bulk,Void}: a
// zero rate takes the site out of the MC population, its list stays.
MSKG_HD void refresh(const Lattice &L, double rtot, const P3 &p, const Site &sp, int coor, const NL &nn, int &err) {
  int cls = 0;
  double r = 0.0;
  get_prob(L, rtot, p, sp, coor, nn, cls, r, err);
  set_rate(L, p, cls, r);
}

MSKG_HD bool put_list(const Lattice &L, Site &s, const P3 &p, const NL &l, EvResult &res) {
  if (!set_list(L.box, s, p.x, p.y, p.z, l)) { res.err = ERR_GEOM; return false; }
  return true;
}

// Deposition, mcevent/deposition.f:24-251.
MSKG_HD void deposition(const Lattice &L, double rtot, const P3 &s, int species, const Choice &ch, EvResult &res) {
  const Geometry &g = *L.geo;
  const int lx = L.box.lx, ly = L.box.ly;
  Site ss = rd(L, s.i);
  int coor = s_coor(ss), coor_old = coor;
  s_set_kind(ss, species);
  wr(L, s.i, ss);
  res.arm = -1;
  NL nn_site = list_of(L, s, ss);
  if (coor == 1) {  // :49-74, FIND_SN + IF_Arm_Choice
    P3 c = entry(L, nn_site, 0);
    Site sc = rd(L, c.i);
    NL lc = list_of(L, c, sc);
    int sn[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, j = 0;
    for (int i = 0; i < 4; i++)
      if (!same(lc, i, s)) {
        if (j < 3) { sn[j][0] = lc.x[i]; sn[j][1] = lc.y[i]; sn[j][2] = lc.z[i]; }
        j++;
      }
    int site[3] = {s.x, s.y, s.z}, cc[3] = {c.x, c.y, c.z}, zig[3][3], arm[3][3];
    find_sn(g, lx, ly, site, cc, sn, zig, arm);
    int carm = 0, czig = 0;  // IF_Arm_Choice.f:35-54
    for (int k = 0; k < 3; k++) {
      carm += s_coor(rd(L, at(L, arm[k][0], arm[k][1], arm[k][2]).i));
      czig += s_coor(rd(L, at(L, zig[k][0], zig[k][1], zig[k][2]).i));
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
    if (!put_list(L, ss, s, nn_site, res)) return;
    wr(L, s.i, ss);
  }
  res.ntouched = 0;
  res.touched[res.ntouched++] = s;
  for (int i = 0; i < 4; i++) {  // :78-235
    P3 p = entry(L, nn_site, i);
    Site sp = rd(L, p.i);
    res.touched[res.ntouched++] = p;
    int kind = s_kind(sp), coornn = s_coor(sp);
    if (!kind_occupied(kind)) {
      if (coornn == 0) {
        coornn = 1;
        s_set_coor(sp, coornn);
        NL nnn;
        for (int j = 0; j < 4; j++) nnn.x[j] = nnn.y[j] = nnn.z[j] = 0;
        nnn.x[0] = s.x; nnn.y[0] = s.y; nnn.z[0] = s.z;
        int cls = 0;
        double r = 0.0;
        get_prob(L, rtot, p, sp, coornn, nnn, cls, r, res.err);
        if (cls && !put_list(L, sp, p, nnn, res)) return;  // AddMC stores the list
        wr(L, p.i, sp);
        set_rate(L, p, cls, r);
      } else if (coornn == 1) {
        NL nnn = list_of(L, p, sp);
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
          for (int j = 2; j < 4; j++) {
            nnn.x[j] = ns[j - 2][0]; nnn.y[j] = ns[j - 2][1]; nnn.z[j] = ns[j - 2][2];
            P3 q = entry(L, nnn, j);
            Site sq = rd(L, q.i);
            int kq = s_kind(sq);
            if (kind_occupied(kq) && s_coor(sq) <= 3 && kq != KIND_WALL) {
              if (recip(list_of(L, q, sq), p)) coornn++;
            }
          }
          s_set_coor(sp, coornn);
          if (!put_list(L, sp, p, nnn, res)) return;
          wr(L, p.i, sp);
          refresh(L, rtot, p, sp, coornn, nnn, res.err);
        }
      } else if (coornn == 2 || coornn == 3) {
        NL nnn = list_of(L, p, sp);
        if (recip(nnn, s)) {
          coornn++;
          s_set_coor(sp, coornn);
          wr(L, p.i, sp);
          refresh(L, rtot, p, sp, coornn, nnn, res.err);
        }
      }
    } else if (kind != KIND_WALL) {
      if (coornn <= 3) {
        NL nnn = list_of(L, p, sp);
        if (recip(nnn, s)) {
          coornn++;
          s_set_coor(sp, coornn);
          wr(L, p.i, sp);
          refresh(L, rtot, p, sp, coornn, nnn, res.err);
          if (coor_old == 1 && i >= 1) coor++;
        }
      }
    } else {
      coor++;  // wall neighbour
    }
  }
  ss = rd(L, s.i);
  if (coor != coor_old) {
    s_set_coor(ss, coor);
    wr(L, s.i, ss);
  }
  refresh(L, rtot, s, ss, coor, nn_site, res.err);
}

// The part of Evaporation.f:82-128 / 225-259 that unfixes a list when a
// site drops to coordination 1: keep only the last occupied reciprocal
// neighbour (NBuff) in slot 1.  nbuff persists across calls within one
// evaporation exactly like the reference's local variable.
MSKG_HD void unfix(const Lattice &L, const P3 &p, const NL &old, NL &nl, int nbuff[3]) {
  for (int j = 0; j < 4; j++) {
    P3 q = entry(L, old, j);
    Site sq = rd(L, q.i);
    int kq = s_kind(sq);
    if (kind_occupied(kq) && s_coor(sq) <= 3 && kq != KIND_WALL) {
      if (recip(list_of(L, q, sq), p)) { nbuff[0] = old.x[j]; nbuff[1] = old.y[j]; nbuff[2] = old.z[j]; }
    }
  }
  for (int j = 0; j < 4; j++) nl.x[j] = nl.y[j] = nl.z[j] = 0;
  nl.x[0] = nbuff[0]; nl.y[0] = nbuff[1]; nl.z[0] = nbuff[2];
}

// Evaporation, mcevent/Evaporation.f:25-262.
MSKG_HD void evaporation(const Lattice &L, double rtot, const P3 &s, EvResult &res) {
  Site ss = rd(L, s.i);
  int coor = s_coor(ss);
  s_set_kind(ss, KIND_EMPTY);
  wr(L, s.i, ss);
  const NL nn_site = list_of(L, s, ss);
  int nbuff[3] = {0, 0, 0};
  res.ntouched = 0;
  res.touched[res.ntouched++] = s;
  for (int i = 0; i < 4; i++) {
    P3 p = entry(L, nn_site, i);
    Site sp = rd(L, p.i);
    res.touched[res.ntouched++] = p;
    int kind = s_kind(sp), coornn = s_coor(sp);
    NL nnn = list_of(L, p, sp);
    if (!kind_occupied(kind)) {
      if (coornn == 1) {
        if (recip(nnn, s)) {
          Site z = rd(L, -1);
          z.temp = sp.temp;  // EraseMC + LattCoo=0, keeping the temperature bits
          wr(L, p.i, z);
          set_rate(L, p, CLS_NONE, 0.0);
        }
      } else if (coornn == 2) {
        if (recip(nnn, s)) {
          NL nl;
          unfix(L, p, nnn, nl, nbuff);
          s_set_coor(sp, 1);
          if (!put_list(L, sp, p, nl, res)) return;
          wr(L, p.i, sp);
          refresh(L, rtot, p, sp, 1, nnn, res.err);  // old list copy, Evaporation.f:121
        }
      } else if (coornn == 3) {
        if (recip(nnn, s)) {
          s_set_coor(sp, 2);
          wr(L, p.i, sp);
          refresh(L, rtot, p, sp, 2, nnn, res.err);
        }
      } else if (coornn == 4) {
        if (recip(nnn, s)) {
          s_set_coor(sp, 3);
          wr(L, p.i, sp);
          refresh(L, rtot, p, sp, 3, nnn, res.err);  // EraseVoid + AddMC
        }
      }
    } else if (kind != KIND_WALL) {
      if (coornn == 4) {
        if (recip(nnn, s)) {
          s_set_coor(sp, 3);
          wr(L, p.i, sp);
          refresh(L, rtot, p, sp, 3, nnn, res.err);  // EraseBulk + AddMC
        }
      } else {
        if (recip(nnn, s)) {
          coornn--;
          s_set_coor(sp, coornn);
          wr(L, p.i, sp);
          refresh(L, rtot, p, sp, coornn, nnn, res.err);
        }
      }
    } else {
      coor--;  // wall neighbour
    }
  }
  ss = rd(L, s.i);
  s_set_coor(ss, coor);
  if (coor == 0) {  // EraseMC zeroes the record
    for (int a = 0; a < 4; a++) ss.d[a][0] = ss.d[a][1] = ss.d[a][2] = 0;
    ss.valid = 0;
    wr(L, s.i, ss);
    set_rate(L, s, CLS_NONE, 0.0);
    return;
  }
  wr(L, s.i, ss);
  int cls = 0;
  double r = 0.0;
  get_prob(L, rtot, s, ss, coor, nn_site, cls, r, res.err);
  if (coor == 1) {
    NL nl;
    unfix(L, s, nn_site, nl, nbuff);
    if (!put_list(L, ss, s, nl, res)) return;
    wr(L, s.i, ss);
  }
  set_rate(L, s, cls, r);
}

// absorption, mcevent/absorption.f:25-78 (slot = ListCov index 1..3).
MSKG_HD void absorption(const Lattice &L, double rtot, const P3 &s, int slot, EvResult &res) {
  Site ss = rd(L, s.i);
  int coor = s_coor(ss);
  s_set_kind(ss, KIND_COV0 + slot - 1);
  wr(L, s.i, ss);
  const NL nn_site = list_of(L, s, ss);
  res.ntouched = 0;
  res.touched[res.ntouched++] = s;
  for (int i = 0; i < 4; i++) {
    P3 p = entry(L, nn_site, i);
    Site sp = rd(L, p.i);
    res.touched[res.ntouched++] = p;
    int k = s_kind(sp);
    if (kind_occupied(k) && k != KIND_WALL && s_coor(sp) <= 3) {
      NL nnn = list_of(L, p, sp);
      if (recip(nnn, s)) refresh(L, rtot, p, sp, s_coor(sp), nnn, res.err);
    }
  }
  refresh(L, rtot, s, ss, coor, nn_site, res.err);
}

// desorption, mcevent/desorption.f:25-117.
MSKG_HD void desorption(const Lattice &L, double rtot, const P3 &s, EvResult &res) {
  Site ss = rd(L, s.i);
  int coor = s_coor(ss);
  s_set_kind(ss, KIND_EMPTY);
  wr(L, s.i, ss);
  const NL nn_site = list_of(L, s, ss);
  res.ntouched = 0;
  res.touched[res.ntouched++] = s;
  int nloop = coor == 1 ? 1 : 4;
  for (int i = 0; i < nloop; i++) {
    P3 p = entry(L, nn_site, i);
    Site sp = rd(L, p.i);
    res.touched[res.ntouched++] = p;
    int k = s_kind(sp);
    if (kind_occupied(k) && k != KIND_WALL) {
      NL nnn = list_of(L, p, sp);
      if (recip(nnn, s)) refresh(L, rtot, p, sp, s_coor(sp), nnn, res.err);
    }
  }
  refresh(L, rtot, s, ss, coor, nn_site, res.err);
}

// The reference event an MC particle of class `cls` executes.  For an
// adsorption site this draws the species / coverage as Get_Prob.f:160-176
// and IndexEvent_Abs (Get_Prob.f:22-52) do, with the draw taken at
// execution instead of at rate-update time (same distribution: the class
// is refreshed whenever the coordination changes).  Returns itrans in the
// reference's numbering (1..3 deposit, 4 evaporate, 5 desorb, 6.. absorb).
MSKG_HD int pick_itrans(const Lattice &L, const P3 &s, int cls, double u) {
  const Tables &t = *L.tab;
  if (cls == CLS_PLACEHOLDER) return 1;
  if (cls == CLS_EVAP || cls == CLS_ISOLATED) return 4;
  if (cls == CLS_DESORB) return 5;
  Site ss = rd(L, s.i);
  int coor = s_coor(ss);
  int temperature = t.la_mode ? (int)ss.temp : 0;
  double p1 = pd_of(t, 1, coor, temperature), p2 = pd_of(t, 2, coor, temperature), p3 = pd_of(t, 3, coor, temperature);
  double pdep = p1 + p2 + p3;
  double pabs = t.pabs[0 + 3 * coor] + t.pabs[1 + 3 * coor] + t.pabs[2 + 3 * coor];
  double pcurr = u * (pdep + pabs);
  if (pcurr <= pdep) {
    if (pcurr <= p1) return 1;
    if (pcurr > p1 + p2) return 3;
    return 2;
  }
  if (t.ncov == 1) return 6;
  int r = 6;
  double lo = pdep;
  for (int i = 1; i <= t.ncov; i++) {
    double hi = lo + t.pabs[(i - 1) + 3 * coor];
    if (pcurr > lo && pcurr <= hi) r = 5 + i;
    lo = hi;
  }
  return r;
}

// Execute one event of type itrans at site s.
MSKG_HD void execute(const Lattice &L, double rtot, const P3 &s, int itrans, const Choice &ch, EvResult &res) {
  res.err = ERR_NONE;
  res.arm = -1;
  res.ntouched = 0;
  if (itrans <= 3) deposition(L, rtot, s, itrans, ch, res);
  else if (itrans == 4) evaporation(L, rtot, s, res);
  else if (itrans == 5) desorption(L, rtot, s, res);
  else absorption(L, rtot, s, itrans - 5, res);
}

}  // namespace mskg
