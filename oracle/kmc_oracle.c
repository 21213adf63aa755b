/*
 * kmc_oracle.c -- CPU restatement of the MulSKIPS KMC event engine.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  See
 * kmc_oracle.h for the scope and for how each part is pinned.
 *
 * Every function below restates a reference routine; the citation is given
 * as file:line relative to the reference repository root.  Where the
 * reference relies on default-REAL(4) literals widened to REAL(8), the same
 * float-rounded constants are used here.
 */
#include "kmc_oracle.h"

#include <math.h>
#include <setjmp.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ---- bit layout of a LattCoo word (modules/defsystem.f:87-91) --------- */
enum {
  POS_INDEX = 0, LEN_INDEX = 8,  /* CovInd: 0 crystal/empty, 113 wall, Z of coverage */
  POS_T = 9, LEN_T = 13,         /* temperature in K (LA / ST modes)              */
  POS_OCC = 23, LEN_OCC = 1,
  POS_COOR = 24, LEN_COOR = 3,
  POS_SIC = 27, LEN_SIC = 3,     /* species 1 Si, 2 C/Ge, 3 B                      */
  POS_SI = 27, POS_C = 28, LEN_SI = 1, LEN_C = 1
};
#define WALL 113
#define NUM_MC_MAX 500000       /* defsystem.f:34 */
#define NUM_AT_MAX 100000000
#define NUM_VD_MAX 100000
#define NUM_BF_MAX 500000

static const int JUMP_DIA[8][3] = { /* defsystem.f:112-113, JumpDia(1:3,i) */
  {3, 3, 3}, {3, -3, -3}, {-3, -3, 3}, {-3, 3, -3},
  {3, 3, -3}, {3, -3, 3}, {-3, -3, -3}, {-3, 3, 3}};

typedef struct { int32_t xyz[3]; int32_t nn[4][3]; int32_t ev; int32_t key; double prob; } mc_rec;
typedef struct { int32_t xyz[3]; int32_t nn[4][3]; } at_rec;
typedef struct { int32_t xyz[3]; int32_t type; } buf_rec;

struct orc_sim {
  int lx, ly, lz;
  int64_t npts;
  int32_t *latt, *lind;
  mc_rec *ad; int32_t n_ad;       /* ListAdAtom(1:NumMcMax), 1-based          */
  at_rec *at; int32_t n_at; int64_t cap_at;  /* ListAtom, grown on demand, zero filled */
  at_rec *vd; int32_t n_vd;       /* ListVoid(1:NumVdMax)                     */
  buf_rec *abuf;                  /* ListAtBuffer(1:NumBfMax)                 */
  double *tree; int levels, size_tree;
  /* tables, Fortran order */
  double pe[3 * 4096], pd[9], pabs[12], pdes[12], ptrans_zig;
  int ncrystal, ncov, list_crystal[3], list_cov[3];
  /* run configuration (main.f:83-212) */
  char init_st[3], run_type[3], exit_strategy[5];
  int len1, len2, len3, len4;
  double sf, alloy;
  int64_t max_iter, out_molmol;
  double max_time, out_time;
  int la_mode; double tm, p0, sigma, tflex;
  /* the laser-annealing build's form: per-species constants, InitCountCrystal
   * and the CountLiquid of the current iteration; la_swap its SiGe rule */
  int la_variant, la_swap;
  double la_p0[3], la_tm[3], la_sigma[3], la_tflex[3];
  int32_t count_init[3], count_liquid[3];
  /* InitSt 'ST': reference temperatures and tables (refPtrans*), the working
   * tables pe/pd/pabs/pdes interpolated at each Get_Prob */
  int st_n;
  int32_t *st_temp;
  double *st_pe, *st_pd, *st_pabs, *st_pdes;
  /* ran3 and the RunType 'T' table */
  orc_ran3 rng; int32_t idum;
  double *numbers; int n_numbers; int64_t counter;
  /* clock and counters */
  double time; int64_t iter;
  int32_t count_crystal[3], count_crystal_old[3], count_cov[3];
  int64_t executed[8];  /* events executed by type: deposits of species 1..3, evaporations,
                           desorptions, absorptions of ListCov 1..3 (the engine's counters) */
  int last_arm;
  int force_arm;  /* >= 0: IF_Arm_Choice returns this instead of comparing/drawing */
  int frozen_below_z;  /* engine option restated: atoms below this z never evaporate */
  int use_f2; double f2[13];  /* engine option restated: evaporation rate factor by the number
                                 of occupied second-neighbour positions */
  int last_key;               /* rate-class key of the last Get_Prob (likelihood accounting) */
  int64_t oob_reads;
  int64_t nonbond_arm, nonbond_nn;  /* geometry results that are not one of the 32 bond vectors */
  /* error handling: the reference STOPs; we unwind to the entry point */
  jmp_buf jb; int have_jb;
  char err[512];
};

static void fail(orc_sim *s, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(s->err, sizeof s->err, fmt, ap);
  va_end(ap);
  if (s->have_jb) longjmp(s->jb, 1);
  abort();
}

static inline int ibits(int32_t v, int pos, int len) {
  return (int)(((uint32_t)v >> pos) & ((1u << len) - 1u));
}
static inline void mvbits(int val, int len, int32_t *dst, int pos) {
  uint32_t m = ((1u << len) - 1u) << pos;
  *dst = (int32_t)(((uint32_t)*dst & ~m) | (((uint32_t)val << pos) & m));
}

/* LattCoo(x,y,z), 0<=x<LenX, 0<=y<LenY, 0<=z<=LenZ.  The reference reads
 * outside the array for neighbours below z=0 of a wall-less slab
 * (e.g. SetSiC3C.f:36-50); here such reads return 0 and are counted. */
static inline int64_t gidx(const orc_sim *s, int x, int y, int z) {
  return (int64_t)x + (int64_t)s->lx * ((int64_t)y + (int64_t)s->ly * z);
}
static inline int inbox(const orc_sim *s, const int32_t p[3]) {
  return p[0] >= 0 && p[0] < s->lx && p[1] >= 0 && p[1] < s->ly && p[2] >= 0 && p[2] <= s->lz;
}
static int32_t LC(orc_sim *s, const int32_t p[3]) {
  if (!inbox(s, p)) { s->oob_reads++; return 0; }
  return s->latt[gidx(s, p[0], p[1], p[2])];
}
static int32_t *LCp(orc_sim *s, const int32_t p[3]) {
  static int32_t sink;
  if (!inbox(s, p)) { s->oob_reads++; sink = 0; return &sink; }
  return &s->latt[gidx(s, p[0], p[1], p[2])];
}
static int32_t LI(orc_sim *s, const int32_t p[3]) {
  if (!inbox(s, p)) { s->oob_reads++; return 0; }
  return s->lind[gidx(s, p[0], p[1], p[2])];
}
static void set_LI(orc_sim *s, const int32_t p[3], int32_t v) {
  if (inbox(s, p)) s->lind[gidx(s, p[0], p[1], p[2])] = v;
}
static inline int eq3(const int32_t a[3], const int32_t b[3]) {
  return a[0] == b[0] && a[1] == b[1] && a[2] == b[2];
}
static inline void cp3(int32_t d[3], const int32_t a[3]) { d[0] = a[0]; d[1] = a[1]; d[2] = a[2]; }

/* Lists beyond their fill are zero in the reference (static arrays, zeroed
 * on erase), so reads past the end return zero records here as well. */
static const mc_rec ZERO_MC;
static const at_rec ZERO_AT;
static const mc_rec *adr(orc_sim *s, int32_t i) { return (i >= 1 && i < NUM_MC_MAX) ? &s->ad[i] : &ZERO_MC; }
static const at_rec *atr(orc_sim *s, int32_t i) { return (i >= 1 && i < s->cap_at) ? &s->at[i] : &ZERO_AT; }
static const at_rec *vdr(orc_sim *s, int32_t i) { return (i >= 1 && i < NUM_VD_MAX) ? &s->vd[i] : &ZERO_AT; }

static int recip(const int32_t nn[4][3], const int32_t site[3]) {
  for (int j = 0; j < 4; j++) if (eq3(nn[j], site)) return 1;
  return 0;
}

/* ---- ran3: mcevent/random.f:22-69 ------------------------------------- */
double orc_ran3_next(orc_ran3 *r, int32_t *idum) {
  const int32_t mbig = 1000000000, mseed = 161803398, mz = 0;
  const double fac = (double)(1.0f / (float)mbig); /* fac=1./Mbig is REAL(4) */
  int32_t mj, mk;
  if (*idum < 0 || r->iff == 0) {
    r->iff = 1;
    mj = mseed - abs(*idum);
    mj = mj % mbig;
    r->ma[55] = mj;
    mk = 1;
    for (int i = 1; i <= 54; i++) {
      int ii = (21 * i) % 55;
      r->ma[ii] = mk;
      mk = mj - mk;
      if (mk < mz) mk += mbig;
      mj = r->ma[ii];
    }
    for (int k = 1; k <= 4; k++)
      for (int i = 1; i <= 55; i++) {
        r->ma[i] = r->ma[i] - r->ma[1 + (i + 30) % 55];
        if (r->ma[i] < mz) r->ma[i] += mbig;
      }
    r->inext = 0;
    r->inextp = 31;
    *idum = 1;
  }
  if (++r->inext == 56) r->inext = 1;
  if (++r->inextp == 56) r->inextp = 1;
  mj = r->ma[r->inext] - r->ma[r->inextp];
  if (mj < mz) mj += mbig;
  r->ma[r->inext] = mj;
  return (double)mj * fac;
}
static double rnd(orc_sim *s) { return orc_ran3_next(&s->rng, &s->idum); }

/* ---- sum tree: AllocateArrays.f:24-45, UpdateTree.f:25-66 ------------- */
static void allocate_tree(orc_sim *s) {
  int levels = 0;
  for (int b = 0; b <= 31; b++) if ((NUM_MC_MAX >> b) & 1) levels = b;
  if (NUM_MC_MAX > (1 << levels)) levels += 2; else levels += 1;
  s->levels = levels;
  s->size_tree = (1 << levels) - 1;
  free(s->tree);
  s->tree = (double *)calloc((size_t)s->size_tree + 2, sizeof(double));
}

static void update_tree(orc_sim *s, int32_t index, double patom) {
  if (patom == 0.0 && index < s->n_ad)
    fail(s, "error stabilizzato da dentro Update! index=%d", index);
  else if (index > s->n_ad)
    fail(s, "Index > NumAdAtom (%d > %d)", index, s->n_ad);
  int32_t pos = index + (1 << (s->levels - 1)) - 1;
  double *t = s->tree;
  t[pos] = patom;
  for (int parent = s->levels - 1; parent >= 1; parent--) {
    int32_t neigh = (pos & 1) ? pos - 1 : pos + 1;
    int32_t np = pos >> 1;
    t[np] = t[pos] + t[neigh];
    pos = np;
  }
}

/* PickTreeEvent.f:25-70 */
static int32_t pick_tree_event(orc_sim *s, double *patom) {
  double *t = s->tree;
  int32_t index = 1;
  for (int parent = 1; parent <= s->levels - 1; parent++) {
    double u;
    if (s->run_type[0] == 'T') {
      s->counter++;
      u = s->numbers[(s->counter - 1) % s->n_numbers];
    } else {
      u = rnd(s);
    }
    *patom = u * t[index];
    index <<= 1;
    if (t[index] < *patom) index++;
  }
  *patom = t[index];
  if (index - (1 << (s->levels - 1)) + 1 > s->n_ad) fail(s, "write Index> NumAdAtom");
  return index - (1 << (s->levels - 1)) + 1;
}

/* ---- lists: Add.f:24-101, Erase.f:24-179 ------------------------------ */
static void add_mc(orc_sim *s, const int32_t site[3], const int32_t nn[4][3], int32_t ev, double prob) {
  if (prob == 0.0) fail(s, "provo ad aggiungere prob nulla");
  s->n_ad++;
  if (s->n_ad >= NUM_MC_MAX) fail(s, "troppe particelle cinetiche");
  set_LI(s, site, s->n_ad);
  mc_rec *r = &s->ad[s->n_ad];
  cp3(r->xyz, site);
  memcpy(r->nn, nn, sizeof r->nn);
  r->ev = ev;
  r->prob = prob;
  r->key = s->last_key;
  update_tree(s, s->n_ad, prob);
}

static void grow_at(orc_sim *s, int64_t need) {
  if (need < s->cap_at) return;
  int64_t nc = s->cap_at ? s->cap_at : 1 << 16;
  while (nc <= need) nc *= 2;
  if (nc > (int64_t)NUM_AT_MAX + 1) nc = (int64_t)NUM_AT_MAX + 1;
  at_rec *p = (at_rec *)realloc(s->at, (size_t)nc * sizeof(at_rec));
  if (!p) fail(s, "out of memory (bulk list)");
  memset(p + s->cap_at, 0, (size_t)(nc - s->cap_at) * sizeof(at_rec));
  s->at = p;
  s->cap_at = nc;
}

static void add_bulk(orc_sim *s, const int32_t site[3], const int32_t nn[4][3]) {
  s->n_at++;
  if (s->n_at >= NUM_AT_MAX) fail(s, "troppe particelle bulk");
  grow_at(s, s->n_at);
  set_LI(s, site, s->n_at);
  cp3(s->at[s->n_at].xyz, site);
  memcpy(s->at[s->n_at].nn, nn, sizeof s->at[0].nn);
}

static void add_void(orc_sim *s, const int32_t site[3], const int32_t nn[4][3]) {
  s->n_vd++;
  if (s->n_vd >= NUM_VD_MAX) fail(s, "troppe vacanze");
  set_LI(s, site, s->n_vd);
  cp3(s->vd[s->n_vd].xyz, site);
  memcpy(s->vd[s->n_vd].nn, nn, sizeof s->vd[0].nn);
}

static void erase_mc(orc_sim *s, const int32_t site[3]) {
  if (site[2] == -1) fail(s, "accesso a latt fuori bond");
  int32_t ind = LI(s, site);
  if (ind > NUM_MC_MAX || ind == 0) fail(s, "err: EraseMC %d %d %d %d", ind, site[0], site[1], site[2]);
  if (ind >= s->n_ad) {
    memset(&s->ad[ind], 0, sizeof(mc_rec));
    set_LI(s, site, 0);
    update_tree(s, ind, 0.0);
    s->n_ad--;
  } else {
    s->ad[ind] = s->ad[s->n_ad];
    int32_t leaf = s->n_ad + (1 << (s->levels - 1)) - 1;
    update_tree(s, ind, s->tree[leaf]);
    update_tree(s, s->n_ad, 0.0);
    int32_t tmp[3];
    cp3(tmp, s->ad[s->n_ad].xyz);
    set_LI(s, tmp, ind);
    memset(&s->ad[s->n_ad], 0, sizeof(mc_rec));
    set_LI(s, site, 0);
    s->n_ad--;
  }
}

static void erase_at_list(orc_sim *s, at_rec *list, int32_t *n, int64_t maxn, const int32_t site[3], const char *who) {
  if (site[2] == -1) fail(s, "accesso a latt fuori bond");
  int32_t ind = LI(s, site);
  if (ind > maxn || ind == 0) fail(s, "err: %s %d %d %d %d", who, ind, site[0], site[1], site[2]);
  if (ind == *n) {
    memset(&list[*n], 0, sizeof(at_rec));
    set_LI(s, site, 0);
    (*n)--;
  } else {
    list[ind] = list[*n];
    int32_t tmp[3];
    cp3(tmp, list[*n].xyz);
    set_LI(s, tmp, ind);
    memset(&list[*n], 0, sizeof(at_rec));
    set_LI(s, site, 0);
    (*n)--;
  }
}
static void erase_bulk(orc_sim *s, const int32_t site[3]) { erase_at_list(s, s->at, &s->n_at, NUM_AT_MAX, site, "EraseBulk"); }
static void erase_void(orc_sim *s, const int32_t site[3]) { erase_at_list(s, s->vd, &s->n_vd, NUM_VD_MAX, site, "EraseVoid"); }

/* ---- geometry: lattice/FIND_NN.f:25-77, lattice/FIND_SN.f:26-84 ------- */
static int isign(int a, int b) { return b >= 0 ? abs(a) : -abs(a); }

void orc_find_nn(const int32_t s1[3], const int32_t s2[3], const int32_t sc[3],
                 int32_t out[2][3], int lx, int ly) {
  const double pi = (double)3.14159265358979323f; /* REAL(4) literal */
  int32_t a[3], b[3];
  cp3(a, s1);
  cp3(b, s2);
  if (abs(sc[0] - a[0]) > lx / 2) a[0] += isign(lx, sc[0] - a[0]);
  if (abs(sc[1] - a[1]) > ly / 2) a[1] += isign(ly, sc[1] - a[1]);
  if (abs(sc[0] - b[0]) > lx / 2) b[0] += isign(lx, sc[0] - b[0]);
  if (abs(sc[1] - b[1]) > ly / 2) b[1] += isign(ly, sc[1] - b[1]);
  double ax[3];
  for (int i = 0; i < 3; i++) ax[i] = (double)sc[i] - (double)a[i];
  double snorm = sqrt(ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2]);
  for (int i = 0; i < 3; i++) ax[i] = ax[i] / snorm;
  for (int irot = 1; irot <= 2; irot++) {
    double ang = (double)(2 * irot) * pi / (double)3.0f;
    double cth = cos(ang), sth = sin(ang);
    double rot[3][3]; /* rot[i][j] = Rot(i+1,j+1) */
    for (int i = 0; i < 3; i++) rot[i][i] = ax[i] * ax[i] * (1 - cth) + cth;
    rot[1][0] = ax[0] * ax[1] * (1 - cth) - ax[2] * sth;
    rot[0][1] = ax[0] * ax[1] * (1 - cth) + ax[2] * sth;
    rot[2][0] = ax[0] * ax[2] * (1 - cth) + ax[1] * sth;
    rot[0][2] = ax[0] * ax[2] * (1 - cth) - ax[1] * sth;
    rot[2][1] = ax[1] * ax[2] * (1 - cth) - ax[0] * sth;
    rot[1][2] = ax[1] * ax[2] * (1 - cth) + ax[0] * sth;
    double r[3];
    for (int k = 0; k < 3; k++) {
      r[k] = 0.0;
      for (int l = 0; l < 3; l++) r[k] = r[k] + rot[l][k] * ((double)b[l] - (double)sc[l]);
    }
    for (int k = 0; k < 3; k++) out[irot - 1][k] = (int32_t)lround(r[k]) + sc[k];
    if (out[irot - 1][0] >= lx) out[irot - 1][0] -= lx;
    if (out[irot - 1][0] < 0) out[irot - 1][0] += lx;
    if (out[irot - 1][1] >= ly) out[irot - 1][1] -= ly;
    if (out[irot - 1][1] < 0) out[irot - 1][1] += ly;
  }
}

void orc_find_sn(const int32_t site[3], const int32_t sitec_in[3], const int32_t sn_in[3][3],
                 int32_t zig[3][3], int32_t arm[3][3], int lx, int ly) {
  int32_t sc[3], sn[3][3];
  cp3(sc, sitec_in);
  memcpy(sn, sn_in, sizeof sn);
  if (abs(site[0] - sc[0]) > lx / 2) sc[0] += isign(lx, site[0] - sc[0]);
  if (abs(site[1] - sc[1]) > ly / 2) sc[1] += isign(ly, site[1] - sc[1]);
  for (int i = 0; i < 3; i++) {
    if (abs(site[0] - sn[i][0]) > lx / 2) sn[i][0] += isign(lx, site[0] - sn[i][0]);
    if (abs(site[1] - sn[i][1]) > ly / 2) sn[i][1] += isign(ly, site[1] - sn[i][1]);
  }
  double cent[3], ax[3];
  for (int k = 0; k < 3; k++) cent[k] = 0.5 * ((double)sc[k] + (double)(float)site[k]);
  for (int k = 0; k < 3; k++) ax[k] = (double)site[k] - (double)sc[k];
  double snorm = sqrt(ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2]);
  for (int k = 0; k < 3; k++) ax[k] = ax[k] / snorm;
  const double tr = (double)8.6602254037f; /* REAL(4) literal, FIND_SN.f:67 */
  for (int i = 0; i < 3; i++) {
    for (int k = 0; k < 3; k++) {
      double rz = 2 * cent[k] - (double)sn[i][k];
      double ra = (double)sn[i][k] + tr * ax[k];
      zig[i][k] = (int32_t)lround(rz);
      arm[i][k] = (int32_t)lround(ra);
    }
    if (zig[i][0] >= lx) zig[i][0] -= lx;
    if (zig[i][0] < 0) zig[i][0] += lx;
    if (zig[i][1] >= ly) zig[i][1] -= ly;
    if (zig[i][1] < 0) zig[i][1] += ly;
    if (arm[i][0] >= lx) arm[i][0] -= lx;
    if (arm[i][0] < 0) arm[i][0] += lx;
    if (arm[i][1] >= ly) arm[i][1] -= ly;
    if (arm[i][1] < 0) arm[i][1] += ly;
  }
}

static int is_bond_disp(const orc_sim *s, const int32_t a[3], const int32_t b[3]) {
  int d[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
  if (d[0] > s->lx / 2) d[0] -= s->lx;
  if (d[0] < -s->lx / 2) d[0] += s->lx;
  if (d[1] > s->ly / 2) d[1] -= s->ly;
  if (d[1] < -s->ly / 2) d[1] += s->ly;
  return d[0] * d[0] + d[1] * d[1] + d[2] * d[2] == 27;
}

/* IF_Arm_Choice.f:21-56 */
static int if_arm_choice(orc_sim *s, int32_t zig[3][3], int32_t arm[3][3]) {
  if (s->force_arm >= 0) return s->force_arm;
  int carm = 0, czig = 0;
  for (int i = 0; i < 3; i++) {
    carm += ibits(LC(s, arm[i]), POS_COOR, LEN_COOR);
    czig += ibits(LC(s, zig[i]), POS_COOR, LEN_COOR);
  }
  if (czig > carm) return 0;
  if (czig == carm) return rnd(s) <= s->ptrans_zig ? 0 : 1;
  return 1;
}

/* ---- rates ------------------------------------------------------------- */
static inline double PE(const orc_sim *s, int sp, int a, int b, int c, int d, int e, int f) {
  return s->pe[(sp - 1) + 3 * (a + 4 * (b + 4 * (c + 4 * (d + 4 * (e + 4 * f)))))];
}
static inline double PD(const orc_sim *s, int sp, int coor) { return s->pd[(sp - 1) + 3 * (coor - 1)]; }
static inline double PABS(const orc_sim *s, int k, int coor) { return s->pabs[(k - 1) + 3 * coor]; }
static inline double PDES(const orc_sim *s, int k, int coor) { return s->pdes[(k - 1) + 3 * coor]; }

/* LA rates, Get_Prob.f:127-149: the reference re-evaluates the whole
 * table; only the entry that is read is needed and the value is the same
 * elementwise expression. */
static const double KB = 0.00008617330350; /* Get_Prob.f:70 */
static double la_pe(const orc_sim *s, double e, int temperature) {
  return s->p0 * exp(-e / (KB * temperature)) * (e != 0.0 ? 1.0 : 0.0);
}
static double la_pd(const orc_sim *s, double e, int temperature) {
  return 0.5 * s->p0 * exp(-e / (KB * s->tm)) * (1.0 + erf((temperature - s->tflex) / s->sigma)) *
         (e != 0.0 ? 1.0 : 0.0);
}

typedef struct { double pd[9]; int la_t; } pdtab;
/* mulskips-source-LaserAnnealing/mcevent/Get_Prob.f:143-151, the entries of
 * species k: PE = P0(k) exp(-PtransE/(kb T)) [/= 0]; PD = 0.5d0 * P0(k) *
 * FLOAT(CountLiquid(k)) / SUM(CountLiquid) * EXP(-PtransD/(kb Tm(k))) *
 * (1 + ERF((T - Tflex(k)) / sigma(k))) [/= 0], left to right, FLOAT a REAL(4)
 * conversion of the integer count.  SUM(CountLiquid) = 0 (a division by
 * zero there) gives no solidification here. */
static double la_pd_variant(const orc_sim *s, int sp, double e, int temperature) {
  if (sp > s->ncrystal) return 0.0;  /* the build sets PD(k,:) for k <= NCrystal only */
  const int32_t sum = s->count_liquid[0] + s->count_liquid[1] + s->count_liquid[2];
  if (sum <= 0) return 0.0;
  return 0.5 * s->la_p0[sp - 1] * (double)(float)s->count_liquid[sp - 1] / (double)sum *
         exp(-e / (KB * s->la_tm[sp - 1])) * (1.0 + erf((temperature - s->la_tflex[sp - 1]) / s->la_sigma[sp - 1])) *
         (e != 0.0 ? 1.0 : 0.0);
}
static double pdv(orc_sim *s, int sp, int coor, int temperature) {
  if (!s->la_mode) return PD(s, sp, coor);
  if (s->la_variant) return la_pd_variant(s, sp, PD(s, sp, coor), temperature);
  return la_pd(s, PD(s, sp, coor), temperature);
}
static double pev(orc_sim *s, int sp, int a, int b, int c, int d, int e, int f, int temperature) {
  if (!s->la_mode) return PE(s, sp, a, b, c, d, e, f);
  if (s->la_variant) {
    const double v = PE(s, sp, a, b, c, d, e, f);
    return s->la_p0[sp - 1] * exp(-v / (KB * temperature)) * (v != 0.0 ? 1.0 : 0.0);
  }
  return la_pe(s, PE(s, sp, a, b, c, d, e, f), temperature);
}
/* InterpProbFromField.f:22-57: the working tables at temperature T, from
 * the reference tables (temperatures increasing).  Outside the references
 * the reference indexes out of its arrays; here that fails. */
static void st_interp(orc_sim *s, int temperature) {
  int iu = -1;
  for (int k = 0; k < s->st_n; k++)
    if (s->st_temp[k] - temperature == 0) { iu = k; break; }  /* FINDLOC(diff, 0) */
  if (iu >= 0) {
    memcpy(s->pe, s->st_pe + (size_t)iu * 3 * 4096, sizeof s->pe);
    memcpy(s->pd, s->st_pd + (size_t)iu * 9, sizeof s->pd);
    memcpy(s->pabs, s->st_pabs + (size_t)iu * 12, sizeof s->pabs);
    memcpy(s->pdes, s->st_pdes + (size_t)iu * 12, sizeof s->pdes);
    return;
  }
  for (int k = 0; k < s->st_n; k++) {  /* MINLOC(diff, MASK = diff > 0) */
    const int d = s->st_temp[k] - temperature;
    if (d > 0 && (iu < 0 || d < s->st_temp[iu] - temperature)) iu = k;
  }
  if (iu <= 0) fail(s, "temperature %d K outside the 'ST' reference temperatures", temperature);
  const int il = iu - 1;
  const int wl = abs(s->st_temp[iu] - temperature), wu = abs(s->st_temp[il] - temperature);
#define ST_BLEND(dst, ref, n)                                                                          \
  for (int i = 0; i < (n); i++)                                                                        \
    (dst)[i] = ((double)wl * (ref)[(size_t)il * (n) + i] + (double)wu * (ref)[(size_t)iu * (n) + i]) / \
               (double)(wl + wu);
  ST_BLEND(s->pe, s->st_pe, 3 * 4096)
  ST_BLEND(s->pd, s->st_pd, 9)
  ST_BLEND(s->pabs, s->st_pabs, 12)
  ST_BLEND(s->pdes, s->st_pdes, 12)
#undef ST_BLEND
}

/* main.f:637 of the laser-annealing build, at the top of every iteration. */
static void la_liquid(orc_sim *s) {
  if (s->la_variant)
    for (int k = 0; k < 3; k++) s->count_liquid[k] = s->count_init[k] - s->count_crystal[k];
}
/* Its main.f:676-681: a deposition of a species whose liquid is used up
 * deposits the other one (SiGe). */
static int32_t la_swap(const orc_sim *s, int32_t itrans) {
  if (s->la_variant && s->la_swap && (itrans == 1 || itrans == 2) &&
      s->count_crystal[itrans - 1] >= s->count_init[itrans - 1])
    return itrans % 2 + 1;
  return itrans;
}

/* IndexEvent_Abs, Get_Prob.f:22-52 */
static int index_event_abs(orc_sim *s, int coor, double ptmp, double pdepo) {
  int r = 0; /* unassigned in the reference when no segment matches */
  if (s->ncov == 1) return 6;
  for (int i = 1; i <= s->ncov; i++) {
    double pmin, pmax;
    if (i == 1) {
      pmin = pdepo;
      pmax = pdepo + PABS(s, 1, coor);
    } else {
      double a = 0, b = 0;
      for (int k = 1; k <= i - 1; k++) a += PABS(s, k, coor);
      for (int k = 1; k <= i; k++) b += PABS(s, k, coor);
      pmin = pdepo + a;
      pmax = pdepo + b;
    }
    if (ptmp > pmin && ptmp <= pmax) r = 5 + i;
  }
  return r;
}

/* The 12 second-neighbour positions of a site of the diamond super-lattice
 * (the fcc shell, (+-6,+-6,0) and permutations in grid units).  No reference
 * counterpart: the engine's optional second-neighbour energy term counts
 * the occupied (non-wall) points among them. */
static const int SECOND[12][3] = {
  {6, 6, 0}, {6, -6, 0}, {-6, 6, 0}, {-6, -6, 0}, {6, 0, 6}, {6, 0, -6},
  {-6, 0, 6}, {-6, 0, -6}, {0, 6, 6}, {0, 6, -6}, {0, -6, 6}, {0, -6, -6}};
static void second_pos(const orc_sim *s, const int32_t site[3], int k, int32_t q[3]) {
  q[0] = ((site[0] + SECOND[k][0]) % s->lx + s->lx) % s->lx;
  q[1] = ((site[1] + SECOND[k][1]) % s->ly + s->ly) % s->ly;
  q[2] = site[2] + SECOND[k][2];
}
static int second_count(orc_sim *s, const int32_t site[3]) {
  int n = 0;
  for (int k = 0; k < 12; k++) {
    int32_t q[3];
    second_pos(s, site, k, q);
    if (q[2] < 0 || q[2] > s->lz) continue;
    int32_t w = s->latt[gidx(s, q[0], q[1], q[2])];
    if (ibits(w, POS_OCC, LEN_OCC) && ibits(w, POS_INDEX, LEN_INDEX) != WALL) n++;
  }
  return n;
}

/* Get_Prob, Get_Prob.f:54-331.  When pure != 0 no random number is drawn
 * and *iev receives the category code of orc_pure_rate instead. */
/* Rate-class keys of the trajectory likelihood (models without coverage
 * species): evaporation (nsic-1) + 3*(n1 + 4*(n2 + 4*n3)) + 192*m with m the
 * occupied second-neighbour positions (0 without that term); an adsorption
 * site of coordination c: KEY_ADS + c - 1; placeholder KEY_PH; isolated
 * atom KEY_ISO; -1 no rate; -2 a class the likelihood does not cover.  The
 * same numbering as the engine's mskg_adjoint.  No reference counterpart. */
enum { KEY_ADS = 2496, KEY_PH = 2499, KEY_ISO = 2500, KEY_N = 2501, ADJ_NS = 2510 };

static void get_prob_impl(orc_sim *s, int occ, int nsic, int covind, int coor, const int32_t site[3],
                          const int32_t nn[4][3], int *iev, double *prob, int pure) {
  int temperature = 0;
  s->last_key = -1;
  if (s->la_mode) {
    temperature = ibits(LC(s, site), POS_T, LEN_T);
    if (temperature == 0) fail(s, "Temperature is zero in get_prob()");
  }
  if (s->st_n) { /* Get_Prob.f:77-90 */
    const int t_st = ibits(LC(s, site), POS_T, LEN_T);
    if (t_st == 0) fail(s, "Temperature is zero in get_prob()");
    st_interp(s, t_st);
  }
  if (covind > 0 && covind < WALL) { /* desorption, :93-123 */
    if (occ != 0) fail(s, "ERROR Covind!=0 with Occ!=0 in get_prob");
    if (coor <= 3) {
      int icov = 0;
      for (int j = 1; j <= s->ncov; j++)
        if (s->list_cov[j - 1] == covind) { icov = j; break; }
      if (icov == 0) fail(s, "ERROR: could not find CovInd=%d inside ListCov", covind);
      *iev = 5;
      *prob = PDES(s, icov, coor);
      s->last_key = -2;
    } else {
      *iev = 0;
      *prob = 0.0;
    }
    return;
  }
  if (occ == 0) { /* deposition / absorption, :155-214 */
    if (coor == 1 || coor == 2 || coor == 3) {
      int ok = 1;
      if (coor == 1) {
        int coornn = ibits(LC(s, nn[0]), POS_COOR, LEN_COOR);
        ok = coornn >= 2;
      }
      if (ok) {
        double pdep = pdv(s, 1, coor, temperature) + pdv(s, 2, coor, temperature) + pdv(s, 3, coor, temperature);
        double pabs = PABS(s, 1, coor) + PABS(s, 2, coor) + PABS(s, 3, coor);
        double ptot = pdep + pabs;
        if (pure) {
          *iev = 1;
        } else {
          double pcurr = rnd(s) * ptot;
          if (pcurr <= pdep) {
            if (pcurr <= pdv(s, 1, coor, temperature)) *iev = 1;
            else if (pcurr > pdv(s, 1, coor, temperature) + pdv(s, 2, coor, temperature)) *iev = 3;
            else *iev = 2;
          } else {
            *iev = index_event_abs(s, coor, pcurr, pdep);
          }
        }
        *prob = ptot;
        s->last_key = s->ncov ? -2 : KEY_ADS + coor - 1;
      } else {
        *iev = pure ? 2 : 1;
        *prob = (double)1.e-8f; /* REAL(4) literal, Get_Prob.f:183 */
        s->last_key = KEY_PH;
      }
    } else if (coor >= 4) {
      *iev = 0;
      *prob = 0.0;
    } /* coor == 0: the reference leaves both outputs unassigned */
    return;
  }
  /* evaporation, :215-324 */
  int n_si = 0, n_c = 0, n_b = 0, n_wall = 0, covnn[3] = {0, 0, 0};
  for (int i = 0; i < 4; i++) {
    int32_t w = LC(s, nn[i]);
    int occn = ibits(w, POS_OCC, LEN_OCC), covn = ibits(w, POS_INDEX, LEN_INDEX);
    int coornn = ibits(w, POS_COOR, LEN_COOR);
    if (occn != 0) {
      if (covn != WALL) {
        int32_t ind = LI(s, nn[i]);
        const int32_t(*nnn)[3] = (coornn >= 1 && coornn <= 3) ? adr(s, ind)->nn : atr(s, ind)->nn;
        if (recip(nnn, site)) {
          int sic = ibits(w, POS_SIC, LEN_SIC);
          if (sic == 1) n_si++;
          if (sic == 2) n_c++;
          if (sic == 3) n_b++;
        }
      } else {
        n_wall++;
      }
    } else if (covn != 0) {
      int32_t ind = LI(s, nn[i]);
      if (recip(adr(s, ind)->nn, site)) {
        int icov = 0;
        for (int j = 1; j <= s->ncov; j++)
          if (s->list_cov[j - 1] == covn) { icov = j; break; }
        if (icov == 0) fail(s, "ERROR: could not find CovIndN=%d inside ListCov", covn);
        covnn[icov - 1]++;
      }
    }
  }
  int tot = n_si + n_c + n_b + n_wall;
  if (tot != coor) fail(s, "%d %d %d %d %d Get Prob Error N_Si+N_C+N_B+N_wall.NE.Coor", coor, n_si, n_c, n_b, n_wall);
  if (tot > 3) {
    *iev = 0;
    *prob = 0.0;
  } else if (tot == 0) {
    *iev = pure ? 6 : 4;
    *prob = (double)100000.0f * s->tree[1];
    s->last_key = KEY_ISO;
  } else {
    *iev = 4;
    int c1 = s->ncov >= 1 ? covnn[0] : 0, c2 = s->ncov >= 2 ? covnn[1] : 0, c3 = s->ncov >= 3 ? covnn[2] : 0;
    *prob = pev(s, nsic, n_si + n_wall, n_c, n_b, c1, c2, c3, temperature);
    int m = 0;
    if (s->use_f2) {
      m = second_count(s, site);
      *prob *= s->f2[m];
    }
    s->last_key = (s->ncov || s->la_mode) ? -2 : (nsic - 1) + 3 * (n_si + n_wall + 4 * (n_c + 4 * n_b)) + 192 * m;
  }
  /* frozen substrate: the atom stays an MC particle (so the reference's
   * list bookkeeping is unchanged) but is never picked */
  if (*iev == 4 && site[2] < s->frozen_below_z) {
    *prob = 1e-300;
    s->last_key = -1;
  }
}
static void get_prob(orc_sim *s, int occ, int nsic, int covind, int coor, const int32_t site[3],
                     const int32_t nn[4][3], int *iev, double *prob) {
  get_prob_impl(s, occ, nsic, covind, coor, site, nn, iev, prob, 0);
}

/* Get_Prob_Ini, Get_Prob_Ini.f:23-161 */
static void get_prob_ini(orc_sim *s, int occ, int nsic, int coor, const int32_t site[3], const int32_t nn[4][3],
                         int *iev, double *prob) {
  int temperature = 0;
  s->last_key = -2; /* initial rates: refreshed before any likelihood accounting */
  if (s->la_mode) {
    temperature = ibits(LC(s, site), POS_T, LEN_T);
    if (temperature == 0) fail(s, "Temperature is zero in get_prob_ini()");
  }
  if (s->st_n) { /* Get_Prob_Ini.f:42-55 */
    const int t_st = ibits(LC(s, site), POS_T, LEN_T);
    if (t_st == 0) fail(s, "Temperature is zero in get_prob_ini()");
    st_interp(s, t_st);
  }
  if (occ == 0) {
    if (coor == 1 || coor == 2 || coor == 3) {
      int ok = 1;
      if (coor == 1) ok = ibits(LC(s, nn[0]), POS_COOR, LEN_COOR) >= 2;
      if (ok) {
        double p1 = pdv(s, 1, coor, temperature), p2 = pdv(s, 2, coor, temperature), p3 = pdv(s, 3, coor, temperature);
        double ptot = p1 + p2 + p3;
        double pcurr = rnd(s) * ptot;
        if (pcurr <= p1) *iev = 1;
        else if (pcurr > p1 + p2) *iev = 3;
        else *iev = 2;
        *prob = ptot;
      } else {
        *iev = 1;
        *prob = (double)1.e-6f; /* REAL(4) literal, Get_Prob_Ini.f:104 */
      }
    } else if (coor == 4) {
      *iev = 0;
      *prob = 0.0;
    }
    return;
  }
  int n_si = 0, n_c = 0, n_b = 0, n_cov = 0;
  for (int i = 0; i < 4; i++) {
    int32_t w = LC(s, nn[i]);
    if (ibits(w, POS_INDEX, LEN_INDEX) == 0) {
      int sic = ibits(w, POS_SIC, LEN_SIC);
      if (sic == 1) n_si++;
      if (sic == 2) n_c++;
      if (sic == 3) n_b++;
    } else {
      n_cov++;
    }
  }
  int tot = n_si + n_c + n_b + n_cov;
  if (tot != coor) fail(s, "Get Prob Error N_Si+N_C++N_B+N_cov.NE.Coor");
  if (tot > 3) {
    *iev = 0;
    *prob = 0.0;
  } else if (tot == 0) {
    *iev = 4;
    *prob = (double)100000.0f * s->tree[1];
  } else {
    *iev = 4;
    *prob = pev(s, nsic, n_si + n_cov, n_c, n_b, 0, 0, 0, temperature);
  }
}

/* ---- events ------------------------------------------------------------ */
/* Deposition, mcevent/deposition.f:24-251 */
static void deposition(orc_sim *s, int32_t ind_void, int iatom) {
  int32_t site[3], nn_site[4][3];
  cp3(site, s->ad[ind_void].xyz);
  int32_t *ws = LCp(s, site);
  int coor = ibits(*ws, POS_COOR, LEN_COOR), coor_old = coor;
  mvbits(1, LEN_OCC, ws, POS_OCC);
  mvbits(iatom, LEN_SIC, ws, POS_SIC);
  s->last_arm = -1;
  if (coor == 1) { /* :49-74 */
    int32_t sitec[3], sn[3][3] = {{0}}, zig[3][3], arm[3][3];
    cp3(sitec, s->ad[ind_void].nn[0]);
    int32_t indnn = LI(s, sitec);
    const mc_rec *c = adr(s, indnn);
    int j = 0;
    for (int i = 0; i < 4; i++)
      if (!eq3(c->nn[i], site)) {
        if (j < 3) cp3(sn[j], c->nn[i]); /* the reference writes past SNSite(3,3) for j=4 */
        j++;
      }
    orc_find_sn(site, sitec, (const int32_t(*)[3])sn, zig, arm, s->lx, s->ly);
    int is_arm = if_arm_choice(s, zig, arm);
    s->last_arm = is_arm;
    for (int i = 0; i < 3; i++) {
      cp3(s->ad[ind_void].nn[i + 1], is_arm ? arm[i] : zig[i]);
      if (!is_bond_disp(s, site, is_arm ? arm[i] : zig[i])) s->nonbond_arm++;
    }
  }
  memcpy(nn_site, s->ad[ind_void].nn, sizeof nn_site);
  for (int i = 0; i < 4; i++) { /* :78-235 */
    int32_t *p = nn_site[i];
    int32_t w = LC(s, p);
    int occ = ibits(w, POS_OCC, LEN_OCC);
    int32_t indnn = LI(s, p);
    int coornn = ibits(w, POS_COOR, LEN_COOR);
    int covind = ibits(w, POS_INDEX, LEN_INDEX);
    int32_t nnn[4][3];
    memset(nnn, 0, sizeof nnn);
    int iev = 0;
    double prob = 0.0;
    if (occ == 0) {
      if (coornn == 0) {
        coornn = 1;
        mvbits(coornn, LEN_COOR, LCp(s, p), POS_COOR);
        cp3(nnn[0], site);
        get_prob(s, 0, 0, covind, coornn, p, nnn, &iev, &prob);
        if (iev != 0) add_mc(s, p, nnn, iev, prob);
      } else if (coornn == 1) {
        memcpy(nnn, adr(s, indnn)->nn, sizeof nnn);
        int dx = abs(nnn[0][0] - site[0]);
        dx = dx < s->lx - dx ? dx : s->lx - dx;
        int dy = abs(nnn[0][1] - site[1]);
        dy = dy < s->ly - dy ? dy : s->ly - dy;
        int dz = nnn[0][2] - site[2];
        if (dx * dx + dy * dy + dz * dz == 72) {
          int32_t ns[2][3];
          coornn = 2;
          cp3(nnn[1], site);
          orc_find_nn(nnn[0], nnn[1], p, ns, s->lx, s->ly);
          if (!is_bond_disp(s, p, ns[0]) || !is_bond_disp(s, p, ns[1])) s->nonbond_nn++;
          for (int j = 2; j < 4; j++) {
            cp3(nnn[j], ns[j - 2]);
            int32_t w3 = LC(s, nnn[j]);
            if (ibits(w3, POS_OCC, LEN_OCC) != 0) {
              int c3 = ibits(w3, POS_COOR, LEN_COOR), cv3 = ibits(w3, POS_INDEX, LEN_INDEX);
              if (c3 <= 3 && cv3 != WALL) {
                if (recip(adr(s, LI(s, nnn[j]))->nn, p)) coornn++;
              }
            }
          }
          mvbits(coornn, LEN_COOR, LCp(s, p), POS_COOR);
          memcpy(s->ad[indnn].nn, nnn, sizeof nnn);
          get_prob(s, 0, 0, covind, coornn, p, nnn, &iev, &prob);
          if (iev == 0) {
            erase_mc(s, p);
            add_void(s, p, nnn);
          } else {
            s->ad[indnn].ev = iev;
            s->ad[indnn].prob = prob; s->ad[indnn].key = s->last_key;
            update_tree(s, indnn, prob);
          }
        }
      } else if (coornn == 2 || coornn == 3) {
        memcpy(nnn, adr(s, indnn)->nn, sizeof nnn);
        if (recip(nnn, site)) {
          coornn++;
          mvbits(coornn, LEN_COOR, LCp(s, p), POS_COOR);
          get_prob(s, 0, 0, covind, coornn, p, nnn, &iev, &prob);
          if (iev == 0) {
            erase_mc(s, p);
            add_void(s, p, nnn);
          } else {
            s->ad[indnn].ev = iev;
            s->ad[indnn].prob = prob; s->ad[indnn].key = s->last_key;
            update_tree(s, indnn, prob);
          }
        }
      }
    } else {
      if (covind != WALL) {
        if (covind != 0) fail(s, "ERROR in deposition: NN with CovInd not in {0,113} and Occ=1");
        if (coornn <= 3) {
          memcpy(nnn, adr(s, indnn)->nn, sizeof nnn);
          if (recip(nnn, site)) {
            coornn++;
            mvbits(coornn, LEN_COOR, LCp(s, p), POS_COOR);
            int sic = ibits(LC(s, p), POS_SIC, LEN_SIC);
            get_prob(s, 1, sic, covind, coornn, p, nnn, &iev, &prob);
            if (iev == 0) {
              erase_mc(s, p);
              add_bulk(s, p, nnn);
            } else {
              s->ad[indnn].ev = iev;
              s->ad[indnn].prob = prob; s->ad[indnn].key = s->last_key;
              update_tree(s, indnn, prob);
            }
            if (coor_old == 1 && i >= 1) coor++;
          }
        }
      } else {
        coor++;
      }
    }
  }
  ind_void = LI(s, site); /* :236 */
  if (coor != coor_old) mvbits(coor, LEN_COOR, LCp(s, site), POS_COOR);
  int iev = 0;
  double prob = 0.0;
  get_prob(s, 1, iatom, 0, coor, site, nn_site, &iev, &prob);
  if (iev == 0) {
    erase_mc(s, site);
    add_bulk(s, site, nn_site);
  } else {
    s->ad[ind_void].ev = iev;
    s->ad[ind_void].prob = prob; s->ad[ind_void].key = s->last_key;
    update_tree(s, ind_void, prob);
  }
}

/* Evaporation, mcevent/Evaporation.f:25-262 */
static void evaporation(orc_sim *s, int32_t ind_atom) {
  int32_t site[3], nn_site[4][3], nbuff[3] = {0, 0, 0};
  cp3(site, s->ad[ind_atom].xyz);
  int32_t *ws = LCp(s, site);
  int iatom = ibits(*ws, POS_SIC, LEN_SIC);
  if (iatom >= 1 && iatom <= 3) s->count_crystal[iatom - 1]--;
  int coor = ibits(*ws, POS_COOR, LEN_COOR);
  mvbits(0, LEN_OCC, ws, POS_OCC);
  mvbits(0, LEN_SIC, ws, POS_SIC);
  memcpy(nn_site, s->ad[ind_atom].nn, sizeof nn_site);
  for (int i = 0; i < 4; i++) {
    int32_t *p = nn_site[i];
    int32_t w = LC(s, p);
    int occ = ibits(w, POS_OCC, LEN_OCC), coornn = ibits(w, POS_COOR, LEN_COOR);
    int32_t indnn = LI(s, p);
    int covind = ibits(w, POS_INDEX, LEN_INDEX);
    int32_t nnn[4][3];
    memset(nnn, 0, sizeof nnn);
    int iev = 0;
    double prob = 0.0;
    if (occ == 0) {
      if (coornn == 1) {
        memcpy(nnn, adr(s, indnn)->nn, sizeof nnn);
        if (recip(nnn, site)) {
          erase_mc(s, p);
          int32_t *wp = LCp(s, p);
          if (s->la_mode || s->st_n || (s->init_st[0] == 'S' && s->init_st[1] == 'T')) {
            int t = ibits(*wp, POS_T, LEN_T);
            *wp = 0;
            mvbits(t, LEN_T, wp, POS_T);
          } else {
            *wp = 0;
          }
        }
      } else if (coornn == 2) {
        memcpy(nnn, adr(s, indnn)->nn, sizeof nnn);
        if (recip(nnn, site)) {
          coornn = 1;
          mvbits(coornn, LEN_COOR, LCp(s, p), POS_COOR);
          for (int j = 0; j < 4; j++) {
            int32_t w3 = LC(s, nnn[j]);
            if (ibits(w3, POS_OCC, LEN_OCC) == 0) {
              memset(s->ad[indnn].nn[j], 0, sizeof(int32_t) * 3);
            } else {
              int c3 = ibits(w3, POS_COOR, LEN_COOR), cv3 = ibits(w3, POS_INDEX, LEN_INDEX);
              if (c3 <= 3 && cv3 != WALL) {
                if (recip(adr(s, LI(s, nnn[j]))->nn, p)) cp3(nbuff, s->ad[indnn].nn[j]);
                memset(s->ad[indnn].nn[j], 0, sizeof(int32_t) * 3);
              } else {
                memset(s->ad[indnn].nn[j], 0, sizeof(int32_t) * 3);
              }
            }
          }
          cp3(s->ad[indnn].nn[0], nbuff);
          get_prob(s, 0, 0, covind, coornn, p, nnn, &iev, &prob); /* old list, :121 */
          s->ad[indnn].ev = iev;
          s->ad[indnn].prob = prob; s->ad[indnn].key = s->last_key;
          update_tree(s, indnn, prob);
        }
      } else if (coornn == 3) {
        memcpy(nnn, adr(s, indnn)->nn, sizeof nnn);
        if (recip(nnn, site)) {
          coornn = 2;
          mvbits(coornn, LEN_COOR, LCp(s, p), POS_COOR);
          get_prob(s, 0, 0, covind, coornn, p, nnn, &iev, &prob);
          s->ad[indnn].ev = iev;
          s->ad[indnn].prob = prob; s->ad[indnn].key = s->last_key;
          update_tree(s, indnn, prob);
        }
      } else if (coornn == 4) {
        memcpy(nnn, vdr(s, indnn)->nn, sizeof nnn);
        if (recip(nnn, site)) {
          coornn = 3;
          mvbits(coornn, LEN_COOR, LCp(s, p), POS_COOR);
          get_prob(s, 0, 0, covind, coornn, p, nnn, &iev, &prob);
          erase_void(s, p);
          add_mc(s, p, nnn, iev, prob);
        }
      }
    } else {
      if (covind != WALL) {
        if (covind != 0) fail(s, "ERROR in evaporation: NN with CovInd not in {0,113} and Occ=1");
        if (coornn == 4) {
          memcpy(nnn, atr(s, indnn)->nn, sizeof nnn);
          if (recip(nnn, site)) {
            coornn = 3;
            mvbits(coornn, LEN_COOR, LCp(s, p), POS_COOR);
            int sic = ibits(LC(s, p), POS_SIC, LEN_SIC);
            get_prob(s, 1, sic, covind, coornn, p, nnn, &iev, &prob);
            erase_bulk(s, p);
            add_mc(s, p, nnn, iev, prob);
          }
        } else {
          memcpy(nnn, adr(s, indnn)->nn, sizeof nnn);
          if (recip(nnn, site)) {
            coornn--;
            mvbits(coornn, LEN_COOR, LCp(s, p), POS_COOR);
            int sic = ibits(LC(s, p), POS_SIC, LEN_SIC);
            get_prob(s, 1, sic, covind, coornn, p, nnn, &iev, &prob);
            s->ad[indnn].ev = iev;
            s->ad[indnn].prob = prob; s->ad[indnn].key = s->last_key;
            update_tree(s, indnn, prob);
          }
        }
      } else {
        coor--;
      }
    }
  }
  ind_atom = LI(s, site);
  mvbits(coor, LEN_COOR, LCp(s, site), POS_COOR);
  if (coor == 0) {
    erase_mc(s, site);
  } else {
    int iev = 0;
    double prob = 0.0;
    get_prob(s, 0, 0, 0, coor, site, nn_site, &iev, &prob);
    s->ad[ind_atom].ev = iev;
    s->ad[ind_atom].prob = prob; s->ad[ind_atom].key = s->last_key;
    if (coor == 1) {
      int32_t cur[4][3];
      memcpy(cur, s->ad[ind_atom].nn, sizeof cur);
      for (int i = 0; i < 4; i++) {
        int32_t w = LC(s, cur[i]);
        if (ibits(w, POS_OCC, LEN_OCC) == 0) {
          memset(s->ad[ind_atom].nn[i], 0, sizeof(int32_t) * 3);
        } else {
          int cnn = ibits(w, POS_COOR, LEN_COOR), cv = ibits(w, POS_INDEX, LEN_INDEX);
          if (cnn <= 3 && cv != WALL) {
            if (recip(adr(s, LI(s, cur[i]))->nn, site)) cp3(nbuff, s->ad[ind_atom].nn[i]);
            memset(s->ad[ind_atom].nn[i], 0, sizeof(int32_t) * 3);
          } else {
            memset(s->ad[ind_atom].nn[i], 0, sizeof(int32_t) * 3);
          }
        }
      }
      cp3(s->ad[ind_atom].nn[0], nbuff);
    }
    update_tree(s, ind_atom, prob);
  }
}

/* absorption, mcevent/absorption.f:25-78 */
static void absorption(orc_sim *s, int32_t ind_void, int covind_new) {
  int32_t site[3], nn_site[4][3];
  cp3(site, s->ad[ind_void].xyz);
  int32_t *ws = LCp(s, site);
  int coor = ibits(*ws, POS_COOR, LEN_COOR);
  mvbits(covind_new, LEN_INDEX, ws, POS_INDEX);
  memcpy(nn_site, s->ad[ind_void].nn, sizeof nn_site);
  for (int i = 0; i < 4; i++) {
    int32_t *p = nn_site[i];
    int32_t w = LC(s, p);
    int32_t indnn = LI(s, p);
    int coorn = ibits(w, POS_COOR, LEN_COOR);
    if (ibits(w, POS_OCC, LEN_OCC) != 0) {
      int covn = ibits(w, POS_INDEX, LEN_INDEX);
      if (covn != WALL && coorn <= 3) {
        int32_t nnn[4][3];
        memcpy(nnn, adr(s, indnn)->nn, sizeof nnn);
        if (recip(nnn, site)) {
          int sic = ibits(w, POS_SIC, LEN_SIC), iev = 0;
          double prob = 0.0;
          get_prob(s, 1, sic, covn, coorn, p, nnn, &iev, &prob);
          s->ad[indnn].ev = iev;
          s->ad[indnn].prob = prob; s->ad[indnn].key = s->last_key;
          update_tree(s, indnn, prob);
        }
      }
    }
  }
  int iev = 0;
  double prob = 0.0;
  get_prob(s, 0, 0, covind_new, coor, site, nn_site, &iev, &prob);
  s->ad[ind_void].ev = iev;
  s->ad[ind_void].prob = prob; s->ad[ind_void].key = s->last_key;
  update_tree(s, ind_void, prob);
}

/* desorption, mcevent/desorption.f:25-117 */
static void desorption(orc_sim *s, int32_t ind_atom) {
  int32_t site[3], nn_site[4][3];
  cp3(site, s->ad[ind_atom].xyz);
  int32_t *ws = LCp(s, site);
  int covind = ibits(*ws, POS_INDEX, LEN_INDEX);
  if (covind == s->list_cov[0]) s->count_cov[0]--;
  else if (covind == s->list_cov[1]) s->count_cov[1]--;
  else if (covind == s->list_cov[2]) s->count_cov[2]--;
  int coor = ibits(*ws, POS_COOR, LEN_COOR);
  mvbits(0, LEN_INDEX, ws, POS_INDEX);
  memcpy(nn_site, s->ad[ind_atom].nn, sizeof nn_site);
  int nloop = (coor == 1) ? 1 : 4;
  for (int i = 0; i < nloop; i++) {
    int32_t *p = nn_site[i];
    int32_t w = LC(s, p);
    int32_t indnn = LI(s, p);
    int coorn = ibits(w, POS_COOR, LEN_COOR);
    if (ibits(w, POS_OCC, LEN_OCC) != 0) {
      int covn = ibits(w, POS_INDEX, LEN_INDEX);
      if (covn != WALL) {
        int32_t nnn[4][3];
        memcpy(nnn, adr(s, indnn)->nn, sizeof nnn);
        if (recip(nnn, site)) {
          int sic = ibits(w, POS_SIC, LEN_SIC), iev = 0;
          double prob = 0.0;
          get_prob(s, 1, sic, covn, coorn, p, nnn, &iev, &prob);
          s->ad[indnn].ev = iev;
          s->ad[indnn].prob = prob; s->ad[indnn].key = s->last_key;
          update_tree(s, indnn, prob);
        }
      }
    }
  }
  int iev = 0;
  double prob = 0.0;
  get_prob(s, 0, 0, 0, coor, site, nn_site, &iev, &prob);
  s->ad[ind_atom].ev = iev;
  s->ad[ind_atom].prob = prob; s->ad[ind_atom].key = s->last_key;
  update_tree(s, ind_atom, prob);
}

/* ---- initial builders (mulskips-source/initial) ----------------------- */
/* Diamond site classification used by every builder: grid coordinates
 * that are multiples of 3 with x/3, y/3, z/3 of equal parity; sum%4==0 is
 * sublattice A (bonds JumpDia(:,1:4)), sum%4==3 sublattice B
 * (JumpDia(:,5:8)).  Fortran MOD truncates like C's %. */
static int dia_type(int x, int y, int z) {
  if (x % 3 || y % 3 || z % 3) return 0;
  int x1 = x / 3, y1 = y / 3, z1 = z / 3;
  if (!((x1 % 2) == (y1 % 2) && (x1 % 2) == (z1 % 2))) return 0;
  int m = (x1 + y1 + z1) % 4;
  return m == 0 ? 1 : (m == 3 ? 2 : 0);
}
static void dia_nn(const orc_sim *s, int type, const int32_t site[3], int32_t nn[4][3]) {
  for (int i = 0; i < 4; i++) {
    const int *j = JUMP_DIA[(type == 1 ? 0 : 4) + i];
    for (int k = 0; k < 3; k++) nn[i][k] = site[k] + j[k];
    if (nn[i][0] >= s->lx) nn[i][0] -= s->lx;
    if (nn[i][0] < 0) nn[i][0] += s->lx;
    if (nn[i][1] >= s->ly) nn[i][1] -= s->ly;
    if (nn[i][1] < 0) nn[i][1] += s->ly;
  }
}
static void set_wall(orc_sim *s, int z0, int z1) {
  for (int z = z0; z <= z1; z++)
    for (int y = 0; y < s->ly; y++)
      for (int x = 0; x < s->lx; x++) {
        int32_t *w = &s->latt[gidx(s, x, y, z)];
        *w |= (1 << POS_OCC);
        mvbits(WALL, LEN_INDEX, w, POS_INDEX);
      }
}

/* Flat (001) slab: SetSi.f:25-220 (species=0: Si everywhere, walls) and
 * SetSiC3C.f:25-198 (species=1: Si on A, C on B, no walls). */
static void set_flat(orc_sim *s, int len, int sic) {
  int32_t site[3], nn[4][3];
  int zb = 0;
  if (!sic) { set_wall(s, 0, 5); zb = 6; }
  for (int z = zb; z <= len - 4; z++)
    for (int y = 0; y < s->ly; y++)
      for (int x = 0; x < s->lx; x++) {
        int t = dia_type(x, y, z);
        if (!t) continue;
        int32_t *w = &s->latt[gidx(s, x, y, z)];
        *w |= (1 << ((sic && t == 2) ? POS_C : POS_SI));
        *w |= (1 << POS_OCC);
        mvbits(4, LEN_COOR, w, POS_COOR);
        site[0] = x; site[1] = y; site[2] = z;
        dia_nn(s, t, site, nn);
        add_bulk(s, site, nn);
      }
  for (int z = len - 3; z <= len - 1; z++)
    for (int y = 0; y < s->ly; y++)
      for (int x = 0; x < s->lx; x++) {
        int t = dia_type(x, y, z);
        if (!t) continue;
        int32_t *w = &s->latt[gidx(s, x, y, z)];
        int sp = (sic && t == 2) ? 2 : 1;
        *w |= (1 << (sp == 2 ? POS_C : POS_SI));
        *w |= (1 << POS_OCC);
        mvbits(2, LEN_COOR, w, POS_COOR);
        site[0] = x; site[1] = y; site[2] = z;
        dia_nn(s, t, site, nn);
        s->last_key = -2;
        add_mc(s, site, nn, 4, PE(s, sp, 2, 0, 0, 0, 0, 0));
      }
  for (int y = 0; y < s->ly; y++)
    for (int x = 0; x < s->lx; x++) {
      int z = len;
      int t = dia_type(x, y, z);
      if (!t) continue;
      mvbits(2, LEN_COOR, &s->latt[gidx(s, x, y, z)], POS_COOR);
      site[0] = x; site[1] = y; site[2] = z;
      dia_nn(s, t, site, nn);
      int iev = 0;
      double prob = 0.0;
      get_prob(s, 0, 0, 0, 2, site, nn, &iev, &prob);
      add_mc(s, site, nn, iev, prob);
    }
  if (!sic) set_wall(s, s->lz - 6, s->lz - 1);
}

/* Coordination then rate pass shared by SetNC.f:60-232, SetNCSp.f,
 * SetInvPOff.f:113-284 and SetSiFINFET.f (walls=1). */
static void coord_pass(orc_sim *s, int x0, int x1, int y0, int y1, int z0, int z1, int walls) {
  int32_t site[3], nn[4][3];
  for (int z = z0; z <= z1; z++)
    for (int y = y0; y <= y1; y++)
      for (int x = x0; x <= x1; x++) {
        if (walls && ibits(s->latt[gidx(s, x, y, z)], POS_INDEX, LEN_INDEX) == WALL) continue;
        int t = dia_type(x, y, z);
        if (!t) continue;
        site[0] = x; site[1] = y; site[2] = z;
        dia_nn(s, t, site, nn);
        int occ = ibits(LC(s, site), POS_OCC, LEN_OCC);
        int coor = 0;
        for (int i = 0; i < 4; i++) {
          int32_t w = LC(s, nn[i]);
          int occn = ibits(w, POS_OCC, LEN_OCC);
          if (walls && occ == 0) {
            if (occn > 0 && ibits(w, POS_INDEX, LEN_INDEX) != WALL) coor++;
          } else if (occn > 0) {
            coor++;
          }
        }
        if (occ != 0 && coor == 0) {
          *LCp(s, site) = 0;
          for (int i = 0; i < 4; i++) {
            int32_t *wn = LCp(s, nn[i]);
            int c = ibits(*wn, POS_COOR, LEN_COOR) - 1;
            if (c < 0) c = 0;
            mvbits(c, LEN_COOR, wn, POS_COOR);
          }
        } else {
          mvbits(coor, LEN_COOR, LCp(s, site), POS_COOR);
        }
      }
}
static void prob_pass_v(orc_sim *s, int x0, int x1, int y0, int y1, int z0, int z1, int walls, int voids);
static void prob_pass(orc_sim *s, int x0, int x1, int y0, int y1, int z0, int z1, int walls) {
  prob_pass_v(s, x0, x1, y0, y1, z0, z1, walls, 0);
}
/* voids=1: an empty site of coordination 4 joins the void list (SetCAD.f:330-332) */
static void prob_pass_v(orc_sim *s, int x0, int x1, int y0, int y1, int z0, int z1, int walls, int voids) {
  int32_t site[3], nn[4][3], nbuff[3] = {0, 0, 0};
  for (int z = z0; z <= z1; z++)
    for (int y = y0; y <= y1; y++)
      for (int x = x0; x <= x1; x++) {
        if (walls && ibits(s->latt[gidx(s, x, y, z)], POS_INDEX, LEN_INDEX) == WALL) continue;
        int t = dia_type(x, y, z);
        if (!t) continue;
        site[0] = x; site[1] = y; site[2] = z;
        dia_nn(s, t, site, nn);
        int32_t w = LC(s, site);
        int occ = ibits(w, POS_OCC, LEN_OCC), coor = ibits(w, POS_COOR, LEN_COOR);
        int iev = 0;
        double prob = 0.0;
        if (occ == 0) {
          if (coor > 0) {
            if (coor == 1) {
              for (int i = 0; i < 4; i++) {
                int32_t wn = LC(s, nn[i]);
                int drop = ibits(wn, POS_OCC, LEN_OCC) == 0 ||
                           (walls && ibits(wn, POS_INDEX, LEN_INDEX) == WALL);
                if (!drop) cp3(nbuff, nn[i]);
                memset(nn[i], 0, sizeof(int32_t) * 3);
              }
              cp3(nn[0], nbuff);
            }
            if (voids && coor == 4) {
              add_void(s, site, nn);
            } else {
              get_prob_ini(s, 0, 0, coor, site, nn, &iev, &prob);
              add_mc(s, site, nn, iev, prob);
            }
          }
        } else {
          if (coor <= 3) {
            get_prob_ini(s, 1, ibits(w, POS_SIC, LEN_SIC), coor, site, nn, &iev, &prob);
            add_mc(s, site, nn, iev, prob);
          } else {
            add_bulk(s, site, nn);
          }
        }
      }
}

/* SetNC.f:25-235 (shape 0: box) and SetNCSp.f:25-237 (shape 1: sphere). */
static void set_nc(orc_sim *s, int l1, int l2, int l3, int sphere) {
  int cx = s->lx / 2, cy = s->ly / 2, cz = s->lz / 2;
  for (int z = cz - l3 / 2; z <= cz + l3 / 2 - 1; z++)
    for (int y = cy - l2 / 2; y <= cy + l2 / 2 - 1; y++)
      for (int x = cx - l1 / 2; x <= cx + l1 / 2 - 1; x++) {
        if (sphere) {
          int d2 = (x - cx) * (x - cx) + (y - cy) * (y - cy) + (z - cz) * (z - cz);
          if (!(d2 <= l1 * l1 / 4)) continue;
        }
        int t = dia_type(x, y, z);
        if (!t) continue;
        int32_t *w = &s->latt[gidx(s, x, y, z)];
        *w |= (1 << (t == 1 ? POS_SI : POS_C));
        *w |= (1 << POS_OCC);
      }
  coord_pass(s, cx - l1 / 2 - 10, cx + l1 / 2 + 9, cy - l2 / 2 - 10, cy + l2 / 2 + 9, cz - l3 / 2 - 10,
             cz + l3 / 2 + 9, 0);
  prob_pass(s, cx - l1 / 2 - 10, cx + l1 / 2 + 9, cy - l2 / 2 - 10, cy + l2 / 2 + 9, cz - l3 / 2 - 10,
            cz + l3 / 2 + 9, 0);
}

/* SetInvPOff.f:21-286 */
static void set_invp_off(orc_sim *s, int len, int len3) {
  int32_t site[3], nn[4][3];
  int lenh = s->lx / 2 - 1;
  for (int z = 0; z <= 20; z++)
    for (int y = 0; y < s->ly; y++)
      for (int x = 0; x < s->lx; x++) {
        int t = dia_type(x, y, z);
        if (!t) continue;
        int32_t *w = &s->latt[gidx(s, x, y, z)];
        *w |= (1 << (t == 1 ? POS_SI : POS_C));
        *w |= (1 << POS_OCC);
        mvbits(4, LEN_COOR, w, POS_COOR);
        site[0] = x; site[1] = y; site[2] = z;
        dia_nn(s, t, site, nn);
        add_bulk(s, site, nn);
      }
  for (int z = 21; z <= len3 - 1; z++)
    for (int y = 0; y < s->ly; y++)
      for (int x = 0; x < s->lx; x++) {
        int t = dia_type(x, y, z);
        if (!t) continue;
        /* single-precision pyramid test: integer sums minus 1.10*(z-Len3)
         * in REAL(4), compared against Len converted to REAL(4). */
        float dz = 1.10f * (float)(z - len3), fl = (float)len;
        int in = ((float)(abs(x - lenh) + abs(y - lenh)) - dz <= fl) ||
                 ((float)(abs(x) + abs(y)) - dz <= fl) ||
                 ((float)(abs(x - s->lx) + abs(y)) - dz <= fl) ||
                 ((float)(abs(x) + abs(y - s->ly)) - dz <= fl) ||
                 ((float)(abs(x - s->lx) + abs(y - s->ly)) - dz <= fl);
        if (in) continue;
        int32_t *w = &s->latt[gidx(s, x, y, z)];
        *w |= (1 << (t == 1 ? POS_SI : POS_C));
        *w |= (1 << POS_OCC);
      }
  coord_pass(s, 0, s->lx - 1, 0, s->ly - 1, 21, len3 + 9, 0);
  prob_pass(s, 0, s->lx - 1, 0, s->ly - 1, 21, len3 + 9, 0);
}

/* SetSiFINFET.f:25-312 */
static void set_si_finfet(orc_sim *s, int len, int lenww, int lenwwtop, int lap) {
  set_wall(s, 0, 5);
  int xa = (int)(0.5f * (float)(s->lx - lap)), xb = (int)(0.5f * (float)(s->lx + lap));
  for (int z = lenww; z <= s->lz - 7; z++)
    for (int y = 0; y < s->ly; y++) {
      for (int x = 0; x <= xa - 1; x++) {
        int32_t *w = &s->latt[gidx(s, x, y, z)];
        *w |= (1 << POS_OCC);
        mvbits(WALL, LEN_INDEX, w, POS_INDEX);
      }
      for (int x = xb; x <= s->lx - 1; x++) {
        int32_t *w = &s->latt[gidx(s, x, y, z)];
        *w |= (1 << POS_OCC);
        mvbits(WALL, LEN_INDEX, w, POS_INDEX);
      }
    }
  set_wall(s, lenwwtop, s->lz - 7);
  set_wall(s, s->lz - 6, s->lz - 1);
  for (int z = 6; z <= len - 1; z++)
    for (int y = 0; y < s->ly; y++)
      for (int x = 0; x < s->lx; x++) {
        int32_t *w = &s->latt[gidx(s, x, y, z)];
        if (ibits(*w, POS_INDEX, LEN_INDEX) == WALL) continue;
        if (!dia_type(x, y, z)) continue;
        *w |= (1 << POS_SI);
        *w |= (1 << POS_OCC);
      }
  coord_pass(s, 0, s->lx - 1, 0, s->ly - 1, 6, lenwwtop + 9, 1);
  prob_pass(s, 0, s->lx - 1, 0, s->ly - 1, 6, lenwwtop + 9, 1);
}

/* SetCAD.f:84-418 for InitSt 'IN': wall planes z = 0..3 and LenZ-4..LenZ-1
 * on every grid point (:84-104); the geometry ids of planes z = 4..LenZ-5
 * (:106-190), ids[x + LenX*(y + LenY*z)]: 10 a wall on that grid point, 1 an
 * atom on a diamond site, of species 1, or of species 2 where species2[] is
 * set (the reference draws RAND() < alloyfraction there; NULL: all species
 * 1, alloyfraction 0); then the coordination pass over the whole box
 * (:194-303, isolated atoms removed) and the probability pass (:309-416),
 * which puts an empty site of coordination 4 in the void list. */
static void set_cad(orc_sim *s, const int8_t *ids, const int8_t *species2) {
  set_wall(s, 0, 3);
  set_wall(s, s->lz - 4, s->lz - 1);
  for (int z = 4; z <= s->lz - 5; z++)
    for (int y = 0; y < s->ly; y++)
      for (int x = 0; x < s->lx; x++) {
        const int64_t g = gidx(s, x, y, z);
        const int id = ids[g];
        int32_t *w = &s->latt[g];
        if (id == 10) {
          *w |= (1 << POS_OCC);
          mvbits(WALL, LEN_INDEX, w, POS_INDEX);
        } else if (id == 1 && dia_type(x, y, z)) {
          *w |= (1 << POS_OCC);
          *w |= (1 << ((species2 && species2[g]) ? POS_C : POS_SI));
        }
      }
  coord_pass(s, 0, s->lx - 1, 0, s->ly - 1, 0, s->lz - 1, 1);
  prob_pass_v(s, 0, s->lx - 1, 0, s->ly - 1, 0, s->lz - 1, 1, 1);
}

/* ---- start.dat (main.f:83-212, SetProbability.f:21-143) -------------- */
typedef struct { FILE *f; char line[4096]; } rdr;
static char *next_line(orc_sim *s, rdr *r) {
  if (!fgets(r->line, sizeof r->line, r->f)) fail(s, "start.dat: unexpected end of file");
  return r->line;
}
/* list-directed read of n integers from one record */
static void read_ints(orc_sim *s, rdr *r, int n, int64_t *v) {
  char *p = next_line(s, r), *e;
  for (int i = 0; i < n; i++) {
    while (*p == ' ' || *p == '\t' || *p == ',') p++;
    v[i] = strtoll(p, &e, 10);
    if (e == p) fail(s, "start.dat: integer expected in '%s'", r->line);
    p = e;
  }
}
static void read_reals(orc_sim *s, rdr *r, int n, double *v) {
  char *p = next_line(s, r), *e;
  for (int i = 0; i < n; i++) {
    while (*p == ' ' || *p == '\t' || *p == ',') p++;
    char tmp[128];
    int k = 0;
    while (p[k] && p[k] != ' ' && p[k] != '\t' && p[k] != ',' && p[k] != '\n' && p[k] != '\r' && k < 127) {
      tmp[k] = (p[k] == 'd' || p[k] == 'D') ? 'e' : p[k];
      k++;
    }
    tmp[k] = 0;
    v[i] = strtod(tmp, &e);
    if (e == tmp) fail(s, "start.dat: real expected in '%s'", r->line);
    p += k;
  }
}
static void read_word(orc_sim *s, rdr *r, char *out, int len) {
  char *p = next_line(s, r);
  while (*p == ' ' || *p == '\t') p++;
  int k = 0;
  while (p[k] && p[k] != ' ' && p[k] != '\t' && p[k] != ',' && p[k] != '\n' && p[k] != '\r' && k < len) {
    out[k] = p[k];
    k++;
  }
  out[k] = 0;
}

int orc_load_deck(orc_sim *s, const char *path) {
  rdr r = {0};
  if (setjmp(s->jb)) { s->have_jb = 0; if (r.f) fclose(r.f); return 2; }
  s->have_jb = 1;
  r.f = fopen(path, "r");
  if (!r.f) fail(s, "cannot open %s", path);
  int64_t v[8];
  read_ints(s, &r, 2, v);
  s->ncrystal = (int)v[0];
  s->ncov = (int)v[1];
  if (s->ncrystal > 3 || s->ncrystal < 1) fail(s, "ERR NOT IMPLEMENTED: NCrystal must be <=3");
  read_ints(s, &r, s->ncrystal, v);
  for (int i = 0; i < s->ncrystal; i++) s->list_crystal[i] = (int)v[i];
  if (s->ncov > 0 && s->ncov <= 3) {
    read_ints(s, &r, s->ncov, v);
    for (int i = 0; i < s->ncov; i++) s->list_cov[i] = (int)v[i];
  } else if (s->ncov != 0) {
    fail(s, "ERR NOT IMPLEMENTED: NCov must be <=3");
  }
  read_word(s, &r, s->init_st, 2);
  const char *is = s->init_st;
  if (!strcmp(is, "FF")) {
    read_ints(s, &r, 4, v);
    s->len1 = (int)v[0]; s->len2 = (int)v[1]; s->len3 = (int)v[2]; s->len4 = (int)v[3];
  } else if (!strcmp(is, "ST") || !strcmp(is, "IN") || !strcmp(is, "SH") || !strcmp(is, "LA")) {
    fail(s, "InitSt '%s' is not restated in the oracle", is);
  } else {
    read_ints(s, &r, 3, v);
    s->len1 = (int)v[0]; s->len2 = (int)v[1]; s->len3 = (int)v[2];
  }
  /* main.f:142-143 compares ListCrystal(:NCrystal) with (/14,32/): only a
   * two- or three-species list starting 14,32 carries an alloy line. */
  if ((s->ncrystal == 2 && s->list_crystal[0] == 14 && s->list_crystal[1] == 32) ||
      (s->ncrystal == 3 && s->list_crystal[0] == 14 && s->list_crystal[1] == 32 && s->list_crystal[2] == 5)) {
    double a;
    read_reals(s, &r, 1, &a);
    s->alloy = a;
  }
  read_reals(s, &r, 1, &s->ptrans_zig);
  read_reals(s, &r, 1, &s->sf);
  read_word(s, &r, s->exit_strategy, 4);
  if (!strcmp(s->exit_strategy, "Iter")) {
    read_ints(s, &r, 1, v); s->max_iter = v[0];
    read_ints(s, &r, 1, v); s->out_molmol = v[0];
  } else if (!strcmp(s->exit_strategy, "Time")) {
    read_reals(s, &r, 1, &s->max_time);
    read_reals(s, &r, 1, &s->out_time);
  }
  read_word(s, &r, s->run_type, 2);
  read_ints(s, &r, 1, v);
  s->idum = (int32_t)v[0];
  char flag[8];
  read_word(s, &r, flag, 2);
  if (!strcmp(flag, "T")) next_line(s, &r);
  read_word(s, &r, flag, 2);
  if (!strcmp(flag, "T")) next_line(s, &r);
  /* SetProbability.f:40-140 */
  static const int max_e[4][3] = {{3, 18, 57}, {9, 50, 150}, {19, 100, 288}, {34, 172, 480}};
  int maxe = max_e[s->ncov][s->ncrystal - 1];
  memset(s->pe, 0, sizeof s->pe);
  for (int i = 0; i < maxe; i++) {
    char *p = next_line(s, &r), *e;
    long indsp = strtol(p, &e, 10);
    long elind = strtol(e, &e, 10);
    char *q = e;
    while (*q == ' ' || *q == '\t') q++;
    double val = strtod(q, NULL);
    if (indsp > s->ncrystal) fail(s, "ERROR: 1st index in evap prob must be <= NCrystal");
    int ind[6] = {0, 0, 0, 0, 0, 0};
    int nd = s->ncrystal + s->ncov;
    long pw = 1;
    for (int j = 1; j <= s->ncrystal; j++) {
      pw = 1;
      for (int k = 0; k < nd - j; k++) pw *= 10;
      ind[j - 1] = (int)((elind / pw) % 10);
    }
    for (int j = s->ncrystal + 1; j <= nd; j++) {
      pw = 1;
      for (int k = 0; k < nd - j; k++) pw *= 10;
      ind[3 + j - s->ncrystal - 1] = (int)((elind / pw) % 10);
    }
    s->pe[(indsp - 1) + 3 * (ind[0] + 4 * (ind[1] + 4 * (ind[2] + 4 * (ind[3] + 4 * (ind[4] + 4 * ind[5])))))] = val;
  }
  for (int j = 0; j < s->ncrystal * 3; j++) {
    char *p = next_line(s, &r), *e;
    long indsp = strtol(p, &e, 10);
    long elind = strtol(e, &e, 10);
    if (indsp > s->ncrystal || indsp < 1 || elind < 1 || elind > 3) fail(s, "bad deposition line");
    s->pd[(indsp - 1) + 3 * (elind - 1)] = strtod(e, NULL);
  }
  for (int blk = 0; blk < 2; blk++)
    for (int j = 0; j < s->ncov * 3; j++) {
      char *p = next_line(s, &r), *e;
      long indsp = strtol(p, &e, 10);
      long elind = strtol(e, &e, 10);
      if (indsp <= s->ncrystal || indsp > s->ncrystal + s->ncov || elind < 0 || elind > 3)
        fail(s, "ERROR: 1st index in absorption/desorption prob block");
      double *t = blk == 0 ? s->pabs : s->pdes;
      t[(indsp - s->ncrystal - 1) + 3 * elind] = strtod(e, NULL);
    }
  fclose(r.f);
  r.f = NULL;
  allocate_tree(s);
  /* main.f:224-294 */
  if (!strcmp(is, "F")) set_flat(s, s->len1 - s->len1 % 24, 1);
  else if (!strcmp(is, "FS")) set_flat(s, s->len1 - s->len1 % 24, 0);
  else if (!strcmp(is, "C")) set_nc(s, s->len1, s->len2, s->len3, 0);
  else if (!strcmp(is, "S")) set_nc(s, s->len1, s->len1, s->len1, 1);
  else if (!strcmp(is, "I")) set_invp_off(s, s->len1, s->len2);
  else if (!strcmp(is, "FF"))
    set_si_finfet(s, s->len1 - s->len1 % 24, s->len2 - s->len2 % 24, s->len3 - s->len3 % 24, s->len4 - s->len4 % 24);
  else fail(s, "initialization not_implemented in the oracle: %s", is);
  memcpy(s->count_crystal_old, s->count_crystal, sizeof s->count_crystal);
  s->iter = 1;
  s->have_jb = 0;
  return 0;
}

/* ---- xyz frame writer: inpout/WriteMolMolXYZFile.f:22-406 ------------- */
/* Fortran Ew.d edit descriptor with the default 0.ddddE+xx form. */
static void fort_e(char *out, double v, int w, int d) {
  char buf[64];
  if (v == 0.0) {
    snprintf(buf, sizeof buf, "0.%0*dE+00", d, 0);
  } else {
    char m[64];
    snprintf(m, sizeof m, "%.*E", d - 1, fabs(v)); /* d significant digits */
    char *ep = strchr(m, 'E');
    int ex = atoi(ep + 1) + 1;
    char digits[40];
    int k = 0;
    for (char *q = m; q < ep; q++) if (*q != '.') digits[k++] = *q;
    digits[k] = 0;
    snprintf(buf, sizeof buf, "%s0.%sE%c%02d", v < 0 ? "-" : "", digits, ex < 0 ? '-' : '+', abs(ex));
  }
  size_t n = strlen(buf);
  if ((int)n >= w) { memcpy(out, buf, n + 1); return; }
  memset(out, ' ', (size_t)w - n);
  memcpy(out + (w - (int)n), buf, n + 1);
}

static const char *sym_of(int z) {
  switch (z) {
    case 6: return "C ";
    case 32: return "Ge ";
    case 5: return "B ";
    default: return NULL;
  }
}

/* The vacancy frame (inpout/WriteMolMolXYZFile.f:222-260): NumVoids, then
 * every vacancy in ListVoid order -- for SiC PVD (NCov 0, NCrystal 2,
 * ListCrystal(2) = 6) classified by the species sum of its four NextNXYZ
 * points (4 CV, 8 SV, 7 SAV, 5 CAV, 6 XV, anything else stops), otherwise as
 * an 'O' line. */
static void write_vacancy_frame(orc_sim *s, FILE *f, const double alat[3]) {
  char e1[32], e2[32], e3[32];
  fprintf(f, " %8d angstroem\n", s->n_vd);
  fort_e(e1, alat[0], 15, 8); fort_e(e2, alat[1], 15, 8); fort_e(e3, alat[2], 15, 8);
  fprintf(f, " periodic %s %s %s\n", e1, e2, e3);
  const int sic_pvd = s->ncov == 0 && s->ncrystal == 2 && s->list_crystal[1] == 6;
  for (int i = 1; i <= s->n_vd; i++) {
    const int32_t *p = s->vd[i].xyz;
    const double fx = s->sf * p[0], fy = s->sf * p[1], fz = s->sf * p[2];
    if (!sic_pvd) {
      fprintf(f, "O %10.5f %10.5f %10.5f  \n", fx, fy, fz);
      continue;
    }
    int iat = 0;
    for (int j = 0; j < 4; j++) iat += ibits(LC(s, s->vd[i].nn[j]), POS_SIC, LEN_SIC);
    switch (iat) {
      case 4: fprintf(f, "Mg %10.5f %10.5f %10.5f # CV (C vacancy)\n", fx, fy, fz); break;
      case 8: fprintf(f, "He %10.5f %10.5f %10.5f # SV (Si vacancy)\n", fx, fy, fz); break;
      case 7: fprintf(f, "Ne %10.5f %10.5f %10.5f # SAV (Si antisite vac)\n", fx, fy, fz); break;
      case 5: fprintf(f, "Be %10.5f %10.5f %10.5f # CAV (C antisite vac)\n", fx, fy, fz); break;
      case 6: fprintf(f, "Ar %10.5f %10.5f %10.5f # XV (near 2C and 2Si)\n", fx, fy, fz); break;
      default: fail(s, "ERROR: Weird Defect type was formed...ABORTING");
    }
  }
}

/* NCoorCorr of WriteMolMolXYZFile.f:284-296: the first neighbours of a bulk
 * atom of the other species (C bits around a Si atom, Si bits around a C). */
static int ncoor_corr(orc_sim *s, int32_t w, const int32_t site[3]) {
  const at_rec *b = atr(s, s->lind[gidx(s, site[0], site[1], site[2])]);
  const int iat = ibits(w, POS_SIC, LEN_SIC);
  int corr = 0;
  if (iat == 1) for (int i = 0; i < 4; i++) corr += ibits(LC(s, b->nn[i]), POS_C, LEN_C);
  else if (iat == 2) for (int i = 0; i < 4; i++) corr += ibits(LC(s, b->nn[i]), POS_SI, LEN_SI);
  return corr;
}

int orc_write_frames(orc_sim *s, const char *xyz_path, const char *d_path, const char *v_path, const char *w_path,
                     int64_t iter);

int orc_write_frame(orc_sim *s, const char *xyz_path, const char *d_path, int64_t iter) {
  return orc_write_frames(s, xyz_path, d_path, NULL, NULL, iter);
}

int orc_write_frames(orc_sim *s, const char *xyz_path, const char *d_path, const char *v_path, const char *w_path,
                     int64_t iter) {
  if (setjmp(s->jb)) { s->have_jb = 0; return 2; }
  s->have_jb = 1;
  int lx = s->lx, ly = s->ly, lz = s->lz;
  double alat[3] = {s->sf * (double)lx, s->sf * (double)ly, s->sf * (double)lz};
  int64_t nat = 0, ncov = 0, noepi = 0;
  for (int z = 1; z <= lz - 1; z++)
    for (int y = 0; y < ly; y++)
      for (int x = 0; x < lx; x++) {
        int32_t w = s->latt[gidx(s, x, y, z)];
        int cv = ibits(w, POS_INDEX, LEN_INDEX);
        if (cv == WALL) continue;
        int occ = ibits(w, POS_OCC, LEN_OCC), coo = ibits(w, POS_COOR, LEN_COOR);
        if (occ == 1 && coo <= 3) nat++;
        else if (occ == 0 && cv != 0) ncov++;
      }
  char e1[32], e2[32], e3[32];
  FILE *fd = d_path ? fopen(d_path, "w") : NULL;
  for (int z = 0; z <= lz - 1; z++)
    for (int y = 0; y < ly; y++)
      for (int x = 0; x < lx; x++) {
        int32_t w = s->latt[gidx(s, x, y, z)];
        if (ibits(w, POS_OCC, LEN_OCC) == 1 && ibits(w, POS_COOR, LEN_COOR) == 4 && !dia_type(x, y, z)) noepi++;
      }
  if (fd) {
    fort_e(e1, s->time, 15, 8);
    fprintf(fd, " %8lld angstroem KMC-time: %s Iter: %12lld\n", (long long)noepi, e1, (long long)iter);
    fort_e(e1, alat[0], 15, 8); fort_e(e2, alat[1], 15, 8); fort_e(e3, alat[2], 15, 8);
    fprintf(fd, " periodic %s %s %s\n", e1, e2, e3);
    for (int z = 0; z <= lz - 1; z++)
      for (int y = 0; y < ly; y++)
        for (int x = 0; x < lx; x++) {
          int32_t w = s->latt[gidx(s, x, y, z)];
          if (ibits(w, POS_OCC, LEN_OCC) == 1 && ibits(w, POS_COOR, LEN_COOR) == 4 && !dia_type(x, y, z)) {
            int iat = ibits(w, POS_SIC, LEN_SIC);
            const char *sym = iat == 1 ? "Ge " : iat == 2 ? "B " : iat == 3 ? "Cl " : NULL;
            if (sym) fprintf(fd, "%s%10.5f %10.5f %10.5f  \n", sym, s->sf * x, s->sf * y, s->sf * z);
          }
        }
    fclose(fd);
  }
  FILE *f = xyz_path ? fopen(xyz_path, "w") : NULL;
  if (xyz_path && !f) fail(s, "cannot write %s", xyz_path);
  if (f) {
    fort_e(e1, s->time, 15, 8);
    fprintf(f, " %8lld angstroem KMC-time: %s Iter: %12lld\n", (long long)(nat + s->n_vd + ncov), e1,
            (long long)iter);
    fort_e(e1, alat[0], 15, 8); fort_e(e2, alat[1], 15, 8); fort_e(e3, alat[2], 15, 8);
    fprintf(f, " periodic %s %s %s\n", e1, e2, e3);
    for (int z = 0; z <= lz - 1; z++)
      for (int y = 0; y < ly; y++)
        for (int x = 0; x < lx; x++) {
          int32_t w = s->latt[gidx(s, x, y, z)];
          int iat = ibits(w, POS_SIC, LEN_SIC), coo = ibits(w, POS_COOR, LEN_COOR);
          int cv = ibits(w, POS_INDEX, LEN_INDEX);
          double fx = s->sf * x, fy = s->sf * y, fz = s->sf * z;
          if (cv == 1) fprintf(f, "H %10.5f %10.5f %10.5f  \n", fx, fy, fz);
          if (cv == 17) fprintf(f, "Cl %10.5f %10.5f %10.5f  \n", fx, fy, fz);
          if (iat == 1 && coo <= 3) fprintf(f, "Si %10.5f %10.5f %10.5f  \n", fx, fy, fz);
          if ((iat == 2 || iat == 3) && coo <= 3) {
            if (iat == 2 && s->ncrystal == 1) fail(s, "ERR: IAt=2 should not happen, NCrystal=1");
            if (iat == 3 && s->ncrystal <= 2) fail(s, "ERR: IAt=3 should not happen if NCrystal<3");
            const char *sym = sym_of(s->list_crystal[iat - 1]);
            if (!sym) fail(s, "ERROR: 2nd crystal specie can only be C/Ge/B");
            fprintf(f, "%s%10.5f %10.5f %10.5f  \n", sym, fx, fy, fz);
          }
        }
    for (int i = 1; i <= s->n_vd; i++)
      fprintf(f, "O %10.5f %10.5f %10.5f  \n", s->sf * s->vd[i].xyz[0], s->sf * s->vd[i].xyz[1],
              s->sf * s->vd[i].xyz[2]);
    fclose(f);
  }
  if (v_path) {
    FILE *fv = fopen(v_path, "w");
    if (!fv) fail(s, "cannot write %s", v_path);
    write_vacancy_frame(s, fv, alat);
    fclose(fv);
  }
  /* SiC "wrong" analysis and its side effect on LattCoo (:262-398): the
   * restore loop reads ListAtBuffer(NumIsWrong) instead of entry i, so
   * the species flip is only undone correctly when all flipped atoms had
   * the same original species. */
  int has_c = 0;
  for (int i = 0; i < 3; i++) if (s->list_crystal[i] == 6) has_c = 1;
  if (s->ncrystal == 2 && has_c) {
    int32_t nwrong = 0;
    for (int z = 1; z <= lz - 1; z++)
      for (int y = 0; y < ly; y++)
        for (int x = 0; x < lx; x++) {
          int32_t *w = &s->latt[gidx(s, x, y, z)];
          if (ibits(*w, POS_INDEX, LEN_INDEX) == WALL) continue;
          if (!(ibits(*w, POS_OCC, LEN_OCC) == 1 && ibits(*w, POS_COOR, LEN_COOR) == 4)) continue;
          int iat = ibits(*w, POS_SIC, LEN_SIC);
          const at_rec *b = atr(s, s->lind[gidx(s, x, y, z)]);
          int corr = 0;
          if (iat == 1) for (int i = 0; i < 4; i++) corr += ibits(LC(s, b->nn[i]), POS_C, LEN_C);
          else if (iat == 2) for (int i = 0; i < 4; i++) corr += ibits(LC(s, b->nn[i]), POS_SI, LEN_SI);
          if (corr == 0) {
            nwrong++;
            if (nwrong >= NUM_BF_MAX) fail(s, "ListAtBuffer overflow");
            if (iat == 1) {
              *w &= ~(1 << POS_SI);
              *w |= (1 << POS_C);
            } else {
              *w &= ~(1 << POS_C);
              *w |= (1 << POS_SI);
            }
            s->abuf[nwrong].xyz[0] = x;
            s->abuf[nwrong].xyz[1] = y;
            s->abuf[nwrong].xyz[2] = z;
            s->abuf[nwrong].type = iat == 1 ? 1 : 2;
          }
        }
    /* the wrong-occupation frame (:300-380), read while the isolated
     * wrong atoms are flipped: bulk atoms with 3 neighbours of the other
     * species, written as Si or C */
    if (w_path) {
      int64_t nat_wrong = 0;
      for (int pass = 0; pass < 2; pass++) {
        FILE *fw = NULL;
        if (pass == 1) {
          fw = fopen(w_path, "w");
          if (!fw) fail(s, "cannot write %s", w_path);
          fprintf(fw, " %8lld angstroem\n", (long long)nat_wrong);
          fort_e(e1, alat[0], 15, 8); fort_e(e2, alat[1], 15, 8); fort_e(e3, alat[2], 15, 8);
          fprintf(fw, " periodic %s %s %s\n", e1, e2, e3);
        }
        for (int z = 1; z <= lz - 1; z++)
          for (int y = 0; y < ly; y++)
            for (int x = 0; x < lx; x++) {
              const int32_t w = s->latt[gidx(s, x, y, z)];
              if (!(ibits(w, POS_OCC, LEN_OCC) == 1 && ibits(w, POS_COOR, LEN_COOR) == 4)) continue;
              const int32_t site[3] = {x, y, z};
              if (ncoor_corr(s, w, site) != 3) continue;
              if (pass == 0) {
                nat_wrong++;
              } else {
                const int iat = ibits(w, POS_SIC, LEN_SIC);
                if (iat == 1) fprintf(fw, "Si %10.5f %10.5f %10.5f  \n", s->sf * x, s->sf * y, s->sf * z);
                if (iat == 2) fprintf(fw, "C %10.5f %10.5f %10.5f  \n", s->sf * x, s->sf * y, s->sf * z);
              }
            }
        if (fw) fclose(fw);
      }
    }
    for (int i = 1; i <= nwrong; i++) {
      int32_t *w = LCp(s, s->abuf[i].xyz);
      if (s->abuf[nwrong].type == 1) {
        *w &= ~(1 << POS_C);
        *w |= (1 << POS_SI);
      } else {
        *w &= ~(1 << POS_SI);
        *w |= (1 << POS_C);
      }
    }
  }
  s->have_jb = 0;
  return 0;
}

/* WriteMolMolSource.f:22-112 side effects + file naming. */
static int frame(orc_sim *s, const char *outdir, int64_t iter, int64_t iout) {
  memcpy(s->count_crystal_old, s->count_crystal, sizeof s->count_crystal);
  char px[4096], pd[4096];
  if (outdir) {
    snprintf(px, sizeof px, "%s/I%08lld.xyz", outdir, (long long)iout);
    snprintf(pd, sizeof pd, "%s/I%08lld_d.xyz", outdir, (long long)iout);
  }
  return orc_write_frame(s, outdir ? px : NULL, outdir ? pd : NULL, iter);
}

/* ---- main loop: main.f:304-371 ---------------------------------------- */
static int32_t do_pick(orc_sim *s, int32_t *itrans) {
  double patom;
  int32_t ind = pick_tree_event(s, &patom);
  s->time = s->time + 1.0 / s->tree[1];
  *itrans = s->ad[ind].ev;
  return ind;
}
/* With the second-neighbour term, a change of occupancy at site changes the
 * rate of every evaporating MC particle (an atom of coordination 1..3) on
 * its 12 second-neighbour positions: they get a fresh Get_Prob after the
 * event's own updates (no reference counterpart). */
static void refresh_second(orc_sim *s, const int32_t site[3]) {
  for (int k = 0; k < 12; k++) {
    int32_t q[3];
    second_pos(s, site, k, q);
    if (q[2] < 0 || q[2] > s->lz) continue;
    int32_t w = s->latt[gidx(s, q[0], q[1], q[2])];
    int coor = ibits(w, POS_COOR, LEN_COOR);
    if (!ibits(w, POS_OCC, LEN_OCC) || ibits(w, POS_INDEX, LEN_INDEX) == WALL || coor < 1 || coor > 3) continue;
    int32_t ind = LI(s, q);
    if (ind < 1 || ind > s->n_ad || !eq3(s->ad[ind].xyz, q)) fail(s, "second neighbour (%d,%d,%d) is not listed", q[0], q[1], q[2]);
    int iev = 0;
    double prob = 0.0;
    get_prob(s, 1, ibits(w, POS_SIC, LEN_SIC), 0, coor, q, s->ad[ind].nn, &iev, &prob);
    s->ad[ind].ev = iev;
    s->ad[ind].prob = prob; s->ad[ind].key = s->last_key;
    update_tree(s, ind, prob);
  }
}

static void do_execute(orc_sim *s, int32_t ind, int32_t itrans) {
  s->last_arm = -1;
  {
    const int slot = itrans <= 3 ? itrans - 1 : (itrans == 4 ? 3 : (itrans == 5 ? 4 : itrans - 1));
    if (slot >= 0 && slot < 8) s->executed[slot]++;
  }
  int32_t site[3];
  cp3(site, s->ad[ind].xyz);
  if (itrans <= 3) {
    deposition(s, ind, itrans);
    if (itrans >= 1) s->count_crystal[itrans - 1]++;
    if (s->use_f2) refresh_second(s, site);
  } else if (itrans == 4) {
    evaporation(s, ind);
    if (s->use_f2) refresh_second(s, site);
  } else if (itrans == 5) {
    desorption(s, ind);
  } else {
    absorption(s, ind, s->list_cov[itrans - 6]);
    s->count_cov[itrans - 6]++;
  }
}

int orc_step(orc_sim *s, orc_event *ev) {
  if (setjmp(s->jb)) { s->have_jb = 0; return 2; }
  s->have_jb = 1;
  if (s->n_ad <= 20) { s->have_jb = 0; return 1; }
  double r = s->tree[1];
  int32_t itrans;
  la_liquid(s);
  int32_t ind = do_pick(s, &itrans);
  itrans = la_swap(s, itrans);
  if (ev) {
    ev->iter = s->iter;
    cp3(ev->site, s->ad[ind].xyz);
    ev->itrans = itrans;
    ev->rtot = r;
    ev->coor = ibits(LC(s, ev->site), POS_COOR, LEN_COOR);
  }
  do_execute(s, ind, itrans);
  if (ev) { ev->arm = s->last_arm; ev->time = s->time; }
  s->iter++;
  s->have_jb = 0;
  return 0;
}

int orc_run(orc_sim *s, int64_t max_events, double max_time, int64_t *done) {
  int64_t n = 0;
  int rc = 0;
  while (n < max_events && s->time < max_time) {
    rc = orc_step(s, NULL);
    if (rc) break;
    n++;
  }
  if (done) *done = n;
  return rc;
}

/* The state at the KMC time t_end: every event whose clock (Time + 1/R,
 * main.f:313) stays at or below t_end is executed, the first one past it is
 * not -- unlike the reference's loop, which exits after executing that
 * event (main.f:363-366).  Used to compare states sampled at one time. */
int orc_run_to(orc_sim *s, int64_t max_events, double t_end, int64_t *done) {
  int64_t n = 0;
  int rc = 0;
  while (n < max_events && s->tree && s->tree[1] > 0.0 && s->time + 1.0 / s->tree[1] <= t_end) {
    rc = orc_step(s, NULL);
    if (rc) break;
    n++;
  }
  if (done) *done = n;
  return rc;
}

int orc_apply(orc_sim *s, const int32_t site[3], int32_t itrans, int32_t arm) {
  if (setjmp(s->jb)) { s->have_jb = 0; s->force_arm = -1; return 2; }
  s->have_jb = 1;
  int32_t ind = LI(s, site);
  if (ind < 1 || ind > s->n_ad || !eq3(s->ad[ind].xyz, site))
    fail(s, "orc_apply: (%d,%d,%d) is not an MC particle", site[0], site[1], site[2]);
  s->force_arm = arm;
  la_liquid(s);
  do_execute(s, ind, itrans);
  s->force_arm = -1;
  s->iter++;
  s->have_jb = 0;
  return 0;
}

/* A state handed over from elsewhere (the engine's store_state layout, the
 * same hand-over as the reference's builders leave at main.f:298): LattCoo,
 * then the three lists with their NextNXYZ, added in the given order with
 * AddMC / AddBulk / AddVoid (Add.f:24-101) so LattInd and the tree follow.
 * A listed MC particle without a rate (an atom of a frozen substrate) takes
 * the vanishing rate 1e-300, as orc_set_frozen gives it; orc_refresh_all
 * re-evaluates every rate afterwards. */
int orc_load_state(orc_sim *s, const int32_t *latt, int32_t n_ad, const int32_t *ad_xyz, const int32_t *ad_nn,
                   const int32_t *ad_ev, const double *ad_prob, int32_t n_bulk, const int32_t *bulk_xyz,
                   const int32_t *bulk_nn, int32_t n_void, const int32_t *void_xyz, const int32_t *void_nn,
                   double time) {
  if (setjmp(s->jb)) { s->have_jb = 0; return 2; }
  s->have_jb = 1;
  if (!s->tree) allocate_tree(s);
  if (n_ad >= NUM_MC_MAX || n_void >= NUM_VD_MAX || n_bulk >= NUM_AT_MAX) fail(s, "orc_load_state: list too long");
  memcpy(s->latt, latt, (size_t)s->npts * sizeof(int32_t));
  memset(s->lind, 0, (size_t)s->npts * sizeof(int32_t));
  memset(s->tree, 0, ((size_t)s->size_tree + 2) * sizeof(double));
  memset(s->ad, 0, (size_t)(NUM_MC_MAX + 1) * sizeof(mc_rec));
  memset(s->vd, 0, (size_t)(NUM_VD_MAX + 1) * sizeof(at_rec));
  if (s->at) memset(s->at, 0, (size_t)s->cap_at * sizeof(at_rec));
  s->n_ad = s->n_at = s->n_vd = 0;
  s->last_key = -1;
  int32_t nn[4][3];
  for (int32_t i = 0; i < n_ad; i++) {
    memcpy(nn, ad_nn + 12 * (int64_t)i, sizeof nn);
    add_mc(s, ad_xyz + 3 * (int64_t)i, nn, ad_ev[i], ad_prob[i] > 0.0 ? ad_prob[i] : 1e-300);
  }
  for (int32_t i = 0; i < n_bulk; i++) {
    memcpy(nn, bulk_nn + 12 * (int64_t)i, sizeof nn);
    add_bulk(s, bulk_xyz + 3 * (int64_t)i, nn);
  }
  for (int32_t i = 0; i < n_void; i++) {
    memcpy(nn, void_nn + 12 * (int64_t)i, sizeof nn);
    add_void(s, void_xyz + 3 * (int64_t)i, nn);
  }
  s->time = time;
  s->iter = 1;
  s->have_jb = 0;
  return 0;
}

int orc_build_cad(orc_sim *s, const int8_t *ids, const int8_t *species2) {
  if (setjmp(s->jb)) { s->have_jb = 0; return 2; }
  s->have_jb = 1;
  if (!s->tree) allocate_tree(s);
  memset(s->latt, 0, (size_t)s->npts * sizeof(int32_t));
  memset(s->lind, 0, (size_t)s->npts * sizeof(int32_t));
  memset(s->tree, 0, ((size_t)s->size_tree + 2) * sizeof(double));
  memset(s->ad, 0, (size_t)(NUM_MC_MAX + 1) * sizeof(mc_rec));
  memset(s->vd, 0, (size_t)(NUM_VD_MAX + 1) * sizeof(at_rec));
  if (s->at) memset(s->at, 0, (size_t)s->cap_at * sizeof(at_rec));
  s->n_ad = s->n_at = s->n_vd = 0;
  s->last_key = -1;
  set_cad(s, ids, species2);
  s->time = 0.0;
  s->iter = 1;
  s->have_jb = 0;
  return 0;
}

int orc_main(orc_sim *s, const char *outdir, orc_event *log, int64_t log_cap, int64_t *log_n, int *stop_reason) {
  volatile int64_t nlog = 0;
  int64_t framecounter = 0;
  *stop_reason = 0;
  if (frame(s, outdir, 0, 0)) { *stop_reason = 2; return 2; }
  int is_iter = !strcmp(s->exit_strategy, "Iter");
  int64_t iterout = 0;
  if (setjmp(s->jb)) { s->have_jb = 0; *stop_reason = 2; if (log_n) *log_n = nlog; return 2; }
  s->have_jb = 1;
  for (;;) {
    iterout = s->iter;
    if (s->n_ad <= 20) { *stop_reason = 1; break; }
    double r = s->tree[1];
    int32_t itrans;
    la_liquid(s);
    int32_t ind = do_pick(s, &itrans);
    itrans = la_swap(s, itrans);
    if (is_iter) {
      if (s->iter % s->out_molmol == 0) {
        framecounter = iterout / s->out_molmol;
        s->have_jb = 0;
        if (frame(s, outdir, iterout, framecounter)) { *stop_reason = 2; return 2; }
        if (setjmp(s->jb)) { s->have_jb = 0; *stop_reason = 2; if (log_n) *log_n = nlog; return 2; }
        s->have_jb = 1;
      }
    } else {
      if (s->time >= (double)(framecounter + 1) * s->out_time) {
        s->have_jb = 0;
        if (frame(s, outdir, iterout, framecounter + 1)) { *stop_reason = 2; return 2; }
        if (setjmp(s->jb)) { s->have_jb = 0; *stop_reason = 2; if (log_n) *log_n = nlog; return 2; }
        s->have_jb = 1;
        framecounter++;
      }
    }
    orc_event *ev = (log && nlog < log_cap) ? &log[nlog] : NULL;
    if (ev) {
      ev->iter = s->iter;
      cp3(ev->site, s->ad[ind].xyz);
      ev->itrans = itrans;
      ev->rtot = r;
      ev->coor = ibits(LC(s, ev->site), POS_COOR, LEN_COOR);
    }
    do_execute(s, ind, itrans);
    if (ev) { ev->arm = s->last_arm; ev->time = s->time; nlog++; }
    s->iter++;
    if (is_iter ? (s->iter > s->max_iter) : (s->time > s->max_time)) break;
  }
  s->have_jb = 0;
  if (log_n) *log_n = nlog;
  int64_t fout = is_iter ? iterout / s->out_molmol : framecounter;
  if (frame(s, outdir, iterout, fout)) { *stop_reason = 2; return 2; }
  return 0;
}

/* ---- lifecycle and accessors ------------------------------------------ */
orc_sim *orc_create(int lx, int ly, int lz) {
  orc_sim *s = (orc_sim *)calloc(1, sizeof(orc_sim));
  if (!s) return NULL;
  s->lx = lx; s->ly = ly; s->lz = lz;
  s->npts = (int64_t)lx * ly * (lz + 1);
  s->latt = (int32_t *)calloc((size_t)s->npts, sizeof(int32_t));
  s->lind = (int32_t *)calloc((size_t)s->npts, sizeof(int32_t));
  s->ad = (mc_rec *)calloc(NUM_MC_MAX + 1, sizeof(mc_rec));
  s->vd = (at_rec *)calloc(NUM_VD_MAX + 1, sizeof(at_rec));
  s->abuf = (buf_rec *)calloc(NUM_BF_MAX + 1, sizeof(buf_rec));
  if (!s->latt || !s->lind || !s->ad || !s->vd || !s->abuf) { orc_destroy(s); return NULL; }
  s->time = 0.0;
  s->iter = 1;
  s->force_arm = -1;
  strcpy(s->run_type, "R");
  strcpy(s->exit_strategy, "Iter");
  s->max_iter = 2100000000;
  s->out_molmol = 100000000;
  return s;
}

void orc_destroy(orc_sim *s) {
  if (!s) return;
  free(s->latt); free(s->lind); free(s->ad); free(s->at); free(s->vd); free(s->abuf);
  free(s->tree); free(s->numbers);
  free(s->st_temp); free(s->st_pe); free(s->st_pd); free(s->st_pabs); free(s->st_pdes);
  free(s);
}

const char *orc_last_error(const orc_sim *s) { return s->err; }

int orc_set_numberslist(orc_sim *s, const double *v, int n) {
  if (n <= 0) return 1;
  free(s->numbers);
  s->numbers = (double *)malloc(sizeof(double) * (size_t)n);
  memcpy(s->numbers, v, sizeof(double) * (size_t)n);
  s->n_numbers = n;
  return 0;
}

void orc_anomalies(const orc_sim *s, int64_t out[3]) {
  out[0] = s->nonbond_arm;
  out[1] = s->nonbond_nn;
  out[2] = s->oob_reads;
}

int orc_set_frozen(orc_sim *s, int z) {
  if (setjmp(s->jb)) { s->have_jb = 0; return 2; }
  s->have_jb = 1;
  s->frozen_below_z = z;
  /* atoms already listed as MC particles below z keep their place but get
   * the vanishing rate Get_Prob now gives them */
  for (int32_t i = 1; i <= s->n_ad; i++)
    if (s->ad[i].xyz[2] < z && s->ad[i].ev == 4) {
      s->ad[i].prob = 1e-300;
      s->ad[i].key = -1;
      update_tree(s, i, 1e-300);
    }
  s->have_jb = 0;
  return 0;
}

int orc_set_second_factor(orc_sim *s, const double *f2) {
  if (setjmp(s->jb)) { s->have_jb = 0; return 2; }
  s->have_jb = 1;
  s->use_f2 = f2 != NULL;
  if (f2) memcpy(s->f2, f2, sizeof s->f2);
  /* evaporating MC particles take the factor now */
  for (int32_t i = 1; i <= s->n_ad; i++) {
    int32_t w = LC(s, s->ad[i].xyz);
    int coor = ibits(w, POS_COOR, LEN_COOR);
    if (!ibits(w, POS_OCC, LEN_OCC) || coor < 1 || coor > 3) continue;
    int iev = 0;
    double prob = 0.0;
    get_prob(s, 1, ibits(w, POS_SIC, LEN_SIC), 0, coor, s->ad[i].xyz, s->ad[i].nn, &iev, &prob);
    s->ad[i].ev = iev;
    s->ad[i].prob = prob; s->ad[i].key = s->last_key;
    update_tree(s, i, prob);
  }
  s->have_jb = 0;
  return 0;
}

/* ---- trajectory likelihood (no reference counterpart) ----------------------
 * The restatement the engine's mskg_adjoint is checked against, computed by
 * brute force: before every event the whole MC list is scanned. */

/* Get_Prob for every MC particle (Get_Prob.f:54-331 with its side effects:
 * an adsorption site redraws its species), so that every stored rate and
 * class key comes from the current state. */
/* Laser-annealing rates (Get_Prob.f:127-149): the tables hold energies,
 * and each rate follows from the site's temperature bits.  The reference
 * enters this mode for InitSt 'LA', whose CAD-file initialisation is not
 * restated here; this switch puts a deck of another InitSt, written with
 * energies in its tables, into the same mode.  Rates are re-evaluated by
 * orc_refresh_all once the temperatures are set. */
int orc_set_la(orc_sim *s, double tm, double p0, double sigma, double tflex) {
  s->la_mode = 1;
  s->tm = tm;
  s->p0 = p0;
  s->sigma = sigma;
  s->tflex = tflex;
  return 0;
}

int orc_set_st_tables(orc_sim *s, int ntemp, const int32_t *temps, const double *pe, const double *pd,
                      const double *pabs, const double *pdes) {
  free(s->st_temp); free(s->st_pe); free(s->st_pd); free(s->st_pabs); free(s->st_pdes);
  s->st_temp = NULL; s->st_pe = s->st_pd = s->st_pabs = s->st_pdes = NULL;
  s->st_n = 0;
  if (ntemp <= 0) return 0;
  s->st_temp = malloc((size_t)ntemp * sizeof(int32_t));
  s->st_pe = malloc((size_t)ntemp * 3 * 4096 * sizeof(double));
  s->st_pd = malloc((size_t)ntemp * 9 * sizeof(double));
  s->st_pabs = malloc((size_t)ntemp * 12 * sizeof(double));
  s->st_pdes = malloc((size_t)ntemp * 12 * sizeof(double));
  if (!s->st_temp || !s->st_pe || !s->st_pd || !s->st_pabs || !s->st_pdes) return -1;
  memcpy(s->st_temp, temps, (size_t)ntemp * sizeof(int32_t));
  memcpy(s->st_pe, pe, (size_t)ntemp * 3 * 4096 * sizeof(double));
  memcpy(s->st_pd, pd, (size_t)ntemp * 9 * sizeof(double));
  memcpy(s->st_pabs, pabs, (size_t)ntemp * 12 * sizeof(double));
  memcpy(s->st_pdes, pdes, (size_t)ntemp * 12 * sizeof(double));
  s->st_n = ntemp;
  return 0;
}

int orc_set_la_variant(orc_sim *s, const double p0[3], const double tm[3], const double sigma[3],
                       const double tflex[3], const int32_t init_count[3], const int32_t count[3], int swap) {
  s->la_mode = 1;
  s->la_variant = 1;
  s->la_swap = swap ? 1 : 0;
  for (int k = 0; k < 3; k++) {
    s->la_p0[k] = p0[k];
    s->la_tm[k] = tm[k];
    s->la_sigma[k] = sigma[k];
    s->la_tflex[k] = tflex[k];
    s->count_init[k] = init_count[k];
    s->count_crystal[k] = count[k];
  }
  la_liquid(s);
  return 0;
}

void orc_la_counts(const orc_sim *s, int32_t count[3], int32_t liquid[3]) {
  for (int k = 0; k < 3; k++) {
    count[k] = s->count_crystal[k];
    liquid[k] = s->count_init[k] - s->count_crystal[k];
  }
}

int orc_refresh_all(orc_sim *s) {
  if (setjmp(s->jb)) { s->have_jb = 0; return 2; }
  s->have_jb = 1;
  for (int32_t i = 1; i <= s->n_ad; i++) {
    const int32_t *x = s->ad[i].xyz;
    int32_t w = LC(s, x);
    int iev = 0;
    double prob = 0.0;
    get_prob(s, ibits(w, POS_OCC, LEN_OCC), ibits(w, POS_SIC, LEN_SIC), ibits(w, POS_INDEX, LEN_INDEX),
             ibits(w, POS_COOR, LEN_COOR), x, s->ad[i].nn, &iev, &prob);
    if (iev == 0 || !(prob > 0.0)) fail(s, "orc_refresh_all: MC particle (%d,%d,%d) without a rate", x[0], x[1], x[2]);
    s->ad[i].ev = iev;
    s->ad[i].prob = prob;
    s->ad[i].key = s->last_key;
    update_tree(s, i, prob);
  }
  s->have_jb = 0;
  return 0;
}

/* Replays log[0..n) (as orc_apply) with waiting times dt[k] and returns,
 * with the numbering of mskg_adjoint: S[ADJ_NS] selections per class,
 * A[KEY_N] = sum_k (MC particles of class c before event k) * dt[k], and
 * scal = (sum dt, sum_k (isolated-atom rates before event k) * dt[k],
 * sum of log rates of selected isolated atoms, sum_k [log r(event k) -
 * Tree(1) dt[k]]), the last the log-likelihood evaluated directly from the
 * reference's tree. */
/* Neumaier's compensated summation: the finite-difference checks subtract
 * two log-likelihoods of ~1e5 that differ in the ~1e-4. */
typedef struct { double sum, c; } ksum;
static void kadd(ksum *k, double v) {
  double t = k->sum + v;
  if (fabs(k->sum) >= fabs(v)) k->c += (k->sum - t) + v;
  else k->c += (v - t) + k->sum;
  k->sum = t;
}
static double kval(const ksum *k) { return k->sum + k->c; }

int orc_adjoint_stats(orc_sim *s, const orc_event *evs, int64_t n, const double *dt, double *S, double *A,
                      double *scal) {
  return orc_adjoint_checkpoints(s, evs, n, dt, S, A, scal, 0, NULL);
}

/* The same with the class populations after every `every` events
 * (every > 0): hist holds n / every rows of KEY_N counts, the MC particles
 * by class key (isolated atoms included) after events every, 2*every, ... */
int orc_adjoint_checkpoints(orc_sim *s, const orc_event *evs, int64_t n, const double *dt, double *S, double *A,
                            double *scal, int64_t every, int32_t *hist) {
  if (setjmp(s->jb)) { s->have_jb = 0; s->force_arm = -1; return 2; }
  s->have_jb = 1;
  memset(S, 0, sizeof(double) * ADJ_NS);
  memset(A, 0, sizeof(double) * KEY_N);
  ksum ttot = {0, 0}, iso_int = {0, 0}, iso_log = {0, 0}, logl = {0, 0};
  for (int64_t k = 0; k < n; k++) {
    const double d = dt[k];
    kadd(&ttot, d);
    for (int32_t i = 1; i <= s->n_ad; i++) {
      const int key = s->ad[i].key;
      if (key == KEY_ISO) kadd(&iso_int, s->ad[i].prob * d);
      if (key >= 0) A[key] += d;
      else if (key == -2) fail(s, "orc_adjoint_stats: a rate class the likelihood does not cover");
    }
    kadd(&logl, -s->tree[1] * d);
    const int32_t *site = evs[k].site;
    int32_t ind = LI(s, site);
    if (ind < 1 || ind > s->n_ad || !eq3(s->ad[ind].xyz, site))
      fail(s, "orc_adjoint_stats: event %lld at (%d,%d,%d) is not an MC particle", (long long)k, site[0], site[1],
           site[2]);
    const int key = s->ad[ind].key, it = evs[k].itrans;
    double r;
    if (key >= KEY_ADS && key < KEY_PH) {
      const int c = key - KEY_ADS + 1;
      if (it < 1 || it > 3) fail(s, "orc_adjoint_stats: adsorption site without a deposition event");
      S[KEY_N + 3 * (it - 1) + (c - 1)] += 1.0;
      r = PD(s, it, c);
    } else if (key == KEY_ISO) {
      S[KEY_ISO] += 1.0;
      r = s->ad[ind].prob;
      kadd(&iso_log, log(r));
    } else if (key >= 0) {
      S[key] += 1.0;
      r = s->ad[ind].prob;
    } else {
      fail(s, "orc_adjoint_stats: event %lld on a site without a rate class", (long long)k);
      r = 0.0;
    }
    kadd(&logl, log(r));
    s->force_arm = evs[k].arm;
    do_execute(s, ind, it);
    s->force_arm = -1;
    s->iter++;
    if (every > 0 && (k + 1) % every == 0) {
      int32_t *row = hist + ((k + 1) / every - 1) * KEY_N;
      memset(row, 0, sizeof(int32_t) * KEY_N);
      for (int32_t i = 1; i <= s->n_ad; i++)
        if (s->ad[i].key >= 0) row[s->ad[i].key]++;
    }
  }
  scal[0] = kval(&ttot);
  scal[1] = kval(&iso_int);
  scal[2] = kval(&iso_log);
  scal[3] = kval(&logl);
  s->have_jb = 0;
  return 0;
}

/* ---- laser-annealing coupling maps ------------------------------------ */
/* Index of grid point (x, y, z), z < LenZ, in a coupling map: layout 0 the
 * socket arrays DIMENSION(LenZ, LenX*LenY) (mulskips-source-LaserAnnealing/
 * modules/defsystem.f:88-91: element (1+z, 1+x+LenX*y)), layout 1 the
 * temperature / Coor file planes (x fastest, then y, then z). */
static int64_t map_index(const orc_sim *s, int layout, int x, int y, int z) {
  if (layout == 0) return (int64_t)z + (int64_t)s->lz * ((int64_t)x + (int64_t)s->lx * y);
  return (int64_t)x + (int64_t)s->lx * ((int64_t)y + (int64_t)s->ly * z);
}

/* SetT, mulskips-source-LaserAnnealing/initial/SetT.f:38-48 (socket map b)
 * and :56-66 (file planes, read as INT): every grid point of the planes
 * z = 0..LenZ-1 whose map value is not 10 takes the value's low LenT bits
 * at PosT (MVBITS). */
int orc_set_temperature_field(orc_sim *s, const int32_t *field, int layout) {
  if (layout != 0 && layout != 1) return -1;
  for (int z = 0; z < s->lz; z++)
    for (int y = 0; y < s->ly; y++)
      for (int x = 0; x < s->lx; x++) {
        const int32_t t = field[map_index(s, layout, x, y, z)];
        if (t != 10) mvbits(t, LEN_T, &s->latt[gidx(s, x, y, z)], POS_T);
      }
  return 0;
}

/* The phase maps of mulskips-source-LaserAnnealing/main.f:796-818: c and d
 * zeroed, then for every grid point of the planes z = 0..LenZ-1 that is not
 * a wall (IQStat /= 113) and occupied, c = Coo and d = IAt (the species);
 * 10 and 10 on walls.  The laser-annealing build keeps the species one bit
 * higher (PosSiC = 28, its defsystem.f:99); IAt is the species index in
 * both layouts. */
int orc_phases(const orc_sim *s, int32_t *c, int32_t *d, int layout) {
  if (layout != 0 && layout != 1) return -1;
  const int64_t n = (int64_t)s->lx * s->ly * s->lz;
  for (int64_t i = 0; i < n; i++) c[i] = d[i] = 0;
  for (int k = 0; k < s->lz; k++)
    for (int j = 0; j < s->ly; j++)
      for (int i = 0; i < s->lx; i++) {
        const int32_t w = s->latt[gidx(s, i, j, k)];
        const int64_t o = map_index(s, layout, i, j, k);
        if (ibits(w, POS_INDEX, LEN_INDEX) != WALL) {
          if (ibits(w, POS_OCC, LEN_OCC) == 1) {
            c[o] = ibits(w, POS_COOR, LEN_COOR);
            d[o] = ibits(w, POS_SIC, LEN_SIC);
          }
        } else {
          c[o] = 10;
          d[o] = 10;
        }
      }
  return 0;
}

int32_t *orc_lattcoo(orc_sim *s) { return s->latt; }
int32_t *orc_lattind(orc_sim *s) { return s->lind; }
void orc_dims(const orc_sim *s, int *lx, int *ly, int *lz) { *lx = s->lx; *ly = s->ly; *lz = s->lz; }
void orc_counts(const orc_sim *s, int32_t *a, int32_t *b, int32_t *c) { *a = s->n_ad; *b = s->n_at; *c = s->n_vd; }
double orc_time(const orc_sim *s) { return s->time; }
void orc_executed(const orc_sim *s, int64_t out[8]) { memcpy(out, s->executed, sizeof s->executed); }
double orc_total_rate(const orc_sim *s) { return s->tree ? s->tree[1] : 0.0; }
int64_t orc_iter(const orc_sim *s) { return s->iter; }
void orc_count_crystal(const orc_sim *s, int32_t cc[3], int32_t cov[3]) {
  memcpy(cc, s->count_crystal, sizeof s->count_crystal);
  memcpy(cov, s->count_cov, sizeof s->count_cov);
}
int orc_ncrystal(const orc_sim *s, int *ncov) { *ncov = s->ncov; return s->ncrystal; }
void orc_tables(const orc_sim *s, double *pe, double *pd, double *pabs, double *pdes, double *pz,
                int32_t lc[3], int32_t lcov[3]) {
  if (pe) memcpy(pe, s->pe, sizeof s->pe);
  if (pd) memcpy(pd, s->pd, sizeof s->pd);
  if (pabs) memcpy(pabs, s->pabs, sizeof s->pabs);
  if (pdes) memcpy(pdes, s->pdes, sizeof s->pdes);
  if (pz) *pz = s->ptrans_zig;
  if (lc) for (int i = 0; i < 3; i++) lc[i] = s->list_crystal[i];
  if (lcov) for (int i = 0; i < 3; i++) lcov[i] = s->list_cov[i];
}

int orc_get_list(const orc_sim *s, int which, int32_t *xyz, int32_t *nn, int32_t *ev, double *prob) {
  int32_t n = which == 0 ? s->n_ad : which == 1 ? s->n_at : s->n_vd;
  for (int32_t i = 1; i <= n; i++) {
    const int32_t *x, (*q)[3];
    if (which == 0) { x = s->ad[i].xyz; q = (const int32_t(*)[3])s->ad[i].nn; }
    else if (which == 1) { x = s->at[i].xyz; q = (const int32_t(*)[3])s->at[i].nn; }
    else { x = s->vd[i].xyz; q = (const int32_t(*)[3])s->vd[i].nn; }
    if (xyz) memcpy(xyz + 3 * (i - 1), x, 12);
    if (nn) memcpy(nn + 12 * (i - 1), q, 48);
    if (which == 0) {
      if (ev) ev[i - 1] = s->ad[i].ev;
      if (prob) prob[i - 1] = s->ad[i].prob;
    }
  }
  return (int)n;
}

int orc_pure_rate(orc_sim *s, const int32_t xyz[3], int32_t *category, double *rate) {
  if (setjmp(s->jb)) { s->have_jb = 0; return 2; }
  s->have_jb = 1;
  int32_t w = LC(s, xyz);
  int32_t ind = LI(s, xyz);
  int occ = ibits(w, POS_OCC, LEN_OCC), coor = ibits(w, POS_COOR, LEN_COOR);
  int cv = ibits(w, POS_INDEX, LEN_INDEX);
  const int32_t(*nn)[3] = adr(s, ind)->nn;
  int iev = 0;
  double prob = 0.0;
  get_prob_impl(s, occ, ibits(w, POS_SIC, LEN_SIC), cv, coor, xyz, nn, &iev, &prob, 1);
  *category = iev;
  *rate = prob;
  s->have_jb = 0;
  return 0;
}
