// mskg_lattice.h -- device-resident lattice layout, geometry and Philox RNG of
// the MulSKIPS KMC engine, written for CDNA4.
//
// Layout (one GPU, all state resident in HBM):
//   The address space is the reference's super-lattice grid (x, y, z) with
//   0 <= x < LX, 0 <= y < LY (periodic) and 0 <= z <= LZ
//   (mulskips-source/modules/defsystem.f:49).  Every reachable lattice point
//   satisfies x = y = z (mod 2) because all bond vectors have odd components
//   and the geometry routines only combine and reflect them, so only those
//   points are stored: point index
//       p = (z * (LY/2) + y/2) * (LX/2) + x/2.
//   Per point one 16-byte record `Site` (one dwordx4 load):
//     bytes 0..11  four neighbour displacements (dx, dy, dz) as int8: the
//                  reference's NextNXYZ(1:3,1:4) relative to the point
//     byte  12     kind (bits 0..2): 0 empty, 1..3 species (Si, C/Ge, B),
//                  4 wall (CovInd 113), 5..7 coverage slot 1..3 (ListCov);
//                  coordination (bits 3..5)
//     byte  13     bit i set: entry i is a displacement; clear: entry i is
//                  the absolute point (0,0,0), which is how the reference
//                  stores an unfixed neighbour (Evaporation.f:95,113-118)
//     bytes 14..15 temperature in K (laser-annealing mode; LattCoo bits 9-21)
//   so an all-zero record is exactly the reference's zeroed list record.
//   Beside it: rate[p] (float, the reference's ProbTrans rounded once from
//   FP64) and cls[p] (uint8 event class, 0 = not an MC particle).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define MSKG_HD __host__ __device__ __forceinline__

namespace mskg {

enum : int {
  KIND_EMPTY = 0, KIND_WALL = 4, KIND_COV0 = 5,
  CLS_NONE = 0,
  CLS_ADSORB = 1,       // empty site: deposit a species or absorb coverage, drawn at execution
  CLS_PLACEHOLDER = 2,  // empty Coor-1 site whose neighbour has Coor<2: deposit species 1 at 1e-8
  CLS_EVAP = 4,         // evaporation, rate from the PTransE table
  CLS_DESORB = 5,       // coverage desorption
  CLS_ISOLATED = 6,     // isolated atom, rate 1e5*R (Get_Prob.f:303-306)
};

struct __align__(16) Site {
  int8_t d[4][3];
  uint8_t kc;     // kind | coord << 3
  uint8_t valid;  // bit i: entry i is a displacement
  uint16_t temp;
};
static_assert(sizeof(Site) == 16, "Site must be one dwordx4");

// A neighbour list as absolute positions (the reference's NextNXYZ).
struct NL {
  int x[4], y[4], z[4];
};

MSKG_HD int s_kind(const Site &s) { return s.kc & 7; }
MSKG_HD int s_coor(const Site &s) { return (s.kc >> 3) & 7; }
MSKG_HD void s_set_kind(Site &s, int k) { s.kc = (uint8_t)((s.kc & ~7) | k); }
MSKG_HD void s_set_coor(Site &s, int c) { s.kc = (uint8_t)((s.kc & 7) | (c << 3)); }
MSKG_HD bool kind_occupied(int k) { return k >= 1 && k <= 4; }  // species or wall: Occ bit set
MSKG_HD bool kind_species(int k) { return k >= 1 && k <= 3; }
MSKG_HD int kind_cov_slot(int k) { return k >= KIND_COV0 ? k - KIND_COV0 + 1 : 0; }

// Geometry constants: cos/sin of the two FIND_NN rotation angles
// (DFLOAT(2*irot)*pi/3. with the REAL(4) pi literal of FIND_NN.f:27),
// evaluated once on the host with the C library so the device uses the very
// same values.
struct Geometry {
  double cth[2], sth[2];
  double arm_shift;  // 8.6602254037 as the REAL(4) literal of FIND_SN.f:67
};

// Rate tables in FP64, the reference's own arrays (defsystem.f:60-68).
struct Tables {
  double pe[3 * 4096];  // PTransE(3, 0:3 x6), Fortran order
  double pd[9];         // PTransD(3, 1:3)
  double pabs[12];      // PTransAbs(3, 0:3)
  double pdes[12];      // PTransDes(3, 0:3)
  double ptrans_zig;
  int ncrystal, ncov;
  int la_mode;          // laser-annealing rates (Get_Prob.f:127-149)
  double p0, tm, sigma, tflex;
  // the laser-annealing build's form (mulskips-source-LaserAnnealing/mcevent/
  // Get_Prob.f:140-153): per-species constants and the liquid fraction
  // CountLiquid(k) / SUM(CountLiquid) on solidification; la_swap: its SiGe
  // species swap when a species' liquid is used up (main.f:676-681 of that build)
  int la_variant, la_swap;
  double la_p0[3], la_tm[3], la_sigma[3], la_tflex[3];
  // InitSt 'ST' (SetProbabilityFromField.f:21-171, InterpProbFromField.f:22-57):
  // st_n reference temperatures st_temp[] with their tables (device memory,
  // table-major: st_pe[k * 12288 + i] ...), interpolated at each site's
  // temperature; 0: off
  int st_n;
  const int *st_temp;
  const double *st_pe, *st_pd, *st_pabs, *st_pdes;
  int frozen_below_z;   // atoms with z below this never evaporate (0: off)
  int use_f2;           // second-neighbour factor on evaporation rates
  double f2[13];        // by the number of occupied second-neighbour positions
};

struct Box {
  int lx, ly, lz, hx, hy;
  int64_t plane, npts;
};

MSKG_HD Box make_box(int lx, int ly, int lz) {
  Box b;
  b.lx = lx; b.ly = ly; b.lz = lz; b.hx = lx / 2; b.hy = ly / 2;
  b.plane = (int64_t)b.hx * b.hy;
  b.npts = b.plane * (lz + 1);
  return b;
}
MSKG_HD bool valid_parity(int x, int y, int z) { return ((x ^ y) & 1) == 0 && ((x ^ z) & 1) == 0; }
MSKG_HD int64_t pidx(const Box &b, int x, int y, int z) {
  return ((int64_t)z * b.hy + (y >> 1)) * b.hx + (x >> 1);
}
MSKG_HD void pxyz(const Box &b, int64_t p, int &x, int &y, int &z) {
  z = (int)(p / b.plane);
  int64_t r = p - (int64_t)z * b.plane;
  int j = (int)(r / b.hx), i = (int)(r - (int64_t)j * b.hx);
  int par = z & 1;
  x = 2 * i + par;
  y = 2 * j + par;
}
// Point index of an absolute position, -1 when it is outside the box or off
// the lattice parity (the reference reads outside LattCoo for z<0; the CPU
// oracle and this engine both read such points as zero records).
MSKG_HD int64_t pidx_checked(const Box &b, int x, int y, int z) {
  if (x < 0 || x >= b.lx || y < 0 || y >= b.ly || z < 0 || z > b.lz || !valid_parity(x, y, z)) return -1;
  return pidx(b, x, y, z);
}

// Absolute position of entry i of a site at (x,y,z), with the reference's
// single x/y wrap of every neighbour-producing routine.
MSKG_HD void entry_pos(const Box &b, const Site &s, int x, int y, int z, int i, int &ox, int &oy, int &oz) {
  if (!((s.valid >> i) & 1)) { ox = oy = oz = 0; return; }
  ox = x + s.d[i][0];
  oy = y + s.d[i][1];
  oz = z + s.d[i][2];
  if (ox >= b.lx) ox -= b.lx;
  if (ox < 0) ox += b.lx;
  if (oy >= b.ly) oy -= b.ly;
  if (oy < 0) oy += b.ly;
}
MSKG_HD NL site_list(const Box &b, const Site &s, int x, int y, int z) {
  NL l;
  for (int i = 0; i < 4; i++) entry_pos(b, s, x, y, z, i, l.x[i], l.y[i], l.z[i]);
  return l;
}
// Store absolute neighbour positions into a record relative to (x,y,z);
// returns false when a displacement does not fit in int8.
MSKG_HD bool set_list(const Box &b, Site &s, int x, int y, int z, const NL &l) {
  uint8_t v = 0;
  for (int i = 0; i < 4; i++) {
    if (l.x[i] == 0 && l.y[i] == 0 && l.z[i] == 0) {
      s.d[i][0] = s.d[i][1] = s.d[i][2] = 0;
      continue;
    }
    int dx = l.x[i] - x, dy = l.y[i] - y, dz = l.z[i] - z;
    if (dx > b.lx / 2) dx -= b.lx;
    if (dx < -b.lx / 2) dx += b.lx;
    if (dy > b.ly / 2) dy -= b.ly;
    if (dy < -b.ly / 2) dy += b.ly;
    if (dx < -128 || dx > 127 || dy < -128 || dy > 127 || dz < -128 || dz > 127) return false;
    s.d[i][0] = (int8_t)dx;
    s.d[i][1] = (int8_t)dy;
    s.d[i][2] = (int8_t)dz;
    v |= (uint8_t)(1u << i);
  }
  s.valid = v;
  return true;
}

// ---- the reference's FP64 geometry --------------------------------------
MSKG_HD int isign_(int a, int b) { return b >= 0 ? (a < 0 ? -a : a) : (a < 0 ? a : -a); }
MSKG_HD int iabs_(int a) { return a < 0 ? -a : a; }
MSKG_HD int nint_(double v) { return (int)round(v); }  // NINT: half away from zero

// FIND_NN.f:25-77: the two remaining tetrahedral neighbours of SiteC given
// Site1 and Site2, by rotating Site2-SiteC about unit(SiteC-Site1) by 120 and
// 240 degrees in FP64 and rounding to the grid.
MSKG_HD void find_nn(const Geometry &g, int lx, int ly, const int s1[3], const int s2[3], const int sc[3],
                     int out[2][3]) {
  int a[3] = {s1[0], s1[1], s1[2]}, b[3] = {s2[0], s2[1], s2[2]};
  if (iabs_(sc[0] - a[0]) > lx / 2) a[0] += isign_(lx, sc[0] - a[0]);
  if (iabs_(sc[1] - a[1]) > ly / 2) a[1] += isign_(ly, sc[1] - a[1]);
  if (iabs_(sc[0] - b[0]) > lx / 2) b[0] += isign_(lx, sc[0] - b[0]);
  if (iabs_(sc[1] - b[1]) > ly / 2) b[1] += isign_(ly, sc[1] - b[1]);
  double ax[3];
  for (int i = 0; i < 3; i++) ax[i] = (double)sc[i] - (double)a[i];
  double sn = sqrt(ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2]);
  for (int i = 0; i < 3; i++) ax[i] = ax[i] / sn;
  for (int irot = 0; irot < 2; irot++) {
    const double cth = g.cth[irot], sth = g.sth[irot];
    double R[3][3];
    for (int i = 0; i < 3; i++) R[i][i] = ax[i] * ax[i] * (1 - cth) + cth;
    R[1][0] = ax[0] * ax[1] * (1 - cth) - ax[2] * sth;
    R[0][1] = ax[0] * ax[1] * (1 - cth) + ax[2] * sth;
    R[2][0] = ax[0] * ax[2] * (1 - cth) + ax[1] * sth;
    R[0][2] = ax[0] * ax[2] * (1 - cth) - ax[1] * sth;
    R[2][1] = ax[1] * ax[2] * (1 - cth) - ax[0] * sth;
    R[1][2] = ax[1] * ax[2] * (1 - cth) + ax[0] * sth;
    for (int k = 0; k < 3; k++) {
      double r = 0.0;
      for (int l = 0; l < 3; l++) r = r + R[l][k] * ((double)b[l] - (double)sc[l]);
      out[irot][k] = nint_(r) + sc[k];
    }
    if (out[irot][0] >= lx) out[irot][0] -= lx;
    if (out[irot][0] < 0) out[irot][0] += lx;
    if (out[irot][1] >= ly) out[irot][1] -= ly;
    if (out[irot][1] < 0) out[irot][1] += ly;
  }
}

// FIND_SN.f:26-84: zig-zag (inversion through the bond centre) and
// arm-chair (translation along the bond) candidates for the back bonds of
// Site, whose only fixed neighbour is SiteC with further neighbours SN.
MSKG_HD void find_sn(const Geometry &g, int lx, int ly, const int site[3], const int sc_in[3], const int sn_in[3][3],
                     int zig[3][3], int arm[3][3]) {
  int sc[3] = {sc_in[0], sc_in[1], sc_in[2]}, sn[3][3];
  for (int i = 0; i < 3; i++)
    for (int k = 0; k < 3; k++) sn[i][k] = sn_in[i][k];
  if (iabs_(site[0] - sc[0]) > lx / 2) sc[0] += isign_(lx, site[0] - sc[0]);
  if (iabs_(site[1] - sc[1]) > ly / 2) sc[1] += isign_(ly, site[1] - sc[1]);
  for (int i = 0; i < 3; i++) {
    if (iabs_(site[0] - sn[i][0]) > lx / 2) sn[i][0] += isign_(lx, site[0] - sn[i][0]);
    if (iabs_(site[1] - sn[i][1]) > ly / 2) sn[i][1] += isign_(ly, site[1] - sn[i][1]);
  }
  double cent[3], ax[3];
  for (int k = 0; k < 3; k++) cent[k] = 0.5 * ((double)sc[k] + (double)(float)site[k]);
  for (int k = 0; k < 3; k++) ax[k] = (double)site[k] - (double)sc[k];
  double snorm = sqrt(ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2]);
  for (int k = 0; k < 3; k++) ax[k] = ax[k] / snorm;
  for (int i = 0; i < 3; i++) {
    for (int k = 0; k < 3; k++) {
      zig[i][k] = nint_(2 * cent[k] - (double)sn[i][k]);
      arm[i][k] = nint_((double)sn[i][k] + g.arm_shift * ax[k]);
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

// ---- Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) -----------------
struct U4 { uint32_t x, y, z, w; };
MSKG_HD void mulhilo(uint32_t a, uint32_t b, uint32_t &hi, uint32_t &lo) {
  uint64_t p = (uint64_t)a * b;
  hi = (uint32_t)(p >> 32);
  lo = (uint32_t)p;
}
MSKG_HD U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; r++) {
    uint32_t hi0, lo0, hi1, lo1;
    mulhilo(M0, c.x, hi0, lo0);
    mulhilo(M1, c.z, hi1, lo1);
    U4 n;
    n.x = hi1 ^ c.y ^ k0;
    n.y = lo1;
    n.z = hi0 ^ c.w ^ k1;
    n.w = lo0;
    c = n;
    k0 += W0;
    k1 += W1;
  }
  return c;
}
// Uniform in the open interval (0,1) from 32 random bits.
MSKG_HD double u01_32(uint32_t a) { return ((double)a + 0.5) * (1.0 / 4294967296.0); }
// Uniform in the open interval (0,1) with 53 random bits.
MSKG_HD double u01_open(uint32_t a, uint32_t b) {
  uint64_t k = (((uint64_t)a << 32) | b) >> 11;
  return ((double)k + 0.5) * (1.0 / 9007199254740992.0);
}

}  // namespace mskg
