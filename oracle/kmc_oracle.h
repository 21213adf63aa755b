/*
 * kmc_oracle.h -- CPU restatement of the MulSKIPS KMC event engine.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP engine
 * in mulskips_amd/: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product path never calls it.
 *
 * It restates, in plain C and FP64, the reference's serial BKL engine with
 * the same data layout (packed LattCoo words, LattInd, the three particle
 * lists with swap-with-last deletion, the FP64 binary sum tree), the same
 * random-number consumption order (ran3 + the fixed numberslist table for
 * RunType 'T') and the same floating-point expressions, so that a deck run
 * through it reproduces the reference's xyz output byte for byte.
 *
 * Citations are relative to the reference repository root
 * (mulskips-source/...).  Each function in kmc_oracle.c names the file:line
 * range it restates.
 *
 * Parity pins (see tests/test_oracle_*.py):
 *   - ran3, tree, Get_Prob, Deposition, Evaporation, absorption, desorption,
 *     FIND_NN, FIND_SN, IF_Arm_Choice, the initial builders used by the
 *     regression decks and the xyz writer are pinned by the 8 valid
 *     regression fixtures (regression-tests/input-and-references-files):
 *     both I00000000.xyz and I00000001.xyz byte-identical, plus the final
 *     clock/site/counter lines of the logs.
 *   - FIND_NN/FIND_SN are additionally pinned by closed forms (rotation by
 *     +-120 deg about a bond, point inversion, mirror reflection) over every
 *     pair of the 32 bond vectors.
 *   - ran3 is pinned by its closed integer recurrence and by the Knuth
 *     subtractive-generator identity checked in the tests.
 *   - LA-mode rates (Get_Prob.f:127-149) are pinned by the closed form only
 *     (the reference LA fixture is stale against the current source):
 *     "parity pinned by closed form, not by a reference run".  orc_set_la
 *     enters that mode for a deck of another InitSt.
 *   - The engine options with no reference counterpart: the second-neighbour
 *     factor by a brute-force count of occupied shell points on a grown
 *     surface (rates with the factor / rates without = f2[count]), the
 *     frozen substrate by its rule at the z bound (tests/test_oracle_options.py);
 *     the likelihood statistics (orc_adjoint_stats) by the identity with the
 *     log-likelihood evaluated directly from the sum tree and by central
 *     finite differences of it (tests/test_adjoint.py).
 *   - InitSt 'ST' interpolation (orc_set_st_tables) by known answers: a
 *     temperature at a reference takes that table, and between two it
 *     takes the weighted blend the integer weights give, checked against
 *     decks whose tables are those blends (tests/test_oracle_st.py).
 *   - The laser-annealing coupling maps (orc_set_temperature_field,
 *     orc_phases) by known answers built from SetT's and main.f's rules:
 *     a map whose value names its own grid point (a transposed or shifted
 *     layout fails), the wall marker 10 and values beyond 13 bits, the top
 *     plane left alone; phase counts against the particle lists of a
 *     regression deck (tests/test_oracle_options.py).
 */
#ifndef MSK_KMC_ORACLE_H
#define MSK_KMC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_sim orc_sim;

/* One executed event, as recorded by orc_step/orc_main.  This is the event
 * log format consumed by the engine's replay entry point (mskg_replay). */
typedef struct {
  int64_t iter;       /* 1-based iteration number (main.f Iter)            */
  int32_t site[3];    /* grid coordinates of the picked MC particle        */
  int32_t itrans;     /* Ind_Event at pick time: 1..3 deposit species,
                         4 evaporate, 5 desorb, 6.. absorb coverage (k-5) */
  int32_t arm;        /* deposition at Coor 1: 1 arm-chair, 0 zig-zag;
                         -1 when no stacking choice was made               */
  int32_t coor;       /* coordination of the site before the event        */
  double  rtot;       /* Tree(1) at pick time (total rate R)              */
  double  time;       /* clock after the event (Time += 1/R)              */
} orc_event;

/* Box sizes are run-time here; the reference fixes them at compile time
 * (mulskips-source/modules/defsystem.f:27). */
orc_sim *orc_create(int len_x, int len_y, int len_z);
void     orc_destroy(orc_sim *s);
const char *orc_last_error(const orc_sim *s);

/* RunType 'T' replays this table of uniforms (PickTreeEvent.f:72-1079). */
int orc_set_numberslist(orc_sim *s, const double *values, int n);

/* Parse a start.dat (main.f:83-212 + SetProbability.f:21-143), set up the
 * tree (AllocateArrays.f) and run the initial builder chosen by InitSt
 * (main.f:224-294).  Supported builders: F, FS, C, S, I, FF.
 * Returns 0 on success. */
int orc_load_deck(orc_sim *s, const char *start_dat_path);

/* Replace the state by SetCAD's (initial/SetCAD.f:84-418, InitSt 'IN'):
 * ids holds one geometry id per grid point, ids[x + LenX*(y + LenY*z)] for
 * z = 0..LenZ-1 (the planes of the geometry file; 10 wall, 1 atom on a
 * diamond site, else empty; planes 0..3 and LenZ-4..LenZ-1 are walls
 * whatever they hold).  species2: NULL, or nonzero where the atom is of
 * species 2 (the reference's RAND() < alloyfraction draw).  Tables and
 * options stay those of the loaded deck.  Returns 0 on success. */
int orc_build_cad(orc_sim *s, const int8_t *ids, const int8_t *species2);

/* The whole program main.f:296-379: frame 0 output, the KMC loop with its
 * output cadence and the final frame.  When outdir is non-NULL the
 * I????????.xyz (and _d.xyz) frames are written there.  When log is
 * non-NULL, up to log_cap events are recorded (log_n receives the count).
 * stop_reason: 0 limit reached, 1 "few MC particles", 2 error. */
int orc_main(orc_sim *s, const char *outdir, orc_event *log, int64_t log_cap,
             int64_t *log_n, int *stop_reason);

/* One iteration of main.f:304-371 without frame output.  Returns 0 after an
 * executed event, 1 for the early stop (NumAdAtom<=20), 2 on error. */
int orc_step(orc_sim *s, orc_event *ev);

/* orc_step repeated until max_events events have run or the clock has
 * reached max_time (checked before each event, so the run ends with the
 * first event at or past max_time, as main.f:363-366 does).  Returns as
 * orc_step; *done receives the events executed. */
int orc_run(orc_sim *s, int64_t max_events, double max_time, int64_t *done);
/* As orc_run, but stopping before the first event whose clock would pass
 * t_end: the state at the KMC time t_end. */
int orc_run_to(orc_sim *s, int64_t max_events, double t_end, int64_t *done);
/* Events executed since creation, by type (deposits of species 1..3,
 * evaporations, desorptions, absorptions of ListCov 1..3). */
void orc_executed(const orc_sim *s, int64_t out[8]);

/* Execute one given event (Ind_Event itrans at the MC particle at site, with
 * the stacking choice forced to arm when >= 0) with the reference's update
 * rules: replays an event log produced elsewhere, e.g. by the engine's
 * parallel run.  Returns 0, or 2 if the site is not an MC particle. */
int orc_apply(orc_sim *s, const int32_t site[3], int32_t itrans, int32_t arm);

/* Replace the state by a handed-over one: LattCoo (len_x*len_y*(len_z+1),
 * x fastest) and the MC / bulk / void lists with NextNXYZ (3x4 ints per
 * record), Ind_Event and ProbTrans for MC particles -- the layout of the
 * engine's mskg_store_state.  The tables and options stay as they are;
 * call orc_refresh_all to re-evaluate the rates in FP64. */
int orc_load_state(orc_sim *s, const int32_t *latt, int32_t n_ad, const int32_t *ad_xyz, const int32_t *ad_nn,
                   const int32_t *ad_ev, const double *ad_prob, int32_t n_bulk, const int32_t *bulk_xyz,
                   const int32_t *bulk_nn, int32_t n_void, const int32_t *void_xyz, const int32_t *void_nn,
                   double time);

/* The engine's frozen-substrate option: atoms with z < z never evaporate.
 * Here they stay MC particles with the vanishing rate 1e-300, so the
 * reference's list bookkeeping (which list Get_Prob reads a neighbour's
 * NextNXYZ from) is unchanged; the engine marks them as not active.
 * No reference counterpart. */
int orc_set_frozen(orc_sim *s, int z);

/* The engine's second-neighbour energy option: the evaporation rate of an
 * atom with 1..3 reciprocal neighbours is multiplied by f2[n2], n2 = the
 * number of occupied (non-wall) points among its 12 second-neighbour
 * positions (+-6,+-6,0 and permutations); after a deposition or evaporation
 * at s, the evaporating MC particles on s's 12 second-neighbour positions get
 * a fresh rate.  f2 = NULL switches it off.  No reference counterpart. */
int orc_set_second_factor(orc_sim *s, const double *f2);

/* Get_Prob for every MC particle, storing rate, event and rate-class key
 * (an adsorption site redraws its species, consuming ran3). */
/* Laser-annealing rates (Get_Prob.f:127-149) for a deck whose tables hold
 * energies: (Tm, P0, sigma, Tflex) as the LA start.dat gives them. */
int orc_set_la(orc_sim *s, double tm, double p0, double sigma, double tflex);
/* The laser-annealing build's form (mulskips-source-LaserAnnealing/mcevent/
 * Get_Prob.f:140-153, main.f:637 and :676-681): per-species constants,
 * InitCountCrystal and CountCrystal; swap: its SiGe species rule applies. */
int orc_set_la_variant(orc_sim *s, const double p0[3], const double tm[3], const double sigma[3],
                       const double tflex[3], const int32_t init_count[3], const int32_t count[3], int swap);
void orc_la_counts(const orc_sim *s, int32_t count[3], int32_t liquid[3]);
/* InitSt 'ST' (mcevent/SetProbabilityFromField.f, InterpProbFromField.f):
 * ntemp increasing reference temperatures and their tables, laid out as the
 * oracle's own (pe 12288, pd 9, pabs 12, pdes 12 per temperature, one after
 * another); every Get_Prob interpolates the working tables at the site's
 * temperature.  ntemp = 0 switches it off. */
int orc_set_st_tables(orc_sim *s, int ntemp, const int32_t *temps, const double *pe, const double *pd,
                      const double *pabs, const double *pdes);
int orc_refresh_all(orc_sim *s);

/* Trajectory-likelihood statistics by brute force (the check for the
 * engine's mskg_adjoint, same class numbering): replays evs[0..n) with
 * waiting times dt and fills S[2510], A[2501] and scal[4] = (sum dt,
 * isolated-rate integral, isolated log-rate sum, log-likelihood evaluated
 * directly as sum_k log r(event k) - Tree(1) dt[k]).  Models without
 * coverage species and LA rates.  No reference counterpart. */
int orc_adjoint_stats(orc_sim *s, const orc_event *evs, int64_t n, const double *dt, double *S, double *A,
                      double *scal);
/* ... and the class populations after every `every` events (hist: n / every
 * rows of KEY_N counts, isolated atoms included). */
int orc_adjoint_checkpoints(orc_sim *s, const orc_event *evs, int64_t n, const double *dt, double *S, double *A,
                            double *scal, int64_t every, int32_t *hist);

/* Write the main xyz frame of WriteMolMolXYZFile.f (and its SiC
 * wrong-site side effect on LattCoo) exactly as the reference does. */
int orc_write_frame(orc_sim *s, const char *xyz_path, const char *d_path,
                    int64_t iter);
/* The same, plus the SiC defect frames of WriteMolMolXYZFile.f: the vacancy
 * frame (_v, :222-260) and the wrong-occupation frame (_w, :264-398; only
 * for NCrystal 2 with C).  Either path may be NULL. */
int orc_write_frames(orc_sim *s, const char *xyz_path, const char *d_path, const char *v_path,
                     const char *w_path, int64_t iter);

/* ---- state access (for building device states and comparing) --------- */
/* The laser-annealing coupling maps (layout 0: DIMENSION(LenZ, LenX*LenY),
 * z fastest; layout 1: x fastest, then y, then z; planes z < LenZ):
 * SetT's temperature update and the phase maps c/d sent back to the FEM
 * side.  Return 0, or -1 for a bad layout. */
int orc_set_temperature_field(orc_sim *s, const int32_t *field, int layout);
int orc_phases(const orc_sim *s, int32_t *c, int32_t *d, int layout);
int32_t *orc_lattcoo(orc_sim *s);   /* (LenX, LenY, LenZ+1), x fastest   */
int32_t *orc_lattind(orc_sim *s);
void orc_dims(const orc_sim *s, int *lx, int *ly, int *lz);
void orc_counts(const orc_sim *s, int32_t *n_ad, int32_t *n_at, int32_t *n_vd);
/* Copy list i (0=MC, 1=bulk, 2=void) out: xyz[3n], nn[12n] (Fortran
 * NextNXYZ(3,4) order), ev[n] and prob[n] only for MC (may be NULL). */
int orc_get_list(const orc_sim *s, int which, int32_t *xyz, int32_t *nn,
                 int32_t *ev, double *prob);
double  orc_time(const orc_sim *s);
double  orc_total_rate(const orc_sim *s);   /* Tree(1) */
int64_t orc_iter(const orc_sim *s);
void orc_count_crystal(const orc_sim *s, int32_t cc[3], int32_t cov[3]);
int  orc_ncrystal(const orc_sim *s, int *ncov);
/* Rate tables as parsed (PTransE(3,0:3,...,0:3) Fortran order, PTransD(3,3),
 * PTransAbs(3,0:3), PTransDes(3,0:3)); any pointer may be NULL. */
void orc_tables(const orc_sim *s, double *pe, double *pd, double *pabs,
                double *pdes, double *ptrans_zig, int32_t list_crystal[3],
                int32_t list_cov[3]);

/* Get_Prob (Get_Prob.f:54-331) evaluated on the current state for the MC
 * particle at xyz, without consuming random numbers and without side
 * effects: returns the event category (0 none, 1 adsorption on an empty
 * site [species/coverage drawn at execution], 2 placeholder deposition
 * (rate 1e-8), 4 evaporation, 5 desorption, 6 isolated-atom evaporation
 * (rate 1e5*R)) and the FP64 rate. */
int orc_pure_rate(orc_sim *s, const int32_t xyz[3], int32_t *category,
                  double *rate);

/* FIND_NN.f / FIND_SN.f restated (FP64 Rodrigues rotation / inversion /
 * translation then NINT), exported for the bond-code table proof. */
void orc_find_nn(const int32_t s1[3], const int32_t s2[3], const int32_t sc[3],
                 int32_t out[2][3], int lx, int ly);
void orc_find_sn(const int32_t site[3], const int32_t sitec[3],
                 const int32_t sn[3][3], int32_t zig[3][3], int32_t arm[3][3],
                 int lx, int ly);

/* Geometry anomalies seen so far: [0] arm-chair neighbours placed at a
 * displacement that is not one of the 32 bond vectors, [1] the same for
 * FIND_NN results, [2] reads of LattCoo outside the box. */
void orc_anomalies(const orc_sim *s, int64_t out[3]);

/* ran3 (random.f:22-69) as a stand-alone generator for tests. */
typedef struct { int32_t ma[56]; int32_t inext, inextp, iff; } orc_ran3;
double orc_ran3_next(orc_ran3 *r, int32_t *idum);

#ifdef __cplusplus
}
#endif
#endif
