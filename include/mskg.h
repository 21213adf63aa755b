/*
 * mskg.h -- C ABI of the MI355X KMC event engine for MulSKIPS.
 *
 * The engine replaces the reference's KMC loop (mulskips-source/main.f:
 * 304-371) and the routines it drives: rate evaluation (mcevent/Get_Prob.f,
 * Get_Prob_Ini.f), event selection (mcevent/PickTreeEvent.f, UpdateTree.f,
 * random.f) and event execution (mcevent/deposition.f, Evaporation.f,
 * absorption.f, desorption.f, Add.f, Erase.f, IF_Arm_Choice.f,
 * lattice/FIND_NN.f, FIND_SN.f).  The Fortran host keeps parsing start.dat,
 * building the initial lattice (initial/*.f) and writing frames
 * (inpout/*.f); it hands the lattice to the engine with mskg_load_state,
 * advances it with mskg_run up to each output boundary and takes it back with
 * mskg_store_state.  Binding stubs for iso_c_binding are in
 * mulskips_amd/fortran/mskg_bindings.f90.
 *
 * Conventions: plain C types only; every entry point returns 0 on success
 * and a nonzero MSKG_E* code on failure (the reference STOPs instead);
 * mskg_last_error() gives the message.  One engine owns one GPU stream and
 * its device memory; calls on one engine must be serialised by the caller,
 * exactly as the reference is single-threaded.  Grid coordinates are the
 * reference's 0-based (x, y, z) with 0<=x<len_x, 0<=y<len_y, 0<=z<=len_z;
 * lattice arrays are (len_x, len_y, len_z+1) with x fastest, the layout of
 * LattCoo (modules/defsystem.f:49).
 */
#ifndef MSKG_H
#define MSKG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  MSKG_OK = 0,
  MSKG_EINVAL = 1,      /* bad argument or configuration                        */
  MSKG_ESTATE = 2,      /* lattice state inconsistent (the reference would STOP) */
  MSKG_ECAPACITY = 3,   /* a caller-provided buffer or a tile list is too small  */
  MSKG_EDEVICE = 4,     /* HIP runtime error or no usable GPU                    */
  MSKG_ENOMEM = 5
};

/* stop_reason values of mskg_run (main.f:307-310, 361-366) */
enum { MSKG_STOP_LIMIT = 0, MSKG_STOP_FEW_MC = 1, MSKG_STOP_NO_EVENTS = 3 };

typedef struct mskg_engine mskg_engine;

/* Everything the reference keeps in MODULE DefSystem / Definitions that the
 * engine needs (defsystem.f:27-68, definitions.f:21-46, main.f:83-212). */
typedef struct {
  int32_t len_x, len_y, len_z;          /* LenX, LenY, LenZ; len_x, len_y multiples of 12 */
  int32_t n_crystal, n_cov;             /* NCrystal, NCov                                */
  int32_t list_crystal[3], list_cov[3]; /* ListCrystal, ListCov (atomic numbers)         */
  const double *ptrans_e;   /* PTransE(3,0:3,0:3,0:3,0:3,0:3,0:3), Fortran order, 12288 values */
  const double *ptrans_d;   /* PTransD(3,1:3), 9 values                           */
  const double *ptrans_abs; /* PTransAbs(3,0:3), 12 values (may be NULL if n_cov=0) */
  const double *ptrans_des; /* PTransDes(3,0:3), 12 values (may be NULL if n_cov=0) */
  double ptrans_zig;        /* PtransZig                                          */
  int32_t la_mode;          /* 1: laser-annealing rates from a per-site T (Get_Prob.f:127-149) */
  double tm, p0, sigma, tflex;
  uint64_t seed;            /* Philox-4x32-10 key for this replica                */
  int32_t tile;             /* synchronous-sublattice tile edge (grid units, even,
                               >=16, len_x/tile and len_y/tile even); 0 = choose  */
  int32_t tile_capacity;    /* max MC particles per tile held on chip; 0 = 1024,
                               at most 6144 (LDS)                                 */
  double events_per_pass;   /* k of the window policy below; 0 = 1.0                */
  int32_t tau_policy;       /* 0: tau = k / (largest tile rate), every tile does at most
                               about k events per colour pass; 1: tau = k * (tiles with MC
                               particles) / (total rate), k events per tile on average */
  int32_t frozen_below_z;   /* engine option without a reference counterpart: atoms with
                               z < frozen_below_z never evaporate (a frozen substrate);
                               0 = off, the reference behaviour */
  const double *second_factor; /* engine option without a reference counterpart: 13 factors
                               f[n2] multiplying the evaporation rate of an atom with 1..3
                               neighbours, n2 = occupied points among its 12 second-neighbour
                               positions (+-6,+-6,0 and permutations), e.g. exp(-n2*E2/kT)
                               for a second-neighbour bond energy E2; NULL = off.  Needs
                               tile >= 24. */
  int32_t stored_len_z;     /* engine option without a reference counterpart: the grid planes
                               z = 0..stored_len_z-1 are held in device memory, the planes
                               above are vacuum the run never reaches (a film that grows to
                               within 24 planes of the stored top stops the run with an
                               error, before any event reads or writes above it); 0 = all
                               len_z+1 planes.  A box of tall vacuum then costs no memory:
                               a stored plane holds len_x*len_y/4 lattice points of 22 B.
                               With it, mskg_load_state reads (and mskg_store_state writes)
                               only the stored planes of LattCoo, the laser-annealing maps
                               and forms are unavailable, and every run, replay, rate and
                               frame result equals that of the engine storing all planes. */
  int32_t reserved[3];
} mskg_config;

/* One logged event, as written by the reference loop instrumented at
 * main.f:312-357 (or by the CPU oracle): the picked site, Ind_Event and, for a
 * deposition at Coor 1, the IF_Arm_Choice outcome. */
typedef struct {
  int64_t iter;
  int32_t site[3];
  int32_t itrans;   /* 1..3 deposit species, 4 evaporate, 5 desorb, 6.. absorb ListCov(k-5) */
  int32_t arm;      /* 1 arm-chair, 0 zig-zag, -1 not applicable */
  int32_t coor;
  double rtot;      /* Tree(1) before the pick */
  double time;      /* clock after the event   */
} mskg_event;

/* Creates the engine on the current HIP device and builds the geometry
 * tables.  Replaces AllocateArrays.f:24-45 and the table set-up of
 * SetProbability.f:21-143 (the host still parses start.dat). */
int mskg_create(const mskg_config *cfg, mskg_engine **out);
void mskg_destroy(mskg_engine *e);
const char *mskg_last_error(const mskg_engine *e);

/* Uploads the host lattice: LattCoo (len_x*len_y*(len_z+1) words, bit layout
 * of defsystem.f:87-91) and the three particle lists with their NextNXYZ
 * (3x4 ints per record, Fortran order), ListAdAtom's Ind_Event/ProbTrans.
 * Replaces nothing in the reference (its lists are module globals); it is
 * the hand-over point after the initial builders (main.f:298). */
int mskg_load_state(mskg_engine *e, const int32_t *latt_coo,
                    int32_t n_ad, const int32_t *ad_xyz, const int32_t *ad_nn,
                    const int32_t *ad_event, const double *ad_prob,
                    int32_t n_bulk, const int32_t *bulk_xyz, const int32_t *bulk_nn,
                    int32_t n_void, const int32_t *void_xyz, const int32_t *void_nn,
                    double kmc_time);

/* Device-side initial state, without a host LattCoo: the flat (001) slab
 * of initial/SetSi.f:25-220 (mode 0: Si on both fcc sublattices, wall
 * layers z=0..5 and len_z-6..len_z-1) or initial/SetSiC3C.f:25-198
 * (mode 1: Si on sublattice A, C on B, no walls) of height len (reduced to
 * a multiple of 24 by the caller, as main.f:239,248 do).  The result equals
 * those builders followed by mskg_load_state, except that the species of an
 * empty site is drawn when it is filled rather than when the slab is built. */
int mskg_build_slab(mskg_engine *e, int32_t mode, int32_t len);

/* The same slab as mode 0 (the SetSi structure with its walls) as a random
 * substitutional alloy: every lattice site holds species 2 (Ge in SiGe)
 * with probability `fraction`, drawn from the engine's Philox stream --
 * the random alloy of initial/SetSiGex.f, which draws from the compiler's
 * RAND().  Needs n_crystal >= 2.  The atoms' rates come from Get_Prob. */
int mskg_build_alloy_slab(mskg_engine *e, int32_t len, double fraction);

/* The shaped initial states of main.f:224-294, built on the device, with
 * InitSt and Len1..Len4 as start.dat gives them (main.f's reductions to
 * multiples of 24 applied where main.f applies them): 'C' the box of
 * initial/SetNC.f:25-235 and 'S' the sphere of SetNCSp.f:25-237 (Si on fcc
 * sublattice A, C on B, centred in the box), 'I' the inverted pyramids of
 * SetInvPOff.f:21-286, 'FF' the fin between walls of SetSiFINFET.f:25-312,
 * and 'F' / 'FS' the slabs of mskg_build_slab.  As the reference: atoms and
 * walls placed, atoms with no occupied neighbour removed, coordination and
 * lists of every site, rates from Get_Prob_Ini (placeholder 1e-6).  The
 * result equals the reference builder followed by mskg_load_state, except
 * that a deposition's species is drawn when it happens.  A temperature set
 * before (laser-annealing mode) stays.  Needs every plane stored. */
int mskg_build_shape(mskg_engine *e, const char *init_st, const int32_t lens[4]);

/* InitSt 'IN' (initial/SetCAD.f:84-418) on the device: ids holds the
 * geometry file, one id per grid point, ids[x + len_x*(y + len_y*z)] for
 * z = 0..len_z-1 (10 a wall on that grid point, 1 an atom on a diamond site,
 * anything else empty; planes 0..3 and len_z-4..len_z-1 are walls whatever
 * they hold), in host memory or, with on_device, device memory.  Atoms are
 * species 2 with probability alloy_fraction (the reference's RAND() <
 * alloyfraction, drawn here from the engine's Philox stream), else 1.
 * Then as mskg_build_shape; an empty site of coordination 4 is a void. */
int mskg_build_cad(mskg_engine *e, const int8_t *ids, int32_t on_device, double alloy_fraction);

/* Per-site temperature in K for laser-annealing mode, (len_x, len_y,
 * len_z+1) like LattCoo's T bits (SetT.f:25-50).  Every value at a lattice
 * point must lie in 0..8191 (else MSKG_EINVAL and nothing changes).  Rates
 * are not re-evaluated: call mskg_refresh_rates. */
int mskg_set_temperature(mskg_engine *e, const int32_t *temp_k);

/* The laser-annealing build's rate form (mulskips-source-LaserAnnealing/
 * mcevent/Get_Prob.f:140-153), for an engine created with la_mode: per-species
 * P0, Tm, sigma and Tflex (its start.dat's Tm(:NCrystal) ... lines), and
 * solidification scaled by CountLiquid(k) / SUM(CountLiquid), CountLiquid =
 * InitCountCrystal - CountCrystal (its main.f:595-637).  init_count and count
 * are InitCountCrystal and the current CountCrystal.  For SiGe (list_crystal
 * 14, 32) a deposition of a species whose liquid is used up deposits the
 * other one (main.f:676-681).  CountCrystal follows the events: exactly, event
 * by event, in mskg_replay (the reference refreshes CountLiquid once per
 * iteration); once per sweep in mskg_run, whose passes read the counts of the
 * sweep's start.  The log of a run records the species deposited.  Loading
 * or building a new state returns to the main build's form; mskg_run_serial
 * does not take this form.  Call after mskg_load_state / mskg_build_slab. */
int mskg_set_la_variant(mskg_engine *e, const double p0[3], const double tm[3], const double sigma[3],
                        const double tflex[3], const int64_t init_count[3], const int64_t count[3]);
/* InitSt 'ST' (mcevent/SetProbabilityFromField.f:21-171): ntemp reference
 * temperatures (K, increasing) with their rate tables, each laid out as
 * mskg_config's tables and stored one after another (pe: ntemp x 12288,
 * pd: ntemp x 9, pabs and pdes: ntemp x 12).  Every rate evaluation takes
 * the tables at the site's temperature (the T bits) as
 * InterpProbFromField.f:22-57 interpolates them (Get_Prob.f:77-90); a
 * temperature outside the references is an error (the reference indexes
 * out of its arrays there).  The 'ST' run with one temperature throughout
 * (fixedT, main.f:114) is an ordinary run with the tables interpolated on
 * the host.  ntemp = 0 switches it off.  Excludes la_mode; rates are left as
 * they are (mskg_refresh_rates). */
int mskg_set_st_tables(mskg_engine *e, int32_t ntemp, const int32_t *temps, const double *pe, const double *pd,
                       const double *pabs, const double *pdes);
/* CountCrystal and CountLiquid as the engine holds them now (either may be null). */
int mskg_la_counts(mskg_engine *e, int64_t count[3], int64_t liquid[3]);

/* Layouts of the whole-grid int32 maps the laser-annealing build exchanges
 * with its FEM driver (mulskips-source-LaserAnnealing/modules/defsystem.f:88-91,
 * main.f:211-334):
 *   MSKG_FIELD_SOCKET  DIMENSION(LenZ, LenX*LenY): element z + LenZ*(x + LenX*y)
 *   MSKG_FIELD_PLANES  the temperature / Coor files: x + LenX*(y + LenY*z)
 * both for the planes z = 0..LenZ-1. */
#define MSKG_FIELD_SOCKET 0
#define MSKG_FIELD_PLANES 1

/* SetT (mulskips-source-LaserAnnealing/initial/SetT.f:38-48 for the socket
 * map, :56-66 for the file), on the device:
 * every point of the planes z < LenZ takes the low 13 bits of its map value
 * as its temperature, except where the value is 10 (the FEM side's wall
 * marker), which leaves the point as it is.  `field` is a host pointer, or a
 * device pointer on the engine's device when on_device != 0 (ready before the
 * call; it is read on the engine's stream).  Replaces the socket branch of
 * SetT, whose array b arrives with SENDFIELD (main.f:211-230); like SetT it
 * leaves the rates as they are (mskg_refresh_rates re-evaluates them). */
int mskg_set_temperature_field(mskg_engine *e, const int32_t *field, int32_t layout, int32_t on_device);

/* The phase maps the laser-annealing build returns to its FEM driver
 * (mulskips-source-LaserAnnealing/main.f:796-818, sent with GETPHASES /
 * GETXGE, :273-335): for the grid points of the
 * planes z < LenZ, coor = coordination and species = species index (1..3) of
 * an occupied point, 10 and 10 on a wall (index 113), 0 and 0 elsewhere.
 * Either output may be null; host pointers, or device pointers when
 * on_device != 0 (complete when the call returns). */
int mskg_phases(mskg_engine *e, int32_t *coor, int32_t *species, int32_t layout, int32_t on_device);

/* Rate-update sweep: re-evaluates Get_Prob (Get_Prob.f:54-331) for every
 * MC particle of the lattice from the current state, reading every stored
 * point.  */
int mskg_refresh_rates(mskg_engine *e);

/* The same sweep over the tile layers that hold MC particles only (no
 * reference counterpart: Get_Prob.f:54-331 over the MC particles, which
 * lie in those layers).  The engine knows those layers while its tile
 * lists are current (after mskg_run, mskg_run_serial or a previous
 * mskg_refresh_active); otherwise -- after an upload, a builder, a replay
 * or mskg_refresh_rates -- it sweeps every point.  The rates and classes
 * it leaves equal mskg_refresh_rates' bit for bit, and mskg_last_timing
 * reports the points it read. */
int mskg_refresh_active(mskg_engine *e);

/* Advances the lattice with the parallel synchronous-sublattice KMC until
 * max_events events have executed or the clock passes max_time (whichever
 * first; the check is made between sweeps).  Replaces main.f:304-371. */
int mskg_run(mskg_engine *e, int64_t max_events, double max_time,
             int64_t *events_done, double *kmc_time, int32_t *stop_reason);

/* How mskg_run orders the eight colour passes of a sweep (no reference
 * counterpart; the reference's loop is serial).  MSKG_SCHEDULE_COLOURS (the
 * default): one launch per colour, each waiting for the whole previous
 * colour.  MSKG_SCHEDULE_TILES: one launch per sweep in which a tile's pass
 * starts as soon as the neighbour tiles of lower colours have finished
 * theirs (a queue of tiles in colour order, per-tile completion flags with
 * agent-scope release / acquire).  Both give the same lattice, clock and
 * counters bit for bit: a tile sees the same state, and draws the same
 * random numbers, in either order.  On MI355X the tile schedule removes the
 * idle time at the colour barriers but each acquire invalidates the L2 of
 * the waiting wave's XCD, which the other waves of that XCD then refill:
 * the barrier schedule is the faster one there. */
#define MSKG_SCHEDULE_TILES 1
#define MSKG_SCHEDULE_COLOURS 0
int mskg_set_schedule(mskg_engine *e, int32_t schedule);

/* Record every event mskg_run executes (capacity events; 0 switches the log
 * off).  The order is a valid serial order of the parallel run: events of
 * one tile are in execution order, a tile's events follow those of its
 * neighbour tiles that ran before it in the sweep, and events of tiles that
 * are not neighbours touch disjoint sites, so they commute.  mskg_get_event_log
 * copies up to cap events and reports how many were recorded. */
int mskg_set_event_log(mskg_engine *e, int64_t capacity);
int mskg_get_event_log(mskg_engine *e, mskg_event *out, int64_t cap, int64_t *n);

/* Applies a logged event sequence in order with the reference's update rules
 * (no random numbers are drawn: species and stacking come from the log).
 * clock_mode 0 advances the clock by 1/rtot per event (main.f:313). */
int mskg_replay(mskg_engine *e, const mskg_event *log, int64_t n, int32_t clock_mode);

/* Downloads the lattice back into LattCoo and the three lists (ordered by
 * grid position; capacities in records).  Ind_Event of an adsorption site is
 * reported as 1; ProbTrans is the FP32 rate widened to FP64 (0 for the atoms
 * of a frozen substrate, which are listed as MC particles like any atom of
 * coordination <= 3). */
int mskg_store_state(mskg_engine *e, int32_t *latt_coo,
                     int32_t cap_ad, int32_t *n_ad, int32_t *ad_xyz, int32_t *ad_nn,
                     int32_t *ad_event, double *ad_prob,
                     int32_t cap_bulk, int32_t *n_bulk, int32_t *bulk_xyz, int32_t *bulk_nn,
                     int32_t cap_void, int32_t *n_void, int32_t *void_xyz, int32_t *void_nn);

/* Event class and FP32 rate of n sites (class codes: 0 none, 1 adsorption,
 * 2 placeholder deposition, 4 evaporation, 5 desorption, 6 isolated atom). */
int mskg_site_rates(mskg_engine *e, int64_t n, const int32_t *xyz, int32_t *cls, float *rate);

/* Counters since load: executed events by reference event type
 * (counts[0..2] deposition of species 1..3, [3] evaporation, [4] desorption,
 * [5..7] absorption of ListCov 1..3), the KMC clock, and list sizes
 * (NumAdAtom, NumAtoms, NumVoids of defsystem.f:36). */
int mskg_counters(mskg_engine *e, int64_t counts[8], double *kmc_time,
                  int64_t *n_adatom, int64_t *n_atoms, int64_t *n_voids);

/* Evaporations by species (1..3) and desorptions by coverage species
 * (ListCov 1..3) since the state was loaded: with the depositions and
 * absorptions of mskg_counters, the host keeps CountCrystal and CountCov
 * (main.f:341,356, Evaporation.f:41, desorption.f:43-47), which
 * WriteMolMolSource.f:43-54 prints.  They follow from a census of the
 * lattice (one pass over the stored points) against the one taken at load. */
int mskg_species_counts(mskg_engine *e, int64_t evaporated[3], int64_t desorbed[3]);

/* Timing of the last mskg_run / mskg_refresh_rates (device milliseconds,
 * measured with HIP events on the engine's stream) and the number of lattice
 * points the last rate sweep read. */
int mskg_last_timing(mskg_engine *e, double *run_ms, double *sweep_ms, int64_t *sweep_points,
                     int64_t *sweeps_done);

/* Diagnostics of the synchronous-sublattice runs since the state was loaded
 * (no reference counterpart; the reference's serial loop has no windows):
 * guard_exits counts tile windows that ended at the cap of 65536 events per
 * tile and window instead of at the window's end -- a nonzero count means the
 * window policy lets one tile run far ahead of the others and the run is not
 * a faithful splitting of the master equation (lower events_per_pass or use
 * tau_policy 0); sweeps counts the sweeps that advanced the clock, tau is the
 * window the next sweep will use, n_mc the MC particles at the last sweep.
 * Any output may be null. */
int mskg_run_diagnostics(mskg_engine *e, int64_t *guard_exits, int64_t *sweeps, double *tau, int64_t *n_mc);

/* Serial BKL on a batch of independent lattices in one launch (one
 * workgroup per engine): the reference's algorithm (mcevent/PickTreeEvent.f,
 * main.f:304-371) with the selection over all MC particles of the lattice
 * by an FP64 three-level rate sum and the clock advanced by -ln(u)/R per
 * event (Philox-4x32-10, key = the engine's seed).  Engines keep their own
 * state, counters and event logs (the log's time is the clock after each
 * event and rtot the R of its selection).  All engines must live on the
 * current device; the launch uses engines[0]'s stream.  max_events[i] and
 * max_time[i] bound replica i.  stop_reason as mskg_run; the few-MC stop
 * does not apply.  A replica's trajectory does not depend on the batch it
 * runs in. */
int mskg_run_serial(mskg_engine *const *engines, int32_t n, const int64_t *max_events, const double *max_time,
                    int64_t *events_done, double *kmc_time, int32_t *stop_reason);

/* The batched ensemble (no reference counterpart; the reference runs one
 * lattice per process, main.f:304-371): the synchronous-sublattice sweep of
 * mskg_run on n independent lattices at once.  Each colour pass is one
 * launch over the tiles of that colour of every replica still running,
 * taken from one queue by persistent waves, so small replicas share the
 * GPU with large ones and the passes' slowest tiles overlap other
 * replicas' work; the per-replica window, clock and stop follow after each
 * sweep in one launch, and a sweep (8 passes, tile sums, sweep end, the
 * next sweep's plan) is one captured graph.  Replica i runs to
 * max_events[i] events or max_time[i] (as mskg_run) with its own tables,
 * tile edge, window policy, seed and event log; its lattice, lists,
 * counters and clock after the call equal, bit for bit, those of mskg_run
 * on that engine alone with the same limits.  Engines must be distinct,
 * live on the current device and share the rate-model kernel (all or none
 * laser-annealing / 'ST', all or none with coverage rates); at most 1024.
 * The launches use engines[0]'s stream.  Stop reasons as mskg_run's; on a
 * replica's error the others finish and the call returns MSKG_ESTATE with
 * that replica's message in mskg_last_error(engines[0]). */
int mskg_run_ensemble(mskg_engine *const *engines, int32_t n, const int64_t *max_events, const double *max_time,
                      int64_t *events_done, double *kmc_time, int32_t *stop_reason);

/* Trajectory likelihood statistics (no reference counterpart; for parameter
 * calibration).  For each engine i, loaded with the initial state of a
 * recorded trajectory logs[i][0..n_events[i]) with waiting times dts[i][k]
 * (the time spent in the state before event k), the engine replays the log
 * and returns
 *   S[i][c]  the number of events drawn from rate class c,
 *   A[i][c]  sum over events k of (points in class c before event k) * dt_k,
 * so that log L = sum_c S_c log r_c - A_c r_c + scalars[i][2] - scalars[i][1]
 * for any rate tables r, and d log L / d theta = sum_c (S_c/r_c - A_c) dr_c/dtheta.
 * An engine that tracks class keys (mskg_track_classes) starts from its
 * current rates and keys; any other engine first refreshes every rate (as
 * mskg_refresh_rates).
 * Classes (MSKG_KEY_*): evaporation of species s with n1 (Si or wall), n2,
 * n3 reciprocal neighbours of species 1..3 and m occupied second-neighbour
 * positions: key (s-1) + 3*(n1 + 4*(n2 + 4*n3)) + 192*m; an adsorption site
 * of coordination c: MSKG_KEY_ADS + c - 1 (its A; its selections are split
 * by deposited species in S[MSKG_KEY_N + 3*(s-1) + c-1]); the placeholder
 * deposition MSKG_KEY_PH (rate 1e-8); isolated atoms MSKG_KEY_ISO, whose
 * rates 1e5*R are per point and enter only through scalars (their A is the
 * integral of their number).
 * S: n x MSKG_ADJ_NS, A: n x MSKG_KEY_N, scalars: n x 4 = (sum of dt,
 * integral of the isolated-atom rates, sum of their log rates, events
 * replayed).  Models without coverage species and without LA rates only.
 * The engines end in the trajectory's final state. */
enum { MSKG_KEY_ADS = 2496, MSKG_KEY_PH = 2499, MSKG_KEY_ISO = 2500, MSKG_KEY_N = 2501, MSKG_ADJ_NS = 2510 };

/* Keep every point's rate-class key current through runs (on != 0; a keyed
 * rate sweep sets them first when a state is loaded).  A tracked engine's
 * state can be copied mid-trajectory (mskg_copy_state) and handed to
 * mskg_adjoint as the start of a trajectory segment: the adjoint then uses
 * the copied rates and keys instead of refreshing them, so the segments of
 * one trajectory replay independently and their statistics add up to the
 * whole trajectory's.  No reference counterpart. */
int mskg_track_classes(mskg_engine *e, int32_t on);

/* Copy src's lattice (records, rates, classes, class keys when both track
 * them) and clock into dst, an engine of the same box and tile. */
int mskg_copy_state(mskg_engine *dst, const mskg_engine *src);
int mskg_adjoint(mskg_engine *const *engines, int32_t n, const mskg_event *const *logs, const double *const *dts,
                 const int64_t *n_events, double *S, double *A, double *scalars);

/* mskg_adjoint with checkpoints (no reference counterpart): with every > 0,
 * hist[i] receives, for replica i, the class populations n_c (MSKG_KEY_N
 * int32 each, isolated atoms included) after events every, 2*every, ...:
 * n_events[i] / every rows.  The populations at 1e4-event intervals are the
 * rate tables of a class-keyed model at those times (R = sum_c n_c r_c). */
int mskg_adjoint_checkpointed(mskg_engine *const *engines, int32_t n, const mskg_event *const *logs,
                              const double *const *dts, const int64_t *n_events, double *S, double *A, double *scalars,
                              int64_t every, int32_t *const *hist);

/* Frame compaction for the xyz output (inpout/WriteMolMolXYZFile.f:77-260):
 * one pass over the lattice writes, as (x, y, z, code) int32 quadruples in
 * the reference's grid order (z, then y, then x; points with z <= LenZ-1):
 *   main_rec  the points the main frame lists -- atoms of coordination <= 3
 *             (code = species 1..3) and coverage points (code = -k for
 *             ListCov(k));
 *   void_rec  vacancies (empty, coordination 4; code 0);
 *   noepi_rec bulk atoms off the cubic diamond sites, the _d frame
 *             (code = species);
 * and counts = (main, voids, non-epitaxial records, atoms of coordination
 * <= 3 with 1 <= z <= LenZ-1, covered points with 1 <= z <= LenZ-1), the
 * last two for the frame header.  Call with zero capacities to size the
 * buffers.  The reference lists the vacancies in ListVoid order, which
 * depends on the run's history; here they come in grid order.  The
 * reference's writer also flips and mis-restores the species of SiC
 * "wrong" sites (:262-398); that side effect on the lattice is not
 * reproduced. */
int mskg_frame(mskg_engine *e, int32_t *main_rec, int64_t cap_main, int32_t *void_rec, int64_t cap_void,
               int32_t *noepi_rec, int64_t cap_noepi, int64_t counts[5]);

/* mskg_frame without a single allocation: no reference counterpart; for
 * callers that cannot call hipMalloc -- code under HIP graph capture, tight
 * serving loops, harnesses that forbid allocation.  The records go to the
 * caller's DEVICE buffers (int32 quadruples, capacities in records; a list
 * longer than its capacity is cut there, records beyond it are not
 * written), and counts is a DEVICE array of 5 int64 receiving the full
 * totals (main, voids, non-epitaxial records, the two header counts), so
 * counts[k] > cap_k tells the caller the list was cut.  The block counts and
 * offsets of the compaction live in a fixed device array of the library
 * (16384 blocks of at most 2^30 lattice points each; larger lattices are
 * rejected), one per GPU, shared by every call on that GPU: calls that
 * overlap on one GPU race.  Returns when the records are written. */
int mskg_frame_device(mskg_engine *e, int32_t *main_rec, int64_t cap_main, int32_t *void_rec, int64_t cap_void,
                      int32_t *noepi_rec, int64_t cap_noepi, int64_t *counts);

/* The SiC defect frames of the xyz output (inpout/WriteMolMolXYZFile.f),
 * gathered on the device as (x, y, z, code) int32 quadruples in the
 * reference's grid order (points with z <= LenZ-1):
 *   vac_rec    every vacancy (an empty or covered point of coordination 4),
 *              code = the sum of the species indices (1 Si, 2 C/Ge, 3 B; 0
 *              for walls and empty points) of its four NextNXYZ points -- the
 *              _v frame's classification (:222-260: 4 CV, 8 SV, 7 SAV, 5 CAV,
 *              6 XV for SiC PVD);
 *   wrong_rec  the _w frame (:264-398), for two crystal species one of which
 *              is C: bulk atoms with 1 <= z <= LenZ-1 and three neighbours of
 *              the other species, read while the isolated wrong atoms (bulk
 *              atoms with no such neighbour, taken in grid order as the
 *              reference flips them) are flipped; code = the species (1 Si,
 *              2 C) as flipped.
 * counts = (vacancies, wrong-occupation records).  Zero capacities size the
 * buffers.  The reference lists vacancies in ListVoid order, which depends on
 * the run's history; here they come in grid order.  The reference's writer
 * leaves some flipped species behind (its restore loop reads one buffer
 * entry for all); the engine restores the lattice exactly.  Exact when the
 * neighbour lists of neighbouring bulk atoms are reciprocal, as the event
 * rules keep them. */
int mskg_frame_defects(mskg_engine *e, int32_t *vac_rec, int64_t cap_vac, int32_t *wrong_rec, int64_t cap_wrong,
                       int64_t counts[2]);

/* Host evaluation of the engine's FP64 geometry -- the same source the
 * kernels run -- for checking it against the reference formulas without a
 * GPU.  FIND_NN (lattice/FIND_NN.f:25-77): out = the two new neighbours of
 * sc given s1 and s2.  FIND_SN (lattice/FIND_SN.f:26-84): zig/arm = the
 * zig-zag and arm-chair candidates (3x3, row per candidate). */
int mskg_geometry_find_nn(const int32_t s1[3], const int32_t s2[3], const int32_t sc[3], int32_t len_x,
                          int32_t len_y, int32_t out[6]);
int mskg_geometry_find_sn(const int32_t site[3], const int32_t sitec[3], const int32_t sn[9], int32_t len_x,
                          int32_t len_y, int32_t zig[9], int32_t arm[9]);

#ifdef __cplusplus
}
#endif
#endif
