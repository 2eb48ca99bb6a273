/* C ABI of libpetite_hip.so: the MI355X (gfx950) shower-transport hot path of
 * PETITE (kjkellyphys/PETITE).
 *
 * The reference has no native boundary (it is pure Python); these entry points
 * replace the per-particle Python loops listed next to each declaration and
 * are bound from Python with ctypes (petite_amd/_lib.py).  All array arguments
 * are DEVICE pointers unless marked "host"; the caller owns every buffer it
 * passes.  A context owns only its copy of the tables and its scratch queues.
 * Functions return PETITE_OK (0) or an error code; petite_last_error() gives
 * the message of the calling thread's last failure.  A context is not
 * re-entrant: calls on one context must not overlap.
 */
#ifndef PETITE_HIP_H
#define PETITE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PETITE_OK 0
#define PETITE_ERR_ARG 1
#define PETITE_ERR_HIP 2
#define PETITE_ERR_NO_SAMPLE 3 /* draw_sample's "No Sample Found" (src/PETITE/shower.py:460-461) */
#define PETITE_ERR_CAPACITY 4  /* particle store too small for the shower output */
#define PETITE_ERR_INTERNAL 5  /* a device queue overflowed its region (a library bug, never the caller's input) */

#define PETITE_NPROC 12 /* Brem Ann PairProd Comp Moller Bhabha MuonE MuonBrem DarkBrem DarkAnn DarkComp DarkMuonBrem */

typedef struct PetiteCtx PetiteCtx;
typedef void* PetiteStream; /* a hipStream_t (0 = null stream) */

/* VEGAS grid of one process (nodes per E-index: ndim blocks of ninc[d]+1). */
typedef struct {
  int32_t n_E;
  int32_t ndim;
  int32_t ninc[4];
  int64_t E_off;    /* offset into the float64 blob: n_E grid energies */
  int64_t node_off; /* offset into the float32 blob: n_E * stride nodes */
  int64_t stride;   /* nodes per energy = sum_d (ninc[d] + 1) */
  double lnE0;      /* geometric energy grid: ln E[0] and (n_E-1)/ln(E[n-1]/E[0]); */
  double inv_dlog;  /* inv_dlog = 0 for a grid that is not geometric (binary search) */
} PetiteGridDesc;

/* Linear n*sigma table [1/cm], zero outside [E[0], E[n-1]]. */
typedef struct {
  int32_t n;
  int32_t pad;
  int64_t E_off; /* float64 blob offsets */
  int64_t v_off;
  double lnE0;     /* as in PetiteGridDesc */
  double inv_dlog;
} PetiteSigmaDesc;

/* Material, cuts and tables (src/PETITE/shower.py:100-142, 189-246, 273-295). */
typedef struct {
  double Z, A, rho, dEdx;  /* dEdx in GeV/m */
  double mT;               /* target mass put in event_info: A_T (shower.py:435) */
  double min_energy;       /* tracking cutoff [GeV] */
  double Eg_min, Ee_min;   /* from the Brem grid parameters (shower.py:214-215) */
  double rescale_MCS;
  double maxF_fudge;
  double min_calc[5];      /* minimum calculable energy for PID 11, -11, 22, 13, -13 */
  double mV;               /* dark vector mass [GeV] (0 for the SM path) */
  double Zeff;             /* bound-electron effective charge (dark_shower.py:78) */
  int64_t max_trials;      /* cap on accept/reject candidates (max_n_integrators * neval) */
  int32_t MS;              /* multiple scattering on charged particles (generate_shower GlobalMS) */
  int32_t pad;
  PetiteGridDesc grid[PETITE_NPROC];
  PetiteSigmaDesc sigma[PETITE_NPROC];
} PetiteTableDesc;

/* One primary particle (generate_shower's p0; src/PETITE/particle.py:53-95). */
typedef struct {
  double E, px, py, pz; /* four-momentum [GeV] */
  double x, y, z;       /* start position [m] */
  double mass;          /* particle mass [GeV] */
  double weight;
  int32_t pid;
  int32_t shower;       /* shower index written to the output records */
  uint64_t sid;         /* Philox stream id of the primary */
} PetitePrimary;

/* Device particle store (structure of arrays, capacity entries each).  One
 * entry per output Particle: p0 (E,px,py,pz), r0 (x,y,z,weight), pf, rf
 * (x,y,z,steps), pid, parent slot (-1 for primaries), meta (generation |
 * process+1 << 16 | child index << 24), shower index, stream id. */
typedef struct {
  float* p0;
  float* r0;
  float* pf;
  float* rf;
  int32_t* pid;
  int32_t* parent;
  int32_t* meta;
  int32_t* shower;
  uint64_t* sid;
  int64_t capacity;
} PetiteStore;

/* Counters of one run (host struct, filled on return). */
typedef struct {
  int64_t n_particles;  /* records written */
  int64_t n_steps;      /* particle-steps (charged substeps + closing increments + photon flights) */
  int64_t n_draws;      /* accept/reject draws */
  int64_t n_trials;     /* candidates evaluated */
  int64_t n_maxf_viol;  /* accepted candidates with J*f > max_F */
  int64_t n_no_sample;  /* draws that hit max_trials */
  int64_t n_iterations; /* generation sweeps */
  int64_t n_dropped;    /* children below min_energy */
  /* kernel times measured with HIP events on the run's stream (0 unless
   * timing is enabled with petite_set_timing) */
  double ms_transport, ms_draw, ms_emit;
  int64_t launches_transport, launches_draw, launches_emit;
  int64_t n_draws_proc[PETITE_NPROC];  /* draws per process code */
  int64_t n_trials_proc[PETITE_NPROC]; /* candidates per process code */
  int64_t n_viol_proc[PETITE_NPROC];   /* accepted draws with J*f above max_F per process code (the
                                        * accept probability min(1, J*f / max_F) clips them, as the
                                        * reference's test does) */
  double ms_tail;          /* depth-first walk time (HIP events, with timing enabled) */
  int64_t launches_tail;   /* depth-first walk launches */
  int64_t n_steps_tail;    /* the particle-steps of n_steps taken in depth-first walks */
} PetiteStats;

/* Create a context: copies the table blobs (host pointers) to the device. */
int petite_ctx_create(const PetiteTableDesc* desc, const double* blob64, int64_t n64, const float* blob32,
                      int64_t n32, PetiteCtx** out);
void petite_ctx_destroy(PetiteCtx* ctx);
const char* petite_last_error(void);
/* Enable (1) or disable (0) per-kernel HIP-event timing in petite_run_showers.
 * No counterpart in the reference: it feeds the roofline figures. */
int petite_set_timing(PetiteCtx* ctx, int32_t on);

/* Launch and queue sizes of petite_run_showers.  No counterpart in the
 * reference, whose driver is a sequential Python FIFO (src/PETITE/shower.py:
 * 634-706); the values change speed only, never results: every particle draws
 * from its own Philox stream.  A field <= 0 selects the built-in default
 * (given in brackets).  Block counts are in waves per compute unit (4 waves
 * per block). */
typedef struct {
  int32_t batch_lanes;       /* transport: lanes of a wave that must wait for a refill / finish
                              * before the wave runs one (1..64) [44] */
  int32_t yield_steps;       /* transport: substeps of one particle in one launch before it is handed
                              * on to the next sweep [40]; a value above 2^30 never hands on */
  int32_t transport_waves;   /* transport waves per CU [20] */
  int32_t emit_waves;        /* emission waves per CU [32] */
  int32_t draw_waves[8];     /* draw waves per CU for SM process code p [Brem 16, PairProd 8, others 2] */
  int32_t min_group_lanes;   /* draws: the fewest lanes that share one request (1, 2, 4, .., 64) [4] */
  int32_t sweep_group;       /* sweeps queued per host check (1..64) [16] */
  int32_t group_requests;    /* draws with fewer requests than lanes: requests a lane group is sized to
                              * serve in one launch (1..16) [4] */
  int32_t small_sweep;       /* sweeps with fewer particles than this (seen at a host check) run the
                              * draws of every process in one launch on the caller's stream, with
                              * launch sizes scaled to the sweep [2^20]; 1: never */
  int64_t sweep_bound;      /* particles per sweep, resumed ones included (>= 64) [2^23] */
  int64_t resume_cap;        /* hand-on records per sweep (1..sweep_bound) [min(sweep_bound, 2^20)] */
  int32_t small_yield_steps; /* yield_steps of the small sweeps [40; measured: 2^30 (no hand-on)
                              * is 4 ms per step slower on the 100k lead workload] */
  int32_t graph;             /* 1: a sweep group with per-process draw streams is captured once per
                              * run as a HIP graph and replayed; 2: direct launches [2] */
  int32_t tail_sweep;        /* once fewer particles than this wait for the next sweep (seen at a host
                              * check), each remaining particle's subtree is walked depth-first by one
                              * wave (transport, draw and children in one kernel) [16384]; 1: never */
  int32_t tail_waves;        /* depth-first walk waves per CU [8] */
  int32_t tail_depth;        /* depth-first walk: stack entries per wave (1..128); a child past a full
                              * stack waits for the next launch [128] */
  int32_t tail_budget;       /* depth-first walk: particles a wave walks from one claimed particle before
                              * it leaves the rest of its stack to the next launch [16]; a value above
                              * 2^30 never leaves it */
} PetiteTuning;

/* Set (tuning != NULL) or reset to the defaults (NULL) the context's launch
 * sizes; petite_get_tuning reports the values in effect (defaults filled in). */
int petite_set_tuning(PetiteCtx* ctx, const PetiteTuning* tuning);
int petite_get_tuning(const PetiteCtx* ctx, PetiteTuning* tuning);

/* max_F table of one process for the context material: the largest J*f over
 * n_samples map candidates at each grid energy (utilities/find_maxes.py:55-119).
 * Writes the device copy used by the draws; maxf_host/mean_host (n_E doubles,
 * may be NULL) receive the maxima and the sigma estimates (mean J*f). */
int petite_maxf_scan(PetiteCtx* ctx, int32_t process, int32_t n_samples, uint64_t seed, double* maxf_host,
                     double* mean_host, PetiteStream stream);
/* Replace the max_F table of one process (host array of n_E doubles). */
int petite_set_maxf(PetiteCtx* ctx, int32_t process, const double* maxf_host, int32_t n_E);

/* Batched accept/reject draws: replaces Shower.draw_sample
 * (src/PETITE/shower.py:401-465) and DarkShower.draw_dark_sample
 * (src/PETITE/dark_shower.py:649-704).  Request i draws process proc[i] at
 * energy E[i] from the Philox stream sid[i]; x_out[4i..4i+3] receives the
 * accepted point, trials_out[i] the candidate count (the reference's VB
 * count) and w_out[i] its J*f.  A request with no accepted candidate within
 * max_trials gets trials_out = -max_trials (0 for a process without a grid)
 * and makes the call return PETITE_ERR_NO_SAMPLE. */
int petite_vegas_draw(PetiteCtx* ctx, const int32_t* proc, const double* E, const uint64_t* sid, int64_t n,
                      uint64_t seed, float* x_out, int32_t* trials_out, float* w_out, PetiteStream stream);

/* petite_vegas_draw without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, results and error codes as petite_vegas_draw; the request
 * regions and counters live in one static device array per GPU (2.4 MB, held
 * until the process ends) instead of context scratch, and the draw kernels run
 * one after another on `stream`.  At most PETITE_NOALLOC_MAX_DRAWS requests
 * per call (larger n: PETITE_ERR_ARG).  Every call on a GPU shares that GPU's
 * array: calls that overlap on one GPU (other streams, other threads, other
 * contexts) race and must not be issued. */
#define PETITE_NOALLOC_MAX_DRAWS 16384
int petite_vegas_draw_noalloc(PetiteCtx* ctx, const int32_t* proc, const double* E, const uint64_t* sid, int64_t n,
                              uint64_t seed, float* x_out, int32_t* trials_out, float* w_out, PetiteStream stream);

/* J*f of single map candidates: for request i, the map of proc[i] at
 * LU_Key(E[i]) applied to the four Philox words words[4i..4i+3] (the same
 * words a draw consumes), and the integrand at the true energy E[i].  The
 * per-candidate quantity compared in draw_sample's accept test
 * (src/PETITE/shower.py:453-457).  x_out: n x 4 doubles, w_out: n doubles. */
int petite_eval_points(PetiteCtx* ctx, const int32_t* proc, const double* E, const uint32_t* words, int64_t n,
                       double* x_out, double* w_out, PetiteStream stream);

/* Whole showers: replaces Shower.generate_shower's particle loop
 * (src/PETITE/shower.py:603-708) with propagate_particle (509-601),
 * sample_scattering (467-507) and draw_sample (401-465) for every particle.
 * prim: device array of n primaries.  The store receives every output
 * particle (primaries first, then children by generation).  shower_energy
 * (device, 4 doubles per shower, may be NULL) accumulates deposited dE/dx,
 * final energy of particles that end, energy of dropped children and the
 * number of atomic-electron targets (in units of m_e). */
int petite_run_showers(PetiteCtx* ctx, const PetitePrimary* prim, int64_t n, uint64_t seed, PetiteStore* store,
                       double* shower_energy, PetiteStats* stats, PetiteStream stream);

/* Single steps of the shower, for the reference's lower-level methods.
 *
 * One particle at a hard interaction: its final four-momentum (pf), position
 * (rf), weight, species, generation and Philox stream id. */
typedef struct {
  double E, px, py, pz;
  double x, y, z;
  double weight;
  int32_t pid;
  int32_t generation;
  uint64_t sid;
} PetiteVertex;

/* Shower.sample_scattering (src/PETITE/shower.py:467-507) for n vertices
 * (device arrays): vertex i interacts by SM process proc[i] (codes 0..7; the
 * caller has checked the threshold of shower.py:469-470): draw_sample at
 * E = v[i].E from stream v[i].sid, the final state rotated onto (px, py, pz),
 * both children written to out entries 2i and 2i+1 (p0 = pf, r0 = rf =
 * the vertex position with its weight, parent = i, meta as in
 * petite_run_showers, stream ids derived from v[i].sid) whatever their
 * energy.  trials_out[i] (device) receives the candidate count (<= 0: no
 * sample; that vertex's entries are zero); PETITE_ERR_NO_SAMPLE when any
 * draw found none. */
int petite_sample_scattering(PetiteCtx* ctx, const PetiteVertex* v, const int32_t* proc, int64_t n, uint64_t seed,
                             PetiteStore* out, int32_t* trials_out, PetiteStream stream);

/* Shower.propagate_particle (src/PETITE/shower.py:509-601) as generate_shower
 * calls it (606-660): photons one exponential flight without scattering,
 * charged leptons substeps with the material's dE/dx and, if the context was
 * created with MS, multiple scattering.  The first n records of `recs` (p0,
 * r0 with the weight, pid, meta generation, shower, sid: a store laid out as
 * petite_run_showers writes it) receive pf and rf (rf.w = the particle-steps
 * taken); nothing else is written. */
int petite_propagate(PetiteCtx* ctx, PetiteStore* recs, int64_t n, uint64_t seed, PetiteStream stream);

/* DarkShower.draw_pe_sample (src/PETITE/dark_shower.py:706-719): the bound
 * electron's momentum [GeV] from its 1s density (effective charge Zeff), one
 * per stream id (device arrays of n); -1 where 65536 accept/reject pairs
 * found none.  No context: the sampler needs no tables. */
int petite_draw_pe(const uint64_t* sid, int64_t n, uint64_t seed, double Zeff, double* pe_out, PetiteStream stream);

/* ----------------------------------------------------------- dark sector --
 * Weight families (GetBSMWeights, src/PETITE/dark_shower.py:595-647):
 * 0 e- DarkBrem, 1 e+ DarkBrem, 2 mu DarkMuonBrem, 3 e+ DarkAnn (bound electron). */
#define PETITE_DARK_FAMILIES 4
#define PETITE_DARK_DRATE_BINS 10

/* Track-length quadrature of the dark tables.  Replaces the scipy quad
 * integrals of DarkShower's weight arrays (construct_brem_weight_array,
 * construct_annihilation_weight_array, construct_muon_brem_weight_array,
 * src/PETITE/dark_shower.py:337-399) and rate rows (_d_rate_d_E_*, 454-533),
 * with the survival factor of the SM interaction integrals
 * (src/PETITE/shower.py:298-354).  Each job integrates
 *   n sigma(E) / (dE/dx) * exp(-sum_c [II_c(Ei) - II_c(E)] / (dE/dx))
 * over E (cm units as the reference: n sigma in 1/cm, dE/dx in GeV/cm), with
 * II_c the running integral of survival channel c's n sigma on its grid,
 * either from `lo` to each end energy Ei (mode 0, the weight W(Ei)) or over
 * ten equal bins of [max(Ei - 10 mfp(Ei) dE/dx, lo), Ei] (mode 1, the drate
 * row: bin centre and integral per bin).  Tables are offsets into one float64
 * blob: x values at off, y values at off + n.  Gauss-Legendre panels (8
 * points) break at every node of `nodes` and at geometric distances below Ei
 * that follow the survival factor; results agree with adaptive quadrature to
 * the reference tables' precision. */
typedef struct {
  int64_t off; /* offset (doubles) of x in the blob; y (if any) at off + n */
  int32_t n;   /* points */
  int32_t pad;
} PetiteQTable;

typedef struct {
  PetiteQTable ns;       /* log10 E, log10 n*sigma [1/cm] of the dark process (log-log, 1e-20 outside) */
  PetiteQTable loss[3];  /* n*sigma (x, y) of each survival channel (linear, 0 outside) */
  PetiteQTable grid[3];  /* grid (x only) each channel's running integral is kept on */
  PetiteQTable mfp[3];   /* n*sigma (x, y) tables summed for the mean free path (get_mfp) */
  PetiteQTable nodes;    /* panel nodes (x only), ascending and distinct */
  PetiteQTable energies; /* end energies Ei (x only) */
  int32_t n_loss, n_mfp; /* channels used, each <= 3 */
  int32_t mode;          /* 0: weights; 1: drate rows */
  int32_t zero_small;    /* 1: n*sigma below 1e-18 counts as 0 (all but bound-electron DarkAnn) */
  double lo;             /* mode 0: lower limit; mode 1: floor of the bins */
  double dEdx;           /* GeV/m */
  int64_t out_off;       /* mode 0: n values W(Ei); mode 1: n x 10 x (bin centre, rate) */
} PetiteDarkQuadJob;

/* Runs the jobs on the current device (synchronous; blob and out are host
 * arrays).  PETITE_ERR_ARG for tables outside the blob or results outside
 * out, PETITE_ERR_INTERNAL if an integral needs more than 512 geometric panel
 * edges. */
int petite_dark_quadrature(const PetiteDarkQuadJob* jobs, int32_t n_jobs, const double* blob, int64_t n_blob,
                           double* out, int64_t n_out);

/* Dark tables of one (material, m_V) pair, built by DarkShower.__init__
 * (src/PETITE/dark_shower.py:85-593).  Offsets index the float64 blob given to
 * petite_dark_setup. */
typedef struct {
  int32_t w_n[PETITE_DARK_FAMILIES];   /* weight arrays: E[n] at w_off, W[n] at w_off + n (linear, 0 outside) */
  int64_t w_off[PETITE_DARK_FAMILIES];
  int32_t d_n[PETITE_DARK_FAMILIES];   /* drate tables (_d_rate_d_E_*, 311-593): saved energies E[n] at d_off, */
  int64_t d_off[PETITE_DARK_FAMILIES]; /* then bin counts cnt[n], then rows [n][10][2] (bin energy, rate) */
  int32_t comp_n;                      /* DarkComp n*sigma [1/cm]: log10 E[n] at comp_off, log10 value[n] after */
  int32_t n_active;                    /* active processes, in the caller's order (dark process codes 8..11, and
                                          * 12 = TwoBody_BSMDecay of pi0 / eta / eta' records) */
  int64_t comp_off;
  int32_t active[8];
  double eps2;                         /* g_e^2 / (4 pi alpha) = kinetic_mixing^2 */
  double E_res, E_C;                   /* bound-electron thresholds (dark_shower.py:234-235) */
  double min_brem, min_ann, min_comp, min_muon_brem; /* _minimum_calculable_dark_energy (243-245) */
  /* geometric weight grids: ln E[0] and (n-1)/ln(E[n-1]/E[0]) (0: search the grid);
   * DarkComp: log10 E[0] and (n-1)/(log10 E[n-1] - log10 E[0]) of its uniform log grid */
  double w_lnE0[PETITE_DARK_FAMILIES];
  double w_inv_dlog[PETITE_DARK_FAMILIES];
  double comp_l0, comp_inv_dl;
  /* TwoBody_BSMDecay (dark_shower.py:633-640): 2 kappa^2 BR(M -> gamma gamma)
   * of pi0, eta, eta' (0 unless code 12 is active); the weight of a meson of
   * mass m_M is meson_w (1 - (m_V/m_M)^2)^3 for m_V < m_M */
  double meson_w[3];
} PetiteDarkDesc;

/* Dark vectors written by petite_dark_reweight, one entry per (SM particle,
 * active process) with a positive weight: p (E, px, py, pz), r (x, y, z,
 * weight), parent (SM store slot) and meta (process code | 0x100 when the
 * entry holds a vector; entries whose production returned None in the
 * reference keep 0 there). */
typedef struct {
  float* p;
  float* r;
  int32_t* parent;
  int32_t* meta;
  int64_t capacity;
} PetiteDarkOut;

typedef struct {
  int64_t n_entries;    /* entries written (SM particle, process pairs with weight > 0) */
  int64_t n_vectors;    /* entries that hold a vector */
  int64_t n_draws;      /* draw_dark_sample calls */
  int64_t n_trials;     /* their candidates */
  int64_t n_no_sample;  /* draws that hit max_trials ("No Sample Found", dark_shower.py:697-700) */
  int64_t n_zero_rate;  /* drate rows that sum to zero (722-750) */
  int64_t n_pe_fail;    /* bound-electron momentum draws that found no sample in 65536 pairs */
  int64_t pad;
  double yield[5];      /* sum of vector weights per dark process (DarkBrem, DarkAnn, DarkComp, DarkMuonBrem,
                         * TwoBody_BSMDecay) */
  double ms_select, ms_draw, ms_emit;
  int64_t launches;
  double ms_meson;      /* k_meson_decay time (HIP events; with petite_set_timing on) */
} PetiteDarkStats;

/* Attach the dark tables to a context whose table descriptor carries the dark
 * grids, m_V and Zeff (blob: host float64 array of n doubles). */
int petite_dark_setup(PetiteCtx* ctx, const PetiteDarkDesc* desc, const double* blob, int64_t n);

/* Dark vectors of a shower record: replaces the loop of
 * DarkShower.generate_dark_shower (src/PETITE/dark_shower.py:806-849) with
 * GetBSMWeights (595-647), produce_bsm_particle (721-804), draw_dark_sample
 * (649-704) and draw_pe_sample (706-719) for every (SM particle, active
 * process) pair.  sm: the first n_sm records of a store filled by
 * petite_run_showers (or by the caller).  PETITE_ERR_CAPACITY when out holds
 * fewer entries than needed (stats->n_entries holds the number). */
int petite_dark_reweight(PetiteCtx* ctx, const PetiteStore* sm, int64_t n_sm, uint64_t seed, PetiteDarkOut* out,
                         PetiteDarkStats* stats, PetiteStream stream);

/* ------------------------------------------------- m_V-continuous yield --
 * No counterpart in the reference, which only evaluates the trained masses.
 * The yield of a shower record per unit eps^2 is
 *   Y(m_V) = sum_i w_i sum_P W_P(E0_i; m_V),
 * W_P = GetBSMWeights (src/PETITE/dark_shower.py:595-647) with eps^2 = 1.  For
 * m_V between trained masses m_j < m_{j+1}, with t = ln(m_V/m_j)/ln(m_{j+1}/m_j),
 * each particle's weight is W_j^(1-t) W_{j+1}^t when both are positive and
 * (1-t) W_j + t W_{j+1} otherwise (DarkAnn and DarkComp: tables at m_V
 * itself with petite_yield_bound).  Masses outside [lightest, heaviest
 * trained mass] are rejected (PETITE_ERR_ARG): there is nothing to
 * interpolate. */
#define PETITE_MAX_MASSES 16

/* Dark tables of n_mass trained masses (ascending), one descriptor each; all
 * offsets index the one float64 blob (host, n doubles).  descs[j].eps2 should
 * be 1.  The context must hold the SM n*sigma tables of the same material. */
int petite_yield_setup(PetiteCtx* ctx, int32_t n_mass, const double* masses, const PetiteDarkDesc* descs,
                       const double* blob, int64_t n);

/* Bound-electron processes of the yield.  With these inputs set (after
 * petite_yield_setup, which clears them), the DarkAnn and DarkComp weights at
 * each requested m_V come from tables built at that m_V on the device, as
 * DarkShower builds them for its own mass (src/PETITE/dark_shower.py:254-292
 * with the closed forms of src/PETITE/atomic_annihilation.py:70-104 and
 * src/PETITE/atomic_compton.py:94-109, the DarkAnn track-length weights as
 * petite_dark_quadrature computes them), instead of interpolation between
 * trained masses: their weights are resonances and thresholds that move
 * with m_V.  Their dY/dm_V is the central difference of tables built at
 * m_V (1 +- rel_step).  Offsets index the float64 blob (host, n doubles). */
typedef struct {
  double Zeff;                         /* effective charge of the bound-electron forms */
  double ne_conv;                      /* n_e * GeVsqcm2: n*sigma [1/cm] per sigma [1/GeV^2] */
  double m_e, alpha;                   /* m_electron [GeV] and alpha_em (Compton form, thresholds) */
  double ann_floor, comp_floor;        /* minimum calculable e+ / photon energies: lower grid ends */
  double rel_step;                     /* relative step of the central difference (0 < h < 0.01) */
  double dEdx;                         /* GeV/m */
  double top_ann[PETITE_MAX_MASSES];   /* per trained mass: top energy of its grids, used for m_V at or */
  double top_comp[PETITE_MAX_MASSES];  /* above it (DarkBrem and free DarkComp tables' last energies) */
  PetiteQTable loss[3];                /* positron survival channels Brem, Ann, Bhabha: n*sigma (x, y) */
  PetiteQTable mfp[3];                 /* mean-free-path tables (x, y): Brem, Bhabha, Ann */
  PetiteQTable nodes;                  /* panel nodes of the survival grids (x only; ascending, distinct) */
} PetiteBoundDesc;

int petite_yield_bound(PetiteCtx* ctx, const PetiteBoundDesc* desc, const double* blob, int64_t n);

/* Y and dY/dm_V at n_points masses (host arrays), accumulated in float64 over
 * the first n_sm records of the store.  dyield_dmV_host may be NULL: the
 * yield alone, without the derivative work. */
int petite_dark_yield(PetiteCtx* ctx, const PetiteStore* sm, int64_t n_sm, int32_t n_points, const double* mV_host,
                      double* yield_host, double* dyield_dmV_host, PetiteStream stream);

/* VEGAS map training (the reference trains its maps with the vegas package:
 * utilities/generate_integrators.py:49-100, options at all_processes.py:
 * 1085-1096).  petite_map_accumulate adds, for every grid energy of process
 * proc, n_per_energy samples uniform in y mapped through the context's
 * current grid: sum_f[ie*stride + node_offset(d) + i] += (J f)^2 and
 * n_f[...] += 1 for the increment i of each dimension d the sample falls in
 * (device arrays of n_E * stride, zeroed by the caller; non-finite J f are
 * skipped).  Sample k of iteration it uses Philox block it*n_per_energy + k
 * of stream (proc << 32 | ie), purpose PURPOSE_MAP_TRAIN = 9.
 * petite_set_grid replaces the context's nodes of proc (host array of
 * n_E * stride floats, the layout of PetiteGridDesc); rescan max_F after. */
int petite_map_accumulate(PetiteCtx* ctx, int32_t proc, int64_t n_per_energy, uint64_t seed, uint32_t iteration,
                          double* sum_f, uint32_t* n_f, PetiteStream stream);
int petite_set_grid(PetiteCtx* ctx, int32_t proc, const float* nodes_host, int64_t n);

/* Detector acceptance of straight-line extrapolations: replaces
 * transverse_position and detector_cut (src/PETITE/shower.py:815-865).
 * p, r: n records of float4 (E, px, py, pz) and (x, y, z, weight) on the
 * device -- the production momentum and point (a PetiteStore's p0/r0, or
 * PetiteDarkOut's p/r).  For detector plane d at z = det_z[d] (host array,
 * 1 <= n_det <= PETITE_MAX_DETECTORS):
 *   T = (z_d - z0) / pz,  (x, y) = (x0 + T px, y0 + T py),  r_T = |(x, y)|,
 * and the record passes when inner_radius < r_T < radius and, if has_ecut,
 * emin < E < emax.  Arithmetic in float64 without contraction, like numpy.
 * mask (device, n uint64, may be NULL): bit d = passes detector d.
 * wsum (device, n_det + 1 float64, may be NULL): wsum[d] = sum of the weights
 * passing detector d, wsum[n_det] = sum of the weights inside the energy
 * window (the denominator of the reference's "Efficiency"); the call zeroes
 * it first.  No context: the kernel needs no tables. */
#define PETITE_MAX_DETECTORS 64
int petite_detector_cut(const float* p, const float* r, int64_t n, const double* det_z_host, int32_t n_det,
                        double radius, double inner_radius, int32_t has_ecut, double emin, double emax,
                        uint64_t* mask, double* wsum, PetiteStream stream);

#ifdef __cplusplus
}
#endif

#endif /* PETITE_HIP_H */
