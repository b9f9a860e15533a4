/*
 * rbn.h — C-ABI of librbn, the MI355X (gfx950) Rayleigh–Bénard lattice-Boltzmann step
 * with Lagrangian particles and trajectory nudging.
 *
 * The reference (Lokahith-Agasthya/LagrangianRBNudging) has no library interface: its
 * "interface" is the call sequence of two C mains over process-global state, configured by
 * `#define` blocks.  Every entry point below replaces one step of that sequence; the
 * comment on each names the reference code it stands in for (paths relative to the
 * reference repository root).
 *
 * Conventions
 *  - Every function returns 0 (RBN_OK) or a negative rbn_status.  Nothing calls exit().
 *    The message for the last failure on a context is rbn_last_error(ctx); failures that
 *    happen before a context exists are reported by rbn_last_error(NULL).
 *  - Host arrays are caller-owned and only read/written during the call.
 *  - Field arrays on the host use the reference's halo-free dump layout [nx][ny]
 *    (y fastest, element (i,j) at i*ny + j for interior node (i+1, j+1)), as written by
 *    dumpf/snapshottemp (Trajectories/code/temppart_dump.c:52-141, 1183-1262).
 *    Population arrays are [nx][ny][9] in the reference's D2Q9 order (dellar,
 *    Trajectories/code/temppart_dump.c:1363-1401).
 *  - A context owns one GPU stream and all device memory of one lattice.  Calls on one
 *    context must be serialised by the caller; distinct contexts are independent.
 *  - Device state is FP64 or FP32 (shifted populations) per rbn_params.precision.  Host
 *    outputs are always double.
 */
#ifndef RBN_H
#define RBN_H

#ifdef __cplusplus
extern "C" {
#endif

#define RBN_API_VERSION 1

typedef struct rbn_ctx rbn_ctx;

typedef enum rbn_status {
  RBN_OK = 0,
  RBN_EINVAL = -1,    /* bad argument or parameter combination               */
  RBN_EIO = -2,       /* file could not be opened/read/written                */
  RBN_EPARSE = -3,    /* parameter file could not be parsed                   */
  RBN_EDEVICE = -4,   /* no usable GPU, or a HIP call failed                  */
  RBN_ESTATE = -5,    /* call out of order (e.g. step before init)            */
  RBN_ERANGE = -6,    /* size beyond a documented limit, particle out of grid */
  RBN_ENOMEM = -7,    /* host or device allocation failed                     */
  RBN_ENUMERIC = -8,  /* non-finite value detected                            */
  RBN_ENOTFOUND = -9, /* observation id missing from a record                 */
  RBN_ECALLBACK = -10 /* a user callback returned non-zero                    */
} rbn_status;

/* Which of the reference's two programs the step restates. */
enum {
  RBN_VARIANT_TRAJECTORIES = 0, /* truth run: Trajectories/code/temppart_dump.c:1674-1759   */
  RBN_VARIANT_NUDGING = 1       /* nudged run: Nudging/code/tempvel_nudge.c:2346-2411         */
};
enum { RBN_FP64 = 0, RBN_FP32 = 1 };
enum { RBN_DEPOSIT_BOX7 = 0, RBN_DEPOSIT_BILINEAR = 1 };

/*
 * Parameters.  The first group carries the reference's #define names
 * (Trajectories/code/temppart_dump.c:15-31, Trajectories/code/mympi2.h:14-21,
 *  Nudging/code/tempvel_nudge.c:13-31) with the same meaning; rbn_params_parse reads them
 * from "#define NAME value" lines.  The second group holds options the reference fixes
 * in code; each defaults to the reference's behaviour.
 */
typedef struct rbn_params {
  /* --- reference #define names --- */
  int nx, ny;              /* global lattice (interior nodes)                          */
  int scratch;             /* 1: initialise from scratch                               */
  int contin;              /* nudged run: continue from dumps (only 0 is supported)    */
  int nstart, nsteps;      /* first step is nstart+1                                   */
  int nout;                /* diagnostic cadence (evolution.dat)                       */
  int noutconfig;          /* trajectory-record cadence                                */
  int ntempread;           /* nudged run: observation-record spacing                   */
  double rhom1;            /* initial density                                          */
  double relax, relaxg;    /* BGK relaxation times of f and g                          */
  double Tup, Tdown;       /* top / bottom plate temperatures                          */
  double gravity;          /* buoyancy: F_y = -gravity * rho * T                       */
  double alpha;            /* nudging strength per step                                */
  int particle;            /* 1: advance particles                                     */
  int particle_scratch;    /* 1: particles initialised from scratch                    */
  int nprtcl;              /* number of particles / observers                          */
  int tracer, heavyprtcl;  /* default particle kind when no per-particle tau is given  */
  double taulag, beta;     /* Maxey–Riley relaxation time and added-mass coefficient   */
  int prtcl_writefull;     /* 1: records also carry dtu..dyv, fu, fv at (xo,yo)        */
  /* --- options (reference behaviour by default) --- */
  int variant;             /* RBN_VARIANT_*                                            */
  int precision;           /* RBN_FP64 / RBN_FP32                                      */
  int deposit;             /* RBN_DEPOSIT_BOX7 (reference) or _BILINEAR                */
  int fix_q3;              /* 0: keep the reference's `xi > nx-2` wrap test (misses
                              xi = nx-2, Nudging/code/tempvel_nudge.c:1601); 1: fixed  */
  int ndiag;               /* energy cadence; 0 = reference (nout / 5000)              */
  double T0_shift;         /* FP32: g stored as g_k - w_k*T0_shift                     */
  double rho0_shift;       /* FP32: f stored as f_k - w_k*rho0_shift                   */
  double heavy_init_factor;/* inertial initial velocity = factor * fluid (0.98 in ref)  */
} rbn_params;

/* Particle state in id order (the reference's prtcl_type, Trajectories/code/mympi2.h:2-7). */
typedef struct rbn_particles {
  int n;
  int *id;
  double *x, *y, *u, *v, *xo, *yo, *uo, *vo;
} rbn_particles;

/*
 * One trajectory record (prtcl_write, Trajectories/code/temppart_dump.c:1467-1571, and the
 * Writefull block, Trajectories/code/mympi2.h:441-556): positions after the step's update
 * and before the periodic/reflecting wrap, in ascending id order.  The Writefull arrays
 * are NULL unless prtcl_writefull is set.  Pointers are valid during the callback only.
 */
typedef struct rbn_record {
  int step, n;
  const int *id;
  const double *x, *y, *u, *v, *xo, *yo, *uo, *vo, *temp;
  const double *dtu, *dtv, *dxu, *dxv, *dyu, *dyv, *fu, *fv;
} rbn_record;

typedef int (*rbn_record_fn)(void *user, const rbn_record *rec);
/* Fill id/x/y/T (length n, record storage order) with trajectory record `record_step`. */
typedef int (*rbn_obs_fn)(void *user, int record_step, int n, int *id, double *x, double *y,
                          double *T);
/* Called every diagnostic step with the evolution.dat quantities. */
typedef int (*rbn_diag_fn)(void *user, int step, double mean_energy, double mean_T);

/* ---------------------------------------------------------------- library */
const char *rbn_version(void);
int rbn_api_version(void);
/* Message of the last failure on ctx (or of the calling thread when ctx is NULL). */
const char *rbn_last_error(const rbn_ctx *ctx);
/* Number of visible GPUs (0 on a host without one). */
int rbn_device_count(int *n);

/* ---------------------------------------------------------------- parameters */
/* Reference defaults: the #define blocks of Trajectories/code/temppart_dump.c:15-31 and
 * Trajectories/code/mympi2.h:14-21. */
void rbn_params_default(rbn_params *p);
/* Read "#define NAME value" lines (C expression values) from a file, updating the fields
 * whose names appear; may be called on several files in turn.  A file that defines
 * `ntempread` or `alpha` selects the nudged variant unless rbn_variant is also given. */
int rbn_params_parse(const char *path, rbn_params *p);
/* Same as rbn_params_parse on an in-memory text. */
int rbn_params_parse_text(const char *text, rbn_params *p);
/* Set one field by its #define (or option) name. */
int rbn_params_set(rbn_params *p, const char *name, double value);
/* Rayleigh number of a parameter set (media(), Nudging/code/tempvel_nudge.c:1262-1283). */
double rbn_params_rayleigh(const rbn_params *p);
/* Inverse of media(): choose relax, relaxg and gravity for (Ra, Pr) with free-fall
 * velocity u_ff = sqrt(|gravity| dT L), L = ny-1, keeping nx, ny, Tup, Tdown, rhom1. */
int rbn_params_from_ra(rbn_params *p, double Ra, double Pr, double u_ff);
/* Validate a parameter set (sizes, cadences, variant constraints); 0 when usable. */
int rbn_params_check(const rbn_params *p);

/* ---------------------------------------------------------------- lifetime */
/* Allocate the lattice on GPU `device`; replaces the malloc block of main()
 * (Trajectories/code/temppart_dump.c:1598-1640, Nudging/code/tempvel_nudge.c:2271-2306)
 * and dellar() (:1363-1401). */
int rbn_create(const rbn_params *p, int device, rbn_ctx **out);
void rbn_destroy(rbn_ctx *ctx);
int rbn_get_params(const rbn_ctx *ctx, rbn_params *out);

/* ---------------------------------------------------------------- initial state */
/* Scratch initialisation: inithydro + equilif/initpop + equilig/initpop
 * (Trajectories/code/temppart_dump.c:680-742, 1643-1648).  NULL rho -> rhom1, NULL u/v ->
 * 0, NULL T -> the reference's tanh front with srand(0)/rand() jitter (rank-0 stream). */
int rbn_init_fields(rbn_ctx *ctx, const double *rho, const double *T, const double *u,
                    const double *v);
/* Linear conduction profile Tdown..Tup (the reference's commented-out alternative,
 * Trajectories/code/temppart_dump.c:697) plus uniform noise in [-amp, amp] from a
 * splitmix64 stream seeded by `seed`; u = v = 0, rho = rhom1. */
int rbn_init_conduction(rbn_ctx *ctx, double noise_amp, unsigned long long seed);
/* Restart: post-collision populations plus the previous step's moments (restoref/g and
 * inithydro's restart branch, Trajectories/code/temppart_dump.c:348-668, 701-727).
 * rho_prev/T_prev feed the lagged force, u_prev/v_prev the inertial d/dt. */
int rbn_set_populations(rbn_ctx *ctx, const double *f, const double *g,
                        const double *rho_prev, const double *T_prev,
                        const double *u_prev, const double *v_prev);
/* Post-collision populations (dumpf/dumpg layout [nx][ny][9]). */
int rbn_get_populations(rbn_ctx *ctx, double *f, double *g);
/* Moments of the latest hydrovar (rhopre, temppre, upre, vpre); any pointer may be NULL. */
int rbn_get_fields(rbn_ctx *ctx, double *rho, double *T, double *u, double *v);
/* evolution.dat quantities <rho|u|^2/2>, <T> of the latest hydrovar
 * (Trajectories/code/temppart_dump.c:1739-1757). */
int rbn_energy(rbn_ctx *ctx, double *mean_energy, double *mean_T);
/* Row profiles of the latest step (profile, Nudging/code/tempvel_nudge.c:1104-1250):
 * out[(j-1)*9 + q] for rows j = 1..ny, q = x averages of rho, u, v, u^2, v^2, T, vT,
 * dzT = T(j+1) - T(j) (row ny reads row 1, the periodic ghost row), vorticity^2 (computed
 * but not printed by the reference; periodic x, one-sided y differences at the walls).
 * Sums in a fixed order on the device. */
int rbn_profiles(rbn_ctx *ctx, double *out);
/* Bilinear interpolation (bilinear, Trajectories/code/mympi2.h:292-311, periodic ghosts of
 * communication, :313-348) of one field of the latest hydrovar at n positions, e.g. the
 * initial particle temperatures init_prtcl writes (:196-201).  field: 0 rho, 1 T, 2 u, 3 v.
 * Positions must lie in (-1, nx+1) x (-1, ny+1) (RBN_ERANGE otherwise). */
int rbn_sample(rbn_ctx *ctx, int field, int n, const double *x, const double *y, double *out);
/* rbn_sample without any device allocation: the positions and samples are staged in a fixed
 * device array of RBN_SAMPLE_NOALLOC_MAX positions (6 MiB, held by every GPU the library runs
 * on until the process ends) instead of a buffer allocated per call; n above the maximum is
 * RBN_ERANGE.  No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops).  All contexts of one GPU share
 * the array, so calls that overlap on one GPU race; sequential calls do not (each returns
 * after its samples are back).  Results equal rbn_sample's bit for bit. */
#define RBN_SAMPLE_NOALLOC_MAX (1 << 18)
int rbn_sample_noalloc(rbn_ctx *ctx, int field, int n, const double *x, const double *y, double *out);

/* ---------------------------------------------------------------- particles (truth run) */
/* init_prtcl (Trajectories/code/mympi2.h:116-290) with caller-chosen positions.
 * tau_p[i] > 0 makes particle i inertial (Maxey–Riley, :373-433) with 1/tau_p[i] in place
 * of onebytaulag; tau_p[i] <= 0 makes it a tracer (AB2, :358-372).  tau_p NULL -> the
 * params' tracer/heavyprtcl/taulag; beta_p NULL -> params.beta.  up/vp NULL -> initial
 * velocity from the fluid (tracer: I[u]; inertial: heavy_init_factor*I[u], :196-207). */
int rbn_init_particles(rbn_ctx *ctx, int n, const int *id, const double *x, const double *y,
                       const double *tau_p, const double *beta_p, const double *up,
                       const double *vp);
/* Current particle state in id order; caller provides arrays of length n. */
int rbn_get_particles(rbn_ctx *ctx, rbn_particles *out);
int rbn_num_particles(const rbn_ctx *ctx, int *n);
/* Called on every record step (istep % noutconfig == 0) — prtcl_write's replacement. */
int rbn_set_record_callback(rbn_ctx *ctx, rbn_record_fn fn, void *user);

/* ---------------------------------------------------------------- nudging (nudged run) */
/* set_nudge_vol's particle part (Nudging/code/tempvel_nudge.c:2083-2122): the observers'
 * ids and their state at nstart (the truth run's init_prtcl{x,y,T,_id}0.in). */
int rbn_init_observers(rbn_ctx *ctx, int n, const int *id, const double *x, const double *y,
                       const double *T);
/* Source of trajectory records for read_prtcl (Nudging/code/tempvel_nudge.c:1380-1476):
 * asked for record istep+ntempread-1 on steps with istep % ntempread == 1. */
int rbn_set_obs_provider(rbn_ctx *ctx, rbn_obs_fn fn, void *user);
/* The same records from caller-owned host arrays, served without a callback (the
 * reference reads them from Prtcl_Veldata1/Particles%d.h5): record r of nrec is for step
 * steps[r] (ascending) and holds n entries id[r*n + k], x[r*n + k], y[r*n + k], T[r*n + k]
 * in record storage order.  The arrays must stay valid while the context steps; replaces
 * the provider.  With the tuning "obs_device" (default 1), rbn_step copies the table to the
 * device at its first step after this call or rbn_init_observers, with the observer map of
 * every record, and reads the records there, when the records hold one entry per observer,
 * the observers' ids are distinct and span at most 4n + 1024 values, and the table takes at
 * most 8 GiB (x, y, T and the map: 28 B per entry); later changes to the arrays are seen
 * after rbn_set_obs_table is called again.  A missing record or observer fails the step
 * that reads it either way. */
int rbn_set_obs_table(rbn_ctx *ctx, int nrec, int n, const int *steps, const int *id, const double *x,
                      const double *y, const double *T);
/* Nudging state of the latest step: nudge_vol [nx][ny] counts, tempreadint [nx][ny]
 * running means, nud_term [nx][ny]; any pointer may be NULL. */
int rbn_get_nudging(rbn_ctx *ctx, int *nudge_vol, double *tempreadint, double *nud_term);
/* Per-node deposit weight W [nx][ny] of the latest step's bilinear deposit
 * (RBN_DEPOSIT_BILINEAR: each observer gives the four corners of its cell its bilinear
 * interpolation weights, Tbar = sum w T / W, and the nudged collision uses
 * Q = alpha * (min(W, 1) * (T - Tbar))); zero for the reference 7x7 deposit.  No reference
 * counterpart: the bilinear deposit is an option beside the reference's running mean. */
int rbn_get_nudging_weights(rbn_ctx *ctx, double *weight);

/* ---------------------------------------------------------------- stepping */
/* Diagnostics callback (evolution.dat line), every ndiag steps. */
int rbn_set_diag_callback(rbn_ctx *ctx, rbn_diag_fn fn, void *user);
/* Advance nsteps time steps in the reference's order (one iteration of the main loops,
 * Trajectories/code/temppart_dump.c:1674-1759 or Nudging/code/tempvel_nudge.c:2346-2411).
 * Returns RBN_ENUMERIC when a step of the call met a node whose rho or T is not finite (the
 * reference stops on a non-finite state, Trajectories/code/temppart_dump.c:414-425); the
 * state is then that of the last step, and the call's records are not delivered.  Every call
 * with nsteps > 0 waits for its steps to finish to read the check. */
int rbn_step(rbn_ctx *ctx, int nsteps);
/* Index of the last completed step (istep). */
int rbn_get_step(const rbn_ctx *ctx, int *istep);
/* Wait for all queued GPU work of ctx. */
int rbn_synchronize(rbn_ctx *ctx);

/* ---------------------------------------------------------------- batched ensemble */
/* Many independent contexts of uneven sizes and parameters advanced together: one ensemble
 * step is a persistent launch over a work list of every member's particle chunks (the update
 * of the previous step) and lattice tiles, and cycles of launches are captured in one
 * hipGraph that is replayed.  Nudged members add, per step, one launch per observer/deposit
 * pass over all of them and the upload of the records that start an interval (fetched from
 * their providers before the graph replay that uploads them), in cycles of lcm(6, record
 * intervals) steps with one launch per step, lcm(12, record intervals) with two.  With one
 * launch per step, eager steps build the next step's deposit (a second set of deposit arrays
 * per member, owned by the ensemble) on a second stream beside the current lattice launch;
 * with two, both steps' deposits precede their launch.  Graph replays run the passes in
 * order on one stream.  No reference counterpart (the
 * reference runs one lattice per job).  Members must share device and precision, be
 * initialised, and stay owned by the caller; results equal stepping each member with
 * rbn_step, except that no trajectory records or diagnostics are produced during ensemble
 * steps.  A member meeting non-finite moments makes rbn_ensemble_step return RBN_ENUMERIC. */
typedef struct rbn_ensemble rbn_ensemble;
int rbn_ensemble_create(rbn_ctx *const *members, int n, rbn_ensemble **out);
void rbn_ensemble_destroy(rbn_ensemble *e);
const char *rbn_ensemble_last_error(const rbn_ensemble *e);
/* Advance every member by nsteps (graph replays for whole cycles of six launches; the first
 * and last launches of a call, and the last launch's particle updates, run eagerly; with
 * nudged members the first 31 steps with one launch per step, 32 with two, run eagerly while
 * the first replay's records are fetched and uploaded).  After
 * an error (a provider failure, a particle leaving its lattice, a device error) the
 * members' states are unspecified: initialise them again before further steps. */
int rbn_ensemble_step(rbn_ensemble *e, int nsteps);
/* Tile shape (tile_w in {64,128,256} nodes, tile_w*tile_h a multiple of 256), persistent
 * workgroups per CU (0: one workgroup per work item) and work order (bit 0 clear: member by
 * member, set: members interleaved, largest first; bit 1: odd steps visit the lattice work
 * items last first; bit 2: a lattice launch and a particle launch per step, the particle
 * launch on a second stream beside the next lattice launch, instead of one launch; bit 3
 * without bit 2: two steps per launch on 64-node-wide tiles of 8 rows in FP32, 4 in FP64,
 * whatever tile_w and tile_h say, except for a single step at the end of a call; bit 4: no
 * graph capture, every launch eager, e.g. for tracing tools). */
int rbn_ensemble_set_tuning(rbn_ensemble *e, int tile_w, int tile_h, int wg_per_cu, int order);
int rbn_ensemble_get_tuning(const rbn_ensemble *e, int *tile_w, int *tile_h, int *wg_per_cu,
                            int *order);
/* Time every candidate launch shape on trial_steps real steps, keep the fastest, and restore
 * every member's state afterwards, also when a step fails (the error is returned and every
 * member is back at its starting state); *best_ms_per_step receives the winner's time.  With
 * nudged members every candidate steps from the starting state (a warm-up call of G steps
 * and a timed call of G * max(1, ceil(trial_steps / G)) steps; G = 6 for one launch per
 * step, 12 for two), and the fastest shape is timed once more in both forms, graph replays
 * and eager steps (order bit 4), keeping the faster: over steps 31 .. 2L + 30 of a call of
 * 2 L + 32 steps with one launch per step (L = lcm(6, record intervals)), steps 32 .. 2L + 31
 * of a call of 2 L + 33 steps with two (L = lcm(12, record intervals)), L <= 120.  The members' records
 * must cover the longest of these spans from the starting state. */
int rbn_ensemble_autotune(rbn_ensemble *e, int trial_steps, double *best_ms_per_step);

/* ---------------------------------------------------------------- adjoint window */
/* Gradient of the trajectory-temperature misfit of a nudged run.  No reference counterpart
 * (the reference has no gradient).  The window starts from the state rbn_init_fields set
 * (no step since) and runs nsteps nudged steps exactly as rbn_step does; the misfit is
 *   J = sum over steps s with s % ntempread == 0, over observers k, of (T_s(x_k,y_k) - T_k)^2
 * with (x_k, y_k, T_k) the record observer k snaps to at step s and T_s(x, y) the bilinear
 * interpolation of the step's temperature.  The state entering every ckpt_every-th step is
 * kept on the device (20 lattice planes per checkpoint) with every fetched record.
 * rbn_adj_forward leaves the context at the end of the window and stores J. */
int rbn_adj_forward(rbn_ctx *ctx, int nsteps, int ckpt_every, double *J);
/* Reverse sweep of the last rbn_adj_forward: recomputes each checkpoint segment (ckpt_every
 * steps of populations held on the device) and runs the adjoint step over it backwards.
 * dJ_dT0 (nx*ny, [i][j] as the T argument of rbn_init_fields) and dJ_dalpha may be NULL.
 * The context's lattice and observer state are not modified. */
int rbn_adj_backward(rbn_ctx *ctx, double *dJ_dT0, double *dJ_dalpha);

/* ---------------------------------------------------------------- launch tuning */
/* Launch shape of the fused lattice step; results do not depend on it.  No reference
 * counterpart (the reference has no launch configuration).  Keys:
 *   "lbm_block"  nodes per workgroup along x: 64, 128, 256 (default) or 512
 *   "lbm_npt"    nodes per thread along x: 1 (default), 2, 4 or 8 (row segments of
 *                lbm_block * lbm_npt nodes per workgroup)
 *   "lbm_nt"     1: non-temporal stores of the step's outputs (default 1, 0 for lattices whose
 *                working set fits in 160 MB: about 52 reals per node)
 *   "lbm_ntload" 1: non-temporal loads of the step's inputs (default 0)
 *   "lbm_uv_keep" 1: ordinary (cached) stores of the velocity under lbm_nt, for the particle
 *                update that gathers from it next (default 1)
 *   "lbm_xcd"    1: give each of the 8 XCDs a contiguous band of rows; 2: give each XCD one
 *                row in eight, its segments visited in order (default 0, 1 for lattices whose
 *                working set fits in 160 MB)
 *   "lbm_reverse" 1: sweep every odd step last row first, so that it starts on the rows the
 *                previous step wrote last (default 1)
 *   "lbm_steps"  2: advance two steps per lattice launch, the first step's populations kept
 *                in LDS (temporal blocking), for every step that writes no moments and is
 *                followed by another step of the same rbn_step call; truth runs with
 *                particles only under particle_fuse; nudged runs without dep_overlap, the
 *                observer updates and deposits of both steps run before the launch (the
 *                deposits depend on the records only) (default 2); 1: one step per launch;
 *                3 (truth runs): three steps per launch by row sweeps wherever at most one
 *                of the three writes moments and neither of the first two is a diagnostics
 *                step, so that a record step may be any of them (the first two steps'
 *                sideways-moving populations passed through LDS, strips of 256 threads, 128
 *                in FP64, overlapping by four columns), else two or one as with 2 (default 3
 *                for FP32 truth runs of at least 2^21 nodes)
 *   "lbm3_seg"   rows per segment of a three-step sweep, 4..4096; 0 (default): the fewest
 *                rows that put every workgroup of the launch on the device at once, at least 16
 *   "lbm3_nt"    1: non-temporal stores of a three-step sweep's populations and force;
 *                0 (default): cached stores, whose completion the sweep's next row waits
 *                for sooner (4096 x 2048 FP32: 0.443 against 0.467 ms per launch)
 *   "lbm2_seg"  (truth runs) 8..64: two steps per launch by row sweeps, every workgroup
 *                sweeping a strip of 256 columns over a segment of lbm2_seg rows, the first step's
 *                sideways-moving populations passed between its threads through LDS; 0: by
 *                tiles of 64 columns and lbm2_tile_h rows (default 0)
 *   "lbm2_xcd"   1: the row-sweep workgroups one XCD receives take consecutive strips and
 *                segments (default 1)
 *   "lbm2_variant" 1: the row sweep issues the loads of a row one row ahead (default 0)
 *   "lbm2_tile_h" rows of a two-step tile of 64 columns (lbm2_seg 0): 4, 8 or 16 (FP32
 *                only), 0 for the default (8 in FP32, 4 in FP64)
 *   "dep_overlap" nudged run, 1: the observer update and deposit of step s+1 run on a second
 *                stream beside the lattice update of step s (a second set of deposit arrays,
 *                allocated on first use); 0: in turn on one stream (default 0: with one
 *                lattice the hand-over between the streams costs more than it hides, e.g.
 *                1024x512 FP32 with 1e4 observers 54 against 50 us per step; ensembles,
 *                whose lattice launch is much longer than their deposits, gain from it)
 *   "dep_in_tile" nudged run with the bilinear deposit, two-step launches of 64 x 8 tiles: 1:
 *                the second step's deposit is summed inside the launch, each tile from its
 *                own bins, instead of by a separate pass (not in a call's last launch, whose
 *                deposit the getters read) (default 1)
 *   "particle_overlap"  1: run the particle update of step s on its own stream concurrently
 *                with the lattice update of step s+1 (default 0)
 *   "particle_fuse" 1: run the particle update of step s as extra workgroups of the lattice
 *                launch of step s+1 (the last step of an rbn_step call, and a record step
 *                followed by a step that rewrites the moments, launch it on its own); ignored
 *                under particle_overlap (default 1)
 *   "adj_store_deposit" 1: rbn_adj_forward keeps every step's deposit (counts, means, weights:
 *                12-20 B per node and step) when that fits in a quarter of the free device
 *                memory, so that rbn_adj_backward does not replay the observers (default 1);
 *                0: always replay
 *   "obs_device" nudged run over an observation table (rbn_set_obs_table): 1: the table is
 *                held on the device and a record becomes current without a host copy or an
 *                upload (1024 x 512 FP32, 1e4 observers, records every 5 steps: 0.0193-0.0197
 *                against 0.0242-0.0272 ms per step); 0: each record fetched, mapped and
 *                uploaded on the step that reads it (default 1) */
int rbn_set_tuning(rbn_ctx *ctx, const char *key, int value);
int rbn_get_tuning(const rbn_ctx *ctx, const char *key, int *value);

/* ---------------------------------------------------------------- measurement */
/* Per-launch HIP-event timing on ctx's stream: every = 0 off, 1 every launch, n > 1 every
 * n-th launch of each kernel class (the events between two launches widen the gap between
 * them by a few microseconds; sampling keeps that out of a timed run). */
int rbn_set_kernel_timing(rbn_ctx *ctx, int every);
/* Launch count and summed duration (ms) of the timed launches of kernel `name`
 * ("lbm", "particles", "deposit": a nudged step's observer update and deposit, "adjoint",
 * "lbm_fused": a one-step lattice launch carrying earlier steps' particle updates, "lbm2": a
 * two-step lattice launch, "lbm3": a three-step lattice launch); resets the
 * counters when reset != 0. */
int rbn_kernel_stats(rbn_ctx *ctx, const char *name, int reset, long long *launches,
                     double *total_ms);
/* Time `nsteps` steps with HIP events on ctx's stream (inputs resident on the GPU). */
int rbn_time_steps(rbn_ctx *ctx, int nsteps, double *elapsed_ms);

#ifdef __cplusplus
}
#endif
#endif /* RBN_H */
