// rbn_internal.h — shared declarations of librbn's host code and HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/rbn.h"

namespace rbn {

const char *thread_error();
int thread_fail(int code, const char *msg);

// D2Q9 lattice of the reference (dellar, Trajectories/code/temppart_dump.c:1363-1401).
// The reverse directions are negations of the forward ones, hence the signed zeros.
constexpr double kW[9] = {16. / 36., 4. / 36., 4. / 36., 4. / 36., 4. / 36.,
                          1. / 36.,  1. / 36., 1. / 36., 1. / 36.};
constexpr double kCx[9] = {0., 1., 0., -1., -0., 1., -1., -1., 1.};
constexpr double kCy[9] = {0., 0., 1., -0., -1., 1., 1., -1., -1.};

// Device lattice geometry.  Populations are structure-of-arrays planes
// [k][j][i] (x fastest, rows of `pitch` elements), interior nodes only: the periodic x
// direction is an index wrap and the walls are handled in-kernel, so no halo is stored.
struct Geom {
  int nx, ny, pitch;
  size_t plane;  // pitch * ny
};

// Per-step constants of the fused step, precomputed on the host with the reference's
// expressions so that the FP64 kernel reproduces them bit for bit.
template <typename R>
struct StepConst {
  R om, one_m_om;     // 1./relax, 1.-om
  R omg, one_m_omg;   // 1./relaxg, 1.-omg
  R src[9];           // (1.0-0.5*om)*w_k        (collisf, T:1320)
  R bw[9];            // (1.0-0.5*omg)*w_k       (nudged collisg, N:1931-1933)
  R g0;               // -gravity                (forcingconstructWW, T:752)
  R wtop[3], wbot[3]; // 2.0*w_k*Tup, 2.0*w_k*Tdown for k = 2,5,6 / 4,7,8 (T:896-921)
  R alpha;
  R rho0, T0;         // FP32 shifts (zero in FP64)
};

struct LbmPtrs {
  const void *fA, *gA;
  void *fB, *gB;
  void *F;           // lagged F_y: read F(s), write F(s+1) (one value per node)
  void *uv;          // hydrovar velocity of this step, (u, v) interleaved per node
  void *rho, *T;     // written when non-null (record/diagnostic/last steps)
  const int *nvol;   // nudged run: nudge_vol counts (nullptr: not nudged, unless tw is set)
  const double *Tbar;// nudged run: tempreadint
  void *nud;         // nudged run: persistent nud_term
  const double *wq;  // nudged run, bilinear deposit: per-node weight W (Q = alpha*(W*(T-Tbar)))
  const void *tw;    // nudged run: the deposit packed per node (DepPtrs::tw); read instead of
                     // nvol, Tbar and wq when set
  int *bad;          // set to 1 when a node's rho or T is not finite (or null: no check)
};

// Strength of the bilinear deposit's nudging at a node whose observer weights sum to W:
// min(W, 1), so that alpha <= 1 never over-relaxes a node that several observers reach.
__host__ __device__ inline double nudge_weight(double W) { return W < 1.0 ? W : 1.0; }

// Two reals of the lattice's precision.
template <typename R>
using Real2 = typename std::conditional<sizeof(R) == 4, float2, double2>::type;

// The nudging input of node o: from the packed deposit p.tw when set, else from the counts,
// means and (bilinear deposit) weights; false where the node has no contributions.  Tb and w
// are the mean and weight rounded to R, w = 1 for the 7x7 deposit.
template <typename R>
__device__ __forceinline__ bool nudge_input(const LbmPtrs &p, size_t o, R &Tb, R &w) {
  if (p.tw) {
    const Real2<R> v = ((const Real2<R> *)p.tw)[o];
    Tb = v.x;
    w = -v.y;
    return __builtin_signbit(v.y);
  }
  if (p.nvol[o] <= 0) return false;
  Tb = (R)p.Tbar[o];
  w = p.wq ? (R)nudge_weight(p.wq[o]) : R(1);
  return true;
}

// Particle kernel arguments (device pointers).
struct PartPtrs {
  int n;
  double *x, *y, *u, *v, *xo, *yo, *uo, *vo;
  const double *inv_tau, *beta;  // inv_tau <= 0: tracer
  double *rec;                   // 17 * n record arrays, or nullptr
  int *err;
};

// Observer (nudging) arrays on the device, observer order.
struct ObsPtrs {
  int n;
  const int *chk;
  int *bcy, *xidx, *yidx;
  double *dx, *dy, *dT, *xint, *yint, *Tint;
  const double *rx, *ry, *rT;  // current record, record order
};

// Deposit work arrays.
struct DepPtrs {
  int *cnt;    // per interior node: contributions (== nudge_vol), [j][i] with pitch
  int *fill;   // per node fill cursor
  int *off;    // per node exclusive prefix of cnt (+1 total at the end)
  int *list;   // contributions, grouped per node
  int *blocksum;
  double *Tbar;  // at nodes with contributions (elsewhere not written)
  double *wq;    // bilinear deposit: per-node weight sum, likewise (nullptr for the 7x7 deposit)
  int mode;      // RBN_DEPOSIT_BOX7 or RBN_DEPOSIT_BILINEAR
  long long list_cap;
  int *overflow;
  int *active;   // the nodes with contributions, in no particular order (active_cap entries)
  int *nactive;  // their number
  long long active_cap;
  int dense;     // many contributions per node: fold every node in row order, no active list
  // The deposit as the nudged collision reads it, one Real2 per node in the lattice's
  // precision (tw64: double2, else float2): x = Tbar, y = -(the weight w, min(W, 1) for the
  // bilinear deposit, 1 for the 7x7 one) where the node has contributions, so that the sign
  // bit of y marks them (w = 0 gives -0.0); both +0.0 where it has none (zeroed memory is an
  // empty deposit).  Q = alpha * (w * (T - Tbar)) then equals the unpacked expressions.
  void *tw;
  int tw64;
  // bilinear deposit: contributions binned by deposit tile (kDepTileW x kDepTileH nodes):
  // per tile the number binned (tcnt[tile * kTcntStride], zero between deposits) and bin_cap
  // slots of two long longs
  // (w T and, shifted left by kDepTileBits, w in fixed point, with the node's index in the
  // tile in the low bits).  Contributions past a full tile's slots go to the fixed-point sums
  // S, Wi (2^-40 units) and the count ocnt at their node, zero between deposits.  Null for the
  // 7x7 deposit.
  int *tcnt;
  long long *bin;
  int bin_cap;
  long long *S, *Wi;
  int *ocnt;
};

// Deposit tiles of the bilinear deposit: 64 x 8 nodes, one workgroup each.  Their counters
// lie kTcntStride ints apart (256 bytes), so that the atomics on different tiles' counters do
// not queue on a few lines.
constexpr int kDepTileW = 64, kDepTileH = 8, kDepTileBits = 9, kTcntStride = 64;
static_assert(kDepTileW * kDepTileH == 1 << kDepTileBits, "a tile index fits kDepTileBits");
__host__ __device__ inline int dep_tiles_x(const Geom &g) { return (g.nx + kDepTileW - 1) / kDepTileW; }
__host__ __device__ inline int dep_tiles(const Geom &g) {
  return dep_tiles_x(g) * ((g.ny + kDepTileH - 1) / kDepTileH);
}
// Slots per tile for n observers (four contributions each) spread over the lattice: twice the
// mean and 64 more, a multiple of 64, at most 4096; a tile that receives more takes its
// overflow through the fixed-point sums.
inline int dep_bin_cap(const Geom &g, long long n) {
  const long long mean = 4 * n / dep_tiles(g) + 1;
  const long long cap = (2 * mean + 64 + 63) / 64 * 64;
  return (int)(cap < 4096 ? cap : 4096);
}

// One nudged member of a batched observer/deposit launch (ensembles): its lattice, the
// deposit's right limit, record interval, the step count before the launching call, its
// observers and deposit arrays.
struct NudgeMember {
  Geom g;
  int right, ntempread, istep0;
  ObsPtrs o;
  DepPtrs d;
};
// One record of a batched upload: 3n doubles (x, y, T) at byte offset src of the staging
// buffer, then (when chk is set) the n-entry observer map, copied to the observers' current
// record rx (rx, ry, rT consecutive) and map.
struct ObsJob {
  double *rx;
  int *chk;
  long long src;
  int n;
};

// Nodes per tile of the deposit's count scan (DepPtrs::blocksum holds one total per tile).
constexpr int kDepScanTile = 1024;
// Passes with per-member block counts (the prefix rows of a batched launch): observer
// update, zero pass, candidates (count or accumulate), count scan and list fill (7x7
// deposit only), fold or finalize.
enum { NB_OBS, NB_ZERO, NB_CAND, NB_SCAN, NB_FILL, NB_FOLD, NB_TILES, kNudgeBatches };

// Launch shape of the fused step: block width (64/128/256/512 nodes), non-temporal stores of
// the outputs / loads of the inputs, XCD-banded block order, alternating sweep direction.
// Defaults: the fastest measured on MI355X at 4096x2048 in both precisions (256 wide,
// non-temporal stores, cached loads, hardware block order, every odd step swept last row
// first: 0.233 ms against 0.238 ms alone and 0.239 against 0.271 ms after the particle
// update in FP32).
struct LbmTuning {
  int block = 256;
  int npt = 1;      // nodes per thread along x (1, 2, 4, 8): row segments of block * npt nodes
  int nt = 1;
  int ntload = 0;   // non-temporal loads of the step's inputs
  int uv_keep = 1;  // ordinary (cached) velocity stores under non-temporal stores
  int xcd = 0;      // 1: a contiguous band of rows per XCD; 2: one row in eight per XCD
  int reverse = 1;  // sweep every other step last row first
  int flip = 0;     // this launch sweeps last row first (set per launch from reverse)
  int overlap = 0;  // particle update of step s concurrent with the lattice step s+1
  int fuse = 1;     // particle update of step s in the grid of the lattice step s+1
  int steps2 = 1;   // truth run: two steps per lattice launch (temporal blocking in LDS)
  int steps3 = 0;   // truth run: three steps per lattice launch by row sweeps (with steps2)
  int seg3 = 0;     // three-step sweeps: rows per segment (0: one wave of workgroups)
  int nt3 = 0;      // three-step sweeps: non-temporal stores of the populations and F
  int tile_h = 0;   // rows of a two-step tile (4, 8, 16; 0: 8 in FP32, 4 in FP64)
  int seg = 0;      // two steps by row sweeps over segments of seg rows (8..64); 0: by tiles
  int xcd2 = 1;     // row sweeps: consecutive strips and segments on one XCD
  int variant = 0;  // row sweeps: 1 pulls one row ahead
  int dep_overlap = 0;  // nudged run: observer update and deposit of step s+1 beside the lattice step s
  int dep_in_tile = 1;  // nudged two-step launches: the second step's bilinear deposit summed in the tiles
  int obs_device = 1;   // nudged run over an observation table: the table held on the device
};

// The particle updates of one to three consecutive steps, in step order, carried by the next
// lattice launch: uv[u] / uvp[u] are the velocity slots of update u's step and the one
// before, rec[u] its record slot (or null), T the temperature of the one record step among
// them (a record update is never carried past a launch that rewrites the moments).
constexpr int kMaxCarried = 3;
struct PartJob {
  int n = 0;
  const void *uv[kMaxCarried] = {}, *uvp[kMaxCarried] = {};
  double *rec[kMaxCarried] = {};
  int record[kMaxCarried] = {}, full[kMaxCarried] = {};
  const void *T = nullptr;
  PartPtrs pp{};
};

// Non-finite state: a node whose moments are not finite sets *bad (a NaN or infinity in any
// population it pulls makes its rho or T NaN or infinite).  Only the store is conditional, so
// the check costs two compares per node and no traffic.
template <typename R>
__device__ __forceinline__ void flag_nonfinite(int *bad, R rho, R T) {
  if (bad && !(__builtin_isfinite(rho) && __builtin_isfinite(T))) *bad = 1;
}

struct AdjState;  // rbn_adjoint.cpp
void adj_free(rbn_ctx *c);

// Launchers (defined in the .hip files).
hipError_t launch_lbm(hipStream_t s, int prec, int variant, const Geom &g, const LbmPtrs &p,
                      const void *step_const, const LbmTuning &t);
// The second step's deposit (packed, DepPtrs::tw) of a two-step nudged launch, and the plane
// that receives the persisted nudging term after it (the first step's deposit and the term
// entering it are the LbmPtrs' tw and nud).
struct NudgePair {
  const void *tw2;
  void *nud_out;
  // or (non-null, bilinear deposit, tiles of kDepTileH rows) the second step's set, whose tile
  // pass was deferred: its deposit is summed from the bins inside the launch
  const DepPtrs *bins2 = nullptr;
  // (non-null, bilinear deposit) a later step's observer update and binning into head_dep,
  // carried as extra workgroups after the tiles (branch of the state machine at that step)
  const ObsPtrs *head_obs = nullptr;
  const DepPtrs *head_dep = nullptr;
  int head_branch = 0, head_ntempread = 1;
};
// Rows of the two-step tiles launch_lbm2 uses for precision prec and tuning tile_h.
int lbm2_tile_rows(int prec, int tile_h);
// Three steps of K1 in one launch by row sweeps (rbn_lbm2.hip, rbn_lbm3_wave.h): p holds the
// first step's inputs and the third step's outputs (fB, gB), p.F = F(s), p.uv = the first
// step's velocity slot; Fout receives F(s+3), uv2 / uv3 the velocities of steps s+1 / s+2,
// rho3 / T3 (or null) the moments of step s+mlev (mlev 0, 1 or 2).  seg: rows per segment (0: the fewest that
// put every workgroup on the device at once); xcd: consecutive strips and segments on one
// XCD; nt: non-temporal stores of the populations and F.  job (or null): particle updates
// carried as extra workgroups ahead of the strips.
hipError_t launch_lbm3(hipStream_t s, int prec, const Geom &g, const LbmPtrs &p, void *Fout, void *uv2, void *uv3,
                       void *rho3, void *T3, int mlev, const void *step_const, int seg, int xcd, int flip,
                       int nt, const PartJob *job);
// Rows per segment launch_lbm3 uses for tuning seg.
int lbm3_segment_rows(int prec, const Geom &g, int seg);
// Two steps of K1 in one launch (rbn_lbm2.hip): p holds the first step's inputs and the second
// step's outputs (fB, gB), p.F = F(s), p.uv = the first step's velocity slot; Fout receives
// F(s+2), uv2 the second step's velocity, rho2 / T2 (or null) its moments.  job (or null):
// particle updates carried as extra workgroups ahead of the tiles.  np (or null): the nudged
// run's pair (tiles only).  nt: non-temporal stores of the tiles' outputs.
hipError_t launch_lbm2(hipStream_t s, int prec, const Geom &g, const LbmPtrs &p, void *Fout, void *uv2,
                       void *rho2, void *T2, const void *step_const, int tile_h, int seg, int xcd, int variant,
                       int flip, int nt, const PartJob *job, const NudgePair *np = nullptr);
// K1 of a truth-run step with the particle update of the previous step in the same grid
// (its own workgroups, ahead of the lattice ones).
hipError_t launch_lbm_fused(hipStream_t s, int prec, const Geom &g, const LbmPtrs &p,
                            const void *step_const, const LbmTuning &t, const PartJob &job);
hipError_t launch_init_eq(hipStream_t s, int prec, const Geom &g, const void *rho, const void *T,
                          const void *uv, void *f, void *gp, void *F, const void *step_const);
hipError_t launch_particles(hipStream_t s, int prec, const Geom &g, const void *uv, const void *uvp,
                            const void *T, const PartPtrs &pp, int record, int full);
// Bilinear samples of a field whose node o holds its value at field[o * cs] (cs = 1 for
// scalar planes, 2 for the u or v component of an interleaved velocity slot).
hipError_t launch_sample(hipStream_t s, int prec, const Geom &g, const void *field, int cs, int n,
                         const double *x, const double *y, double *out, int *err);
hipError_t launch_particle_init(hipStream_t s, int prec, const Geom &g, const void *uv,
                                const PartPtrs &pp, double heavy_factor);
// One nudged step's observer update (branch: 1 first step of a record interval, 0 its last,
// 2 the others) and its deposit (counts, scan, contribution lists, fold).
// tail (or null): a bilinear deposit of the same lattice whose tile pass was deferred, run in
// the first launch (a 7x7 set is never one: it has no bins, and is ignored); defer: leave
// this deposit's tile pass (bilinear; a 7x7 deposit always completes) to a later call's
// tail, which must come before anything reads the deposit.
// An observation table on the device (nrec records of n entries): map[r n + i] = the first
// entry of record r whose id (ids[r n + j]) is observer i's (lookup[id - lo] = i for ids in
// lo .. lo + range - 1, -1 for no observer), INT_MAX when none; first_missing[r] = the first
// observer record r misses, n when none.
hipError_t launch_obs_table(hipStream_t s, const int *ids, int nrec, int n, const int *lookup, int lo, int range,
                            int *map, int *first_missing);
hipError_t launch_obs_deposit(hipStream_t s, int branch, int ntempread, const Geom &g, int right_limit,
                              const ObsPtrs &o, const DepPtrs &d, const DepPtrs *tail = nullptr,
                              bool defer = false);
// Blocks member d takes in each pass of a batched observer/deposit launch (kNudgeBlock
// threads each; an observer pass block updates observers lb * kNudgeBlock onward).
constexpr int kNudgeBlock = 256;
void nudge_batch_blocks(const NudgeMember &d, int blocks[kNudgeBatches]);
// The block map of the batched passes: for every block of pass p (entries off[p] ..
// off[p] + tot[p] - 1 of map, device memory, passes in order) its member, its index among
// that member's blocks of the pass, their number, and p.
struct NudgePlan {
  const int4 *map;
  int off[kNudgeBatches], tot[kNudgeBatches];
  int nm;
};
// Observer update and deposit of step t of a call (t = 0 is the first) for plan.nm members
// at once (members: device descriptors of the deposit set of step t).  tail (or null): the
// descriptors of the previous step's set, whose bilinear tile passes were deferred; defer:
// leave this step's bilinear tile passes to the next call's tail.  api: the bilinear tile
// passes also write the counts, means and weights (the last deposit of a call, which the
// getters and the members' own arrays receive); without it, and in a deferred pass, only the
// packed deposit the collisions read.
hipError_t launch_nudge_batch(hipStream_t s, const NudgeMember *members, const NudgePlan &plan, int t,
                              const NudgeMember *tail = nullptr, bool defer = false, bool api = true);
// The bilinear members' tile passes alone (their observer updates ran earlier).
hipError_t launch_nudge_tiles(hipStream_t s, const NudgeMember *members, const NudgePlan &plan, bool api);
// Scatter njobs staged records (jobs and data in the device staging buffer) to their observers.
hipError_t launch_obs_scatter(hipStream_t s, const ObsJob *jobs, int njobs, const unsigned char *staging);
hipError_t launch_energy(hipStream_t s, int prec, const Geom &g, const void *rho, const void *uv,
                         const void *T, double *partial, int nblocks, double *out2);
int energy_blocks(const Geom &g);
// Row profiles of the latest moments into out[ny * 9] (device, FP64).
hipError_t launch_profiles(hipStream_t s, int prec, const Geom &g, const void *rho, const void *T,
                           const void *uv, double *out);

}  // namespace rbn

// Context: one lattice on one GPU.
struct rbn_ctx {
  rbn_params P{};
  int dev = 0;
  hipStream_t stream = nullptr;
  rbn::Geom geom{};
  rbn::LbmTuning tune{};
  int prec = RBN_FP64;
  size_t rs = 8;  // bytes per real
  std::vector<unsigned char> consts;  // StepConst<R>

  // lattice state
  void *fA = nullptr, *fB = nullptr, *gA = nullptr, *gB = nullptr;
  void *F = nullptr;
  void *F2 = nullptr;  // second force plane: a two-step launch reads F and writes F2, then swaps
  // Velocity ring: uv[cur] is the latest hydrovar velocity; the step before it is in the slot
  // that was cur then (uv[(cur+7)%8] while rbn_step advances the ring; an ensemble keeps the
  // ring in the first six slots).  Each slot holds (u, v) interleaved per node (2 * plane
  // reals), so that the particle update fetches one cache line per patch row for both
  // components.  Eight slots let a three-step lattice launch (steps s .. s+2) carry the
  // particle updates of steps s-3 .. s-1, which read the slots of steps s-4 .. s-1.
  static constexpr int kVel = 8;
  void *uv[kVel] = {};
  int cur = 0;
  void *rho = nullptr, *T = nullptr;
  hipStream_t pstream = nullptr;       // particle stream (default priority)
  hipEvent_t ev_lbm = nullptr;         // K1 of the latest step finished
  hipEvent_t ev_part[2] = {};          // K2 of step s finished, by s % 2
  // nudged run with dep_overlap: the second deposit set (allocated on first use) and the
  // events between the deposits (pstream) and the lattice launches (stream), by step parity
  rbn::DepPtrs dep2{};
  hipEvent_t ev_dep[2] = {}, ev_lat[2] = {};
  bool part_record_pending = false;    // K2 of the latest step reads T (record step)
  void *nud = nullptr;
  void *nud2 = nullptr;  // nudged run: a two-step launch reads nud and writes nud2, then they swap
  int istep = 0;
  bool initialised = false;
  bool moments_valid = false;  // rho/T hold the latest hydrovar

  // particles
  int np = 0;
  std::vector<int> pid;  // ids in ascending order
  double *d_part = nullptr;  // 10 arrays of np: x,y,u,v,xo,yo,uo,vo,inv_tau,beta
  // Records: a ring of device/pinned-host slots.  K2 writes a slot on the compute stream,
  // the copy stream moves it to the host while later steps run, and the callback fires
  // (in step order) when the slot is reused or rbn_step returns.
  static constexpr int kRecSlots = 3;
  struct RecSlot {
    double *d = nullptr;  // 17 * np
    double *h = nullptr;  // pinned
    hipEvent_t written = nullptr, copied = nullptr;
    int step = -1;        // pending record step, -1 when free
  } rec[kRecSlots];
  int rec_next = 0;
  // The particle updates waiting to ride in the next lattice launch (particle_fuse): the
  // latest step's, or the latest two or three steps' after a launch of that many steps, in
  // step order; never pending when rbn_step returns.
  struct PendingParticles {
    int istep = 0, cur = 0, prv = 0, record = 0, full = 0;
    RecSlot *slot = nullptr;
  } pend[rbn::kMaxCarried];
  int npend = 0;
  hipStream_t copy_stream = nullptr;
  int *d_err = nullptr;
  rbn_record_fn rec_fn = nullptr;
  void *rec_user = nullptr;

  // observers
  int nobs = 0;
  std::vector<int> oid;
  int *d_obs_i = nullptr;     // chk, bcy, xidx, yidx (4 * nobs)
  double *d_obs_d = nullptr;  // dx, dy, dT, xint, yint, Tint, rx, ry, rT (9 * nobs)
  rbn_obs_fn obs_fn = nullptr;
  void *obs_user = nullptr;
  // rbn_set_obs_table: caller-owned records served by the built-in provider
  struct ObsTable {
    int nrec = 0, n = 0;
    const int *steps = nullptr, *id = nullptr;
    const double *x = nullptr, *y = nullptr, *T = nullptr;
  } obs_table;
  // Record staging: two pinned slots (x, y, T as doubles, then ids and the observer map as
  // ints) so that the upload of a record is queued on the stream without waiting for it;
  // a slot is refilled only after the event of its previous upload.  The observer -> record
  // map is rebuilt and uploaded only when a record's ids differ from the previous record's.
  struct ObsSlot {
    unsigned char *h = nullptr;
    size_t cap = 0;
    hipEvent_t copied = nullptr;
  } obs_slot[2];
  int obs_next = 0;
  std::vector<int> prev_rid;  // ids of the record whose observer map is on the device
  // rbn_step over an observation table held on the device (obs_device): x, y, T of record r
  // at d_tab + 3 r nobs (the layout of rx, ry, rT), its observer map at d_tabmap + r nobs,
  // tab_miss[r] its first missing observer (nobs when none); tab_rec: the table record now
  // current, or -1 (the current record is in the observer arrays' own rx, ry, rT, chk);
  // tab_tried: built or found unsuitable since the table or the observers last changed.
  double *d_tab = nullptr;
  int *d_tabmap = nullptr;
  std::vector<int> tab_miss;
  int tab_rec = -1;
  bool tab_tried = false;
  rbn::DepPtrs dep{};

  // diagnostics
  rbn_diag_fn diag_fn = nullptr;
  void *diag_user = nullptr;
  double *d_energy = nullptr;  // partials + 2 results

  // timing: every timing_every-th launch of each kernel class is bracketed with events
  static constexpr int kKernelClasses = 7;
  bool timing = false;
  int timing_every = 1;
  long long k_seen[kKernelClasses] = {};
  struct Pending {
    int kernel;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> event_pool;
  long long k_launches[kKernelClasses] = {};
  double k_ms[kKernelClasses] = {};

  // adjoint window (rbn_adj_forward / rbn_adj_backward)
  bool at_init = false;             // state is the one rbn_init_fields set (no step since)
  int adj_store_deposit = 1;        // keep every step's deposit when it fits (else replay it)
  rbn::AdjState *adj = nullptr;

  std::string err;
};
