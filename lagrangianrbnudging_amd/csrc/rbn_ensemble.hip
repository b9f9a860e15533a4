// rbn_ensemble.hip — batched ensemble: many independent truth-run lattices of uneven sizes
// and Rayleigh numbers advanced together.  No reference counterpart: the reference runs one
// lattice per MPI job.
//
// Work items:
//   lattice: every (member, tile) item of the fused LBM step (device code of K1,
//            rbn_lbm_node.h);
//   particle: every (member, 256-particle chunk) item of the particle step (K2,
//            rbn_particle_node.h).
// By default one persistent launch per step runs the particle items of the previous step
// followed by the lattice items of this step (the particle update of step k-1 reads the
// velocity slots of steps k-1 and k-2 and writes only particle state; the lattice step k
// writes slot k), so the latency-bound gathers overlap the lattice streaming.  The first
// step of a call runs the lattice items alone and a last launch runs the final step's
// particle items.  With separate phases (order bit 2) each step is a lattice launch and a
// particle launch on a second stream beside the next lattice launch.
// A fixed pool of workgroups (wg_per_cu per CU, chosen by the autotuner; 0 = one workgroup
// per item) pulls items from an atomic counter, so small and large members share the GPU
// instead of each small lattice leaving most CUs idle.  The host captures kCycle = 6
// consecutive steps (after six steps the A/B population planes and the six-deep velocity
// ring of every member are back where they started) into one hipGraph and replays it.
//
// Work counters: a launch at step k uses counter set k % 2 (lattice or fused launches one
// counter, particle launches the other) and clears its counter of set (k+1) % 2 for the
// next step, which only starts after this launch has finished; the counters are therefore
// zero at every launch without any memset node in the graph.
#include "rbn_lbm2_tile.h"
#include "rbn_lbm_node.h"
#include "rbn_particle_node.h"
#include "rbn_ensemble.h"
#include "rbn_nudge_node.h"

namespace rbn {
namespace {

constexpr int kThreads = 256;
static_assert(kThreads == lbm2::kTileThreads, "two-step tiles run on whole workgroups");

// One-step tile of step k: input planes pf[par] / pg[par], force F (read and rewritten in
// place).
template <typename R>
__device__ __forceinline__ void lbm_item(const EnsMember<R> &m, int tile, int k, int par, R *F, R *nud, int dset,
                                         int tile_w, int tile_h, int write_moments) {
  const int tx = tile % m.tiles_x, ty = tile / m.tiles_x;
  LbmPtrs p{};
  p.fA = par ? m.pf[1] : m.pf[0];
  p.fB = par ? m.pf[0] : m.pf[1];
  p.gA = par ? m.pg[1] : m.pg[0];
  p.gB = par ? m.pg[0] : m.pg[1];
  p.F = F;
  const int out = (m.cur0 + 1 + k) % kCycle;
  p.uv = m.uv[out];
  p.rho = write_moments ? m.rho : nullptr;
  p.T = write_moments ? m.T : nullptr;
  p.bad = m.bad;
  const int col = threadIdx.x % tile_w;
  if (m.variant == RBN_VARIANT_NUDGING) {
    p.tw = m.tw[dset];
    p.nud = nud;
    for (int r = threadIdx.x / tile_w; r < tile_h; r += kThreads / tile_w) {
      const int i = tx * tile_w + col, j = ty * tile_h + r;
      if (i < m.g.nx && j < m.g.ny) {
        if constexpr (sizeof(R) == 8)
          lbm::lbm_node64<RBN_VARIANT_NUDGING, 1>(m.g, p, m.c, i, j);
        else
          lbm::lbm_node32<RBN_VARIANT_NUDGING, 1>(m.g, p, m.c, i, j);
      }
    }
    return;
  }
  for (int r = threadIdx.x / tile_w; r < tile_h; r += kThreads / tile_w) {
    const int i = tx * tile_w + col, j = ty * tile_h + r;
    if (i < m.g.nx && j < m.g.ny) {
      if constexpr (sizeof(R) == 8)
        lbm::lbm_node64<RBN_VARIANT_TRAJECTORIES, 1>(m.g, p, m.c, i, j);
      else
        lbm::lbm_node32<RBN_VARIANT_TRAJECTORIES, 1>(m.g, p, m.c, i, j);
    }
  }
}

template <typename R>
__device__ __forceinline__ void particle_item(const EnsMember<R> &m, int chunk, int k) {
  const int p = chunk * kThreads + threadIdx.x;
  if (p >= m.np) return;
  // k is -1 for the previous cycle's last step (carried by step 0)
  const int cur = (m.cur0 + 1 + k + kCycle) % kCycle, prv = (m.cur0 + k + kCycle) % kCycle;
  const int nx = m.g.nx, ny = m.g.ny, pitch = m.g.pitch;
  const part::VelField<R> UV = part::vel_of<R>(m.uv[cur], nx, ny, pitch);
  const part::VelField<R> UVp = part::vel_of<R>(m.uv[prv], nx, ny, pitch);
  const part::Field<R> T = part::scalar_of<R>(m.T, nx, ny, pitch);
  PartPtrs pp{};
  const size_t n = (size_t)m.np;
  double *b = m.part;
  pp.n = m.np;
  pp.x = b;
  pp.y = b + n;
  pp.u = b + 2 * n;
  pp.v = b + 3 * n;
  pp.xo = b + 4 * n;
  pp.yo = b + 5 * n;
  pp.uo = b + 6 * n;
  pp.vo = b + 7 * n;
  pp.inv_tau = b + 8 * n;
  pp.beta = b + 9 * n;
  pp.rec = nullptr;
  pp.err = m.err;
  part::particle_update(UV, UVp, T, pp, p, 0, 0);
}

// Counters: [set][slot], set = step % 2, slot 0 for lattice and fused launches, 1 for
// particle launches.  A launch at step k clears only its counter of set (k+1) % 2: the next
// launch with the same slot is on the same stream and starts after this one ends, while a
// launch with the other slot may be running concurrently.
__device__ __forceinline__ void clear_next(int *counters, int k, int slot) {
  if (blockIdx.x == 0 && threadIdx.x == 0) counters[((k + 1) & 1) * 2 + slot] = 0;
}

// Item `it` of a launch of PHASE: the particle items come first (phases 1, 2; phase 2 runs
// them for step k-1), then the lattice items (phases 0, 2), visited last first when reverse
// is set (odd steps, so that a step starts on the tiles its predecessor wrote last).
template <typename R, int PHASE>
__device__ __forceinline__ void run_item(const EnsMember<R> *__restrict__ M, const int2 *__restrict__ items,
                                         int it, int n_lbm, int n_first, int k, int tile_w, int tile_h,
                                         int write_moments, int reverse) {
  if (it < n_first) {
    const int2 w = items[n_lbm + it];
    particle_item(M[w.x], w.y, PHASE == 2 ? k - 1 : k);
    return;
  }
  const int l = it - n_first;
  const int2 w = items[reverse ? n_lbm - 1 - l : l];
  lbm_item(M[w.x], w.y, k, k & 1, M[w.x].F[0], M[w.x].nud[0], k & 1, tile_w, tile_h, write_moments);
}

template <typename R, int PHASE>
__global__ __launch_bounds__(kThreads) void k_ensemble(const EnsMember<R> *__restrict__ M,
                                                       const int2 *__restrict__ items, int n_lbm,
                                                       int n_part, int *counters, int k, int tile_w,
                                                       int tile_h, int write_moments, int persistent,
                                                       int reverse) {
  constexpr int slot = PHASE == 1 ? 1 : 0;
  const int n_first = PHASE == 0 ? 0 : n_part;
  const int n_items = n_first + (PHASE == 1 ? 0 : n_lbm);
  clear_next(counters, k, slot);
  if (!persistent) {
    const int it = blockIdx.x;
    if (it < n_items) run_item<R, PHASE>(M, items, it, n_lbm, n_first, k, tile_w, tile_h, write_moments, reverse);
    return;
  }
  __shared__ int item;
  int *ctr = &counters[(k & 1) * 2 + slot];
  for (;;) {
    if (threadIdx.x == 0) item = atomicAdd(ctr, 1);
    __syncthreads();
    const int it = __builtin_amdgcn_readfirstlane(item);  // uniform: descriptors in scalar registers
    __syncthreads();
    if (it >= n_items) break;
    run_item<R, PHASE>(M, items, it, n_lbm, n_first, k, tile_w, tile_h, write_moments, reverse);
  }
}

}  // namespace

namespace {
template <typename R, int PHASE>
void launch_phase(hipStream_t s, dim3 g, const void *members, const int2 *items, int n_lbm, int n_part,
                  int *counters, int k, int tile_w, int tile_h, int write_moments, int persistent,
                  int reverse) {
  hipLaunchKernelGGL((k_ensemble<R, PHASE>), g, dim3(kThreads), 0, s, (const EnsMember<R> *)members, items,
                     n_lbm, n_part, counters, k, tile_w, tile_h, write_moments, persistent, reverse);
}

template <typename R>
void launch_prec(hipStream_t s, int phase, dim3 g, const void *members, const int2 *items, int n_lbm,
                 int n_part, int *counters, int k, int tile_w, int tile_h, int write_moments,
                 int persistent, int reverse) {
  if (phase == 0)
    launch_phase<R, 0>(s, g, members, items, n_lbm, n_part, counters, k, tile_w, tile_h, write_moments,
                       persistent, reverse);
  else if (phase == 1)
    launch_phase<R, 1>(s, g, members, items, n_lbm, n_part, counters, k, tile_w, tile_h, write_moments,
                       persistent, reverse);
  else
    launch_phase<R, 2>(s, g, members, items, n_lbm, n_part, counters, k, tile_w, tile_h, write_moments,
                       persistent, reverse);
}
// Kinds of two-step-mode launches: one step or particle updates only (no tile LDS), two-step
// tiles of an ensemble of truth runs, of nudged runs, of both.  Each kind compiles only the
// bodies it runs: both bodies in one kernel take 173 VGPRs in FP32 (two waves per SIMD), the
// nudged one alone 163 (three, as many as the LDS allows).
enum { kOneStep = 0, kPairsTruth = 1, kPairsNudged = 2, kPairsMixed = 3 };
constexpr int pair_ring_g(int kind) { return kind == kPairsTruth ? 3 : 9; }

// Two-step tile of steps k, k+1 (64 x TH nodes, rbn_lbm2_tile.h): the truth-run body, or the
// nudged one (its LDS ring keeps all g directions) reading the deposits of steps k and k+1
// (sets k & 1 and (k + 1) & 1) and moving the nudging term from nud[fpar] to nud[1 - fpar].
template <typename R, int TH, int KIND>
__device__ __forceinline__ void pair_item(const EnsMember<R> &m, int tile, const EnsStep &st,
                                          lbm2::TileLds<R, TH, pair_ring_g(KIND)> &L) {
  constexpr int GR = pair_ring_g(KIND);
  const int tx = tile % m.t2x, ty = tile / m.t2x;
  lbm2::TwoStepPtrs p{};
  p.fA = m.pf[st.lpar];
  p.gA = m.pg[st.lpar];
  p.fB = m.pf[1 - st.lpar];
  p.gB = m.pg[1 - st.lpar];
  p.Fin = m.F[st.fpar];
  p.Fout = m.F[1 - st.fpar];
  p.uv1 = m.uv[(m.cur0 + 1 + st.k) % kCycle];
  p.uv2 = m.uv[(m.cur0 + 2 + st.k) % kCycle];
  p.rho2 = st.moments ? m.rho : nullptr;
  p.T2 = st.moments ? m.T : nullptr;
  p.bad = m.bad;
  if constexpr (KIND != kPairsTruth) {
    if (KIND == kPairsNudged || m.variant == RBN_VARIANT_NUDGING) {
      for (int u = 0; u < 2; u++) p.tw[u] = m.tw[(st.k + u) & 1];
      const int s2 = (st.k + 1) & 1;
      if (st.bins2 && m.bin[s2]) {
        p.bin2 = m.bin[s2];
        p.tcnt2 = m.tcnt[s2];
        p.bin_cap2 = m.bin_cap;
        p.S2 = m.S[s2];
        p.Wi2 = m.Wi[s2];
        p.ocnt2 = m.ocnt[s2];
      }
      p.nud_in = m.nud[st.fpar];
      p.nud_out = m.nud[1 - st.fpar];
      lbm2::two_step_tile<R, TH, RBN_VARIANT_NUDGING, GR>(m.g, p, m.c, tx, ty, L);
      return;
    }
  }
  if constexpr (KIND != kPairsNudged) lbm2::two_step_tile<R, TH, RBN_VARIANT_TRAJECTORIES, GR>(m.g, p, m.c, tx, ty, L);
}

// LDS of a two-step-mode launch: the two-step tile's, none for one-step launches.
template <typename R, int TH, int KIND>
struct PairLds {
  lbm2::TileLds<R, TH, pair_ring_g(KIND)> L;
};
template <typename R, int TH>
struct PairLds<R, TH, kOneStep> {};

// Item `it` of a two-step-mode launch: carried particle items first, then lattice items
// (two-step tiles, or one-step tiles for kOneStep), then the carried observer-update items
// (last: they fill the launch's drain, when lattice items run out).
template <typename R, int TH, int KIND>
__device__ __forceinline__ void run_item2(const EnsMember<R> *__restrict__ M, const int2 *__restrict__ items, int it,
                                          int n_lbm, int n_part, int n_first, int n_lat, const EnsStep &st,
                                          const EnsHead &head, int tile_w, int tile_h, PairLds<R, TH, KIND> &S) {
  if (it < n_first) {
    const int2 w = items[n_lbm + it];
    for (int u = 0; u < st.ncarry; u++) particle_item(M[w.x], w.y, st.k - st.ncarry + u);
    return;
  }
  if (it >= n_first + n_lat) {
    const int4 w = head.map[it - n_first - n_lat];
    obs::head_one(head.M[w.x], w.y * kNudgeBlock + threadIdx.x, head.t);
    return;
  }
  const int l = it - n_first, lr = st.reverse ? n_lat - 1 - l : l;
  if constexpr (KIND != kOneStep) {
    const int2 w = items[n_lbm + n_part + lr];
    pair_item<R, TH, KIND>(M[w.x], w.y, st, S.L);
  } else {
    const int2 w = items[lr];
    lbm_item(M[w.x], w.y, st.k, st.lpar, M[w.x].F[st.fpar], M[w.x].nud[st.fpar], st.k & 1, tile_w, tile_h,
             st.moments);
  }
}

// KIND: kOneStep for the launches of one step or particle updates only (lat 1, 0), a pairs
// kind for those of two-step tiles (lat 2).  The nudged pairs kernel is held to three waves
// per SIMD (as many as its LDS allows): unconstrained, the FP32 build takes 171 VGPRs (two
// waves), constrained 168 without spills.
template <typename R, int TH, int KIND>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(KIND == kPairsNudged ? 3 : 1)))
void k_ensemble2(const EnsMember<R> *__restrict__ M, const int2 *__restrict__ items, int n_lbm, int n_part,
                 int n_lbm2, int *counters, EnsStep st, EnsHead head, int tile_w, int tile_h, int persistent) {
  __shared__ PairLds<R, TH, KIND> L;
  const int slot = st.lat ? 0 : 1;
  const int n_first = st.ncarry > 0 ? n_part : 0;
  const int n_lat = st.lat == 2 ? n_lbm2 : (st.lat == 1 ? n_lbm : 0);
  const int n_items = n_first + n_lat + head.n;
  clear_next(counters, st.cset, slot);
  if (!persistent) {
    const int it = blockIdx.x;
    if (it < n_items)
      run_item2<R, TH, KIND>(M, items, it, n_lbm, n_part, n_first, n_lat, st, head, tile_w, tile_h, L);
    return;
  }
  __shared__ int item;
  int *ctr = &counters[(st.cset & 1) * 2 + slot];
  for (;;) {
    if (threadIdx.x == 0) item = atomicAdd(ctr, 1);
    __syncthreads();
    const int it = __builtin_amdgcn_readfirstlane(item);  // uniform: descriptors in scalar registers
    __syncthreads();
    if (it >= n_items) break;
    run_item2<R, TH, KIND>(M, items, it, n_lbm, n_part, n_first, n_lat, st, head, tile_w, tile_h, L);
  }
}

}  // namespace

hipError_t launch_ensemble_phase(hipStream_t s, int prec, int phase, const void *members,
                                 const int2 *items, int n_lbm, int n_part, int *counters, int k,
                                 int tile_w, int tile_h, int write_moments, int grid, int reverse) {
  const int n_items = (phase == 0 ? 0 : n_part) + (phase == 1 ? 0 : n_lbm);
  if (n_items <= 0) return hipSuccess;
  const int persistent = grid > 0;
  const dim3 g(persistent ? (unsigned)grid : (unsigned)n_items);
  if (prec == RBN_FP64)
    launch_prec<double>(s, phase, g, members, items, n_lbm, n_part, counters, k, tile_w, tile_h,
                        write_moments, persistent, reverse);
  else
    launch_prec<float>(s, phase, g, members, items, n_lbm, n_part, counters, k, tile_w, tile_h,
                       write_moments, persistent, reverse);
  return hipGetLastError();
}

hipError_t launch_ensemble_pairs(hipStream_t s, int prec, const void *members, const int2 *items, int n_lbm,
                                 int n_part, int n_lbm2, int *counters, const EnsStep &st, int tile_w,
                                 int tile_h, int grid, int nudged, const EnsHead &head_in) {
  EnsHead head = head_in;
  if (st.lat != 2 || !head.M) head.n = 0;  // only two-step launches carry observer updates
  const int n_items =
      (st.ncarry > 0 ? n_part : 0) + (st.lat == 2 ? n_lbm2 : (st.lat == 1 ? n_lbm : 0)) + head.n;
  if (n_items <= 0) return hipSuccess;
  const int persistent = grid > 0;
  const dim3 g(persistent ? (unsigned)grid : (unsigned)n_items), b(kThreads);
  const auto *M64 = (const EnsMember<double> *)members;
  const auto *M32 = (const EnsMember<float> *)members;
  const int kind = st.lat != 2 ? kOneStep : (nudged == 0 ? kPairsTruth : (nudged == 1 ? kPairsNudged : kPairsMixed));
#define RBN_ENS2(R, M, TH, KIND)                                                                                     \
  hipLaunchKernelGGL((k_ensemble2<R, TH, KIND>), g, b, 0, s, M, items, n_lbm, n_part, n_lbm2, counters, st, head, \
                     tile_w, tile_h, persistent)
#define RBN_ENS2_KINDS(R, M, TH)           \
  switch (kind) {                          \
    case kPairsTruth:                      \
      RBN_ENS2(R, M, TH, kPairsTruth);     \
      break;                               \
    case kPairsNudged:                     \
      RBN_ENS2(R, M, TH, kPairsNudged);    \
      break;                               \
    case kPairsMixed:                      \
      RBN_ENS2(R, M, TH, kPairsMixed);     \
      break;                               \
    default:                               \
      RBN_ENS2(R, M, TH, kOneStep);        \
  }
  if (prec == RBN_FP64) {
    RBN_ENS2_KINDS(double, M64, kEnsTH64)
  } else {
    RBN_ENS2_KINDS(float, M32, kEnsTH32)
  }
#undef RBN_ENS2_KINDS
#undef RBN_ENS2
  return hipGetLastError();
}

}  // namespace rbn
