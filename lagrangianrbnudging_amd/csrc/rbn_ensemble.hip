// rbn_ensemble.hip — batched ensemble: many independent truth-run lattices of uneven sizes
// and Rayleigh numbers advanced together.  No reference counterpart: the reference runs one
// lattice per MPI job.
//
// One ensemble step is two persistent launches over a work list:
//   E1: every (member, tile) item of the fused LBM step (device code of K1, rbn_lbm_node.h);
//   E2: every (member, 256-particle chunk) item of the particle step (K2, rbn_particle_node.h).
// A fixed pool of workgroups (wg_per_cu per CU, chosen by the autotuner; 0 = one workgroup
// per item) pulls items from an atomic counter, so small and large members share the GPU
// instead of each small lattice leaving most CUs idle.  The host captures four consecutive
// steps (after four steps the A/B population planes and the four-deep velocity ring of
// every member are back where they started) into one hipGraph and replays it.
//
// Work counters: step k of the four-step cycle uses counter set k % 2 and clears set
// (k+1) % 2 for the next step, which only starts after this launch has finished; the
// counters are therefore zero at every launch without any memset node in the graph.
#include "rbn_lbm_node.h"
#include "rbn_particle_node.h"
#include "rbn_ensemble.h"

namespace rbn {
namespace {

constexpr int kThreads = 256;

template <typename R>
__device__ __forceinline__ void lbm_item(const EnsMember<R> &m, int tile, int k, int tile_w,
                                         int tile_h, int write_moments) {
  const int tx = tile % m.tiles_x, ty = tile / m.tiles_x;
  const int par = k & 1;
  LbmPtrs p{};
  p.fA = par ? m.pf[1] : m.pf[0];
  p.fB = par ? m.pf[0] : m.pf[1];
  p.gA = par ? m.pg[1] : m.pg[0];
  p.gB = par ? m.pg[0] : m.pg[1];
  p.F = m.F;
  const int out = (m.cur0 + 1 + k) & 3;
  p.uv = m.uv[out];
  p.rho = write_moments ? m.rho : nullptr;
  p.T = write_moments ? m.T : nullptr;
  const int col = threadIdx.x % tile_w;
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
  const int cur = (m.cur0 + 1 + k) & 3, prv = (m.cur0 + k) & 3;
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

// Counters: [set][phase], set = step % 2.  A launch of phase P at step k clears only
// counter [(k+1)%2][P]: the next launch of the same phase is on the same stream and starts
// after this one ends, while the other phase may be running concurrently.
__device__ __forceinline__ void clear_next(int *counters, int k, int phase) {
  if (blockIdx.x == 0 && threadIdx.x == 0) counters[((k + 1) & 1) * 2 + phase] = 0;
}

// reverse: visit the items last first (lattice phase on odd steps, so that a step starts on
// the tiles its predecessor wrote last).
template <typename R, int PHASE>
__global__ __launch_bounds__(kThreads) void k_ensemble(const EnsMember<R> *__restrict__ M,
                                                       const int2 *__restrict__ items, int n_items,
                                                       int *counters, int k, int tile_w, int tile_h,
                                                       int write_moments, int persistent, int reverse) {
  clear_next(counters, k, PHASE);
  if (!persistent) {
    const int it = blockIdx.x;
    if (it >= n_items) return;
    const int2 w = items[reverse ? n_items - 1 - it : it];
    if (PHASE == 0)
      lbm_item(M[w.x], w.y, k, tile_w, tile_h, write_moments);
    else
      particle_item(M[w.x], w.y, k);
    return;
  }
  __shared__ int item;
  int *ctr = &counters[(k & 1) * 2 + PHASE];
  for (;;) {
    if (threadIdx.x == 0) item = atomicAdd(ctr, 1);
    __syncthreads();
    const int it = item;
    __syncthreads();
    if (it >= n_items) break;
    const int2 w = items[reverse ? n_items - 1 - it : it];
    if (PHASE == 0)
      lbm_item(M[w.x], w.y, k, tile_w, tile_h, write_moments);
    else
      particle_item(M[w.x], w.y, k);
  }
}

}  // namespace

hipError_t launch_ensemble_phase(hipStream_t s, int prec, int phase, const void *members,
                                 const int2 *items, int n_items, int *counters, int k, int tile_w,
                                 int tile_h, int write_moments, int grid, int reverse) {
  if (n_items <= 0) return hipSuccess;
  const int persistent = grid > 0;
  const dim3 g(persistent ? (unsigned)grid : (unsigned)n_items), b(kThreads);
  if (prec == RBN_FP64) {
    const auto *M = (const EnsMember<double> *)members;
    if (phase == 0)
      hipLaunchKernelGGL((k_ensemble<double, 0>), g, b, 0, s, M, items, n_items, counters, k, tile_w,
                         tile_h, write_moments, persistent, reverse);
    else
      hipLaunchKernelGGL((k_ensemble<double, 1>), g, b, 0, s, M, items, n_items, counters, k, tile_w,
                         tile_h, write_moments, persistent, reverse);
  } else {
    const auto *M = (const EnsMember<float> *)members;
    if (phase == 0)
      hipLaunchKernelGGL((k_ensemble<float, 0>), g, b, 0, s, M, items, n_items, counters, k, tile_w,
                         tile_h, write_moments, persistent, reverse);
    else
      hipLaunchKernelGGL((k_ensemble<float, 1>), g, b, 0, s, M, items, n_items, counters, k, tile_w,
                         tile_h, write_moments, persistent, reverse);
  }
  return hipGetLastError();
}

}  // namespace rbn
