// rbn_lbm_fused.hip — K1 of a truth-run step and K2 of the step before in one grid.
//
// The particle update of step s reads the velocity slots of steps s and s-1 (and T of step
// s on record steps) and writes only particle state; the lattice step s+1 reads the
// populations and writes the next velocity slot.  The two touch disjoint memory, so the
// particle update can ride in the launch of the next lattice step: its workgroups come
// first in the grid, and their gathers (latency bound, a few cache lines per particle)
// overlap the streaming of the lattice workgroups instead of running as a launch of their
// own between two lattice launches.  After a two-step launch two updates are pending; a
// thread runs both for its particle, in step order.  The caller keeps the one conflict out:
// a record step's update reads T, so it is not carried by a launch that rewrites the
// moments.
#include "rbn_lbm_node.h"
#include "rbn_particle_node.h"

namespace rbn {
namespace {

using namespace lbm;

constexpr int kMaxBlock = 512;

template <int NT>
__global__ __launch_bounds__(kMaxBlock) void k_lbm_fused64(Geom g, LbmPtrs p, StepConst<double> c, int order,
                                                           int npt, part::PartSteps<double> a) {
  if ((int)blockIdx.x < a.nblocks) {
    part::run_part_steps(a, blockIdx.x * blockDim.x + threadIdx.x);
    return;
  }
  int i0, j;
  segment_node(g, blockIdx.x - a.nblocks, gridDim.x - a.nblocks, order, npt, i0, j);
  for (int k = 0; k < npt; k++) {
    const int i = i0 + k * (int)blockDim.x;
    if (i < g.nx) lbm_node64<RBN_VARIANT_TRAJECTORIES, NT>(g, p, c, i, j);
  }
}

template <int NT>
__global__ __launch_bounds__(kMaxBlock) void k_lbm_fused32(Geom g, LbmPtrs p, StepConst<float> c, int order,
                                                           int npt, part::PartSteps<float> a) {
  if ((int)blockIdx.x < a.nblocks) {
    part::run_part_steps(a, blockIdx.x * blockDim.x + threadIdx.x);
    return;
  }
  int i0, j;
  segment_node(g, blockIdx.x - a.nblocks, gridDim.x - a.nblocks, order, npt, i0, j);
  for (int k = 0; k < npt; k++) {
    const int i = i0 + k * (int)blockDim.x;
    if (i < g.nx) lbm_node32<RBN_VARIANT_TRAJECTORIES, NT>(g, p, c, i, j);
  }
}

template <int NT>
void fused_dispatch(hipStream_t s, int prec, unsigned lattice_blocks, int bw, const Geom &g,
                    const LbmPtrs &p, const void *step_const, int order, int npt, const PartJob &job) {
  if (prec == RBN_FP64) {
    const part::PartSteps<double> a = part::part_steps<double>(g, &job, bw);
    hipLaunchKernelGGL((k_lbm_fused64<NT>), dim3(lattice_blocks + (unsigned)a.nblocks), dim3(bw), 0, s, g, p,
                       *(const StepConst<double> *)step_const, order, npt, a);
  } else {
    const part::PartSteps<float> a = part::part_steps<float>(g, &job, bw);
    hipLaunchKernelGGL((k_lbm_fused32<NT>), dim3(lattice_blocks + (unsigned)a.nblocks), dim3(bw), 0, s, g, p,
                       *(const StepConst<float> *)step_const, order, npt, a);
  }
}

}  // namespace

hipError_t launch_lbm_fused(hipStream_t s, int prec, const Geom &g, const LbmPtrs &p,
                            const void *step_const, const LbmTuning &t, const PartJob &job) {
  const int bw = (t.block == 64 || t.block == 128 || t.block == 512) ? t.block : 256;
  const int npt = (t.npt == 2 || t.npt == 4 || t.npt == 8) ? t.npt : 1;
  const unsigned lattice_blocks = (unsigned)((g.nx + bw * npt - 1) / (bw * npt)) * (unsigned)g.ny;
  const int nt = (t.nt ? 1 : 0) | (t.ntload ? 2 : 0) | (t.uv_keep ? 4 : 0);
  const int order = (t.xcd == 1 ? 1 : 0) | (t.flip ? 2 : 0) | (t.xcd == 2 ? 4 : 0);
  // Bit 2 only changes the velocity store when bit 0 is set (4 and 6 are 0 and 2).
  switch (nt & 7) {
    case 0: case 4: fused_dispatch<0>(s, prec, lattice_blocks, bw, g, p, step_const, order, npt, job); break;
    case 1: fused_dispatch<1>(s, prec, lattice_blocks, bw, g, p, step_const, order, npt, job); break;
    case 2: case 6: fused_dispatch<2>(s, prec, lattice_blocks, bw, g, p, step_const, order, npt, job); break;
    case 3: fused_dispatch<3>(s, prec, lattice_blocks, bw, g, p, step_const, order, npt, job); break;
    case 5: fused_dispatch<5>(s, prec, lattice_blocks, bw, g, p, step_const, order, npt, job); break;
    default: fused_dispatch<7>(s, prec, lattice_blocks, bw, g, p, step_const, order, npt, job); break;
  }
  return hipGetLastError();
}

}  // namespace rbn
