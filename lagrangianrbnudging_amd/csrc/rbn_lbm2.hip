// rbn_lbm2.hip — two truth-run steps of K1 in one launch (the tile code, rbn_lbm2_tile.h,
// keeps the first step's populations in LDS).  A step that must write its moments is never
// the first of a pair (the host launches it alone).  Particle updates of earlier steps ride
// in the launch as workgroups ahead of the tiles (rbn_lbm_fused.hip explains why that is
// safe).
#include "rbn_lbm2_tile.h"
#include "rbn_particle_node.h"

namespace rbn {
namespace {

using namespace lbm2;

struct Lbm2Args {
  TwoStepPtrs p;
  int tiles_x, ntiles, flip;
};

template <typename R, int TH>
__global__ __launch_bounds__(kTileThreads) void k_lbm2(Geom g, Lbm2Args a, StepConst<R> c, part::PartSteps<R> ps) {
  __shared__ TileLds<R, TH> L;
  if ((int)blockIdx.x < ps.nblocks) {
    part::run_part_steps(ps, blockIdx.x * kTileThreads + threadIdx.x);
    return;
  }
  const int b = (int)blockIdx.x - ps.nblocks;
  const int t = a.flip ? a.ntiles - 1 - b : b;
  const int ty = t / a.tiles_x, tx = t - ty * a.tiles_x;
  two_step_tile<R, TH>(g, a.p, c, tx, ty, L);
}

// Tile heights: 4, 8 or 16 rows (LDS 29.5 / 49.6 / 89.6 KB in FP32); FP64 at most 8 rows
// (99 KB; 16 rows would exceed the 160 KB of a CU).  0 picks the measured best: 8 rows in
// FP32 (3 workgroups per CU), 4 in FP64 (2 workgroups per CU).
int tile_rows(int prec, int tile_h) {
  if (tile_h == 0) return prec == RBN_FP64 ? 4 : 8;
  if (tile_h == 4) return 4;
  if (tile_h == 16 && prec != RBN_FP64) return 16;
  return 8;
}

void dispatch32(hipStream_t s, int th, const Geom &g, const Lbm2Args &a, const void *step_const,
                const PartJob *job) {
  const StepConst<float> &c = *(const StepConst<float> *)step_const;
  const part::PartSteps<float> ps = part::part_steps<float>(g, job, kTileThreads);
  const dim3 grid((unsigned)(a.ntiles + ps.nblocks)), block(kTileThreads);
  if (th == 4)
    hipLaunchKernelGGL((k_lbm2<float, 4>), grid, block, 0, s, g, a, c, ps);
  else if (th == 16)
    hipLaunchKernelGGL((k_lbm2<float, 16>), grid, block, 0, s, g, a, c, ps);
  else
    hipLaunchKernelGGL((k_lbm2<float, 8>), grid, block, 0, s, g, a, c, ps);
}

void dispatch64(hipStream_t s, int th, const Geom &g, const Lbm2Args &a, const void *step_const,
                const PartJob *job) {
  const StepConst<double> &c = *(const StepConst<double> *)step_const;
  const part::PartSteps<double> ps = part::part_steps<double>(g, job, kTileThreads);
  const dim3 grid((unsigned)(a.ntiles + ps.nblocks)), block(kTileThreads);
  if (th == 4)
    hipLaunchKernelGGL((k_lbm2<double, 4>), grid, block, 0, s, g, a, c, ps);
  else
    hipLaunchKernelGGL((k_lbm2<double, 8>), grid, block, 0, s, g, a, c, ps);
}

}  // namespace

hipError_t launch_lbm2(hipStream_t s, int prec, const Geom &g, const LbmPtrs &p, void *Fout, void *uv2,
                       void *rho2, void *T2, const void *step_const, int tile_h, int flip,
                       const PartJob *job) {
  const int th = tile_rows(prec, tile_h);
  Lbm2Args a;
  a.p.fA = p.fA;
  a.p.gA = p.gA;
  a.p.fB = p.fB;
  a.p.gB = p.gB;
  a.p.Fin = p.F;
  a.p.Fout = Fout;
  a.p.uv1 = p.uv;
  a.p.uv2 = uv2;
  a.p.rho2 = rho2;
  a.p.T2 = T2;
  a.tiles_x = (g.nx + kTW - 1) / kTW;
  a.ntiles = a.tiles_x * ((g.ny + th - 1) / th);
  a.flip = flip;
  if (prec == RBN_FP64)
    dispatch64(s, th, g, a, step_const, job);
  else
    dispatch32(s, th, g, a, step_const, job);
  return hipGetLastError();
}

}  // namespace rbn
