// rbn_lbm2.hip — two steps of K1 in one launch: truth-run steps by row sweeps over column
// strips (rbn_lbm2_wave.h) or by tiles (rbn_lbm2_tile.h), nudged-run steps by tiles, the first
// step's populations passed through LDS either way.  A step that must write its moments is
// never the first of a pair (the host launches it alone).  Particle updates of earlier steps
// ride in the launch as workgroups ahead of the lattice ones (rbn_lbm_fused.hip explains why
// that is safe).
#include "rbn_lbm2_tile.h"
#include "rbn_lbm2_wave.h"
#include "rbn_lbm3_wave.h"
#include "rbn_nudge_node.h"
#include "rbn_particle_node.h"

namespace rbn {
namespace {

using namespace lbm2;

// head_blocks > 0 (nudged run): workgroup k after the tiles updates observers
// k kTileThreads onward to a later step and bins their contributions (rbn_nudge_node.h).
struct Lbm2Args {
  TwoStepPtrs p;
  int tiles_x, ntiles, flip;
  int head_blocks, head_branch, head_ntempread;
  ObsPtrs head_obs;
  DepPtrs head_dep;
};

// FP32 truth-run tiles of 64 x 8 run at most three workgroups per CU: with four (40.4 KB of
// LDS each) the launch measured 3.6 % slower at 4096 x 2048 (0.361 against 0.349 ms with
// the benchmark's particles), so the pad takes a workgroup's LDS above a quarter of the CU's
// 160 KB.  Nudged-run tiles (47.6 KB) and FP64 tiles of 64 x 4 (43.6 KB) fit three per CU
// without it; FP64 gains from the third (0.556 against 0.594 ms with two).
template <typename R, int TH, int V>
struct KernelLds {
  TileLds<R, TH, ring_g<V>()> L;
  char cap[(sizeof(R) == 4 && TH == 8 && V != RBN_VARIANT_NUDGING) ? 8192 : 1];
};

template <typename R, int TH, int V, int NT>
__global__ __launch_bounds__(kTileThreads) void k_lbm2(Geom g, Lbm2Args a, StepConst<R> c, part::PartSteps<R> ps) {
  __shared__ KernelLds<R, TH, V> S;
  auto &L = S.L;
  if ((int)blockIdx.x < ps.nblocks) {
    part::run_part_steps(ps, blockIdx.x * kTileThreads + threadIdx.x);
    return;
  }
  const int b = (int)blockIdx.x - ps.nblocks;
  if constexpr (V == RBN_VARIANT_NUDGING) {
    if (b >= a.ntiles) {
      const int i = (b - a.ntiles) * kTileThreads + threadIdx.x;
      if (i < a.head_obs.n) {
        obs::obs_update_one(a.head_obs, i, a.head_branch, g.nx, g.ny, a.head_ntempread);
        obs::bin_one(g, a.head_dep, a.head_obs, i);
      }
      return;
    }
  }
  const int t = a.flip ? a.ntiles - 1 - b : b;
  const int ty = t / a.tiles_x, tx = t - ty * a.tiles_x;
  two_step_tile<R, TH, V, ring_g<V>(), NT>(g, a.p, c, tx, ty, L);
}

struct WaveArgs {
  TwoStepPtrs p;
  int strips, seg, nwg, flip, xcd;
};

// Dispatch index -> work index so that the workgroups one XCD receives (every eighth in
// dispatch order) take consecutive strips and segments, whose edge rows and columns they
// share in that XCD's L2.
__device__ __forceinline__ int xcd_order(int b, int n) {
  const int x = b & 7, k = b >> 3, per = n >> 3, rem = n & 7;
  return x * per + min(x, rem) + k;
}

template <typename R, int AHEAD, int SW>
__global__ __launch_bounds__(SW) void k_lbm2w(Geom g, WaveArgs a, StepConst<R> c, part::PartSteps<R> ps) {
  __shared__ WaveLds<R, SW> L;
  if ((int)blockIdx.x < ps.nblocks) {
    part::run_part_steps(ps, blockIdx.x * SW + threadIdx.x);
    return;
  }
  int b = (int)blockIdx.x - ps.nblocks;
  if (a.xcd) b = xcd_order(b, a.nwg);
  const int w = a.flip ? a.nwg - 1 - b : b;
  const int sy = w / a.strips, sx = w - sy * a.strips;
  const int ja = sy * a.seg;
  two_step_strip<R, AHEAD, SW>(g, a.p, c, sx, ja, min(ja + a.seg, g.ny), L);
}

constexpr int kStripW = 256;

struct Sweep3Args {
  lbm3::ThreeStepPtrs p;
  int strips, seg, nwg, flip, xcd;
};

// Threads of a three-step strip: 256 in FP32; 128 in FP64, whose LDS rows are twice as wide
// (49.9 KB at 128: three workgroups per CU, where 256 threads fit one).
template <typename R>
constexpr int strip_w3() { return sizeof(R) == 8 ? 128 : kStripW; }

// Three steps per launch (rbn_lbm3_wave.h): strip sx outputs columns [sx nx / strips,
// (sx + 1) nx / strips), so that every strip has nearly the same number of live threads.
// FP32 holds three waves per SIMD (as many as the LDS allows), FP64 two.
template <typename R, int SW, int NT>
__global__ __launch_bounds__(SW) __attribute__((amdgpu_waves_per_eu(sizeof(R) == 8 ? 2 : 3))) void k_lbm3w(Geom g, Sweep3Args a,
                                                                                    StepConst<R> c,
                                                                                    part::PartSteps<R> ps) {
  __shared__ lbm3::Lds3<R, SW> L;
  if ((int)blockIdx.x < ps.nblocks) {
    // 256 particles per workgroup whatever the strip width (FP64: two per thread), so that
    // the particle workgroups take as many slots as with FP32's 256-thread strips
    constexpr int PPT = 256 / SW;
#pragma unroll
    for (int k = 0; k < PPT; k++) part::run_part_steps(ps, ((int)blockIdx.x * PPT + k) * SW + threadIdx.x);
    return;
  }
  int b = (int)blockIdx.x - ps.nblocks;
  if (a.xcd) b = xcd_order(b, a.nwg);
  const int w = a.flip ? a.nwg - 1 - b : b;
  const int sy = w / a.strips, sx = w - sy * a.strips;
  const int ox0 = (int)((long long)sx * g.nx / a.strips), ox1 = (int)((long long)(sx + 1) * g.nx / a.strips);
  const int ja = sy * a.seg;
  lbm3::three_step_strip<R, SW, NT>(g, a.p, c, ox0, ox1, ja, min(ja + a.seg, g.ny), L);
}

template <typename R>
const void *lbm3_kernel() {
  return reinterpret_cast<const void *>(&k_lbm3w<R, strip_w3<R>(), 1>);
}

// Workgroups of k_lbm3w that fit on the device at once (resident on every CU), per device
// and kernel, measured once.
int lbm3_slots(int prec) {
  static int cache[64][2] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 0;
  int &slot = cache[dev][prec == RBN_FP64];
  if (slot > 0) return slot;
  const void *k = prec == RBN_FP64 ? lbm3_kernel<double>() : lbm3_kernel<float>();
  int per_cu = 0, cus = 0;
  const int sw = prec == RBN_FP64 ? strip_w3<double>() : strip_w3<float>();
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, sw, 0) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  slot = per_cu * cus;
  return slot;
}

template <typename R>
void dispatch_sweep3(hipStream_t s, const Geom &g, const Sweep3Args &a, int nt, const void *step_const,
                     const PartJob *job) {
  const StepConst<R> &c = *(const StepConst<R> *)step_const;
  constexpr int SW = strip_w3<R>();
  const part::PartSteps<R> ps = part::part_steps<R>(g, job, 256);
  const dim3 grid((unsigned)(a.nwg + ps.nblocks)), block(SW);
  if (nt)
    hipLaunchKernelGGL((k_lbm3w<R, SW, 1>), grid, block, 0, s, g, a, c, ps);
  else
    hipLaunchKernelGGL((k_lbm3w<R, SW, 0>), grid, block, 0, s, g, a, c, ps);
}

template <typename R>
void dispatch_wave(hipStream_t s, const Geom &g, const WaveArgs &a, int variant, const void *step_const,
                   const PartJob *job) {
  const StepConst<R> &c = *(const StepConst<R> *)step_const;
  const part::PartSteps<R> ps = part::part_steps<R>(g, job, kStripW);
  const dim3 grid((unsigned)(a.nwg + ps.nblocks)), block(kStripW);
  if (variant & 1)
    hipLaunchKernelGGL((k_lbm2w<R, 1, kStripW>), grid, block, 0, s, g, a, c, ps);
  else
    hipLaunchKernelGGL((k_lbm2w<R, 0, kStripW>), grid, block, 0, s, g, a, c, ps);
}

// Tile heights: 4, 8 or 16 rows in FP32; FP64 at most 8 rows (16 rows would exceed the 160 KB
// of a CU).  0 picks the measured best: 8 rows in FP32, 4 in FP64 (three workgroups per CU
// each).
int tile_rows_of(int prec, int tile_h) {
  if (tile_h == 0) return prec == RBN_FP64 ? 4 : 8;
  if (tile_h == 4) return 4;
  if (tile_h == 16 && prec != RBN_FP64) return 16;
  return 8;
}

template <int V, int NT>
void dispatch32(hipStream_t s, int th, const Geom &g, const Lbm2Args &a, const void *step_const,
                const PartJob *job) {
  const StepConst<float> &c = *(const StepConst<float> *)step_const;
  const part::PartSteps<float> ps = part::part_steps<float>(g, job, kTileThreads);
  const dim3 grid((unsigned)(a.ntiles + a.head_blocks + ps.nblocks)), block(kTileThreads);
  if (th == 4)
    hipLaunchKernelGGL((k_lbm2<float, 4, V, NT>), grid, block, 0, s, g, a, c, ps);
  else if (th == 16)
    hipLaunchKernelGGL((k_lbm2<float, 16, V, NT>), grid, block, 0, s, g, a, c, ps);
  else
    hipLaunchKernelGGL((k_lbm2<float, 8, V, NT>), grid, block, 0, s, g, a, c, ps);
}

template <int V, int NT>
void dispatch64(hipStream_t s, int th, const Geom &g, const Lbm2Args &a, const void *step_const,
                const PartJob *job) {
  const StepConst<double> &c = *(const StepConst<double> *)step_const;
  const part::PartSteps<double> ps = part::part_steps<double>(g, job, kTileThreads);
  const dim3 grid((unsigned)(a.ntiles + a.head_blocks + ps.nblocks)), block(kTileThreads);
  if (th == 4)
    hipLaunchKernelGGL((k_lbm2<double, 4, V, NT>), grid, block, 0, s, g, a, c, ps);
  else
    hipLaunchKernelGGL((k_lbm2<double, 8, V, NT>), grid, block, 0, s, g, a, c, ps);
}

template <int V>
void dispatch(hipStream_t s, int prec, int nt, int th, const Geom &g, const Lbm2Args &a, const void *step_const,
              const PartJob *job) {
  if (prec == RBN_FP64) {
    if (nt)
      dispatch64<V, 1>(s, th, g, a, step_const, job);
    else
      dispatch64<V, 0>(s, th, g, a, step_const, job);
  } else {
    if (nt)
      dispatch32<V, 1>(s, th, g, a, step_const, job);
    else
      dispatch32<V, 0>(s, th, g, a, step_const, job);
  }
}

}  // namespace

int lbm2_tile_rows(int prec, int tile_h) { return tile_rows_of(prec, tile_h); }

int lbm3_strips(int prec, int nx) {
  const int sw = prec == RBN_FP64 ? strip_w3<double>() : strip_w3<float>();
  return (nx + sw - 5) / (sw - 4);
}

int lbm3_segment_rows(int prec, const Geom &g, int seg) {
  if (seg > 0) return min(seg, g.ny);
  // 0: the fewest rows per segment that still put every workgroup on the device at once
  // (one wave of workgroups, no tail), at least 16
  const int slots = lbm3_slots(prec);
  const int strips = lbm3_strips(prec, g.nx);
  const int nseg = slots > 0 ? max(1, slots / strips) : 1;
  return min(g.ny, max(16, (g.ny + nseg - 1) / nseg));
}

hipError_t launch_lbm3(hipStream_t s, int prec, const Geom &g, const LbmPtrs &p, void *Fout, void *uv2, void *uv3,
                       void *rho3, void *T3, int mlev, const void *step_const, int seg, int xcd, int flip,
                       int nt, const PartJob *job) {
  Sweep3Args a{};
  a.p.fA = p.fA;
  a.p.gA = p.gA;
  a.p.fB = p.fB;
  a.p.gB = p.gB;
  a.p.Fin = p.F;
  a.p.Fout = Fout;
  a.p.uv[0] = p.uv;
  a.p.uv[1] = uv2;
  a.p.uv[2] = uv3;
  a.p.rho = rho3;
  a.p.T = T3;
  a.p.mlev = rho3 && T3 ? mlev : -1;
  a.p.bad = p.bad;
  a.seg = lbm3_segment_rows(prec, g, seg);
  a.strips = lbm3_strips(prec, g.nx);
  a.nwg = a.strips * ((g.ny + a.seg - 1) / a.seg);
  a.flip = flip;
  a.xcd = xcd;
  if (prec == RBN_FP64)
    dispatch_sweep3<double>(s, g, a, nt, step_const, job);
  else
    dispatch_sweep3<float>(s, g, a, nt, step_const, job);
  return hipGetLastError();
}

hipError_t launch_lbm2(hipStream_t s, int prec, const Geom &g, const LbmPtrs &p, void *Fout, void *uv2,
                       void *rho2, void *T2, const void *step_const, int tile_h, int seg, int xcd, int variant,
                       int flip, int nt, const PartJob *job, const NudgePair *np) {
  if (seg > 0 && !np) {
    WaveArgs w{};
    w.p.fA = p.fA;
    w.p.gA = p.gA;
    w.p.fB = p.fB;
    w.p.gB = p.gB;
    w.p.Fin = p.F;
    w.p.Fout = Fout;
    w.p.uv1 = p.uv;
    w.p.uv2 = uv2;
    w.p.rho2 = rho2;
    w.p.T2 = T2;
    w.p.bad = p.bad;
    w.seg = min(seg, kMaxSeg);
    w.strips = (g.nx + kStripW - 1) / kStripW;
    w.nwg = w.strips * ((g.ny + w.seg - 1) / w.seg);
    w.flip = flip;
    w.xcd = xcd;
    if (prec == RBN_FP64)
      dispatch_wave<double>(s, g, w, variant, step_const, job);
    else
      dispatch_wave<float>(s, g, w, variant, step_const, job);
    return hipGetLastError();
  }
  const int th = tile_rows_of(prec, tile_h);
  Lbm2Args a{};
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
  a.p.bad = p.bad;
  if (np) {
    a.p.tw[0] = p.tw;
    a.p.nud_in = p.nud;
    a.p.tw[1] = np->tw2;
    a.p.nud_out = np->nud_out;
    if (np->bins2) {
      if (th != kDepTileH || np->bins2->mode != RBN_DEPOSIT_BILINEAR) return hipErrorInvalidValue;
      a.p.bin2 = np->bins2->bin;
      a.p.tcnt2 = np->bins2->tcnt;
      a.p.bin_cap2 = np->bins2->bin_cap;
      a.p.S2 = np->bins2->S;
      a.p.Wi2 = np->bins2->Wi;
      a.p.ocnt2 = np->bins2->ocnt;
    }
    if (np->head_obs && np->head_dep) {
      if (np->head_dep->mode != RBN_DEPOSIT_BILINEAR) return hipErrorInvalidValue;
      a.head_obs = *np->head_obs;
      a.head_dep = *np->head_dep;
      a.head_branch = np->head_branch;
      a.head_ntempread = np->head_ntempread;
      a.head_blocks = (np->head_obs->n + kTileThreads - 1) / kTileThreads;
    }
  }
  a.tiles_x = (g.nx + kTW - 1) / kTW;
  a.ntiles = a.tiles_x * ((g.ny + th - 1) / th);
  a.flip = flip;
  if (np)
    dispatch<RBN_VARIANT_NUDGING>(s, prec, nt, th, g, a, step_const, job);
  else
    dispatch<RBN_VARIANT_TRAJECTORIES>(s, prec, nt, th, g, a, step_const, job);
  return hipGetLastError();
}

}  // namespace rbn
