// rbn_lbm.hip — the plain launch of K1 (one thread per node, device code in
// rbn_lbm_node.h), the initial-equilibrium kernel and the deterministic energy reduction.
#include "rbn_lbm_node.h"

namespace rbn {
namespace {

using namespace lbm;

constexpr int kBlock = 256;     // block width of the init kernels
constexpr int kMaxBlock = 512;  // K1 block widths 64..512 (tuning)

template <int VARIANT, int NT>
__global__ __launch_bounds__(kMaxBlock) void k_lbm64(Geom g, LbmPtrs p, StepConst<double> c, int order,
                                                     int npt) {
  int i0, j;
  segment_node(g, blockIdx.x, gridDim.x, order, npt, i0, j);
  for (int k = 0; k < npt; k++) {
    const int i = i0 + k * (int)blockDim.x;
    if (i < g.nx) lbm_node64<VARIANT, NT>(g, p, c, i, j);
  }
}

template <int VARIANT, int NT>
__global__ __launch_bounds__(kMaxBlock) void k_lbm32(Geom g, LbmPtrs p, StepConst<float> c, int order,
                                                     int npt) {
  int i0, j;
  segment_node(g, blockIdx.x, gridDim.x, order, npt, i0, j);
  for (int k = 0; k < npt; k++) {
    const int i = i0 + k * (int)blockDim.x;
    if (i < g.nx) lbm_node32<VARIANT, NT>(g, p, c, i, j);
  }
}

// ---------------------------------------------------------------------------- init

// f = feq(rho, u, v), g = geq(T, u, v) (inithydro + initpop, T:1643-1648) and the force of
// the first step from the initial moments.
__global__ void k_init64(Geom g, const double *rho, const double *T, const double *uv, double *f,
                         double *gp, double *F, double g0) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= g.nx) return;
  const size_t o = (size_t)j * g.pitch + i;
  double b[9];
  eq_brackets64(uv[2 * o], uv[2 * o + 1], b);
#pragma unroll
  for (int k = 0; k < 9; k++) {
    f[(size_t)k * g.plane + o] = rho[o] * kW[k] * b[k];
    gp[(size_t)k * g.plane + o] = T[o] * kW[k] * b[k];
  }
  F[o] = g0 * rho[o] * (T[o] - 0.0);
}

__global__ void k_init32(Geom g, const float *rho, const float *T, const float *uv, float *f,
                         float *gp, float *F, float g0, float rho0, float T0) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= g.nx) return;
  const size_t o = (size_t)j * g.pitch + i;
  float A[9];
  eq_dev32(uv[2 * o], uv[2 * o + 1], A);
  const float r = rho[o], t = T[o];
  float feq[9], geq[9];
  eq_pops32(r - rho0, r, A, feq);
  eq_pops32(t - T0, t, A, geq);
#pragma unroll
  for (int k = 0; k < 9; k++) {
    f[(size_t)k * g.plane + o] = feq[k];
    gp[(size_t)k * g.plane + o] = geq[k];
  }
  F[o] = g0 * r * t;
}

// ---------------------------------------------------------------------------- energy

constexpr int kRedBlock = 256;

// Per-block partial sums of 0.5*rho*(u^2+v^2) and T over interior rows (evolution.dat,
// T:1739-1757), accumulated in FP64 in a fixed order: the result is deterministic.
template <typename R>
__global__ __launch_bounds__(kRedBlock) void k_energy_partial(Geom g, const R *rho, const R *uv,
                                                              const R *T, double *partial) {
  __shared__ double se[kRedBlock], st[kRedBlock];
  const long long n = (long long)g.nx * g.ny;
  double e = 0.0, t = 0.0;
  for (long long idx = (long long)blockIdx.x * kRedBlock + threadIdx.x; idx < n;
       idx += (long long)gridDim.x * kRedBlock) {
    const int j = (int)(idx / g.nx), i = (int)(idx - (long long)j * g.nx);
    const size_t o = (size_t)j * g.pitch + i;
    const double r = rho[o], uu = uv[2 * o], vv = uv[2 * o + 1];
    e += 0.5 * r * (uu * uu + vv * vv);
    t += (double)T[o];
  }
  se[threadIdx.x] = e;
  st[threadIdx.x] = t;
  __syncthreads();
  for (int s = kRedBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      se[threadIdx.x] += se[threadIdx.x + s];
      st[threadIdx.x] += st[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = se[0];
    partial[2 * blockIdx.x + 1] = st[0];
  }
}

__global__ void k_energy_final(const double *partial, int nb, double nxny, double *out2) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double e = 0.0, t = 0.0;
  for (int b = 0; b < nb; b++) {
    e += partial[2 * b];
    t += partial[2 * b + 1];
  }
  out2[0] = e / nxny;
  out2[1] = t / nxny;
}

// ---------------------------------------------------------------------------- profiles

// One block per row j: x averages of rho, u, v, u^2, v^2, T, vT, dzT, vorticity^2
// (profile, N:1104-1250).  dzT = T(j+1) - T(j) with the top row reading row 1 (the periodic
// ghost row); vorticity = dv/dx - du/dy with periodic x and one-sided y differences at the
// walls.  Each thread sums a fixed column subset in order, then a fixed tree over the block:
// the result does not depend on scheduling.
template <typename R>
__global__ __launch_bounds__(kRedBlock) void k_profiles(Geom g, const R *rho, const R *T, const R *uv,
                                                        double *out) {
  __shared__ double red[9][kRedBlock];
  const int j = blockIdx.x;
  const size_t r0 = (size_t)j * g.pitch;
  const size_t ru = (size_t)(j == g.ny - 1 ? 0 : j + 1) * g.pitch;
  const int jn = j == g.ny - 1 ? j : j + 1, js = j == 0 ? 0 : j - 1;
  const double yscale = (j == 0 || j == g.ny - 1) ? 1.0 : 0.5;
  double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int i = threadIdx.x; i < g.nx; i += kRedBlock) {
    const size_t o = r0 + i;
    const int il = i == 0 ? g.nx - 1 : i - 1, ir = i == g.nx - 1 ? 0 : i + 1;
    const double u = uv[2 * o], v = uv[2 * o + 1], t = T[o];
    const double dvbydx = 0.5 * ((double)uv[2 * (r0 + ir) + 1] - (double)uv[2 * (r0 + il) + 1]);
    const double dubydy = yscale * ((double)uv[2 * ((size_t)jn * g.pitch + i)] - (double)uv[2 * ((size_t)js * g.pitch + i)]);
    const double vor = dvbydx - dubydy;
    acc[0] += (double)rho[o];
    acc[1] += u;
    acc[2] += v;
    acc[3] += u * u;
    acc[4] += v * v;
    acc[5] += t;
    acc[6] += v * t;
    acc[7] += (double)T[ru + i] - t;
    acc[8] += vor * vor;
  }
  for (int q = 0; q < 9; q++) red[q][threadIdx.x] = acc[q];
  __syncthreads();
  for (int s = kRedBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int q = 0; q < 9; q++) red[q][threadIdx.x] += red[q][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < 9) out[(size_t)j * 9 + threadIdx.x] = red[threadIdx.x][0] * (1 / (double)g.nx);
}

}  // namespace

namespace {
template <int V, int NT>
void lbm_dispatch(hipStream_t s, int prec, dim3 grid, dim3 block, const Geom &g, const LbmPtrs &p,
                  const void *step_const, int order, int npt) {
  if (prec == RBN_FP64)
    hipLaunchKernelGGL((k_lbm64<V, NT>), grid, block, 0, s, g, p,
                       *(const StepConst<double> *)step_const, order, npt);
  else
    hipLaunchKernelGGL((k_lbm32<V, NT>), grid, block, 0, s, g, p,
                       *(const StepConst<float> *)step_const, order, npt);
}

template <int V>
void lbm_dispatch_nt(int nt, hipStream_t s, int prec, dim3 grid, dim3 block, const Geom &g,
                     const LbmPtrs &p, const void *step_const, int order, int npt) {
  // Bit 2 only changes the velocity store when bit 0 is set (4 and 6 are 0 and 2).
  switch (nt & 7) {
    case 0: case 4: lbm_dispatch<V, 0>(s, prec, grid, block, g, p, step_const, order, npt); break;
    case 1: lbm_dispatch<V, 1>(s, prec, grid, block, g, p, step_const, order, npt); break;
    case 2: case 6: lbm_dispatch<V, 2>(s, prec, grid, block, g, p, step_const, order, npt); break;
    case 3: lbm_dispatch<V, 3>(s, prec, grid, block, g, p, step_const, order, npt); break;
    case 5: lbm_dispatch<V, 5>(s, prec, grid, block, g, p, step_const, order, npt); break;
    default: lbm_dispatch<V, 7>(s, prec, grid, block, g, p, step_const, order, npt); break;
  }
}
}  // namespace

hipError_t launch_lbm(hipStream_t s, int prec, int variant, const Geom &g, const LbmPtrs &p,
                      const void *step_const, const LbmTuning &t) {
  const int bw = (t.block == 64 || t.block == 128 || t.block == 512) ? t.block : 256;
  const int npt = (t.npt == 2 || t.npt == 4 || t.npt == 8) ? t.npt : 1;
  const unsigned gx = (unsigned)((g.nx + bw * npt - 1) / (bw * npt));
  dim3 grid(gx * (unsigned)g.ny), block(bw);
  const int nt = (t.nt ? 1 : 0) | (t.ntload ? 2 : 0) | (t.uv_keep ? 4 : 0);
  const int order = (t.xcd == 1 ? 1 : 0) | (t.flip ? 2 : 0) | (t.xcd == 2 ? 4 : 0);
  if (variant == RBN_VARIANT_NUDGING)
    lbm_dispatch_nt<RBN_VARIANT_NUDGING>(nt, s, prec, grid, block, g, p, step_const, order, npt);
  else
    lbm_dispatch_nt<RBN_VARIANT_TRAJECTORIES>(nt, s, prec, grid, block, g, p, step_const, order, npt);
  return hipGetLastError();
}

hipError_t launch_init_eq(hipStream_t s, int prec, const Geom &g, const void *rho, const void *T,
                          const void *uv, void *f, void *gp, void *F, const void *step_const) {
  dim3 grid((g.nx + kBlock - 1) / kBlock, g.ny), block(kBlock);
  if (prec == RBN_FP64) {
    const StepConst<double> &c = *(const StepConst<double> *)step_const;
    hipLaunchKernelGGL(k_init64, grid, block, 0, s, g, (const double *)rho, (const double *)T,
                       (const double *)uv, (double *)f, (double *)gp, (double *)F, c.g0);
  } else {
    const StepConst<float> &c = *(const StepConst<float> *)step_const;
    hipLaunchKernelGGL(k_init32, grid, block, 0, s, g, (const float *)rho, (const float *)T,
                       (const float *)uv, (float *)f, (float *)gp, (float *)F, c.g0, c.rho0, c.T0);
  }
  return hipGetLastError();
}

hipError_t launch_profiles(hipStream_t s, int prec, const Geom &g, const void *rho, const void *T,
                           const void *uv, double *out) {
  if (prec == RBN_FP64)
    hipLaunchKernelGGL(k_profiles<double>, dim3(g.ny), dim3(kRedBlock), 0, s, g, (const double *)rho,
                       (const double *)T, (const double *)uv, out);
  else
    hipLaunchKernelGGL(k_profiles<float>, dim3(g.ny), dim3(kRedBlock), 0, s, g, (const float *)rho,
                       (const float *)T, (const float *)uv, out);
  return hipGetLastError();
}

int energy_blocks(const Geom &g) {
  long long n = (long long)g.nx * g.ny;
  long long nb = (n + kRedBlock - 1) / kRedBlock;
  return (int)(nb < 1024 ? nb : 1024);
}

hipError_t launch_energy(hipStream_t s, int prec, const Geom &g, const void *rho, const void *uv,
                         const void *T, double *partial, int nblocks, double *out2) {
  if (prec == RBN_FP64)
    hipLaunchKernelGGL(k_energy_partial<double>, dim3(nblocks), dim3(kRedBlock), 0, s, g,
                       (const double *)rho, (const double *)uv, (const double *)T, partial);
  else
    hipLaunchKernelGGL(k_energy_partial<float>, dim3(nblocks), dim3(kRedBlock), 0, s, g,
                       (const float *)rho, (const float *)uv, (const float *)T, partial);
  hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(64), 0, s, partial, nblocks,
                     (double)g.nx * (double)g.ny, out2);
  return hipGetLastError();
}

}  // namespace rbn
