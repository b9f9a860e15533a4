// rbn_particles.hip — the plain launch of K2 (one thread per particle; device code in
// rbn_particle_node.h), the initial particle velocity and the bilinear sampling kernel.
#include "rbn_particle_node.h"

namespace rbn {
namespace {

using namespace part;

constexpr int kBlock = 256;

template <typename R>
__global__ __launch_bounds__(kBlock) void k_particles(VelField<R> UV, VelField<R> UVp, Field<R> T,
                                                      PartPtrs pp, int record, int full) {
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p < pp.n) particle_update(UV, UVp, T, pp, p, record, full);
}

// Initial particle velocity from the fluid (init_prtcl, M:181-208).
template <typename R>
__global__ void k_particle_init(VelField<R> UV, PartPtrs pp, double factor) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= pp.n) return;
  int err = 0;
  const double xx = pp.x[p], yy = pp.y[p];
  double tu, tv;
  interp_uv(UV, xx, yy, tu, tv, err);
  double pu = tu, pv = tv;
  if (pp.inv_tau[p] > 0.0) {
    pu = factor * tu;
    pv = factor * tv;
  }
  pp.u[p] = pu;
  pp.v[p] = pv;
  pp.uo[p] = pu;
  pp.vo[p] = pv;
  pp.xo[p] = xx;
  pp.yo[p] = yy;
  if (err) atomicOr(pp.err, 1);
}

// Bilinear samples of one field at given positions (observation operator / init records).
template <typename R>
__global__ void k_sample(Field<R> F, int n, const double *x, const double *y, double *out, int *errp) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  int err = 0;
  out[p] = interp(F, x[p], y[p], err);
  if (err) atomicOr(errp, 1);
}

template <typename R>
Field<R> mk(const void *a, const Geom &g, int cs = 1) {
  return Field<R>{(const R *)a, g.nx, g.ny, g.pitch, cs};
}
template <typename R>
VelField<R> mk_uv(const void *uv, const Geom &g) {
  return vel_of<R>(uv, g.nx, g.ny, g.pitch);
}

}  // namespace

hipError_t launch_particles(hipStream_t s, int prec, const Geom &g, const void *uv, const void *uvp,
                            const void *T, const PartPtrs &pp, int record, int full) {
  if (pp.n <= 0) return hipSuccess;
  dim3 grid((pp.n + kBlock - 1) / kBlock), block(kBlock);
  if (prec == RBN_FP64)
    hipLaunchKernelGGL(k_particles<double>, grid, block, 0, s, mk_uv<double>(uv, g), mk_uv<double>(uvp, g),
                       mk<double>(T, g), pp, record, full);
  else
    hipLaunchKernelGGL(k_particles<float>, grid, block, 0, s, mk_uv<float>(uv, g), mk_uv<float>(uvp, g),
                       mk<float>(T, g), pp, record, full);
  return hipGetLastError();
}

hipError_t launch_sample(hipStream_t s, int prec, const Geom &g, const void *field, int cs, int n,
                         const double *x, const double *y, double *out, int *err) {
  if (n <= 0) return hipSuccess;
  dim3 grid((n + kBlock - 1) / kBlock), block(kBlock);
  if (prec == RBN_FP64)
    hipLaunchKernelGGL(k_sample<double>, grid, block, 0, s, mk<double>(field, g, cs), n, x, y, out, err);
  else
    hipLaunchKernelGGL(k_sample<float>, grid, block, 0, s, mk<float>(field, g, cs), n, x, y, out, err);
  return hipGetLastError();
}

hipError_t launch_particle_init(hipStream_t s, int prec, const Geom &g, const void *uv,
                                const PartPtrs &pp, double heavy_factor) {
  if (pp.n <= 0) return hipSuccess;
  dim3 grid((pp.n + kBlock - 1) / kBlock), block(kBlock);
  if (prec == RBN_FP64)
    hipLaunchKernelGGL(k_particle_init<double>, grid, block, 0, s, mk_uv<double>(uv, g), pp,
                       heavy_factor);
  else
    hipLaunchKernelGGL(k_particle_init<float>, grid, block, 0, s, mk_uv<float>(uv, g), pp,
                       heavy_factor);
  return hipGetLastError();
}

}  // namespace rbn
