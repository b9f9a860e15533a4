// rbn_particles.hip — K2: Lagrangian particles of the truth run.
//
// One thread advances one particle through the reference's per-step sequence
// (Trajectories/code/mympi2.h):
//   bilinear interpolation of the fluid velocity (bilinear, :292-311), with the periodic
//   ghost values of communication (:313-348) resolved as an index wrap in x and in y;
//   tracer: second-order Adams–Bashforth position update (:358-372);
//   inertial: the Maxey–Riley velocity relaxation without gravity (:373-433), whose node
//   fields (time difference, centred x/y differences, one-sided at the walls, :384-405)
//   are evaluated on the fly from u, v and the previous step's u, v instead of being
//   stored as six full fields;
//   on record steps the trajectory-record quantities (prtcl_write,
//   Trajectories/code/temppart_dump.c:1467-1509, and the Writefull block, :441-501), taken
//   after the update and before the wrap;
//   the periodic-x wrap and reflecting-y rule of toggle_prtcl (:568-760).
// Particle state is FP64 in both precision modes.  With FP64 fields the arithmetic is
// operand-for-operand the reference's (compiled with -ffp-contract=off).
#include "rbn_internal.h"

namespace rbn {
namespace {

constexpr int kBlock = 256;

template <typename R>
struct Field {
  const R *__restrict__ a;
  int nx, ny, pitch;
  // Value at halo coordinates ic in [0, nx+1], jc in [0, ny+1]: ghosts are periodic copies
  // (communication on one rank, periodic in both directions).
  __device__ __forceinline__ double at(int ic, int jc) const {
    const int i = ic == 0 ? nx : (ic == nx + 1 ? 1 : ic);
    const int j = jc == 0 ? ny : (jc == ny + 1 ? 1 : jc);
    return (double)a[(size_t)(j - 1) * pitch + (i - 1)];
  }
};

// Lower corner of the interpolation cell (C truncation, M:293-296) and whether all four
// corners lie on the halo-extended lattice, i.e. x in (-1, nx+1), y in (-1, ny+1).  The
// position is range-checked before the int conversion, so NaN or infinite positions are
// rejected instead of saturating into range.
__device__ __forceinline__ bool corners(double xx, double yy, int nx, int ny, int &x1, int &y1) {
  if (!(xx > -1.0 && xx < (double)nx + 1.0 && yy > -1.0 && yy < (double)ny + 1.0)) {
    x1 = y1 = 0;
    return false;
  }
  x1 = (int)xx;
  y1 = (int)yy;
  return true;
}

// bilinear (M:292-311) from four corner values.
__device__ __forceinline__ double blend(double u11, double u12, double u21, double u22, int x1,
                                        int y1, double xx, double yy) {
  const int x2 = x1 + 1, y2 = y1 + 1;
  double val1 = u11 * ((double)x2 - xx) * ((double)y2 - yy);
  double val2 = u21 * (xx - (double)x1) * ((double)y2 - yy);
  double val3 = u12 * ((double)x2 - xx) * (yy - (double)y1);
  double val4 = u22 * (xx - (double)x1) * (yy - (double)y1);
  return val1 + val2 + val3 + val4;
}

template <typename R>
__device__ __forceinline__ double interp(const Field<R> &F, double xx, double yy, int &err) {
  int x1, y1;
  if (!corners(xx, yy, F.nx, F.ny, x1, y1)) {
    err = 1;
    return 0.0;
  }
  return blend(F.at(x1, y1), F.at(x1, y1 + 1), F.at(x1 + 1, y1), F.at(x1 + 1, y1 + 1), x1, y1,
               xx, yy);
}

// The six Maxey–Riley node fields at halo node (ic, jc) (M:374-411): ghost nodes take the
// value of the periodic interior node they copy.
template <typename R>
__device__ __forceinline__ void node_derivs(const Field<R> &U, const Field<R> &V,
                                            const Field<R> &Up, const Field<R> &Vp, int ic,
                                            int jc, double d[6]) {
  const int nx = U.nx, ny = U.ny;
  const int I = ic == 0 ? nx : (ic == nx + 1 ? 1 : ic);
  const int J = jc == 0 ? ny : (jc == ny + 1 ? 1 : jc);
  const double u0 = U.at(I, J), v0 = V.at(I, J);
  d[0] = u0 - Up.at(I, J);                        // dtu
  d[1] = v0 - Vp.at(I, J);                        // dtv
  d[2] = (U.at(I + 1, J) - U.at(I - 1, J)) * 0.5;  // dxu
  d[3] = (V.at(I + 1, J) - V.at(I - 1, J)) * 0.5;  // dxv
  if (J == 1) {
    d[4] = (U.at(I, 2) - u0);
    d[5] = (V.at(I, 2) - v0);
  } else if (J == ny) {
    d[4] = (u0 - U.at(I, ny - 1));
    d[5] = (v0 - V.at(I, ny - 1));
  } else {
    d[4] = (U.at(I, J + 1) - U.at(I, J - 1)) * 0.5;
    d[5] = (V.at(I, J + 1) - V.at(I, J - 1)) * 0.5;
  }
}

// Interpolated dtu, dtv, dxu, dxv, dyu, dyv at (xx, yy).
template <typename R>
__device__ __forceinline__ void interp_derivs(const Field<R> &U, const Field<R> &V,
                                              const Field<R> &Up, const Field<R> &Vp,
                                              double xx, double yy, double out[6], int &err) {
  int x1, y1;
  if (!corners(xx, yy, U.nx, U.ny, x1, y1)) {
    err = 1;
    for (int q = 0; q < 6; q++) out[q] = 0.0;
    return;
  }
  double d11[6], d12[6], d21[6], d22[6];
  node_derivs(U, V, Up, Vp, x1, y1, d11);
  node_derivs(U, V, Up, Vp, x1, y1 + 1, d12);
  node_derivs(U, V, Up, Vp, x1 + 1, y1, d21);
  node_derivs(U, V, Up, Vp, x1 + 1, y1 + 1, d22);
#pragma unroll
  for (int q = 0; q < 6; q++) out[q] = blend(d11[q], d12[q], d21[q], d22[q], x1, y1, xx, yy);
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_particles(Field<R> U, Field<R> V, Field<R> Up,
                                                      Field<R> Vp, Field<R> T, PartPtrs pp,
                                                      int record, int full) {
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= pp.n) return;
  const double delta_t = 1.0;
  int err = 0;
  const double xx = pp.x[p], yy = pp.y[p];
  const double uo = pp.u[p], vo = pp.v[p];
  double pu = uo, pv = vo;
  const double up = interp(U, xx, yy, err);
  const double vp = interp(V, xx, yy, err);
  const double onebytau = pp.inv_tau[p];
  const bool heavy = onebytau > 0.0;
  double D[6];
  double x, y;
  if (heavy) {
    interp_derivs(U, V, Up, Vp, xx, yy, D, err);
    const double beta = pp.beta[p];
    pu = pu + delta_t * (beta * (D[0] + up * D[2] + vp * D[4]) - onebytau * (pu - up));
    pv = pv + delta_t * (beta * (D[1] + up * D[3] + vp * D[5]) - onebytau * (pv - vp));
    x = xx + delta_t * pu;
    y = yy + delta_t * pv;
  } else {
    x = xx + 1.5 * delta_t * up - 0.5 * delta_t * uo;
    y = yy + 1.5 * delta_t * vp - 0.5 * delta_t * vo;
    pu = up;
    pv = vp;
  }
  if (record && pp.rec) {
    const size_t n = (size_t)pp.n;
    double *r = pp.rec;
    r[0 * n + p] = x;
    r[1 * n + p] = y;
    r[2 * n + p] = pu;
    r[3 * n + p] = pv;
    r[4 * n + p] = xx;
    r[5 * n + p] = yy;
    r[6 * n + p] = uo;
    r[7 * n + p] = vo;
    r[8 * n + p] = interp(T, x, y, err);
    if (full) {
      // Writefull quantities at (xo, yo) = (xx, yy): the same interpolants as the update.
      if (!heavy) interp_derivs(U, V, Up, Vp, xx, yy, D, err);
#pragma unroll
      for (int q = 0; q < 6; q++) r[(9 + q) * n + p] = D[q];
      r[15 * n + p] = up;
      r[16 * n + p] = vp;
    }
  }
  // toggle_prtcl on one rank (M:598-635, 694-717).
  const double NXd = (double)U.nx, NYd = (double)U.ny;
  if (x >= NXd) {
    x = x - NXd;
  } else if (x <= 0.0) {
    x = x + NXd;
  }
  if (y < 1.0) {
    y = y + 2.0 * (1.0 - y);
    pv = -pv;
  } else if (y >= NYd) {
    y = y + 2.0 * (NYd - y);
    pv = -pv;
  }
  pp.x[p] = x;
  pp.y[p] = y;
  pp.u[p] = pu;
  pp.v[p] = pv;
  pp.xo[p] = xx;
  pp.yo[p] = yy;
  pp.uo[p] = uo;
  pp.vo[p] = vo;
  if (err) atomicOr(pp.err, 1);
}

// Initial particle velocity from the fluid (init_prtcl, M:181-208).
template <typename R>
__global__ void k_particle_init(Field<R> U, Field<R> V, PartPtrs pp, double factor) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= pp.n) return;
  int err = 0;
  const double xx = pp.x[p], yy = pp.y[p];
  const double tu = interp(U, xx, yy, err), tv = interp(V, xx, yy, err);
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
Field<R> mk(const void *a, const Geom &g) {
  return Field<R>{(const R *)a, g.nx, g.ny, g.pitch};
}

}  // namespace

hipError_t launch_particles(hipStream_t s, int prec, const Geom &g, const void *u, const void *v,
                            const void *up, const void *vp, const void *T, const PartPtrs &pp,
                            int need_deriv, int record, int full) {
  (void)need_deriv;
  if (pp.n <= 0) return hipSuccess;
  dim3 grid((pp.n + kBlock - 1) / kBlock), block(kBlock);
  if (prec == RBN_FP64)
    hipLaunchKernelGGL(k_particles<double>, grid, block, 0, s, mk<double>(u, g), mk<double>(v, g),
                       mk<double>(up, g), mk<double>(vp, g), mk<double>(T, g), pp, record, full);
  else
    hipLaunchKernelGGL(k_particles<float>, grid, block, 0, s, mk<float>(u, g), mk<float>(v, g),
                       mk<float>(up, g), mk<float>(vp, g), mk<float>(T, g), pp, record, full);
  return hipGetLastError();
}

hipError_t launch_sample(hipStream_t s, int prec, const Geom &g, const void *field, int n,
                         const double *x, const double *y, double *out, int *err) {
  if (n <= 0) return hipSuccess;
  dim3 grid((n + kBlock - 1) / kBlock), block(kBlock);
  if (prec == RBN_FP64)
    hipLaunchKernelGGL(k_sample<double>, grid, block, 0, s, mk<double>(field, g), n, x, y, out, err);
  else
    hipLaunchKernelGGL(k_sample<float>, grid, block, 0, s, mk<float>(field, g), n, x, y, out, err);
  return hipGetLastError();
}

hipError_t launch_particle_init(hipStream_t s, int prec, const Geom &g, const void *u,
                                const void *v, const PartPtrs &pp, double heavy_factor) {
  if (pp.n <= 0) return hipSuccess;
  dim3 grid((pp.n + kBlock - 1) / kBlock), block(kBlock);
  if (prec == RBN_FP64)
    hipLaunchKernelGGL(k_particle_init<double>, grid, block, 0, s, mk<double>(u, g),
                       mk<double>(v, g), pp, heavy_factor);
  else
    hipLaunchKernelGGL(k_particle_init<float>, grid, block, 0, s, mk<float>(u, g),
                       mk<float>(v, g), pp, heavy_factor);
  return hipGetLastError();
}

}  // namespace rbn
