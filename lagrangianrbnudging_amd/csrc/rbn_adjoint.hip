// rbn_adjoint.hip — hand-written adjoint of the nudged lattice step and of the bilinear
// observation operator.  No reference counterpart: the reference has no gradient.
//
// Forward map of step s at node n of the nudged run (F_x = 0, zero-gradient g walls; the
// deposit counts nv, means Tbar and weights W -- 1 for the 7x7 deposit -- are functions of
// the observation records only):
//   pull f, g                                    rho = sum f_k,  S = sum g_k
//   T = S + [nv>0] nud_in / 2                    u = j_x / rho,  v = (j_y + F_y/2) / rho
//   F_out = g0 rho T                             Q = alpha W (T - Tbar)
//   f'_k = (1-om) f_k + om w_k rho b_k + src_k fac_k
//   g'_k = (1-omg) g_k + omg w_k T b_k - [nv>0] bw_k Q
//   nud_out = [nv>0] ? -Q : nud_in
// with b_k = 1 + 3 c.u + 4.5 (c.u)^2 - 1.5 |u|^2 and fac_k = 3 (c_yk - v) F_y + 9 (c.u) c_yk F_y.
//
// One launch per step in gather form, fused with the reverse of the next step's streaming:
// each thread owns one node, gathers the adjoints of its post-collision populations from
// the pulled-population adjoints of step s+1 through the reverse of the pull map
// (half-way bounce-back for f; for g the zero-gradient wall copies, which make rows 1 and
// ny-2 readable by their wall row and, for the populations leaving the wall, by their own
// row), recomputes its forward locals from the recomputed state entering step s, and
// applies the transposed Jacobian of the map above.  Local arithmetic is FP64 in both
// storage precisions; dJ/dalpha is reduced per block in FP64 and summed in a fixed order.
//
// Misfit sources enter as a fixed-point (2^-k scaled int64) plane so that the scatter of
// 2 r_k times the bilinear weights is order independent and the gradient is reproducible
// bit for bit.
#include "rbn_adjoint.h"
#include "rbn_particle_node.h"

namespace rbn {
namespace {

constexpr int kT = 256;
constexpr int CX[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
constexpr int CY[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
constexpr int OPP[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};

__device__ __forceinline__ int wrapx(int i, int nx) { return i < 0 ? i + nx : (i >= nx ? i - nx : i); }

// Fixed-point scale of one observation: |sum at one node| <= 2 max|r| n < 2^62 after scaling.
__device__ __forceinline__ double fixed_scale(unsigned long long maxbits, int n) {
  const double bound = __longlong_as_double((long long)maxbits) * 2.0 * (double)(n > 0 ? n : 1);
  if (!(bound > 0.0)) return 1.0;
  int e = 0;
  (void)frexp(bound, &e);  // bound < 2^e
  return ldexp(1.0, 62 - e);
}

// Deterministic block sum (fixed shuffle tree within each 64-wide wavefront, then over the
// wavefronts in order); the result is valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double *lds) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) lds[wave] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) s += lds[w];
  return s;
}

template <typename R>
__global__ __launch_bounds__(kT) void k_adjoint(const AdjStepArgs a) {
  __shared__ double lds[kT / 64];
  const Geom g = a.g;
  const int i = blockIdx.x * kT + threadIdx.x;
  const int j = blockIdx.y;
  double dal = 0.0;
  if (i < g.nx) {
    const size_t P = g.plane;
    const size_t r0 = (size_t)j * g.pitch, o = r0 + i;
    const int im = i == 0 ? g.nx - 1 : i - 1, ip = i == g.nx - 1 ? 0 : i + 1;
    const bool bot = j == 0, top = j == g.ny - 1;
    const size_t rm = bot ? r0 : r0 - g.pitch, rp = top ? r0 : r0 + g.pitch;
    const R *fA = (const R *)a.fA, *gA = (const R *)a.gA;
    auto FA = [&](int k, size_t off) { return (double)fA[(size_t)k * P + off] + a.sf[k]; };
    auto GA = [&](int k, size_t off) { return (double)gA[(size_t)k * P + off] + a.sg[k]; };

    // ---- forward locals: the pull of the nudged step
    double f[9], q[9];
    f[0] = FA(0, r0 + i);
    f[1] = FA(1, r0 + im);
    f[3] = FA(3, r0 + ip);
    if (!bot) {
      f[2] = FA(2, rm + i);
      f[5] = FA(5, rm + im);
      f[6] = FA(6, rm + ip);
    } else {
      f[2] = FA(4, r0 + i);
      f[5] = FA(7, r0 + i);
      f[6] = FA(8, r0 + i);
    }
    if (!top) {
      f[4] = FA(4, rp + i);
      f[7] = FA(7, rp + ip);
      f[8] = FA(8, rp + im);
    } else {
      f[4] = FA(2, r0 + i);
      f[7] = FA(5, r0 + i);
      f[8] = FA(6, r0 + i);
    }
    if (bot || top) {
      const size_t rs = bot ? r0 + g.pitch : r0 - g.pitch;
      for (int k = 0; k < 9; k++) q[k] = GA(k, rs + i);
    } else {
      const size_t rsm = (j == 1) ? r0 : rm, rsp = (j == g.ny - 2) ? r0 : rp;
      q[0] = GA(0, r0 + i);
      q[1] = GA(1, r0 + im);
      q[3] = GA(3, r0 + ip);
      q[2] = GA(2, rsm + i);
      q[5] = GA(5, rsm + im);
      q[6] = GA(6, rsm + ip);
      q[4] = GA(4, rsp + i);
      q[7] = GA(7, rsp + ip);
      q[8] = GA(8, rsp + im);
    }
    double rho = 0.0, S = 0.0, jx = 0.0, jy = 0.0;
    for (int k = 0; k < 9; k++) {
      rho += f[k];
      S += q[k];
      jx += CX[k] * f[k];
      jy += CY[k] * f[k];
    }
    const int nv = a.nvol[o];
    const double T = nv > 0 ? S + 0.5 * (double)((const R *)a.nudin)[o] : S;
    const double Fy = (double)((const R *)a.Fin)[o];
    const double ri = 1.0 / rho;
    const double u = jx * ri, v = (jy + 0.5 * Fy) * ri;
    const double usq = u * u + v * v;

    // ---- adjoints of the outputs of step s
    const R *ln = (const R *)a.lam_next;
    double Lf[9], Lg[9];
    if (ln) {
      for (int k = 0; k < 9; k++) {
        const int mj = j + CY[k];
        Lf[k] = (mj >= 0 && mj < g.ny) ? (double)ln[(size_t)k * P + (size_t)mj * g.pitch + wrapx(i + CX[k], g.nx)]
                                       : (double)ln[(size_t)OPP[k] * P + o];
      }
      for (int k = 0; k < 9; k++) {
        double s = 0.0;
        if (j >= 1 && j <= g.ny - 2) {
          const R *L = ln + (size_t)(9 + k) * P;
          const int mi = wrapx(i + CX[k], g.nx), mj = j + CY[k];
          if (mj >= 1 && mj <= g.ny - 2) s += (double)L[(size_t)mj * g.pitch + mi];
          if (j == 1 && CY[k] == 1) s += (double)L[(size_t)1 * g.pitch + mi];
          if (j == g.ny - 2 && CY[k] == -1) s += (double)L[(size_t)(g.ny - 2) * g.pitch + mi];
          if (j == 1) s += (double)L[i];
          if (j == g.ny - 2) s += (double)L[(size_t)(g.ny - 1) * g.pitch + i];
        }
        Lg[k] = s;
      }
    } else {
      for (int k = 0; k < 9; k++) Lf[k] = Lg[k] = 0.0;
    }
    R *lamF = (R *)a.lamF, *lamN = (R *)a.lamN;
    const double LF = (double)lamF[o], LN = (double)lamN[o];
    double lT = 0.0;
    if (a.lamT_fx) {
      const unsigned long long b = a.lamT_fx[o];
      if (b) {
        lT = (double)(long long)b / fixed_scale(*a.maxbits, a.nobs);
        a.lamT_fx[o] = 0ull;
      }
    }

    // ---- transposed Jacobian
    double lrho = a.g0 * T * LF;
    lT += a.g0 * rho * LF;
    double lQ = 0.0, lnud = 0.0;
    if (nv > 0) lQ = -LN;
    else lnud = LN;
    double lu = 0.0, lv = 0.0, lFy = 0.0;
    double lf[9], lq[9];
    for (int k = 0; k < 9; k++) {
      const double cu = CX[k] * u + CY[k] * v;
      const double bk = 1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * usq;
      lf[k] = a.one_m_om * Lf[k];
      lq[k] = a.one_m_omg * Lg[k];
      const double wLf = a.om * kW[k] * Lf[k], wLg = a.omg * kW[k] * Lg[k];
      lrho += wLf * bk;
      lT += wLg * bk;
      if (nv > 0) lQ -= a.bw[k] * Lg[k];
      const double lb = wLf * rho + wLg * T;
      const double db = 3.0 + 9.0 * cu;
      lu += lb * (db * CX[k] - 3.0 * u);
      lv += lb * (db * CY[k] - 3.0 * v);
      const double lfac = a.src[k] * Lf[k];
      lFy += lfac * (3.0 * (CY[k] - v) + 9.0 * cu * CY[k]);
      lv += lfac * (9.0 * CY[k] * CY[k] - 3.0) * Fy;
      lu += lfac * (9.0 * CX[k] * CY[k]) * Fy;
    }
    if (nv > 0) {
      const double W = a.wq ? nudge_weight(a.wq[o]) : 1.0;
      lT += a.alpha * W * lQ;
      dal = W * (T - a.Tbar[o]) * lQ;
    }
    const double ljx = ri * lu, ljy = ri * lv;
    lFy += 0.5 * ri * lv;
    lrho -= ri * ri * (jx * lu + (jy + 0.5 * Fy) * lv);
    if (nv > 0) lnud += 0.5 * lT;
    R *lam = (R *)a.lam;
    for (int k = 0; k < 9; k++) {
      lam[(size_t)k * P + o] = (R)(lf[k] + lrho + CX[k] * ljx + CY[k] * ljy);
      lam[(size_t)(9 + k) * P + o] = (R)(lq[k] + lT);
    }
    lamF[o] = (R)lFy;
    lamN[o] = (R)lnud;
  }
  const double s = block_sum(dal, lds);
  if (threadIdx.x == 0) a.dalpha_part[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = s;
}

__global__ __launch_bounds__(kT) void k_sum_partials(const double *part, int n, double *acc) {
  __shared__ double lds[kT / 64];
  double v = 0.0;
  for (int b = threadIdx.x; b < n; b += kT) v += part[b];
  const double s = block_sum(v, lds);
  if (threadIdx.x == 0) acc[0] += s;
}

template <typename R>
__global__ void k_adj_init(Geom g, const R *lam, const R *lamF, const R *rho0, const R *uv0, double g0,
                           double *dT0) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= g.nx) return;
  const size_t P = g.plane, o = (size_t)j * g.pitch + i;
  const double u = uv0[2 * o], v = uv0[2 * o + 1], usq = u * u + v * v;
  double s = 0.0;
  if (j >= 1 && j <= g.ny - 2) {
    for (int k = 0; k < 9; k++) {
      const R *L = lam + (size_t)(9 + k) * P;
      const int mi = wrapx(i + CX[k], g.nx), mj = j + CY[k];
      double lg = 0.0;
      if (mj >= 1 && mj <= g.ny - 2) lg += (double)L[(size_t)mj * g.pitch + mi];
      if (j == 1 && CY[k] == 1) lg += (double)L[(size_t)1 * g.pitch + mi];
      if (j == g.ny - 2 && CY[k] == -1) lg += (double)L[(size_t)(g.ny - 2) * g.pitch + mi];
      if (j == 1) lg += (double)L[i];
      if (j == g.ny - 2) lg += (double)L[(size_t)(g.ny - 1) * g.pitch + i];
      const double cu = CX[k] * u + CY[k] * v;
      s += kW[k] * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * usq) * lg;
    }
  }
  dT0[o] = s + g0 * (double)rho0[o] * (double)lamF[o];
}

template <typename R>
__global__ __launch_bounds__(kT) void k_misfit(Geom g, const R *T, int n, const double *x, const double *y,
                                               const double *Tobs, double *res, double *part,
                                               unsigned long long *maxbits, int *err) {
  __shared__ double lds[kT / 64];
  const int p = blockIdx.x * kT + threadIdx.x;
  double r2 = 0.0;
  if (p < n) {
    const part::Field<R> F = part::scalar_of<R>(T, g.nx, g.ny, g.pitch);
    int e = 0;
    const double val = part::interp(F, x[p], y[p], e);
    if (e) atomicOr(err, 1);
    const double r = e ? 0.0 : val - Tobs[p];
    res[p] = r;
    r2 = r * r;
    atomicMax(maxbits, (unsigned long long)__double_as_longlong(fabs(r)));
  }
  const double s = block_sum(r2, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void k_misfit_adjoint(Geom g, int n, const double *x, const double *y, const double *res,
                                 const unsigned long long *maxbits, unsigned long long *lamT) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const double xx = x[p], yy = y[p];
  int x1, y1;
  if (!part::corners(xx, yy, g.nx, g.ny, x1, y1)) return;
  const int x2 = x1 + 1, y2 = y1 + 1;
  const double sc = fixed_scale(*maxbits, n);
  const double r2 = 2.0 * res[p];
  auto add = [&](int ic, int jc, double w) {
    const int I = ic == 0 ? g.nx : (ic == g.nx + 1 ? 1 : ic);
    const int J = jc == 0 ? g.ny : (jc == g.ny + 1 ? 1 : jc);
    const long long q = llrint(r2 * w * sc);
    if (q) atomicAdd(&lamT[(size_t)(J - 1) * g.pitch + (I - 1)], (unsigned long long)q);
  };
  add(x1, y1, ((double)x2 - xx) * ((double)y2 - yy));
  add(x2, y1, (xx - (double)x1) * ((double)y2 - yy));
  add(x1, y2, ((double)x2 - xx) * (yy - (double)y1));
  add(x2, y2, (xx - (double)x1) * (yy - (double)y1));
}

}  // namespace

int adj_blocks(const Geom &g) { return ((g.nx + kT - 1) / kT) * g.ny; }
int misfit_blocks(int n) { return (n + kT - 1) / kT; }

hipError_t launch_adjoint_step(hipStream_t s, int prec, const AdjStepArgs &a) {
  const dim3 grid((a.g.nx + kT - 1) / kT, a.g.ny), block(kT);
  if (prec == RBN_FP64)
    hipLaunchKernelGGL(k_adjoint<double>, grid, block, 0, s, a);
  else
    hipLaunchKernelGGL(k_adjoint<float>, grid, block, 0, s, a);
  return hipGetLastError();
}

hipError_t launch_sum_partials(hipStream_t s, const double *part, int n, double *acc) {
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kT), 0, s, part, n, acc);
  return hipGetLastError();
}

hipError_t launch_adj_init(hipStream_t s, int prec, const Geom &g, const void *lam, const void *lamF,
                           const void *rho0, const void *uv0, double g0, double *dT0) {
  const dim3 grid((g.nx + kT - 1) / kT, g.ny), block(kT);
  if (prec == RBN_FP64)
    hipLaunchKernelGGL(k_adj_init<double>, grid, block, 0, s, g, (const double *)lam, (const double *)lamF,
                       (const double *)rho0, (const double *)uv0, g0, dT0);
  else
    hipLaunchKernelGGL(k_adj_init<float>, grid, block, 0, s, g, (const float *)lam, (const float *)lamF,
                       (const float *)rho0, (const float *)uv0, g0, dT0);
  return hipGetLastError();
}

hipError_t launch_misfit(hipStream_t s, int prec, const Geom &g, const void *T, int n, const double *x,
                         const double *y, const double *Tobs, double *res, double *part,
                         unsigned long long *maxbits, int *err) {
  if (n <= 0) return hipSuccess;
  const dim3 grid(misfit_blocks(n)), block(kT);
  if (prec == RBN_FP64)
    hipLaunchKernelGGL(k_misfit<double>, grid, block, 0, s, g, (const double *)T, n, x, y, Tobs, res, part,
                       maxbits, err);
  else
    hipLaunchKernelGGL(k_misfit<float>, grid, block, 0, s, g, (const float *)T, n, x, y, Tobs, res, part,
                       maxbits, err);
  return hipGetLastError();
}

hipError_t launch_misfit_adjoint(hipStream_t s, const Geom &g, int n, const double *x, const double *y,
                                 const double *res, const unsigned long long *maxbits,
                                 unsigned long long *lamT_fx) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_misfit_adjoint, dim3(misfit_blocks(n)), dim3(kT), 0, s, g, n, x, y, res, maxbits,
                     lamT_fx);
  return hipGetLastError();
}

}  // namespace rbn
