// rbn_lbm2_tile.h — device code of a two-step tile (temporal blocking in LDS), shared by the
// plain two-step launch (rbn_lbm2.hip) and the batched ensemble (rbn_ensemble.hip).
//
// A workgroup of kTileThreads threads owns a tile of kTW x TH nodes.  Phase 1 computes step s
// on the tile and a one-node ring around it (the ring is computed again by the neighbouring
// tiles), pulling from the step's input planes in HBM, and keeps the post-collision f and g
// of every ring node in LDS; it writes the step's velocity for the tile and keeps the tile's
// next force F(s+1) in LDS.  Phase 2 computes step s+1 on the tile, pulling from LDS, and
// writes the output planes, the force F(s+2), the velocity of step s+1 and, when asked, its
// moments.  Per node and two steps the HBM traffic is one read and one write of the
// populations (plus the ring reads) instead of two of each.
//
// The node arithmetic is the truth-run path of lbm_node64 / lbm_node32 (rbn_lbm_node.h),
// operand for operand, so a two-step tile gives bit for bit the result of two one-step
// launches.  The force is double-buffered: other tiles still read F(s) at nodes whose F(s+2)
// this tile writes.  Only the truth run (plate walls) has a two-step form.
#pragma once

#include "rbn_lbm_node.h"

namespace rbn {
namespace lbm2 {

using namespace lbm;

constexpr int kTileThreads = 256;
constexpr int kTW = 64;  // tile width: one wavefront per tile row

// Pulled populations of a truth-run node (walls: half-way bounce-back for f from the node's
// own pre-stream values, the plate rule for g), through an accessor a.f(k, di, dj) /
// a.g(k, di, dj) of the input at the node offset (di, dj).
template <typename R, class Acc>
__device__ __forceinline__ void gather_truth(const Acc &a, bool bot, bool top, const StepConst<R> &c, R f[9],
                                             R q[9]) {
  f[0] = a.f(0, 0, 0);
  f[1] = a.f(1, -1, 0);
  f[3] = a.f(3, 1, 0);
  if (!bot) {
    f[2] = a.f(2, 0, -1);
    f[5] = a.f(5, -1, -1);
    f[6] = a.f(6, 1, -1);
  } else {
    f[2] = a.f(4, 0, 0);
    f[5] = a.f(7, 0, 0);
    f[6] = a.f(8, 0, 0);
  }
  if (!top) {
    f[4] = a.f(4, 0, 1);
    f[7] = a.f(7, 1, 1);
    f[8] = a.f(8, -1, 1);
  } else {
    f[4] = a.f(2, 0, 0);
    f[7] = a.f(5, 0, 0);
    f[8] = a.f(6, 0, 0);
  }
  q[0] = a.g(0, 0, 0);
  q[1] = a.g(1, -1, 0);
  q[3] = a.g(3, 1, 0);
  if (!bot) {
    q[2] = a.g(2, 0, -1);
    q[5] = a.g(5, -1, -1);
    q[6] = a.g(6, 1, -1);
  }
  if (!top) {
    q[4] = a.g(4, 0, 1);
    q[7] = a.g(7, 1, 1);
    q[8] = a.g(8, -1, 1);
  }
  if (top) {
    q[4] = -q[2] + c.wtop[0];
    q[7] = -q[5] + c.wtop[1];
    q[8] = -q[6] + c.wtop[2];
  }
  if (bot) {
    q[2] = -q[4] + c.wbot[0];
    q[5] = -q[7] + c.wbot[1];
    q[6] = -q[8] + c.wbot[2];
  }
}

// The input planes in HBM at node (i, j): periodic x, rows j-1 / j+1 (never read across a
// wall: gather_truth does not pull through it).
template <typename R>
struct GlobalAcc {
  const R *__restrict__ fa;
  const R *__restrict__ ga;
  size_t plane, r0, rm, rp;
  int i, im, ip;
  __device__ __forceinline__ size_t off(int di, int dj) const {
    return (dj < 0 ? rm : (dj > 0 ? rp : r0)) + (size_t)(di < 0 ? im : (di > 0 ? ip : i));
  }
  __device__ __forceinline__ R f(int k, int di, int dj) const { return fa[(size_t)k * plane + off(di, dj)]; }
  __device__ __forceinline__ R g(int k, int di, int dj) const { return ga[(size_t)k * plane + off(di, dj)]; }
};

// The step-s populations in LDS at ring-local node (li, lj), rows of RW nodes, planes of RS.
template <typename R, int RW, int RS>
struct LdsAcc {
  const R *lf;
  const R *lg;
  int o;  // lj * RW + li
  __device__ __forceinline__ R f(int k, int di, int dj) const { return lf[k * RS + o + dj * RW + di]; }
  __device__ __forceinline__ R g(int k, int di, int dj) const { return lg[k * RS + o + dj * RW + di]; }
};

// Moments and collision of a truth-run node: lbm_node64's arithmetic.
__device__ __forceinline__ void collide64(const StepConst<double> &c, const double f[9], const double q[9],
                                          double Fy, double fo[9], double go[9], double &rho, double &T,
                                          double &u, double &v, double &Fn) {
  rho = 0.0;
  T = 0.0;
#pragma unroll
  for (int k = 0; k < 9; k++) rho = rho + f[k];
#pragma unroll
  for (int k = 0; k < 9; k++) T = T + q[k];
  const double rhoi = 1. / rho;
  double fl[9];
#pragma unroll
  for (int k = 0; k < 9; k++) fl[k] = f[k] * rhoi;
  const double Fx = 0.0;
  u = fl[1] - fl[3] + fl[5] - fl[6] - fl[7] + fl[8] + 0.5 * rhoi * Fx;
  v = fl[5] + fl[2] + fl[6] - fl[7] - fl[4] - fl[8] + 0.5 * rhoi * Fy;
  Fn = c.g0 * rho * (T - 0.0);
  double b[9], fac[9];
  eq_brackets64(u, v, b);
  guo_factor64(u, v, Fy, fac);
#pragma unroll
  for (int k = 0; k < 9; k++) {
    const double feq = rho * kW[k] * b[k];
    fo[k] = f[k] * c.one_m_om + c.om * feq + c.src[k] * fac[k];
  }
#pragma unroll
  for (int k = 0; k < 9; k++) {
    const double geq = T * kW[k] * b[k];
    go[k] = q[k] * c.one_m_omg + c.omg * geq;
  }
}

// Moments and collision of a truth-run node with shifted FP32 populations: lbm_node32's
// arithmetic.
__device__ __forceinline__ void collide32(const StepConst<float> &c, const float f[9], const float q[9],
                                          float Fy, float fo[9], float go[9], float &rho, float &T, float &u,
                                          float &v, float &Fn) {
  float dr = 0.f, dT = 0.f;
#pragma unroll
  for (int k = 0; k < 9; k++) dr += f[k];
#pragma unroll
  for (int k = 0; k < 9; k++) dT += q[k];
  rho = c.rho0 + dr;
  T = c.T0 + dT;
  const float rhoi = 1.f / rho;
  const float jx = f[1] - f[3] + f[5] - f[6] - f[7] + f[8];
  const float jy = f[2] - f[4] + f[5] + f[6] - f[7] - f[8];
  u = jx * rhoi;
  v = jy * rhoi + 0.5f * rhoi * Fy;
  Fn = c.g0 * rho * T;
  float A[9];
  eq_dev32(u, v, A);
  const float tz = -3.f * v * Fy, tp = 3.f * (1.f - v) * Fy, tm = 3.f * (-1.f - v) * Fy;
  const float d2 = 9.f * v * Fy, d5 = 9.f * (u + v) * Fy, d6 = 9.f * (v - u) * Fy;
  const float fac[9] = {tz, tz, tp + d2, tz, tm + d2, tp + d5, tp + d6, tm + d5, tm + d6};
  float feq[9], geq[9];
  eq_pops32(dr, rho, A, feq);
#pragma unroll
  for (int k = 0; k < 9; k++) fo[k] = f[k] * c.one_m_om + c.om * feq[k] + c.src[k] * fac[k];
  const float Q = 0.f;
  eq_pops32(dT, T, A, geq);
#pragma unroll
  for (int k = 0; k < 9; k++) go[k] = q[k] * c.one_m_omg + c.omg * geq[k] - c.bw[k] * Q;
}

__device__ __forceinline__ void collide(const StepConst<double> &c, const double f[9], const double q[9],
                                        double Fy, double fo[9], double go[9], double &rho, double &T,
                                        double &u, double &v, double &Fn) {
  collide64(c, f, q, Fy, fo, go, rho, T, u, v, Fn);
}
__device__ __forceinline__ void collide(const StepConst<float> &c, const float f[9], const float q[9],
                                        float Fy, float fo[9], float go[9], float &rho, float &T, float &u,
                                        float &v, float &Fn) {
  collide32(c, f, q, Fy, fo, go, rho, T, u, v, Fn);
}

// The planes of a two-step tile.
struct TwoStepPtrs {
  const void *fA, *gA;  // step s input planes
  void *fB, *gB;        // step s+2 input planes (output)
  const void *Fin;      // F(s)
  void *Fout;           // F(s+2)
  void *uv1, *uv2;      // velocity slots of steps s and s+1
  void *rho2, *T2;      // moments of step s+1, or null
  int *bad;             // non-finite flag of the second step's moments, or null
};

// LDS of a two-step tile: the ring's post-collision f and g, and the tile's next force.
template <typename R, int TH>
struct TileLds {
  static constexpr int RW = kTW + 2, RH = TH + 2, RS = RW * RH;
  R f[9 * RS];
  R g[9 * RS];
  R F[kTW * TH];
};

// Both steps on tile (tx, ty) of 64 x TH nodes; every thread of the workgroup calls it.
template <typename R, int TH>
__device__ __forceinline__ void two_step_tile(const Geom &g, const TwoStepPtrs &a, const StepConst<R> &c, int tx,
                                              int ty, TileLds<R, TH> &L) {
  constexpr int RW = TileLds<R, TH>::RW, RS = TileLds<R, TH>::RS;
  R *lf = L.f, *lg = L.g, *lF = L.F;
  const int i0 = tx * kTW, j0 = ty * TH;
  const int ncols = min(kTW, g.nx - i0), nrows = min(TH, g.ny - j0);

  // Phase 1: step s on the tile and its ring.
  for (int qn = threadIdx.x; qn < RS; qn += kTileThreads) {
    const int r = qn / RW, cc = qn - r * RW;
    const int y = j0 - 1 + r;
    if (cc > ncols + 1 || r > nrows + 1 || y < 0 || y >= g.ny) continue;
    int x = i0 - 1 + cc;
    x = x < 0 ? x + g.nx : (x >= g.nx ? x - g.nx : x);
    const bool bot = (y == 0), top = (y == g.ny - 1);
    GlobalAcc<R> acc;
    acc.fa = (const R *)a.fA;
    acc.ga = (const R *)a.gA;
    acc.plane = g.plane;
    acc.r0 = (size_t)y * g.pitch;
    acc.rm = bot ? acc.r0 : acc.r0 - g.pitch;
    acc.rp = top ? acc.r0 : acc.r0 + g.pitch;
    acc.i = x;
    acc.im = x == 0 ? g.nx - 1 : x - 1;
    acc.ip = x == g.nx - 1 ? 0 : x + 1;
    R f[9], q[9], fo[9], go[9], rho, T, u, v, Fn;
    gather_truth<R>(acc, bot, top, c, f, q);
    const R Fy = ((const R *)a.Fin)[acc.r0 + x];
    collide(c, f, q, Fy, fo, go, rho, T, u, v, Fn);
#pragma unroll
    for (int k = 0; k < 9; k++) {
      lf[k * RS + qn] = fo[k];
      lg[k * RS + qn] = go[k];
    }
    if (r >= 1 && r <= nrows && cc >= 1 && cc <= ncols) {
      st_uv<0, R>(a.uv1, acc.r0 + x, u, v);
      lF[(r - 1) * kTW + (cc - 1)] = Fn;
    }
  }
  __syncthreads();

  // Phase 2: step s+1 on the tile, from LDS.
  for (int qn = threadIdx.x; qn < kTW * TH; qn += kTileThreads) {
    const int r = qn / kTW + 1, cc = qn - (r - 1) * kTW + 1;
    if (cc > ncols || r > nrows) continue;
    const int x = i0 - 1 + cc, y = j0 - 1 + r;
    const bool bot = (y == 0), top = (y == g.ny - 1);
    LdsAcc<R, RW, RS> acc{lf, lg, r * RW + cc};
    R f[9], q[9], fo[9], go[9], rho, T, u, v, Fn;
    gather_truth<R>(acc, bot, top, c, f, q);
    collide(c, f, q, lF[qn], fo, go, rho, T, u, v, Fn);
    flag_nonfinite(a.bad, rho, T);
    const size_t o = (size_t)y * g.pitch + x;
    R *fB = (R *)a.fB, *gB = (R *)a.gB;
#pragma unroll
    for (int k = 0; k < 9; k++) st<1>(&fB[(size_t)k * g.plane + o], fo[k]);
#pragma unroll
    for (int k = 0; k < 9; k++) st<1>(&gB[(size_t)k * g.plane + o], go[k]);
    st<1>(&((R *)a.Fout)[o], Fn);
    st_uv<0, R>(a.uv2, o, u, v);
    if (a.rho2) ((R *)a.rho2)[o] = rho;
    if (a.T2) ((R *)a.T2)[o] = T;
  }
}

}  // namespace lbm2
}  // namespace rbn
