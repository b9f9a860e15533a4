// rbn_lbm_node.h — device code of K1, the fused D2Q9 double-population step, for one lattice
// node; shared by the plain launch (rbn_lbm.hip) and the persistent ensemble launch
// (rbn_ensemble.hip).
//
// One thread owns one lattice node and performs, in registers, everything the reference
// does in separate sweeps over an array-of-structs lattice:
//   pull-streaming of f and g (moveplusforcingconstructWW, Trajectories/code/temppart_dump.c:757-922
//   or Nudging/code/tempvel_nudge.c:677-1032) with periodic x as an index wrap (pbc, :234-250),
//   the f half-way bounce-back from the node's own pre-stream populations,
//   the g wall rule (fixed temperatures, or zero gradient in the nudged run),
//   hydrovar (:957-987; nudged form N:1069-1102),
//   the lagged buoyancy force (forcingconstructWW, :744-755: F(s+1) = -gravity*rho(s)*T(s) is
//   stored per node and read back next step),
//   equilif/equilig (:1097-1180), collisf (BGK + Guo, :1314-1328) and collisg
//   (plain :1330-1341, nudged N:1917-1945).
// Populations live in HBM as structure-of-arrays planes [k][j][i] (x fastest); a step reads
// the A planes and writes the B planes (A-B pattern), so every access is a coalesced row
// segment shifted by at most one element.  Traffic per node: 18 population loads + 18
// stores, 1 load + 1 store of the lagged force, 2 stores of (u,v) = 40 reals.
//
// FP64 path: every expression keeps the reference's operand order and the library is
// compiled with -ffp-contract=off, so results equal the CPU restatement bit for bit.  Two
// value-preserving rewrites reduce the divisions: equilif and equilig share the velocity
// terms (they are computed from the same u, v), and the Guo factor uses that F_x = 0 and
// the D2Q9 direction symmetry (e.g. direction 7 gives ((-u)+(-v))*(-F_y) = (u+v)*F_y
// exactly), so 6 distinct quotients replace 18.  Both change at most the sign of a zero.
//
// FP32 path: populations are stored shifted, f_k - w_k*rho0 and g_k - w_k*T0, and the
// step is written in those deviations (mass and temperature are summed from the small
// deviations before the reference level is added back), which keeps the round-off of the
// O(1) reference level out of the O(u^2) physics.
#pragma once

#include "rbn_internal.h"

namespace rbn {
namespace lbm {

constexpr double CS2 = 1. / 3.;
constexpr double CS22 = 2. * CS2;
constexpr double CSSQ = 2. / 9.;

// NT flags of the fused step: bit 0 non-temporal stores of the outputs (not read again
// within the step), bit 1 non-temporal loads of the inputs (read once per step), bit 2
// ordinary stores of the velocity even under bit 0 (the particle update gathers from it
// right after the step, so it is left to the caches).
template <int NT, typename T>
__device__ __forceinline__ void st(T *p, T v) {
  if (NT & 1)
    __builtin_nontemporal_store(v, p);
  else
    *p = v;
}

template <typename R, int NT = 0>
struct Planes {
  const R *__restrict__ a;
  size_t P;
  __device__ __forceinline__ R at(int k, size_t off) const {
    if (NT & 2) return __builtin_nontemporal_load(&a[(size_t)k * P + off]);
    return a[(size_t)k * P + off];
  }
};

// Store of the interleaved (u, v) pair of one node as one vector.
template <int NT, typename T>
__device__ __forceinline__ void st_uv(void *uv, size_t o, T u, T v) {
  typedef T T2 __attribute__((ext_vector_type(2)));
  T2 w;
  w.x = u;
  w.y = v;
  st<(NT & 4) ? 0 : NT>(&((T2 *)uv)[o], w);
}

template <int NT, typename T>
__device__ __forceinline__ T ld(const T *p) {
  if (NT & 2) return __builtin_nontemporal_load(p);
  return *p;
}

// ---------------------------------------------------------------------------- FP64

// Velocity part of the D2Q9 equilibrium, shared by equilif and equilig: b[k] is the
// bracket of feq_k = r1*w_k*b[k] (Trajectories/code/temppart_dump.c:1110-1135).
__device__ __forceinline__ void eq_brackets64(double u, double v, double b[9]) {
  double usq = u * u, vsq = v * v;
  double sumsq = (usq + vsq) / CS22;
  double sumsq2 = sumsq * (1. - CS2) / CS2;
  double u2 = usq / CSSQ, v2 = vsq / CSSQ;
  double ui = u / CS2, vi = v / CS2;
  double uv = ui * vi;
  b[0] = (1. - sumsq);
  b[1] = (1. - sumsq + u2 + ui);
  b[2] = (1. - sumsq + v2 + vi);
  b[3] = (1. - sumsq + u2 - ui);
  b[4] = (1. - sumsq + v2 - vi);
  b[5] = (1. + sumsq2 + ui + vi + uv);
  b[6] = (1. + sumsq2 - ui + vi - uv);
  b[7] = (1. + sumsq2 - ui - vi + uv);
  b[8] = (1. + sumsq2 + ui - vi - uv);
}

// Guo forcing factor of collisf (T:1321-1322) for F_x = 0, grouped by direction class.
__device__ __forceinline__ void guo_factor64(double u, double v, double Fy, double fac[9]) {
  const double tz = ((0. - v) * Fy) / CS2;   // c_y = 0   (k = 0, 1, 3)
  const double tp = ((1. - v) * Fy) / CS2;   // c_y = +1  (k = 2, 5, 6)
  const double tm = ((-1. - v) * Fy) / CS2;  // c_y = -1  (k = 4, 7, 8)
  const double d2 = (v * Fy) / (CS2 * CS2);          // k = 2, 4
  const double d5 = ((u + v) * Fy) / (CS2 * CS2);    // k = 5, 7
  const double d6 = ((v - u) * Fy) / (CS2 * CS2);    // k = 6, 8
  fac[0] = tz;
  fac[1] = tz;
  fac[3] = tz;
  fac[2] = tp + d2;
  fac[4] = tm + d2;
  fac[5] = tp + d5;
  fac[6] = tp + d6;
  fac[7] = tm + d5;
  fac[8] = tm + d6;
}

// The fused step at interior node (i, j) (0-based), FP64.
template <int VARIANT, int NT>
__device__ __forceinline__ void lbm_node64(const Geom &g, const LbmPtrs &p, const StepConst<double> &c,
                                           int i, int j) {
  const int im = (i == 0) ? g.nx - 1 : i - 1;  // periodic x (pbc)
  const int ip = (i == g.nx - 1) ? 0 : i + 1;
  const bool bot = (j == 0), top = (j == g.ny - 1);
  const size_t r0 = (size_t)j * g.pitch;
  const size_t rm = bot ? r0 : r0 - g.pitch;
  const size_t rp = top ? r0 : r0 + g.pitch;
  const Planes<double, NT> fA{(const double *)p.fA, g.plane}, gA{(const double *)p.gA, g.plane};

  double f[9], q[9];
  // f: pull, with half-way bounce-back at the walls from the node's own pre-stream
  // populations (T:859-894).
  f[0] = fA.at(0, r0 + i);
  f[1] = fA.at(1, r0 + im);
  f[3] = fA.at(3, r0 + ip);
  if (!bot) {
    f[2] = fA.at(2, rm + i);
    f[5] = fA.at(5, rm + im);
    f[6] = fA.at(6, rm + ip);
  } else {
    f[2] = fA.at(4, r0 + i);
    f[5] = fA.at(7, r0 + i);
    f[6] = fA.at(8, r0 + i);
  }
  if (!top) {
    f[4] = fA.at(4, rp + i);
    f[7] = fA.at(7, rp + ip);
    f[8] = fA.at(8, rp + im);
  } else {
    f[4] = fA.at(2, r0 + i);
    f[7] = fA.at(5, r0 + i);
    f[8] = fA.at(6, r0 + i);
  }
  if (VARIANT == RBN_VARIANT_TRAJECTORIES) {
    q[0] = gA.at(0, r0 + i);
    q[1] = gA.at(1, r0 + im);
    q[3] = gA.at(3, r0 + ip);
    if (!bot) {
      q[2] = gA.at(2, rm + i);
      q[5] = gA.at(5, rm + im);
      q[6] = gA.at(6, rm + ip);
    }
    if (!top) {
      q[4] = gA.at(4, rp + i);
      q[7] = gA.at(7, rp + ip);
      q[8] = gA.at(8, rp + im);
    }
    // Fixed plate temperatures from the post-stream opposite populations (T:896-921).
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
  } else {
    // Zero-gradient walls (N:684-702, 783-962): wall rows take the pre-stream populations
    // of their inner neighbour row and are not streamed; the rows next to them pull from
    // that copy.
    if (bot || top) {
      const size_t rs = bot ? r0 + g.pitch : r0 - g.pitch;
#pragma unroll
      for (int k = 0; k < 9; k++) q[k] = gA.at(k, rs + i);
    } else {
      const size_t rsm = (j == 1) ? r0 : rm;
      const size_t rsp = (j == g.ny - 2) ? r0 : rp;
      q[0] = gA.at(0, r0 + i);
      q[1] = gA.at(1, r0 + im);
      q[3] = gA.at(3, r0 + ip);
      q[2] = gA.at(2, rsm + i);
      q[5] = gA.at(5, rsm + im);
      q[6] = gA.at(6, rsm + ip);
      q[4] = gA.at(4, rsp + i);
      q[7] = gA.at(7, rsp + ip);
      q[8] = gA.at(8, rsp + im);
    }
  }

  // hydrovar
  double rho = 0.0, T = 0.0;
#pragma unroll
  for (int k = 0; k < 9; k++) rho = rho + f[k];
#pragma unroll
  for (int k = 0; k < 9; k++) T = T + q[k];
  bool nv = false;
  double Tb = 0.0, wn = 0.0;
  if (VARIANT == RBN_VARIANT_NUDGING && (p.tw || p.nvol)) {
    nv = nudge_input<double>(p, r0 + i, Tb, wn);
    if (nv) T = (T + 0.5 * ((const double *)p.nud)[r0 + i]);
  }
  flag_nonfinite(p.bad, rho, T);
  const double rhoi = 1. / rho;
  double fl[9];
#pragma unroll
  for (int k = 0; k < 9; k++) fl[k] = f[k] * rhoi;
  double *F = (double *)p.F;
  const double Fx = 0.0;
  const double Fy = ld<NT>(&F[r0 + i]);
  const double u = fl[1] - fl[3] + fl[5] - fl[6] - fl[7] + fl[8] + 0.5 * rhoi * Fx;
  const double v = fl[5] + fl[2] + fl[6] - fl[7] - fl[4] - fl[8] + 0.5 * rhoi * Fy;
  st<NT>(&F[r0 + i], c.g0 * rho * (T - 0.0));
  st_uv<NT, double>(p.uv, r0 + i, u, v);
  if (p.rho) ((double *)p.rho)[r0 + i] = rho;
  if (p.T) ((double *)p.T)[r0 + i] = T;

  // equilibria + collisions
  double b[9], fac[9];
  eq_brackets64(u, v, b);
  guo_factor64(u, v, Fy, fac);
  double *fB = (double *)p.fB, *gB = (double *)p.gB;
#pragma unroll
  for (int k = 0; k < 9; k++) {
    const double feq = rho * kW[k] * b[k];
    st<NT>(&fB[(size_t)k * g.plane + r0 + i], f[k] * c.one_m_om + c.om * feq + c.src[k] * fac[k]);
  }
  if (VARIANT == RBN_VARIANT_NUDGING && nv) {
    // w = 1 (7x7 deposit) leaves alpha * (T - Tbar) unchanged
    const double Q = c.alpha * (wn * (T - Tb));
#pragma unroll
    for (int k = 0; k < 9; k++) {
      const double geq = T * kW[k] * b[k];
      st<NT>(&gB[(size_t)k * g.plane + r0 + i], q[k] * c.one_m_omg + c.omg * geq - c.bw[k] * Q);
    }
    ((double *)p.nud)[r0 + i] = -1. * Q;
  } else {
#pragma unroll
    for (int k = 0; k < 9; k++) {
      const double geq = T * kW[k] * b[k];
      st<NT>(&gB[(size_t)k * g.plane + r0 + i], q[k] * c.one_m_omg + c.omg * geq);
    }
  }
}

// ---------------------------------------------------------------------------- FP32

__device__ __forceinline__ void eq_dev32(float u, float v, float A[9]) {
  // A_k = (bracket_k - 1) of the equilibrium, with 1/cs2 = 3, 1/(2cs2) = 1.5,
  // (1-cs2)/cs2 = 2, 1/cssq = 4.5.
  float usq = u * u, vsq = v * v;
  float sumsq = 1.5f * (usq + vsq);
  float sumsq2 = 2.f * sumsq;
  float u2 = 4.5f * usq, v2 = 4.5f * vsq;
  float ui = 3.f * u, vi = 3.f * v;
  float uv = ui * vi;
  A[0] = -sumsq;
  A[1] = -sumsq + u2 + ui;
  A[2] = -sumsq + v2 + vi;
  A[3] = -sumsq + u2 - ui;
  A[4] = -sumsq + v2 - vi;
  A[5] = sumsq2 + ui + vi + uv;
  A[6] = sumsq2 - ui + vi - uv;
  A[7] = sumsq2 - ui - vi + uv;
  A[8] = sumsq2 + ui - vi - uv;
}

// FP32 equilibria of the deviations from the reference levels, for f (mass deviation dr,
// density rho) and g (temperature deviation dT, temperature T): eq_k = w_k (d + level A_k)
// for the moving populations and the rest population as the remainder, eq_0 = d - sum of the
// others.  The float weights sum to 1 + 7.5e-9, so eq_0 = w_0 (d + level A_0) would scale the
// deviation sums by that factor through every collision and drift T linearly in time (about
// 7e-9 per step at T - T0 = 0.5); with the remainder the equilibria carry d exactly, up to the
// rounding of the sums.
__device__ __forceinline__ void eq_pops32(float d, float level, const float A[9], float eq[9]) {
  float s = 0.f;
#pragma unroll
  for (int k = 1; k < 9; k++) {
    eq[k] = (float)kW[k] * (d + level * A[k]);
    s += eq[k];
  }
  eq[0] = d - s;
}

// Moments and collision of a truth-run node with shifted FP32 populations, written for short
// dependent chains: pairwise sums for the mass and temperature deviations, the hardware
// reciprocal for 1/rho (at most 1 ulp from the rounded quotient), fused multiply-adds in the
// equilibria and the relaxation.  The same operations as the FP64 path otherwise (the
// deviations' equilibria of eq_pops32, the Guo forcing, the BGK relaxation with 1 - omega
// formed from the rounded omega), so the error against FP64 stays within the stated bound.
// Every FP32 truth-run launch form (one, two or three steps per launch, ensembles) calls
// this function, so they agree bit for bit.
__device__ __forceinline__ void collide_t32(const StepConst<float> &c, const float f[9], const float q[9],
                                            float Fy, float fo[9], float go[9], float &rho, float &T,
                                            float &u, float &v, float &Fn) {
  const float dr = ((f[0] + f[1]) + (f[2] + f[3])) + ((f[4] + f[5]) + (f[6] + f[7])) + f[8];
  const float dT = ((q[0] + q[1]) + (q[2] + q[3])) + ((q[4] + q[5]) + (q[6] + q[7])) + q[8];
  rho = c.rho0 + dr;
  T = c.T0 + dT;
  const float rhoi = __builtin_amdgcn_rcpf(rho);
  const float jx = ((f[1] - f[3]) + (f[5] - f[6])) + (f[8] - f[7]);
  const float jy = ((f[2] - f[4]) + (f[5] + f[6])) - (f[7] + f[8]);
  u = jx * rhoi;
  v = __builtin_fmaf(0.5f, Fy, jy) * rhoi;
  Fn = c.g0 * rho * T;
  // A_k = bracket_k - 1 of the equilibrium: 1/cs2 = 3, 1/(2 cs2) = 1.5, 1/(2 cs2^2) = 4.5
  const float ui = 3.f * u, vi = 3.f * v;
  const float usq = u * u, vsq = v * v;
  const float sumsq = 1.5f * (usq + vsq);
  const float a = __builtin_fmaf(4.5f, usq, -sumsq), b = __builtin_fmaf(4.5f, vsq, -sumsq);
  const float uv = ui * vi, s2 = 2.f * sumsq;
  const float sp = s2 + uv, sm = s2 - uv, pp = ui + vi, pm = ui - vi;
  const float A[9] = {-sumsq, a + ui, b + vi, a - ui, b - vi, sp + pp, sm - pm, sp - pp, sm + pm};
  // Guo factor with F_x = 0: ((c_y - v) F_y)/cs2 + ((c.u)(c_y F_y))/cs2^2
  const float f3 = 3.f * Fy, f9 = 9.f * Fy;
  const float tz = -v * f3, tp = (1.f - v) * f3, tm = (-1.f - v) * f3;
  const float d2 = v * f9, d5 = (u + v) * f9, d6 = (v - u) * f9;
  const float fac[9] = {tz, tz, tp + d2, tz, tm + d2, tp + d5, tp + d6, tm + d5, tm + d6};
  // equilibria of the deviations (eq_pops32's form): the moving populations
  // w_k (d + level A_k), the rest population the remainder
  float feq[9], geq[9];
#pragma unroll
  for (int k = 1; k < 9; k++) {
    feq[k] = (float)kW[k] * __builtin_fmaf(rho, A[k], dr);
    geq[k] = (float)kW[k] * __builtin_fmaf(T, A[k], dT);
  }
  feq[0] = dr - (((feq[1] + feq[2]) + (feq[3] + feq[4])) + ((feq[5] + feq[6]) + (feq[7] + feq[8])));
  geq[0] = dT - (((geq[1] + geq[2]) + (geq[3] + geq[4])) + ((geq[5] + geq[6]) + (geq[7] + geq[8])));
  // the forcing weights (1 - om/2) w_k take three values (w_k does)
#pragma unroll
  for (int k = 0; k < 9; k++) {
    const float src = c.src[k == 0 ? 0 : (k < 5 ? 1 : 5)];
    fo[k] = __builtin_fmaf(f[k], c.one_m_om, __builtin_fmaf(c.om, feq[k], src * fac[k]));
    go[k] = __builtin_fmaf(q[k], c.one_m_omg, c.omg * geq[k]);
  }
}

// The fused step at interior node (i, j) (0-based), FP32 with shifted populations.
template <int VARIANT, int NT>
__device__ __forceinline__ void lbm_node32(const Geom &g, const LbmPtrs &p, const StepConst<float> &c,
                                           int i, int j) {
  const int im = (i == 0) ? g.nx - 1 : i - 1;
  const int ip = (i == g.nx - 1) ? 0 : i + 1;
  const bool bot = (j == 0), top = (j == g.ny - 1);
  const size_t r0 = (size_t)j * g.pitch;
  const size_t rm = bot ? r0 : r0 - g.pitch;
  const size_t rp = top ? r0 : r0 + g.pitch;
  const Planes<float, NT> fA{(const float *)p.fA, g.plane}, gA{(const float *)p.gA, g.plane};

  float f[9], q[9];
  f[0] = fA.at(0, r0 + i);
  f[1] = fA.at(1, r0 + im);
  f[3] = fA.at(3, r0 + ip);
  if (!bot) {
    f[2] = fA.at(2, rm + i);
    f[5] = fA.at(5, rm + im);
    f[6] = fA.at(6, rm + ip);
  } else {
    f[2] = fA.at(4, r0 + i);
    f[5] = fA.at(7, r0 + i);
    f[6] = fA.at(8, r0 + i);
  }
  if (!top) {
    f[4] = fA.at(4, rp + i);
    f[7] = fA.at(7, rp + ip);
    f[8] = fA.at(8, rp + im);
  } else {
    f[4] = fA.at(2, r0 + i);
    f[7] = fA.at(5, r0 + i);
    f[8] = fA.at(6, r0 + i);
  }
  if (VARIANT == RBN_VARIANT_TRAJECTORIES) {
    q[0] = gA.at(0, r0 + i);
    q[1] = gA.at(1, r0 + im);
    q[3] = gA.at(3, r0 + ip);
    if (!bot) {
      q[2] = gA.at(2, rm + i);
      q[5] = gA.at(5, rm + im);
      q[6] = gA.at(6, rm + ip);
    }
    if (!top) {
      q[4] = gA.at(4, rp + i);
      q[7] = gA.at(7, rp + ip);
      q[8] = gA.at(8, rp + im);
    }
    // Shifted form of the plate rule: g~_2 = -g~_4 + 2 w_4 (Tdown - T0), etc.
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
  } else {
    if (bot || top) {
      const size_t rs = bot ? r0 + g.pitch : r0 - g.pitch;
#pragma unroll
      for (int k = 0; k < 9; k++) q[k] = gA.at(k, rs + i);
    } else {
      const size_t rsm = (j == 1) ? r0 : rm;
      const size_t rsp = (j == g.ny - 2) ? r0 : rp;
      q[0] = gA.at(0, r0 + i);
      q[1] = gA.at(1, r0 + im);
      q[3] = gA.at(3, r0 + ip);
      q[2] = gA.at(2, rsm + i);
      q[5] = gA.at(5, rsm + im);
      q[6] = gA.at(6, rsm + ip);
      q[4] = gA.at(4, rsp + i);
      q[7] = gA.at(7, rsp + ip);
      q[8] = gA.at(8, rsp + im);
    }
  }

  float *fB = (float *)p.fB, *gB = (float *)p.gB;
  float *F = (float *)p.F;
  if (VARIANT == RBN_VARIANT_TRAJECTORIES) {
    float fo[9], go[9], rho, T, u, v, Fn;
    collide_t32(c, f, q, ld<NT>(&F[r0 + i]), fo, go, rho, T, u, v, Fn);
    flag_nonfinite(p.bad, rho, T);
    st<NT>(&F[r0 + i], Fn);
    st_uv<NT, float>(p.uv, r0 + i, u, v);
    if (p.rho) ((float *)p.rho)[r0 + i] = rho;
    if (p.T) ((float *)p.T)[r0 + i] = T;
#pragma unroll
    for (int k = 0; k < 9; k++) st<NT>(&fB[(size_t)k * g.plane + r0 + i], fo[k]);
#pragma unroll
    for (int k = 0; k < 9; k++) st<NT>(&gB[(size_t)k * g.plane + r0 + i], go[k]);
    return;
  }
  float dr = 0.f, dT = 0.f;
#pragma unroll
  for (int k = 0; k < 9; k++) dr += f[k];
#pragma unroll
  for (int k = 0; k < 9; k++) dT += q[k];
  bool nv = false;
  float Tb = 0.f, wn = 0.f;
  if (VARIANT == RBN_VARIANT_NUDGING && (p.tw || p.nvol)) {
    nv = nudge_input<float>(p, r0 + i, Tb, wn);
    if (nv) dT += 0.5f * ((const float *)p.nud)[r0 + i];
  }
  const float rho = c.rho0 + dr;
  const float T = c.T0 + dT;
  flag_nonfinite(p.bad, rho, T);
  const float rhoi = 1.f / rho;
  const float Fy = ld<NT>(&F[r0 + i]);
  const float jx = f[1] - f[3] + f[5] - f[6] - f[7] + f[8];
  const float jy = f[2] - f[4] + f[5] + f[6] - f[7] - f[8];
  const float u = jx * rhoi;
  const float v = jy * rhoi + 0.5f * rhoi * Fy;
  st<NT>(&F[r0 + i], c.g0 * rho * T);
  st_uv<NT, float>(p.uv, r0 + i, u, v);
  if (p.rho) ((float *)p.rho)[r0 + i] = rho;
  if (p.T) ((float *)p.T)[r0 + i] = T;

  float A[9];
  eq_dev32(u, v, A);
  // Guo factor with F_x = 0: ((c_y - v) F_y)/cs2 + ((c.u)(c_y F_y))/cs2^2.
  const float tz = -3.f * v * Fy, tp = 3.f * (1.f - v) * Fy, tm = 3.f * (-1.f - v) * Fy;
  const float d2 = 9.f * v * Fy, d5 = 9.f * (u + v) * Fy, d6 = 9.f * (v - u) * Fy;
  const float fac[9] = {tz, tz, tp + d2, tz, tm + d2, tp + d5, tp + d6, tm + d5, tm + d6};
  float feq[9], geq[9];
  eq_pops32(dr, rho, A, feq);
#pragma unroll
  for (int k = 0; k < 9; k++)
    st<NT>(&fB[(size_t)k * g.plane + r0 + i], f[k] * c.one_m_om + c.om * feq[k] + c.src[k] * fac[k]);
  eq_pops32(dT, T, A, geq);
  float Q = 0.f;
  if (nv) {
    Q = c.alpha * (wn * (T - Tb));
    ((float *)p.nud)[r0 + i] = -Q;
  }
#pragma unroll
  for (int k = 0; k < 9; k++)
    st<NT>(&gB[(size_t)k * g.plane + r0 + i], q[k] * c.one_m_omg + c.omg * geq[k] - c.bw[k] * Q);
}

// First node of this thread in lattice block L of N.  The lattice blocks enumerate row
// segments of blockDim.x * npt nodes, row major; the thread owns nodes i, i + blockDim.x,
// ..., npt of them, so the x-neighbour lines at the seams between its block's sub-segments
// stay inside one CU.  order bit 0 (XCD banding): block L goes to segment (L % 8) * (N/8) +
// L/8: the hardware deals consecutive blocks round-robin to the 8 XCDs, so this gives every
// XCD one contiguous band of rows and keeps the column halo of neighbouring segments in the
// same L2.  order bit 1 (reversed sweep): segments are visited last row first, so that a
// step that follows a forward sweep starts on the rows its predecessor wrote last.  order
// bit 2 (rows per XCD, when ny is a multiple of 8): block L sweeps segment (L/8) % gx of row
// 8 * (L/8 / gx) + L % 8, so every XCD owns one row in eight and visits the segments of its
// row in order: the x-neighbour lines at a seam stay in one L2 while all XCDs still work on
// the same eight rows.
__device__ __forceinline__ void segment_node(const Geom &g, int L, int N, int order, int npt, int &i,
                                             int &j) {
  const int seg = blockDim.x * npt;
  const int gx = (g.nx + seg - 1) / seg;
  if ((order & 1) && (N & 7) == 0) L = (L & 7) * (N >> 3) + (L >> 3);
  if (order & 2) L = N - 1 - L;
  if ((order & 4) && (g.ny & 7) == 0) {
    const int q = L >> 3;
    j = (q / gx) * 8 + (L & 7);
    i = (q - (q / gx) * gx) * seg + threadIdx.x;
    return;
  }
  j = L / gx;
  i = (L - j * gx) * seg + threadIdx.x;
}

}  // namespace lbm
}  // namespace rbn
