// rbn_lbm2_tile.h — device code of a two-step tile (temporal blocking in LDS), shared by the
// plain two-step launch (rbn_lbm2.hip) and the batched ensemble (rbn_ensemble.hip).
//
// A workgroup of kTileThreads threads owns a tile of kTW x TH nodes.  Phase 1 computes step s
// on the tile and a one-node ring around it (the ring is computed again by the neighbouring
// tiles), pulling from the step's input planes in HBM, and keeps in LDS the post-collision f
// and g that step s+1 pulls (all directions of the tile, the inward ones of the ring); it
// writes the step's velocity for the tile and keeps each tile node's next force F(s+1) in the
// registers of the thread that updates the node again.  Phase 2 computes step s+1 on the
// tile, pulling from LDS, and
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

// Pulled f of a node (walls: half-way bounce-back from the node's own pre-stream values),
// through an accessor a.f(k, di, dj) of the input at the node offset (di, dj).
template <typename R, class Acc>
__device__ __forceinline__ void gather_f(const Acc &a, bool bot, bool top, R f[9]) {
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
}

// Pulled populations of a truth-run node: f as gather_f, g with the plate rule at the walls.
template <typename R, class Acc>
__device__ __forceinline__ void gather_truth(const Acc &a, bool bot, bool top, const StepConst<R> &c, R f[9],
                                             R q[9]) {
  gather_f<R>(a, bot, top, f);
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

// Pulled populations of a nudged-run node: f as gather_f; g with the zero-gradient walls of
// lbm_node64's nudged branch (Nudging/code/tempvel_nudge.c:684-702, 783-962): a wall row takes
// the pre-stream populations of its inner neighbour row, unshifted, and the rows next to the
// walls (b1: row 1, t1: row ny-2) pull the directions that cross toward the wall from their own
// row, where the wall row's copy came from.
template <typename R, class Acc>
__device__ __forceinline__ void gather_nudged(const Acc &a, bool bot, bool top, bool b1, bool t1, R f[9], R q[9]) {
  gather_f<R>(a, bot, top, f);
  if (bot || top) {
    const int dj = bot ? 1 : -1;
#pragma unroll
    for (int k = 0; k < 9; k++) q[k] = a.g(k, 0, dj);
    return;
  }
  const int dm = b1 ? 0 : -1, dp = t1 ? 0 : 1;
  q[0] = a.g(0, 0, 0);
  q[1] = a.g(1, -1, 0);
  q[3] = a.g(3, 1, 0);
  q[2] = a.g(2, 0, dm);
  q[5] = a.g(5, -1, dm);
  q[6] = a.g(6, 1, dm);
  q[4] = a.g(4, 0, dp);
  q[7] = a.g(7, 1, dp);
  q[8] = a.g(8, -1, dp);
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
// (collide_t32).
__device__ __forceinline__ void collide32(const StepConst<float> &c, const float f[9], const float q[9],
                                          float Fy, float fo[9], float go[9], float &rho, float &T, float &u,
                                          float &v, float &Fn) {
  collide_t32(c, f, q, Fy, fo, go, rho, T, u, v, Fn);
}

// The nudging of one node in a step of the nudged run: its deposit count, mean and (bilinear
// deposit) weight, and the persisted term entering the step.
// A node's nudging input (nudge_input): on where it has contributions, the mean and weight.
template <typename R>
struct NodeNudge {
  bool on;
  R Tb, w;
};

// Moments and collision of a nudged-run node: lbm_node64's nudged arithmetic (half the
// persisted term in T where the node receives a deposit; the nudged collisg), operand for
// operand.  nud: the term entering the step; nud_out: after it (unchanged without a deposit).
__device__ __forceinline__ void collide_n64(const StepConst<double> &c, const double f[9], const double q[9],
                                            double Fy, const NodeNudge<double> &d, double nud, double fo[9], double go[9],
                                            double &rho, double &T, double &u, double &v, double &Fn,
                                            double &nud_out) {
  rho = 0.0;
  T = 0.0;
#pragma unroll
  for (int k = 0; k < 9; k++) rho = rho + f[k];
#pragma unroll
  for (int k = 0; k < 9; k++) T = T + q[k];
  if (d.on) T = (T + 0.5 * nud);
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
  if (d.on) {
    const double Q = c.alpha * (d.w * (T - d.Tb));
#pragma unroll
    for (int k = 0; k < 9; k++) {
      const double geq = T * kW[k] * b[k];
      go[k] = q[k] * c.one_m_omg + c.omg * geq - c.bw[k] * Q;
    }
    nud_out = -1. * Q;
  } else {
#pragma unroll
    for (int k = 0; k < 9; k++) {
      const double geq = T * kW[k] * b[k];
      go[k] = q[k] * c.one_m_omg + c.omg * geq;
    }
    nud_out = nud;
  }
}

// lbm_node32's nudged arithmetic with shifted populations.
__device__ __forceinline__ void collide_n32(const StepConst<float> &c, const float f[9], const float q[9], float Fy,
                                            const NodeNudge<float> &d, float nud, float fo[9], float go[9], float &rho,
                                            float &T, float &u, float &v, float &Fn, float &nud_out) {
  float dr = 0.f, dT = 0.f;
#pragma unroll
  for (int k = 0; k < 9; k++) dr += f[k];
#pragma unroll
  for (int k = 0; k < 9; k++) dT += q[k];
  if (d.on) dT += 0.5f * nud;
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
  float Q = 0.f;
  nud_out = nud;
  if (d.on) {
    Q = c.alpha * (d.w * (T - d.Tb));
    nud_out = -Q;
  }
  eq_pops32(dT, T, A, geq);
#pragma unroll
  for (int k = 0; k < 9; k++) go[k] = q[k] * c.one_m_omg + c.omg * geq[k] - c.bw[k] * Q;
}

__device__ __forceinline__ void collide(const StepConst<double> &c, const double f[9], const double q[9],
                                        double Fy, double fo[9], double go[9], double &rho, double &T,
                                        double &u, double &v, double &Fn) {
  collide64(c, f, q, Fy, fo, go, rho, T, u, v, Fn);
}
__device__ __forceinline__ void collide_n(const StepConst<double> &c, const double f[9], const double q[9],
                                          double Fy, const NodeNudge<double> &d, double nud, double fo[9], double go[9],
                                          double &rho, double &T, double &u, double &v, double &Fn, double &nud_out) {
  collide_n64(c, f, q, Fy, d, nud, fo, go, rho, T, u, v, Fn, nud_out);
}
__device__ __forceinline__ void collide_n(const StepConst<float> &c, const float f[9], const float q[9], float Fy,
                                          const NodeNudge<float> &d, float nud, float fo[9], float go[9], float &rho,
                                          float &T, float &u, float &v, float &Fn, float &nud_out) {
  collide_n32(c, f, q, Fy, d, nud, fo, go, rho, T, u, v, Fn, nud_out);
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
  // nudged run: the deposits of steps s and s+1 (packed, DepPtrs::tw) and the persisted
  // nudging term entering step s and after step s+1 (double-buffered: other tiles still read
  // the input at the nodes of this tile's ring)
  const void *tw[2];
  // or, when bin2 is set (bilinear deposit, tiles of kDepTileH rows: every tile is one
  // deposit tile), step s+1's deposit summed from its bins inside the launch (DepPtrs' tcnt,
  // bin, bin_cap, S, Wi, ocnt of its set), tw[1] unused
  const long long *bin2;
  int *tcnt2;
  int bin_cap2;
  long long *S2, *Wi2;
  int *ocnt2;
  const void *nud_in;
  void *nud_out;
};

// LDS of a two-step tile: the post-collision f and g of step s that step s+1 pulls.  Local
// coordinates: rows 0 .. TH+1 and columns 0 .. kTW+1, the tile at rows 1 .. TH and columns
// 1 .. kTW, the one-node ring around it.  Tile nodes keep all nine directions; a ring node
// only the three that point into the tile: the row below (row 0) directions 2, 5, 6, the row
// above (row TH+1) 4, 7, 8, the left column (column 0) 1, 5, 8, the right column (column
// kTW+1) 3, 6, 7.  In the nudged run the ring rows keep all nine directions of g: a wall row
// copies its inner neighbour row, which for a one-row tile at the top wall is the row below.
// A tile cut short by the lattice edge (nrows < TH or ncols < kTW) has its upper or right
// ring inside the tile rows or columns, where all nine directions are kept.
// FP32 at TH = 8: 40.4 KB for the truth run; FP64 at TH = 4: 43.6 KB.
// GR: g directions kept in the ring rows, 3 (truth run) or 9 (nudged run, or an ensemble
// launch that may hold both).
template <int V>
constexpr int ring_g() { return V == RBN_VARIANT_NUDGING ? 9 : 3; }

template <typename R, int TH, int GR_>
struct TileLds {
  static constexpr int RW = kTW + 2;
  static constexpr int GR = GR_;
  R t[2][9][TH][kTW];  // [f/g][k][row-1][col-1]
  R rbf[3][RW];        // row 0, f: directions 2, 5, 6
  R raf[3][RW];        // row TH+1, f: directions 4, 7, 8
  R rbg[GR][RW];       // row 0, g: 2, 5, 6 (or all nine)
  R rag[GR][RW];       // row TH+1, g: 4, 7, 8 (or all nine)
  R cl[2][3][TH];      // column 0, rows 1 .. TH: directions 1, 5, 8
  R cr[2][3][TH];      // column kTW+1, rows 1 .. TH: directions 3, 6, 7
  // nudged run, tiles of kDepTileH rows: the second step's deposit per tile node (mean, weight
  // or -1) when the launch sums it from the bins (kept here, not in registers, across the
  // first step)
  Real2<R> d2[(GR == 9 && TH == kDepTileH) ? kTW * TH : 1];
};

// Slot of direction k in each ring store (directions that never reach the tile from there
// map to 0: no access to them is ever made).
__device__ __forceinline__ constexpr int slot_below(int k) { return k == 5 ? 1 : (k == 6 ? 2 : 0); }
__device__ __forceinline__ constexpr int slot_above(int k) { return k == 7 ? 1 : (k == 8 ? 2 : 0); }
__device__ __forceinline__ constexpr int slot_left(int k) { return k == 5 ? 1 : (k == 8 ? 2 : 0); }
__device__ __forceinline__ constexpr int slot_right(int k) { return k == 6 ? 1 : (k == 7 ? 2 : 0); }

// The step-s populations in LDS at local node (r, cc) of the tile: a.f(k, di, dj) is
// direction k of the node (r + dj, cc + di).
template <typename R, int TH, int GR>
struct LdsAcc {
  using L_t = TileLds<R, TH, GR>;
  const L_t &L;
  int r, cc;
  __device__ __forceinline__ R at(int w, int k, int di, int dj) const {
    const int rs = r + dj, cs = cc + di;
    if (rs == 0) return w ? L.rbg[L_t::GR == 9 ? k : slot_below(k)][cs] : L.rbf[slot_below(k)][cs];
    if (rs == TH + 1) return w ? L.rag[L_t::GR == 9 ? k : slot_above(k)][cs] : L.raf[slot_above(k)][cs];
    if (cs == 0) return L.cl[w][slot_left(k)][rs - 1];
    if (cs == kTW + 1) return L.cr[w][slot_right(k)][rs - 1];
    return L.t[w][k][rs - 1][cs - 1];
  }
  __device__ __forceinline__ R f(int k, int di, int dj) const { return at(0, k, di, dj); }
  __device__ __forceinline__ R g(int k, int di, int dj) const { return at(1, k, di, dj); }
};

// Store the post-collision populations of local node (r, cc) where step s+1 reads them.
template <typename R, int TH, int GR>
__device__ __forceinline__ void keep_node(TileLds<R, TH, GR> &L, int r, int cc, const R fo[9], const R go[9]) {
  constexpr bool all_g = GR == 9;
  if (r == 0) {
    L.rbf[0][cc] = fo[2], L.rbf[1][cc] = fo[5], L.rbf[2][cc] = fo[6];
    if (all_g) {
#pragma unroll
      for (int k = 0; k < 9; k++) L.rbg[k][cc] = go[k];
    } else {
      L.rbg[0][cc] = go[2], L.rbg[1][cc] = go[5], L.rbg[2][cc] = go[6];
    }
  } else if (r == TH + 1) {
    L.raf[0][cc] = fo[4], L.raf[1][cc] = fo[7], L.raf[2][cc] = fo[8];
    if (all_g) {
#pragma unroll
      for (int k = 0; k < 9; k++) L.rag[k][cc] = go[k];
    } else {
      L.rag[0][cc] = go[4], L.rag[1][cc] = go[7], L.rag[2][cc] = go[8];
    }
  } else if (cc == 0) {
    L.cl[0][0][r - 1] = fo[1], L.cl[0][1][r - 1] = fo[5], L.cl[0][2][r - 1] = fo[8];
    L.cl[1][0][r - 1] = go[1], L.cl[1][1][r - 1] = go[5], L.cl[1][2][r - 1] = go[8];
  } else if (cc == kTW + 1) {
    L.cr[0][0][r - 1] = fo[3], L.cr[0][1][r - 1] = fo[6], L.cr[0][2][r - 1] = fo[7];
    L.cr[1][0][r - 1] = go[3], L.cr[1][1][r - 1] = go[6], L.cr[1][2][r - 1] = go[7];
  } else {
#pragma unroll
    for (int k = 0; k < 9; k++) {
      L.t[0][k][r - 1][cc - 1] = fo[k];
      L.t[1][k][r - 1][cc - 1] = go[k];
    }
  }
}

// The nudging of node o in step u of the launch (0: s, 1: s+1).
template <typename R>
__device__ __forceinline__ NodeNudge<R> node_nudge(const TwoStepPtrs &a, int u, size_t o) {
  const Real2<R> v = ((const Real2<R> *)a.tw[u])[o];
  return {__builtin_signbit(v.y) != 0, v.x, -v.y};
}

// Step s at local node (r, cc) of tile (i0, j0), pulled from the input planes in HBM; its
// populations go to LDS; returns F(s+1) in Fn and (nudged run) the nudging term after the
// step in nud, and writes the velocity of a tile node.
template <typename R, int TH, int V, int GR>
__device__ __forceinline__ void first_step(const Geom &g, const TwoStepPtrs &a, const StepConst<R> &c, int i0,
                                           int j0, int r, int cc, bool tile, TileLds<R, TH, GR> &L, R &Fn, R &nud) {
  static_assert(V != RBN_VARIANT_NUDGING || GR == 9, "the nudged run keeps all g directions of the ring rows");
  const int y = j0 - 1 + r;
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
  const size_t o = acc.r0 + x;
  R f[9], q[9], fo[9], go[9], rho, T, u, v;
  const R Fy = ((const R *)a.Fin)[o];
  if (V == RBN_VARIANT_NUDGING) {
    gather_nudged<R>(acc, bot, top, y == 1, y == g.ny - 2, f, q);
    const NodeNudge<R> d = node_nudge<R>(a, 0, o);
    const R nud_in = d.on ? ((const R *)a.nud_in)[o] : R(0);
    collide_n(c, f, q, Fy, d, nud_in, fo, go, rho, T, u, v, Fn, nud);
    if (tile && !d.on) nud = ((const R *)a.nud_in)[o];  // persists where no deposit lands
  } else {
    gather_truth<R>(acc, bot, top, c, f, q);
    collide(c, f, q, Fy, fo, go, rho, T, u, v, Fn);
  }
  keep_node<R, TH, GR>(L, r, cc, fo, go);
  if (tile) st_uv<0, R>(a.uv1, o, u, v);
}

// Step s+1's bilinear deposit of tile (tx, ty) (kDepTileH rows) from its bins, as the tile
// pass sums it (rbn_nudge.hip: fixed-point sums in any order, the overflow's sums in HBM
// added and cleared, the counter left at zero), in the LDS the first step has not yet
// filled; the mean and weight of each tile node (w = -1 where it has none) go to L.d2.
template <typename R, int TH, int GR>
__device__ __forceinline__ void bin_deposit(const Geom &g, const TwoStepPtrs &a, int tx, int ty,
                                            TileLds<R, TH, GR> &L) {
  constexpr int kNodes = kTW * TH, NQ = kNodes / kTileThreads;
  static_assert(kNodes == kDepTileW * kDepTileH && NQ * kTileThreads == kNodes, "one deposit tile per tile");
  static_assert(sizeof(L.d2) == kNodes * sizeof(Real2<R>), "tile deposit store");
  static_assert(kNodes * (2 * sizeof(long long) + sizeof(int)) <= sizeof(L.t), "LDS scratch");
  auto *S = reinterpret_cast<unsigned long long *>(&L.t[0][0][0][0]);
  auto *W = S + kNodes;
  auto *n = reinterpret_cast<int *>(W + kNodes);
  const int t = ty * dep_tiles_x(g) + tx;
  const longlong2 *e = (const longlong2 *)a.bin2 + (size_t)t * a.bin_cap2;
  const int cnt = a.tcnt2[(size_t)t * kTcntStride];
  const int k0 = threadIdx.x, k1 = threadIdx.x + kTileThreads;
  longlong2 v0 = make_longlong2(0, 0), v1 = make_longlong2(0, 0);
  if (k0 < a.bin_cap2) v0 = e[k0];  // slots past the count: stale, read but not used
  if (k1 < a.bin_cap2) v1 = e[k1];
  for (int k = threadIdx.x; k < kNodes; k += kTileThreads) {
    S[k] = 0;
    W[k] = 0;
    n[k] = 0;
  }
  __syncthreads();
  if (threadIdx.x == 0) a.tcnt2[(size_t)t * kTcntStride] = 0;  // every thread has read it
  const int m = min(cnt, a.bin_cap2);
  auto add = [&](const longlong2 &v) {
    const int local = (int)(v.y & (kNodes - 1));
    atomicAdd(&S[local], (unsigned long long)v.x);
    atomicAdd(&W[local], (unsigned long long)(v.y >> kDepTileBits));
    atomicAdd(&n[local], 1);
  };
  if (k0 < m) add(v0);
  if (k1 < m) add(v1);
  for (int k = k1 + kTileThreads; k < m; k += kTileThreads) add(e[k]);
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    const int k = threadIdx.x + q * kTileThreads;
    const int i = tx * kTW + k % kTW, j = ty * TH + k / kTW;
    long long s = (long long)S[k], wi = (long long)W[k];
    int c = n[k];
    if (cnt > a.bin_cap2 && i < g.nx && j < g.ny) {
      const size_t o = (size_t)j * g.pitch + i;
      if (a.ocnt2[o] != 0) {
        c += a.ocnt2[o];
        s += a.S2[o];
        wi += a.Wi2[o];
        a.ocnt2[o] = 0;
        a.S2[o] = 0;
        a.Wi2[o] = 0;
      }
    }
    Real2<R> d;
    if (c > 0) {
      const double Tbar = wi != 0 ? (double)s / (double)wi : 0.0, Wd = (double)wi * 0x1p-40;
      d.x = (R)Tbar;
      d.y = (R)nudge_weight(Wd);
    } else {
      d.x = R(0);
      d.y = R(-1);
    }
    L.d2[k] = d;
  }
  __syncthreads();  // the first step overwrites the sums
}

// Both steps on tile (tx, ty) of kTW x TH nodes; every thread of the workgroup calls it.
// Phase 1 computes step s on the tile and its ring: first the tile nodes, thread t taking
// nodes t, t + kTileThreads, ... of the tile in row order (the same nodes it updates in phase 2,
// so their next force and nudging term stay in its registers), then the ring nodes, one per
// thread.  Phase 2 computes step s+1 on the tile from LDS and writes the output planes, with
// non-temporal stores (NT = 1: lattices larger than the caches) or ordinary ones (NT = 0:
// the next launch finds them in the caches).
template <typename R, int TH, int V, int GR, int NT = 1>
__device__ __forceinline__ void two_step_tile(const Geom &g, const TwoStepPtrs &a, const StepConst<R> &c, int tx,
                                              int ty, TileLds<R, TH, GR> &L) {
  constexpr int NQ = kTW * TH / kTileThreads;  // tile nodes per thread
  constexpr int NRING = 2 * (kTW + 2) + 2 * TH;
  static_assert(kTW * TH % kTileThreads == 0 && NRING <= kTileThreads, "tile shape");
  const int i0 = tx * kTW, j0 = ty * TH;
  const int ncols = min(kTW, g.nx - i0), nrows = min(TH, g.ny - j0);
  R Fk[NQ], Nk[NQ];
  // step s+1's deposit, when summed here from its bins
  bool binned = false;
  if constexpr (V == RBN_VARIANT_NUDGING && TH == kDepTileH) {
    if (a.bin2) {
      bin_deposit<R, TH, GR>(g, a, tx, ty, L);
      binned = true;
    }
  }

  // Phase 1: step s.  Tile rows / columns up to nrows+1 / ncols+1 (the ring of a cut-short
  // tile lies there).
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    const int n = threadIdx.x + q * kTileThreads;
    const int r = n / kTW + 1, cc = n % kTW + 1;
    Fk[q] = R(0);
    Nk[q] = R(0);
    if (r > nrows + 1 || cc > ncols + 1 || j0 - 1 + r >= g.ny) continue;
    first_step<R, TH, V, GR>(g, a, c, i0, j0, r, cc, r <= nrows && cc <= ncols, L, Fk[q], Nk[q]);
  }
  if (threadIdx.x < NRING) {
    const int t = threadIdx.x;
    int r, cc;
    if (t < kTW + 2) {
      r = 0, cc = t;
    } else if (t < 2 * (kTW + 2)) {
      r = TH + 1, cc = t - (kTW + 2);
    } else if (t < 2 * (kTW + 2) + TH) {
      r = t - 2 * (kTW + 2) + 1, cc = 0;
    } else {
      r = t - 2 * (kTW + 2) - TH + 1, cc = kTW + 1;
    }
    const int y = j0 - 1 + r;
    if (r <= nrows + 1 && cc <= ncols + 1 && y >= 0 && y < g.ny) {
      R Fn, nud;
      first_step<R, TH, V, GR>(g, a, c, i0, j0, r, cc, false, L, Fn, nud);
    }
  }
  __syncthreads();

  // Phase 2: step s+1 on the tile, from LDS.
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    const int n = threadIdx.x + q * kTileThreads;
    const int r = n / kTW + 1, cc = n % kTW + 1;
    if (cc > ncols || r > nrows) continue;
    const int x = i0 - 1 + cc, y = j0 - 1 + r;
    const bool bot = (y == 0), top = (y == g.ny - 1);
    const size_t o = (size_t)y * g.pitch + x;
    LdsAcc<R, TH, GR> acc{L, r, cc};
    R f[9], q9[9], fo[9], go[9], rho, T, u, v, Fn;
    if (V == RBN_VARIANT_NUDGING) {
      gather_nudged<R>(acc, bot, top, y == 1, y == g.ny - 2, f, q9);
      R nud;
      NodeNudge<R> d2;
      if constexpr (V == RBN_VARIANT_NUDGING && TH == kDepTileH) {
        if (binned) {
          const Real2<R> v = L.d2[n];
          d2 = {v.y >= R(0), v.x, v.y};
        } else {
          d2 = node_nudge<R>(a, 1, o);
        }
      } else {
        d2 = node_nudge<R>(a, 1, o);
      }
      collide_n(c, f, q9, Fk[q], d2, Nk[q], fo, go, rho, T, u, v, Fn, nud);
      ((R *)a.nud_out)[o] = nud;
    } else {
      gather_truth<R>(acc, bot, top, c, f, q9);
      collide(c, f, q9, Fk[q], fo, go, rho, T, u, v, Fn);
    }
    flag_nonfinite(a.bad, rho, T);
    R *fB = (R *)a.fB, *gB = (R *)a.gB;
#pragma unroll
    for (int k = 0; k < 9; k++) st<NT>(&fB[(size_t)k * g.plane + o], fo[k]);
#pragma unroll
    for (int k = 0; k < 9; k++) st<NT>(&gB[(size_t)k * g.plane + o], go[k]);
    st<NT>(&((R *)a.Fout)[o], Fn);
    st_uv<0, R>(a.uv2, o, u, v);
    if (a.rho2) ((R *)a.rho2)[o] = rho;
    if (a.T2) ((R *)a.T2)[o] = T;
  }
}

}  // namespace lbm2
}  // namespace rbn
