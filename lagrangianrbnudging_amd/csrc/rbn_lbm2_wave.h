// rbn_lbm2_wave.h — two truth-run steps per launch by a row sweep (wavefront temporal
// blocking), the alternative to the two-step tile of rbn_lbm2_tile.h.
//
// A workgroup of SW threads owns a strip of SW columns and a segment of rows
// [ja, jb).  It first computes step s on the two columns beside the strip for rows ja-1 .. jb
// (the halo, kept in LDS), then sweeps the rows upwards: iteration i computes step s on row
// jr = ja-1+i of the strip (one column per thread, pulling from the input planes in HBM) and
// step s+1 on row jr-1, which pulls from step s on rows jr-2 .. jr.  Each post-collision
// value of step s is pulled by exactly one node of step s+1, so a thread keeps the values its
// own column needs (directions 0, 2 and 4, and the wall rows' bounce-back sources) in
// registers and passes the six directions that move sideways (1, 3, 5, 6, 7, 8) through one
// LDS row, written when the neighbour needs them: 7 and 8 of row jr, 1 and 3 of row jr-1
// (held one iteration), 5 and 6 of row jr-2 (held two).  The row is double-buffered, so an
// iteration has one barrier.
//
// Per node and two steps the HBM traffic is one read and one write of the populations plus
// (SEG+2)/SEG for the rows recomputed at the segment ends and the strip's halo columns,
// against the ring of a 64 x TH tile ((66 (TH+2)) / (64 TH)) in the tile form.  The node
// arithmetic is the tile form's (gather_truth, collide), so the result is bit for bit two
// one-step launches.  At 4096 x 2048 in FP32 the sweep fetches 0.80 GB per launch against
// the tile form's 1.27 GB (both write 0.77 GB) but takes as long: one row of pulls in flight
// per workgroup (19 loads per thread, 3-4 workgroups per CU) does not cover the HBM latency
// that the tile form's phase 1 (up to 50 loads per thread in flight) does.
#pragma once

#include "rbn_lbm2_tile.h"

namespace rbn {
namespace lbm2 {

constexpr int kMaxSeg = 64;  // rows per segment, at most

// Slots of the exchanged directions in the LDS row: f then g.
enum { X1 = 0, X3, X5, X6, X7, X8, kNX };

template <typename R, int SW>
struct WaveLds {
  static constexpr int W = SW + 2;  // slot 0: left halo, 1..ncols: the strip, ncols+1: right halo
  R row[2][2 * kNX][W];
  R halo[2][kMaxSeg + 2][6];  // left: f1 f5 f8 g1 g5 g8; right: f3 f6 f7 g3 g6 g7
};

// The pulled populations of step s at node (x, y) and its force F(s).
template <typename R>
struct Pulled {
  R f[9], q[9], Fy;
};

template <typename R>
__device__ __forceinline__ void pull_first(const Geom &g, const TwoStepPtrs &a, const StepConst<R> &c, int x, int y,
                                           Pulled<R> &p) {
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
  gather_truth<R>(acc, bot, top, c, p.f, p.q);
  p.Fy = ((const R *)a.Fin)[acc.r0 + x];
}

// Step s at node (x, y): pulled from the input planes in HBM, collided; the post-collision
// populations, the velocity and the next force F(s+1).
template <typename R>
__device__ __forceinline__ void first_step_node(const Geom &g, const TwoStepPtrs &a, const StepConst<R> &c, int x,
                                                int y, R fo[9], R go[9], R &u, R &v, R &Fn) {
  Pulled<R> p;
  R rho, T;
  pull_first<R>(g, a, c, x, y, p);
  collide(c, p.f, p.q, p.Fy, fo, go, rho, T, u, v, Fn);
}

// Step s+1 at (x, j): the pulls of gather_truth served from the registers of the thread's
// own column (step s at rows j-1, j, j+1) and from the LDS row (the neighbours' sideways
// directions).  Every (direction, offset) pair gather_truth asks for is one of those below.
template <typename R, int SW>
struct WaveAcc {
  const R (&Pf)[9];  // step s, row j, own column
  const R (&Pg)[9];
  R ppf2, ppg2;      // step s, row j-1, direction 2
  R cf4, cg4;        // step s, row j+1, direction 4
  const R (*row)[WaveLds<R, SW>::W];
  int s;  // this column's slot
  __device__ __forceinline__ R pull(const R (&P)[9], R pp2, R c4, int base, int k, int di, int dj) const {
    if (di == 0 && dj == 0) return P[k];
    if (dj == 0) return di < 0 ? row[base + X1][s - 1] : row[base + X3][s + 1];
    if (dj < 0) {
      if (di == 0) return pp2;
      return di < 0 ? row[base + X5][s - 1] : row[base + X6][s + 1];
    }
    if (di == 0) return c4;
    return di > 0 ? row[base + X7][s + 1] : row[base + X8][s - 1];
  }
  __device__ __forceinline__ R f(int k, int di, int dj) const { return pull(Pf, ppf2, cf4, 0, k, di, dj); }
  __device__ __forceinline__ R g(int k, int di, int dj) const { return pull(Pg, ppg2, cg4, kNX, k, di, dj); }
};

// Both steps on strip sx, segment [ja, jb) (jb - ja <= kMaxSeg); every thread of the
// workgroup calls it.  AHEAD: the pulls of a row are issued one iteration ahead.
template <typename R, int AHEAD, int SW>
__device__ __forceinline__ void two_step_strip(const Geom &g, const TwoStepPtrs &a, const StepConst<R> &c, int sx,
                                               int ja, int jb, WaveLds<R, SW> &L) {
  const int t = threadIdx.x;
  const int c0 = sx * SW;
  const int ncols = min(SW, g.nx - c0);
  const bool live = t < ncols;
  const int x = c0 + t, s = t + 1;

  // The halo: step s on the columns left and right of the strip, rows ja-1 .. jb.
  const int hrows = jb - ja + 2;
  for (int q = t; q < 2 * hrows; q += SW) {
    const int side = q >= hrows, r = q - side * hrows, y = ja - 1 + r;
    if (y < 0 || y >= g.ny) continue;
    int hx = side ? c0 + ncols : c0 - 1;
    hx = hx < 0 ? hx + g.nx : (hx >= g.nx ? hx - g.nx : hx);
    R fo[9], go[9], u, v, Fn;
    first_step_node<R>(g, a, c, hx, y, fo, go, u, v, Fn);
    R *h = L.halo[side][r];
    if (!side) {
      h[0] = fo[1], h[1] = fo[5], h[2] = fo[8], h[3] = go[1], h[4] = go[5], h[5] = go[8];
    } else {
      h[0] = fo[3], h[1] = fo[6], h[2] = fo[7], h[3] = go[3], h[4] = go[6], h[5] = go[7];
    }
  }

  R Pf[9], Pg[9], PF = R(0);             // step s, row jr-1
  R ppf[3], ppg[3];                      // step s, row jr-2: directions 2, 5, 6
#pragma unroll
  for (int k = 0; k < 9; k++) Pf[k] = Pg[k] = R(0);
#pragma unroll
  for (int k = 0; k < 3; k++) ppf[k] = ppg[k] = R(0);
  R *const fB = (R *)a.fB, *const gB = (R *)a.gB;

  // The pulls of row jr are issued one iteration ahead, so that they are in flight across
  // the iteration's barrier.  Rows outside the lattice and columns past the strip's last
  // are pulled from a clamped node and never used.
  const int xl = min(x, g.nx - 1);
  Pulled<R> cur;
  if (AHEAD) pull_first<R>(g, a, c, xl, max(ja - 1, 0), cur);

  for (int i = 0; i < hrows; i++) {
    const int jr = ja - 1 + i, j = jr - 1;
    Pulled<R> nxt;
    if (AHEAD)
      pull_first<R>(g, a, c, xl, min(jr + 1, g.ny - 1), nxt);
    else
      pull_first<R>(g, a, c, xl, min(max(jr, 0), g.ny - 1), cur);
    R Cf[9], Cg[9], CF = R(0);
    if (live && jr >= 0 && jr < g.ny) {
      R u, v, rho, T;
      collide(c, cur.f, cur.q, cur.Fy, Cf, Cg, rho, T, u, v, CF);
      if (jr >= ja && jr < jb) st_uv<0, R>(a.uv1, (size_t)jr * g.pitch + x, u, v);
    }
    constexpr int W = WaveLds<R, SW>::W;
    R(*row)[W] = L.row[i & 1];
    if (live) {
      row[X1][s] = Pf[1], row[X3][s] = Pf[3], row[X5][s] = ppf[1], row[X6][s] = ppf[2];
      row[X7][s] = Cf[7], row[X8][s] = Cf[8];
      row[kNX + X1][s] = Pg[1], row[kNX + X3][s] = Pg[3], row[kNX + X5][s] = ppg[1], row[kNX + X6][s] = ppg[2];
      row[kNX + X7][s] = Cg[7], row[kNX + X8][s] = Cg[8];
    }
    const bool second = i >= 2;  // step s+1 on row j in [ja, jb)
    if (second && t == 0) {
      // left halo: direction 1 of row j, 5 of row j-1, 8 of row j+1 (halo row r = y - ja + 1)
      const int r = j - ja + 1;
      row[X1][0] = L.halo[0][r][0], row[X5][0] = L.halo[0][r - 1][1], row[X8][0] = L.halo[0][r + 1][2];
      row[kNX + X1][0] = L.halo[0][r][3], row[kNX + X5][0] = L.halo[0][r - 1][4];
      row[kNX + X8][0] = L.halo[0][r + 1][5];
    }
    if (second && t == ncols - 1) {
      const int r = j - ja + 1, e = ncols + 1;
      row[X3][e] = L.halo[1][r][0], row[X6][e] = L.halo[1][r - 1][1], row[X7][e] = L.halo[1][r + 1][2];
      row[kNX + X3][e] = L.halo[1][r][3], row[kNX + X6][e] = L.halo[1][r - 1][4];
      row[kNX + X7][e] = L.halo[1][r + 1][5];
    }
    __syncthreads();
    if (second && live) {
      const WaveAcc<R, SW> acc{Pf, Pg, ppf[0], ppg[0], Cf[4], Cg[4], row, s};
      R f[9], q[9], fo[9], go[9], rho, T, u, v, Fn;
      gather_truth<R>(acc, j == 0, j == g.ny - 1, c, f, q);
      collide(c, f, q, PF, fo, go, rho, T, u, v, Fn);
      flag_nonfinite(a.bad, rho, T);
      const size_t o = (size_t)j * g.pitch + x;
#pragma unroll
      for (int k = 0; k < 9; k++) st<1>(&fB[(size_t)k * g.plane + o], fo[k]);
#pragma unroll
      for (int k = 0; k < 9; k++) st<1>(&gB[(size_t)k * g.plane + o], go[k]);
      st<1>(&((R *)a.Fout)[o], Fn);
      st_uv<0, R>(a.uv2, o, u, v);
      if (a.rho2) ((R *)a.rho2)[o] = rho;
      if (a.T2) ((R *)a.T2)[o] = T;
    }
    ppf[0] = Pf[2], ppf[1] = Pf[5], ppf[2] = Pf[6];
    ppg[0] = Pg[2], ppg[1] = Pg[5], ppg[2] = Pg[6];
#pragma unroll
    for (int k = 0; k < 9; k++) Pf[k] = Cf[k], Pg[k] = Cg[k];
    PF = CF;
    if (AHEAD) cur = nxt;
  }
}

}  // namespace lbm2
}  // namespace rbn
