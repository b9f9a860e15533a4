// rbn_lbm3_wave.h — three truth-run steps per launch by a row sweep (wavefront temporal
// blocking three steps deep).
//
// A workgroup of SW threads owns a strip of output columns [ox0, ox1) (at most SW - 4) and a
// segment of rows [ja, jb).  Thread t holds column ox0 - 2 + t (periodic in x): the step
// count of a column shrinks by one on each side per step, so the strip's threads compute
// step s on columns ox0-2 .. ox1+1, step s+1 on ox0-1 .. ox1 and step s+2 on the output
// columns, and neighbouring strips overlap by four columns instead of exchanging halos.
// Iteration i of the upward sweep computes step s on row r = ja-2+i (pulling from the input
// planes in HBM), step s+1 on row r-1 and step s+2 on row r-2.  A level passes its
// post-collision values to the next one: the rows a thread's own column needs stay in its
// registers, the six directions that move sideways go through LDS, written when the row is
// computed (Lds3), so that no thread holds its neighbours' values across iterations; an
// iteration has two barriers, one after each of the first two levels.
//
// Per node and three steps the HBM traffic is one read and one write of the populations and
// the force, three (u, v) stores, plus (seg+6)/seg row reads at the segment ends and four
// columns per strip: about 59 B per FP32 node-step against 84 B with two steps per launch.
// The node arithmetic is the tile form's (gather_truth, collide), so a three-step launch
// gives bit for bit three one-step launches.
#pragma once

#include "rbn_lbm2_wave.h"

namespace rbn {
namespace lbm3 {

using namespace lbm2;

// The planes of a three-step launch: step s reads fA, gA and F(s); step s+2 writes fB, gB
// (step s+3's inputs) and F(s+3); uv[u] receives the velocity of step s+u, rho / T the
// moments of step s+mlev (mlev 0, 1 or 2; -1: none); bad (or null) the non-finite flag of
// step s+2's moments.
struct ThreeStepPtrs {
  const void *fA, *gA;
  void *fB, *gB;
  const void *Fin;
  void *Fout;
  void *uv[3];
  void *rho, *T;
  int mlev;
  int *bad;
};

// The sideways-moving post-collision values a level hands to the next, written by the
// producing thread t into slot t+1 the iteration it computes its row and read by the
// neighbouring columns' threads: directions 7 and 8 in the same iteration (the consumer's
// row j+1), 1 and 3 in the next (row j, by iteration parity), 5 and 6 two iterations later
// (row j-1, by iteration modulo 3).  [e]: the level boundary (0: step s -> s+1, 1: s+1 ->
// s+2); [4]: f of the first and second direction, then g of both.  Every slot is written
// before one of the iteration's barriers and read between that barrier and the next write of
// the slot.  49.5 KB in FP32: three workgroups per CU, as many as the registers allow (four,
// with directions 5 and 6 written a row late into 40 KB and the registers held to 128,
// measured 5 % slower).
template <typename R, int SW>
struct Lds3 {
  static constexpr int W = SW + 2;
  R up[2][4][W];
  R side[2][2][4][W];
  R down[2][3][4][W];
};

// A level's output row in the registers of one column: all nine directions of f and g and
// the force it computed for the next level.  Each level keeps two, written in turn (the
// sweep is unrolled by two), so that the row before stays readable as the next level's own
// row without copies.
template <typename R>
struct Row {
  R f[9], g[9], F;
};

template <typename R>
__device__ __forceinline__ void row_zero(Row<R> &r) {
#pragma unroll
  for (int k = 0; k < 9; k++) r.f[k] = r.g[k] = R(0);
  r.F = R(0);
}

// Hand the sideways directions of a freshly computed row to the neighbours.
template <typename R, int SW>
__device__ __forceinline__ void put_row(Lds3<R, SW> &L, int e, int par, int m3, int s, const Row<R> &C) {
  L.up[e][0][s] = C.f[7], L.up[e][1][s] = C.f[8], L.up[e][2][s] = C.g[7], L.up[e][3][s] = C.g[8];
  auto &sd = L.side[e][par];
  sd[0][s] = C.f[1], sd[1][s] = C.f[3], sd[2][s] = C.g[1], sd[3][s] = C.g[3];
  auto &dn = L.down[e][m3];
  dn[0][s] = C.f[5], dn[1][s] = C.f[6], dn[2][s] = C.g[5], dn[3][s] = C.g[6];
}

// The pulls of gather_truth for a node of level e+1 at row j: its own column from the
// producer's row j (P: direction 0 and the wall rows' bounce-back sources), direction 2 of
// row j-1 (p2) and direction 4 of row j+1 (c4), the sideways directions from the
// neighbours' slots.  Every (direction, offset) pair gather_truth asks for is one of these.
template <typename R, int SW>
struct Acc3 {
  const Row<R> &P;
  R p2f, p2g, c4f, c4g;
  const R (*up)[Lds3<R, SW>::W];
  const R (*side)[Lds3<R, SW>::W];
  const R (*down)[Lds3<R, SW>::W];
  int sl, sr;  // the left and right neighbours' slots (slot t+1 belongs to thread t)
  __device__ __forceinline__ R pull(int w, int k, int di, int dj) const {
    if (di == 0 && dj == 0) return w ? P.g[k] : P.f[k];
    if (dj == 0) return di < 0 ? side[2 * w][sl] : side[2 * w + 1][sr];
    if (dj < 0) {
      if (di == 0) return w ? p2g : p2f;
      return di < 0 ? down[2 * w][sl] : down[2 * w + 1][sr];
    }
    if (di == 0) return w ? c4g : c4f;
    return di > 0 ? up[2 * w][sr] : up[2 * w + 1][sl];
  }
  __device__ __forceinline__ R f(int k, int di, int dj) const { return pull(0, k, di, dj); }
  __device__ __forceinline__ R g(int k, int di, int dj) const { return pull(1, k, di, dj); }
};

// Step s pulls of node (x, y) from the input planes and its force F(s): 19 loads issued
// without branches (a wall row's f loads take the bounce-back source's address; its g loads
// across the wall read its own row and are replaced by the plate rule in finish), so that
// the same number of loads is in flight on every path and a prefetched row stays in flight
// until its first use.
template <typename R>
struct Raw {
  R f[9], q[9], Fy;
};

template <typename R>
__device__ __forceinline__ void issue(const Geom &g, const ThreeStepPtrs &a, int x, int y, Raw<R> &p) {
  const bool bot = (y == 0), top = (y == g.ny - 1);
  const size_t P = g.plane, r0 = (size_t)y * g.pitch;
  const size_t rm = bot ? r0 : r0 - g.pitch, rp = top ? r0 : r0 + g.pitch;
  const size_t i = (size_t)x, im = x == 0 ? g.nx - 1 : x - 1, ip = x == g.nx - 1 ? 0 : x + 1;
  const R *fa = (const R *)a.fA, *ga = (const R *)a.gA;
  p.f[0] = fa[r0 + i];
  p.f[1] = fa[P + r0 + im];
  p.f[3] = fa[3 * P + r0 + ip];
  p.f[2] = fa[bot ? 4 * P + r0 + i : 2 * P + rm + i];
  p.f[5] = fa[bot ? 7 * P + r0 + i : 5 * P + rm + im];
  p.f[6] = fa[bot ? 8 * P + r0 + i : 6 * P + rm + ip];
  p.f[4] = fa[top ? 2 * P + r0 + i : 4 * P + rp + i];
  p.f[7] = fa[top ? 5 * P + r0 + i : 7 * P + rp + ip];
  p.f[8] = fa[top ? 6 * P + r0 + i : 8 * P + rp + im];
  p.q[0] = ga[r0 + i];
  p.q[1] = ga[P + r0 + im];
  p.q[3] = ga[3 * P + r0 + ip];
  p.q[2] = ga[2 * P + rm + i];
  p.q[5] = ga[5 * P + rm + im];
  p.q[6] = ga[6 * P + rm + ip];
  p.q[4] = ga[4 * P + rp + i];
  p.q[7] = ga[7 * P + rp + ip];
  p.q[8] = ga[8 * P + rp + im];
  p.Fy = ((const R *)a.Fin)[r0 + i];
}

// gather_truth's plate rule on the issued g of a wall row (top first, as gather_truth).
template <typename R>
__device__ __forceinline__ void finish(const StepConst<R> &c, bool bot, bool top, R q[9]) {
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

// Three steps on output columns [ox0, ox1) (1 <= ox1 - ox0 <= SW - 4) and rows [ja, jb);
// every thread of the workgroup calls it.  The sweep is unrolled by two so that the two rows
// of each level alternate between two register sets without copies.  (Issuing a row's loads
// one iteration ahead, in flight across the barriers, measured no faster: the sweep is bound
// by the latency of its dependent arithmetic and barriers at three waves per SIMD, about 52 %
// VALU-busy, not by memory latency.)
template <typename R, int SW, int NT>
__device__ __forceinline__ void three_step_strip(const Geom &g, const ThreeStepPtrs &a, const StepConst<R> &c,
                                                 int ox0, int ox1, int ja, int jb, Lds3<R, SW> &L) {
  const int t = threadIdx.x;
  const int live = ox1 - ox0 + 4;         // threads with a column
  const bool has = t < live;               // step s
  const bool mid = t >= 1 && t < live - 1;  // step s+1
  const bool out = t >= 2 && t < live - 2;  // step s+2: the strip's output columns
  // threads past the strip load the last live column's values (the same cache lines): every
  // thread issues the same loads, with no branch around them
  int x = (ox0 - 2 + min(t, live - 1)) % g.nx;
  x = x < 0 ? x + g.nx : x;
  const int sl = t, sr = t + 2;  // the neighbours' slots
  const size_t pitch = g.pitch;
  R *const fB = (R *)a.fB, *const gB = (R *)a.gB;
  auto row_of = [&](int r) { return min(max(r, 0), g.ny - 1); };

  const int nit = jb - ja + 4;
  int m3 = 0;  // i % 3
  // Iteration i: cur receives the loads of row ja-2+i; C0 / C1 receive steps s and s+1 of rows r0 / r1 (they held rows
  // r0-2 / r1-2, whose direction 2 is read first), P0 / P1 hold rows r0-1 / r1-1.  Whether a
  // level runs depends on its row only (the same for the whole workgroup): every thread
  // computes it, those without a column of that level on duplicated or stale inputs whose
  // results no thread with one reads; stores and the non-finite check are masked to the
  // strip's columns.  A level skipped outside the lattice clears its row.
  Raw<R> cur;
  auto body = [&](int i, Row<R> &C0, const Row<R> &P0, Row<R> &C1, const Row<R> &P1) {
    const int r0 = ja - 2 + i, r1 = r0 - 1, r2 = r0 - 2;
    const int par = i & 1, prev = par ^ 1, m3b = m3 == 2 ? 0 : m3 + 1;  // (i - 2) % 3
    issue<R>(g, a, x, row_of(r0), cur);
    // step s on row r0
    const R p0f = C0.f[2], p0g = C0.g[2];
    if (r0 >= 0 && r0 < g.ny) {
      R u, v, rho, T;
      finish(c, r0 == 0, r0 == g.ny - 1, cur.q);
      collide(c, cur.f, cur.q, cur.Fy, C0.f, C0.g, rho, T, u, v, C0.F);
      if (out && r0 >= ja && r0 < jb) {
        st_uv<0, R>(a.uv[0], (size_t)r0 * pitch + x, u, v);
        if (a.mlev == 0) {
          ((R *)a.rho)[(size_t)r0 * pitch + x] = rho;
          ((R *)a.T)[(size_t)r0 * pitch + x] = T;
        }
      }
    } else {
      row_zero(C0);
    }
    if (has) put_row<R, SW>(L, 0, par, m3, t + 1, C0);
    __syncthreads();
    // step s+1 on row r1
    const R p1f = C1.f[2], p1g = C1.g[2];
    if (r1 >= ja - 1 && r1 >= 0 && r1 < g.ny) {
      const Acc3<R, SW> acc{P0, p0f, p0g, C0.f[4], C0.g[4], L.up[0], L.side[0][prev], L.down[0][m3b], sl, sr};
      R f[9], q[9], u, v, rho, T;
      gather_truth<R>(acc, r1 == 0, r1 == g.ny - 1, c, f, q);
      collide(c, f, q, P0.F, C1.f, C1.g, rho, T, u, v, C1.F);
      if (out && r1 >= ja && r1 < jb) {
        st_uv<0, R>(a.uv[1], (size_t)r1 * pitch + x, u, v);
        if (a.mlev == 1) {
          ((R *)a.rho)[(size_t)r1 * pitch + x] = rho;
          ((R *)a.T)[(size_t)r1 * pitch + x] = T;
        }
      }
    } else {
      row_zero(C1);
    }
    if (mid) put_row<R, SW>(L, 1, par, m3, t + 1, C1);
    __syncthreads();
    // step s+2 on row r2
    if (r2 >= ja && r2 < jb) {
      const Acc3<R, SW> acc{P1, p1f, p1g, C1.f[4], C1.g[4], L.up[1], L.side[1][prev], L.down[1][m3b], sl, sr};
      R f[9], q[9], fo[9], go[9], rho, T, u, v, Fn;
      gather_truth<R>(acc, r2 == 0, r2 == g.ny - 1, c, f, q);
      collide(c, f, q, P1.F, fo, go, rho, T, u, v, Fn);
      if (out) {
        flag_nonfinite(a.bad, rho, T);
        const size_t o = (size_t)r2 * pitch + x;
#pragma unroll
        for (int k = 0; k < 9; k++) st<NT>(&fB[(size_t)k * g.plane + o], fo[k]);
#pragma unroll
        for (int k = 0; k < 9; k++) st<NT>(&gB[(size_t)k * g.plane + o], go[k]);
        st<NT>(&((R *)a.Fout)[o], Fn);
        st_uv<0, R>(a.uv[2], o, u, v);
        if (a.mlev == 2) {
          ((R *)a.rho)[o] = rho;
          ((R *)a.T)[o] = T;
        }
      }
    }
    m3 = m3 == 2 ? 0 : m3 + 1;
  };
  Row<R> U0a, U0b, U1a, U1b;
  row_zero(U0a);
  row_zero(U0b);
  row_zero(U1a);
  row_zero(U1b);
  for (int i = 0; i < nit; i += 2) {
    body(i, U0a, U0b, U1a, U1b);
    if (i + 1 < nit) body(i + 1, U0b, U0a, U1b, U1a);
  }
}

}  // namespace lbm3
}  // namespace rbn
