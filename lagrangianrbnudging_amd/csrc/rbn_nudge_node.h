// rbn_nudge_node.h — per-observer device code of the nudged run, shared by the observer
// passes (rbn_nudge.hip) and the launches that carry the next step's observer update
// (rbn_lbm2.hip, rbn_ensemble.hip): the observer state machine (read_prtcl,
// Nudging/code/tempvel_nudge.c:1380-1781) and the bilinear deposit's binning.
#pragma once

#include "rbn_internal.h"

namespace rbn {
namespace obs {

// ---------------------------------------------------------------- observer update

__device__ __forceinline__ void advance_x(const ObsPtrs &o, int i, double nxd) {
  double xposnow = o.xint[i] + o.dx[i];
  if (xposnow > nxd) xposnow = xposnow - nxd;
  if (xposnow < 0.0) xposnow = xposnow + nxd;
  o.xint[i] = xposnow;
  o.xidx[i] = (int)round(xposnow);
}

__device__ __forceinline__ void advance_y_T(const ObsPtrs &o, int i, double nyd) {
  double yposnow = o.yint[i] + o.dy[i];
  int bc = o.bcy[i];
  if (yposnow > nyd) {
    yposnow = yposnow - nyd + 1.0;
    bc = 2;
  }
  if (yposnow < 1.0) {
    yposnow = yposnow + nyd - 1.0;
    bc = 2;
  }
  o.yint[i] = yposnow;
  o.yidx[i] = (int)round(yposnow);
  double Tnow = o.Tint[i] + o.dT[i];
  if (bc == 1) Tnow = o.Tint[i];
  if (bc == 2) {
    Tnow = o.rT[o.chk[i]];
    bc = 1;
  }
  o.bcy[i] = bc;
  o.Tint[i] = Tnow;
}

// branch: 1 = istep%ntempread==1 (new record in rx/ry/rT, N:1489-1552), 0 = snap to the
// record (N:1621-1628), 2 = advance by the slopes (N:1687-1720).
__device__ __forceinline__ void obs_update_one(const ObsPtrs &o, int i, int branch, int nx, int ny, int ntempread) {
  const double nxd = (double)nx, nyd = (double)ny, ntr = (double)ntempread;
  const int c = o.chk[i];
  if (branch == 1) {
    o.bcy[i] = 0;
    double xread = o.rx[c], yread = o.ry[c];
    const double xint = o.xint[i], yint = o.yint[i];
    double dx = (xread - xint) / (ntr);
    if (abs((int)(xread - xint)) > (nxd / 2.0)) {  // abs() of a double: int truncation
      if (xread > xint) {
        xread = xread - nxd;
        dx = (xread - xint) / (ntr);
      } else {
        xread = xread + nxd;
        dx = (xread - xint) / (ntr);
      }
    }
    o.dx[i] = dx;
    advance_x(o, i, nxd);
    double dy = (yread - yint) / (ntr);
    if (abs((int)(yread - yint)) > (nyd / 2.0)) {
      o.bcy[i] = 1;
      if (yread > yint) {
        yread = yread - nyd + 1.0;
        dy = (yread - yint) / (ntr);
      } else {
        yread = yread + nyd - 1.0;
        dy = (yread - yint) / (ntr);
      }
    }
    o.dy[i] = dy;
    o.dT[i] = (o.rT[c] - o.Tint[i]) / (ntr);
    advance_y_T(o, i, nyd);
  } else if (branch == 0) {
    o.xint[i] = o.rx[c];
    o.yint[i] = o.ry[c];
    o.Tint[i] = o.rT[c];
    o.xidx[i] = (int)round(o.rx[c]);
    o.yidx[i] = (int)round(o.ry[c]);
  } else {
    advance_x(o, i, nxd);
    advance_y_T(o, i, nyd);
  }
}

// Bilinear deposit (RBN_DEPOSIT_BILINEAR, no reference counterpart): observer k gives each
// corner c of its interpolation cell the record-interpolation weight (M:292-311, same
// operand order), ghost corners folded onto the interior node they copy; zero weights are
// not contributions.  Returns the node offset and the weight, or -1.
constexpr int kBilCells = 4;

__device__ __forceinline__ long long bil_cell(const Geom &g, double xx, double yy, int c, double &w) {
  if (!(xx > -1.0 && xx < (double)g.nx + 1.0 && yy > -1.0 && yy < (double)g.ny + 1.0)) return -1;
  const int x1 = (int)xx, y1 = (int)yy, x2 = x1 + 1, y2 = y1 + 1;
  const int ic = (c & 1) ? x2 : x1, jc = (c & 2) ? y2 : y1;
  const double wx = (c & 1) ? (xx - (double)x1) : ((double)x2 - xx);
  const double wy = (c & 2) ? (yy - (double)y1) : ((double)y2 - yy);
  w = wx * wy;
  if (w == 0.0) return -1;
  const int I = ic == 0 ? g.nx : (ic == g.nx + 1 ? 1 : ic);
  const int J = jc == 0 ? g.ny : (jc == g.ny + 1 ? 1 : jc);
  return (long long)(J - 1) * g.pitch + (I - 1);
}

// Fixed point of the bilinear deposit's sums: v * 2^40 rounded to the nearest integer, ties
// to even (the scaling is exact).
__device__ __forceinline__ long long dep_fix(double v) { return (long long)rint(v * 0x1p40); }

// Bilinear deposit: observer k's contributions to its cell's corners, each binned into a slot
// of its node's deposit tile, or past the tile's slots added to the node's sums (DepPtrs).
// The corners that share a tile (all four but near a tile's edge) take their slots with one
// atomic.
__device__ __forceinline__ void bin_one(const Geom &g, const DepPtrs &d, const ObsPtrs &o, int k) {
  const double xx = o.xint[k], yy = o.yint[k], Tk = o.Tint[k];
  const int tiles_x = dep_tiles_x(g);
  long long node[kBilCells], wf[kBilCells], wTf[kBilCells];
  int tile[kBilCells], local[kBilCells];
#pragma unroll
  for (int c = 0; c < kBilCells; c++) {
    double w = 0.0;
    node[c] = bil_cell(g, xx, yy, c, w);
    wf[c] = dep_fix(w);
    wTf[c] = dep_fix(w * Tk);
    const bool on = node[c] >= 0;
    const int j = on ? (int)(node[c] / g.pitch) : 0, i = on ? (int)(node[c] - (long long)j * g.pitch) : 0;
    tile[c] = on ? (j / kDepTileH) * tiles_x + i / kDepTileW : -1;  // -1: no contribution
    local[c] = (j % kDepTileH) * kDepTileW + i % kDepTileW;
  }
  // the first corner of each tile takes the slots of every corner in that tile (no array is
  // written in this loop, so that it unrolls into registers)
#pragma unroll
  for (int c = 0; c < kBilCells; c++) {
    const int t = tile[c];
    bool first = t >= 0;
#pragma unroll
    for (int c1 = 0; c1 < c; c1++) first = first && tile[c1] != t;
    if (!first) continue;
    int n = 0;
#pragma unroll
    for (int c2 = c; c2 < kBilCells; c2++) n += tile[c2] == t;
    const int slot0 = atomicAdd(&d.tcnt[(size_t)t * kTcntStride], n);
    int rank = 0;
#pragma unroll
    for (int c2 = c; c2 < kBilCells; c2++) {
      if (tile[c2] != t) continue;
      const int slot = slot0 + rank++;
      if (slot < d.bin_cap) {
        // w in the high bits (times 2^kDepTileBits: a weight of either sign survives the
        // arithmetic shift that takes it out), the node's index in the tile in the low ones
        longlong2 e;
        e.x = wTf[c2];
        e.y = wf[c2] * (1LL << kDepTileBits) + local[c2];
        ((longlong2 *)d.bin)[(size_t)t * d.bin_cap + slot] = e;
      } else {
        atomicAdd((unsigned long long *)&d.S[node[c2]], (unsigned long long)wTf[c2]);
        atomicAdd((unsigned long long *)&d.Wi[node[c2]], (unsigned long long)wf[c2]);
        atomicAdd(&d.ocnt[node[c2]], 1);
      }
    }
  }
}

// The branch of the state machine at step istep: 1 = a new record (istep % ntempread == 1),
// 0 = snap to the record, 2 = advance by the slopes.
__device__ __forceinline__ int obs_branch(int istep, int ntempread) {
  const int r = istep % ntempread;
  return r == 1 ? 1 : (r == 0 ? 0 : 2);
}

// Observer i of an ensemble member at step t of a call (its step count istep0 + t + 1): the
// update, and with the bilinear deposit the binning of its contributions.
__device__ __forceinline__ void head_one(const NudgeMember &d, int i, int t) {
  if (i >= d.o.n) return;
  obs_update_one(d.o, i, obs_branch(d.istep0 + t + 1, d.ntempread), d.g.nx, d.g.ny, d.ntempread);
  if (d.d.mode == RBN_DEPOSIT_BILINEAR) bin_one(d.g, d.d, d.o, i);
}

}  // namespace obs
}  // namespace rbn
