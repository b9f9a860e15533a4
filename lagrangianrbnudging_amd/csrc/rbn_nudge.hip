// rbn_nudge.hip — K3: observation time interpolation and the ordered nudging deposit of
// the nudged run (read_prtcl, Nudging/code/tempvel_nudge.c:1380-1781).
//
// Observer update: one thread per observer runs the reference's three-branch state machine
// (new record / snap to record / advance by slope), which is per-observer independent.
//
// Deposit: every observer k adds its temperature to a 7x7 box of nodes around its rounded
// position (plus the x-wrap windows) through the running mean
//     Tbar <- (Tbar * n + T_k) / (n + 1),  n <- n + 1          (N:1578-1579)
// processed in observer order.  The mean is order dependent in floating point, so the
// GPU version restates it deterministically:
//   1. count contributions per node (atomic increments: order free);
//   2. exclusive scan of the counts gives every node a slot range;
//   3. write each contribution's observer index into its node's range (atomic slot
//      allocation: order free);
//   4. per node, sort its observer indices ascending and fold the running mean in that
//      order — the reference's order, so the result is bit-identical.
// Each observer contributes at most once per node when nx >= 8, so observer order is the
// complete per-node order.
#include <algorithm>

#include "rbn_internal.h"

namespace rbn {
namespace {

constexpr int kBlock = 256;
constexpr int kScanItems = 4;
constexpr int kScanTile = kBlock * kScanItems;

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
__global__ void k_obs_update(ObsPtrs o, int branch, int nx, int ny, int ntempread) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= o.n) return;
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

// ---------------------------------------------------------------- deposit

// Visit the nodes observer k deposits into (N:1568-1616, global form), calling fn(offset)
// for each interior node.
template <typename Fn>
__device__ __forceinline__ void visit(const Geom &g, int right, int xi, int yi, Fn &&fn) {
  auto cell = [&](int i, int j) {
    if ((i > 0) && (i <= g.nx) && (j > 0) && (j <= g.ny)) fn((size_t)(j - 1) * g.pitch + (i - 1));
  };
  for (int i1 = -3; i1 < 4; i1++)
    for (int j1 = -3; j1 < 4; j1++) cell(xi + i1, yi + j1);
  if (xi < 4) {
    const int t = g.nx + xi;
    for (int i1 = -3; i1 < 1; i1++)
      for (int j1 = -3; j1 < 4; j1++) cell(t + i1, yi + j1);
  }
  if (xi > right) {
    const int t = xi - g.nx;
    for (int i1 = 0; i1 < 4; i1++)
      for (int j1 = -3; j1 < 4; j1++) cell(t + i1, yi + j1);
  }
}

__global__ void k_dep_zero(size_t n, int *cnt, int *fill) {
  for (size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x; o < n;
       o += (size_t)gridDim.x * blockDim.x) {
    cnt[o] = 0;
    fill[o] = 0;
  }
}

__global__ void k_dep_count(Geom g, int right, ObsPtrs o, int *cnt) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= o.n) return;
  visit(g, right, o.xidx[k], o.yidx[k], [&](size_t off) { atomicAdd(&cnt[off], 1); });
}

// Exclusive scan, pass 1: tile-local scans and tile totals.
__global__ __launch_bounds__(kBlock) void k_scan_tiles(size_t n, const int *in, int *out,
                                                       int *tile_sum) {
  __shared__ int s[kBlock];
  const size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
  int v[kScanItems], t = 0;
#pragma unroll
  for (int q = 0; q < kScanItems; q++) {
    v[q] = (base + q < n) ? in[base + q] : 0;
    t += v[q];
  }
  s[threadIdx.x] = t;
  __syncthreads();
  for (int d = 1; d < kBlock; d <<= 1) {  // Hillis–Steele inclusive scan of thread totals
    int add = threadIdx.x >= (unsigned)d ? s[threadIdx.x - d] : 0;
    __syncthreads();
    s[threadIdx.x] += add;
    __syncthreads();
  }
  int run = s[threadIdx.x] - t;
#pragma unroll
  for (int q = 0; q < kScanItems; q++) {
    if (base + q < n) out[base + q] = run;
    run += v[q];
  }
  if (threadIdx.x == kBlock - 1) tile_sum[blockIdx.x] = s[kBlock - 1];
}

// Pass 2: exclusive scan of the tile totals by one block; total written to *total.
__global__ __launch_bounds__(kBlock) void k_scan_sums(int ntiles, int *tile_sum, int *total) {
  __shared__ int s[kBlock];
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < ntiles; base += kBlock) {
    const int idx = base + threadIdx.x;
    const int v = idx < ntiles ? tile_sum[idx] : 0;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int d = 1; d < kBlock; d <<= 1) {
      int add = threadIdx.x >= (unsigned)d ? s[threadIdx.x - d] : 0;
      __syncthreads();
      s[threadIdx.x] += add;
      __syncthreads();
    }
    if (idx < ntiles) tile_sum[idx] = carry + s[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == kBlock - 1) carry += s[kBlock - 1];
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}

__global__ void k_scan_add(size_t n, int *out, const int *tile_sum) {
  const size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o < n) out[o] += tile_sum[o / kScanTile];
}

__global__ void k_dep_fill(Geom g, int right, ObsPtrs o, const int *off, int *fill, int *list,
                           long long cap, int *overflow) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= o.n) return;
  visit(g, right, o.xidx[k], o.yidx[k], [&](size_t c) {
    const long long slot = (long long)off[c] + atomicAdd(&fill[c], 1);
    if (slot < cap)
      list[slot] = k;
    else
      atomicOr(overflow, 1);
  });
}

// Per node: order the contributing observers and fold the running mean (N:1578-1579).
__global__ void k_dep_fold(Geom g, const int *cnt, const int *off, int *list, const double *Tk,
                           int nobs, double *Tbar, long long cap) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= g.nx) return;
  const size_t c = (size_t)j * g.pitch + i;
  const int n = cnt[c];
  double t = 0.0;
  if (n > 0 && (long long)off[c] + n <= cap) {
    int *seg = list + off[c];
    for (int a = 1; a < n; a++) {  // insertion sort: segments are short
      const int key = seg[a];
      int b = a - 1;
      while (b >= 0 && seg[b] > key) {
        seg[b + 1] = seg[b];
        b--;
      }
      seg[b + 1] = key;
    }
    for (int a = 0; a < n; a++) {
      const int k = seg[a];
      if (k >= 0 && k < nobs) t = (t * (double)a + Tk[k]) / ((double)a + 1.0);
    }
  }
  Tbar[c] = t;
}

}  // namespace

hipError_t launch_obs_update(hipStream_t s, int branch, int nx, int ny, int ntempread,
                             const ObsPtrs &o) {
  if (o.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_obs_update, dim3((o.n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, o,
                     branch, nx, ny, ntempread);
  return hipGetLastError();
}

hipError_t launch_deposit(hipStream_t s, const Geom &g, int right_limit, const ObsPtrs &o,
                          const DepPtrs &d) {
  const size_t n = g.plane;
  const int ntiles = (int)((n + kScanTile - 1) / kScanTile);
  const int zb = (int)std::min<size_t>((n + kBlock - 1) / kBlock, 4096);
  hipLaunchKernelGGL(k_dep_zero, dim3(zb), dim3(kBlock), 0, s, n, d.cnt, d.fill);
  const dim3 ob((o.n + kBlock - 1) / kBlock > 0 ? (o.n + kBlock - 1) / kBlock : 1);
  if (o.n > 0) hipLaunchKernelGGL(k_dep_count, ob, dim3(kBlock), 0, s, g, right_limit, o, d.cnt);
  hipLaunchKernelGGL(k_scan_tiles, dim3(ntiles), dim3(kBlock), 0, s, n, d.cnt, d.off, d.blocksum);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kBlock), 0, s, ntiles, d.blocksum, d.off + n);
  hipLaunchKernelGGL(k_scan_add, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s,
                     n, d.off, d.blocksum);
  if (o.n > 0)
    hipLaunchKernelGGL(k_dep_fill, ob, dim3(kBlock), 0, s, g, right_limit, o, d.off, d.fill,
                       d.list, d.list_cap, d.overflow);
  hipLaunchKernelGGL(k_dep_fold, dim3((g.nx + kBlock - 1) / kBlock, g.ny), dim3(kBlock), 0, s, g,
                     d.cnt, d.off, d.list, o.Tint, o.n, d.Tbar, d.list_cap);
  return hipGetLastError();
}

}  // namespace rbn
