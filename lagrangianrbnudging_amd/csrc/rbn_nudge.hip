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
//   4. per node with contributions, sort its observer indices ascending and fold the
//      running mean in that order — the reference's order, so the result is bit-identical.
//      With fewer candidate contributions than a quarter of the nodes the fill pass lists
//      the nodes that have any and the fold visits only those; otherwise it sweeps the
//      rows (DepPtrs::dense).  The means of the other nodes are not written:
//      every reader (the nudged collision, the adjoint, the getters) looks at them only where
//      the count is positive.
// Each observer contributes at most once per node when nx >= 8, so observer order is the
// complete per-node order.  Count and fill run one thread per (observer, candidate node), so
// the atomics of one observer are independent; the scan's tile offsets are added where the
// ranges are used.
//
// Bilinear deposit (RBN_DEPOSIT_BILINEAR; no reference counterpart, an option beside the
// reference's running mean): each observer gives the four corners of its interpolation cell
// the weights the record interpolation uses, so the deposit is the transpose of that
// operator.  Per node W = sum w_k and S = sum w_k T_k are taken in fixed point (every term
// rounded to a multiple of 2^-40, 64-bit integer sums), so the sums are exact and do not
// depend on the order in which the contributions arrive.  The observer update's thread bins
// each contribution (w T, w and the node) into a slot of its node's deposit tile of 64 x 8
// nodes (one atomic on the tile's counter); one workgroup per tile then sums its slots in LDS
// and writes every node of the tile: the count, the packed deposit, and Tbar = S / W and W
// where contributions arrived.  Contributions past a tile's slots take 64-bit atomics on the
// node's sums in HBM instead, which the tile adds and clears.  The nudged collision uses
// Q = alpha * (min(W, 1) * (T - Tbar)).  The sums must stay below 2^23 in magnitude
// (|T| W < 8.4e6 at a node).  Random contributions thus cost one atomic on an L2-resident
// counter and a 16-byte store each, where per-node atomics on the sums would each move a
// line of HBM.
#include <algorithm>
#include <climits>

#include "rbn_internal.h"
#include "rbn_nudge_node.h"

namespace rbn {
namespace {

using namespace obs;

constexpr int kBlock = kNudgeBlock;
constexpr int kScatterSplit = 16;  // blocks per record of a batched upload
constexpr int kScanItems = 4;
constexpr int kScanTile = kBlock * kScanItems;
static_assert(kScanTile == kDepScanTile, "tile totals are sized by kDepScanTile");

// ---------------------------------------------------------------- deposit

// The nodes observer k deposits into (N:1568-1616, global form): the 7x7 box at its rounded
// position, plus 4 x 7 nodes of the x-wrap window when xi < 4 (columns nx+xi-3 .. nx+xi) or
// xi > right (columns xi-nx .. xi-nx+3).  With nx >= 8 at most one window applies, so an
// observer has at most kCells = 77 candidate nodes, and it reaches each node at most once.
// cell_of(c) maps candidate c to the node offset, or -1 when the candidate is outside the
// lattice or its window is inactive.  Work is spread as one thread per (observer, candidate)
// so that the atomics of one observer are independent and their latencies overlap.
constexpr int kCells = 77;

__device__ __forceinline__ long long cell_of(const Geom &g, int right, int xi, int yi, int c) {
  int i, j;
  if (c < 49) {
    i = xi + c / 7 - 3;
    j = yi + c % 7 - 3;
  } else {
    const int w = c - 49, col = w / 7;
    j = yi + w % 7 - 3;
    if (xi < 4)
      i = g.nx + xi + col - 3;
    else if (xi > right)
      i = xi - g.nx + col;
    else
      return -1;
  }
  if ((i > 0) && (i <= g.nx) && (j > 0) && (j <= g.ny)) return (long long)(j - 1) * g.pitch + (i - 1);
  return -1;
}

// The packed deposit (DepPtrs::tw) at node c: the mean and weight w (stored negated), or
// the empty value.
__device__ __forceinline__ void tw_put(const DepPtrs &d, size_t c, double Tbar, double w) {
  if (d.tw64)
    ((double2 *)d.tw)[c] = make_double2(Tbar, -w);
  else
    ((float2 *)d.tw)[c] = make_float2((float)Tbar, -(float)w);
}
__device__ __forceinline__ void tw_zero(const DepPtrs &d, size_t c) {
  if (d.tw64)
    ((double2 *)d.tw)[c] = make_double2(0.0, 0.0);
  else
    ((float2 *)d.tw)[c] = make_float2(0.f, 0.f);
}

// 7x7 deposit: zero the counts, the fill cursors, the packed deposit and the active-node
// count: block b of nb, grid-stride.
__device__ __forceinline__ void zero_part(size_t n, const DepPtrs &d, int b, int nb) {
  if (b == 0 && threadIdx.x == 0) *d.nactive = 0;
  for (size_t o = (size_t)b * kBlock + threadIdx.x; o < n; o += (size_t)nb * kBlock) {
    d.cnt[o] = 0;
    d.fill[o] = 0;
    tw_zero(d, o);
  }
}

// 7x7 deposit: candidate t (observer t / kCells, cell t % kCells) counts one contribution
// at its node.
__device__ __forceinline__ void count_one(const Geom &g, int right, const ObsPtrs &o, int *cnt, long long t) {
  if (t >= (long long)o.n * kCells) return;
  const long long off = cell_of(g, right, o.xidx[(int)(t / kCells)], o.yidx[(int)(t / kCells)], (int)(t % kCells));
  if (off >= 0) atomicAdd(&cnt[off], 1);
}

__global__ void k_dep_count(Geom g, int right, ObsPtrs o, int *cnt) {
  count_one(g, right, o, cnt, (long long)blockIdx.x * blockDim.x + threadIdx.x);
}

// Exclusive scan, pass 1: tile-local scans and tile totals (tile b; s: kBlock ints of LDS).
__device__ __forceinline__ void scan_tile(size_t n, const int *in, int *out, int *tile_sum, int b, int *s) {
  const size_t base = (size_t)b * kScanTile + (size_t)threadIdx.x * kScanItems;
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
  if (threadIdx.x == kBlock - 1) tile_sum[b] = s[kBlock - 1];
}

__global__ __launch_bounds__(kBlock) void k_scan_tiles(size_t n, const int *in, int *out,
                                                       int *tile_sum) {
  __shared__ int s[kBlock];
  scan_tile(n, in, out, tile_sum, blockIdx.x, s);
}

// Pass 2: exclusive scan of the tile totals by one block; total written to *total (s:
// kBlock ints of LDS, carry: one).
__device__ __forceinline__ void scan_sums(int ntiles, int *tile_sum, int *total, int *s, int &carry) {
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

__global__ __launch_bounds__(kBlock) void k_scan_sums(int ntiles, int *tile_sum, int *total) {
  __shared__ int s[kBlock];
  __shared__ int carry;
  scan_sums(ntiles, tile_sum, total, s, carry);
}

// Start of node c's slot range: tile-local exclusive prefix + the tile's exclusive prefix.
__device__ __forceinline__ long long seg_start(const int *off, const int *tile_sum, size_t c) {
  return (long long)off[c] + tile_sum[c / kScanTile];
}

// Append node to the active list for the threads with first set: the wavefronts count
// their entries in LDS and one thread per workgroup takes the workgroup's range with one
// atomic, so the list's counter sees one atomic per workgroup.  Every thread of the
// workgroup calls it.
__device__ __forceinline__ void append_active(bool first, int node, int *active, int *nactive, long long cap) {
  __shared__ int total, base;
  if (threadIdx.x == 0) total = 0;
  __syncthreads();
  const unsigned long long mask = __ballot(first);
  const int lane = __lane_id(), leader = mask ? __ffsll((long long)mask) - 1 : 0;
  int woff = 0;
  if (mask && lane == leader) woff = atomicAdd(&total, __popcll(mask));
  woff = __shfl(woff, leader);
  __syncthreads();
  if (threadIdx.x == 0) base = total > 0 ? atomicAdd(nactive, total) : 0;
  __syncthreads();
  if (first) {
    const long long a = (long long)base + woff + __popcll(mask & ((1ull << lane) - 1ull));
    if (a < cap) active[a] = node;
  }
}

// 7x7 deposit: contribution t takes the next slot of its node's range (its key is the
// observer); the node's first contribution also lists the node as active.
__device__ __forceinline__ void fill_one(const Geom &g, int right, const ObsPtrs &o, const int *off,
                                         const int *tile_sum, int *fill, int *list, long long cap, int *overflow,
                                         int *active, int *nactive, long long active_cap, long long t) {
  const int key = (int)(t / kCells);
  const long long node = t < (long long)o.n * kCells
                             ? cell_of(g, right, o.xidx[key], o.yidx[key], (int)(t % kCells))
                             : -1;
  const int rank = node >= 0 ? atomicAdd(&fill[node], 1) : -1;
  if (active) append_active(rank == 0, (int)node, active, nactive, active_cap);
  if (node < 0) return;
  const long long slot = seg_start(off, tile_sum, (size_t)node) + rank;
  if (slot < cap)
    list[slot] = key;
  else
    atomicOr(overflow, 1);
}

__global__ void k_dep_fill(Geom g, int right, ObsPtrs o, const int *off, const int *tile_sum, DepPtrs d) {
  fill_one(g, right, o, off, tile_sum, d.fill, d.list, d.list_cap, d.overflow, d.dense ? nullptr : d.active,
           d.nactive, d.active_cap, (long long)blockIdx.x * blockDim.x + threadIdx.x);
}

// Bilinear deposit, deposit tile t: the sums of its binned contributions in LDS (exact, in
// any order), plus the sums in HBM when the tile overflowed (cleared here), then every node
// of the tile: the packed deposit, and with api also the count, and Tbar = S / W and W (each
// sum converted to double) where the count is positive (the arrays the getters, the adjoint
// and the ensemble's copies read; the nudged collision reads only the packed deposit).
// Leaves the tile's counter at zero.  A workgroup runs two tiles, one per half (kTileHalf
// threads): with 10 KB of LDS per tile, twice as many tiles are in flight per CU as with one
// per workgroup, which hides the pass's latency (72 us for the ensemble's 26,880 tiles with
// one per workgroup).  Every thread of the workgroup calls it, with its half's tile (t < 0:
// none; the half still takes part in the barriers).
struct DepTileLds {
  unsigned long long S[kDepTileW * kDepTileH], W[kDepTileW * kDepTileH];
  int n[kDepTileW * kDepTileH];
};
constexpr int kTileHalf = kBlock / 2;
struct DepTilePairLds {
  DepTileLds h[2];
};

__device__ __forceinline__ void dep_tile(const Geom &g, const DepPtrs &d, int t, DepTilePairLds &P, bool api) {
  constexpr int kNodes = kDepTileW * kDepTileH;
  const int tid = threadIdx.x % kTileHalf;
  DepTileLds &L = P.h[threadIdx.x / kTileHalf];
  const bool on = t >= 0;
  const longlong2 *e = on ? (const longlong2 *)d.bin + (size_t)t * d.bin_cap : nullptr;
  // the counter and each thread's first two slots in one round trip (slots past the count
  // hold stale entries, read but not used)
  const int n = on ? d.tcnt[(size_t)t * kTcntStride] : 0;
  const int k0 = tid, k1 = tid + kTileHalf;
  longlong2 v0 = make_longlong2(0, 0), v1 = make_longlong2(0, 0);
  if (on && k0 < d.bin_cap) v0 = e[k0];
  if (on && k1 < d.bin_cap) v1 = e[k1];
  for (int k = tid; k < kNodes; k += kTileHalf) {
    L.S[k] = 0;
    L.W[k] = 0;
    L.n[k] = 0;
  }
  __syncthreads();
  const int m = on ? min(n, d.bin_cap) : 0;
  auto add = [&](const longlong2 &v) {
    const int local = (int)(v.y & (kNodes - 1));
    atomicAdd(&L.S[local], (unsigned long long)v.x);
    atomicAdd(&L.W[local], (unsigned long long)(v.y >> kDepTileBits));
    atomicAdd(&L.n[local], 1);
  };
  if (k0 < m) add(v0);
  if (k1 < m) add(v1);
  for (int k = k1 + kTileHalf; k < m; k += kTileHalf) add(e[k]);
  __syncthreads();
  if (!on) return;
  const int tiles_x = dep_tiles_x(g), ty = t / tiles_x, tx = t - ty * tiles_x;
  for (int k = tid; k < kNodes; k += kTileHalf) {
    const int i = tx * kDepTileW + k % kDepTileW, j = ty * kDepTileH + k / kDepTileW;
    if (i >= g.nx || j >= g.ny) continue;
    const size_t c = (size_t)j * g.pitch + i;
    long long s = (long long)L.S[k], wi = (long long)L.W[k];
    int cnt = L.n[k];
    if (n > d.bin_cap && d.ocnt[c] != 0) {
      cnt += d.ocnt[c];
      s += d.S[c];
      wi += d.Wi[c];
      d.ocnt[c] = 0;
      d.S[c] = 0;
      d.Wi[c] = 0;
    }
    if (api) d.cnt[c] = cnt;
    if (cnt > 0) {
      const double Tbar = wi != 0 ? (double)s / (double)wi : 0.0, W = (double)wi * 0x1p-40;
      if (api) {
        d.Tbar[c] = Tbar;
        d.wq[c] = W;
      }
      tw_put(d, c, Tbar, nudge_weight(W));
    } else {
      tw_zero(d, c);
    }
  }
  if (tid == 0) d.tcnt[(size_t)t * kTcntStride] = 0;  // every thread read it before the first barrier
}

// Tile index of this thread's half in a tile pass of ntiles tiles, block b (two per block).
__device__ __forceinline__ int half_tile(int b, int ntiles) {
  const int t = 2 * b + (int)threadIdx.x / kTileHalf;
  return t < ntiles ? t : -1;
}

// The step's observer update, and per deposit: 7x7, the zero pass (blocks past the
// observers'; the two touch disjoint arrays); bilinear, each observer's contributions binned
// by the thread that updated it (no zero pass: the tile pass leaves every array it reads
// clear).  The tail_blocks blocks after the observers' run the tile pass of `tail`, an
// earlier bilinear deposit whose bins are complete (it reads only its own set's arrays); the
// observer blocks go first, as they wait on atomics and the tile passes fill in around them.
__global__ __launch_bounds__(kBlock) void k_obs_zero(Geom g, ObsPtrs o, int branch, int ntempread, int tail_blocks,
                                                     int obs_blocks, DepPtrs d, DepPtrs tail) {
  __shared__ DepTilePairLds L;
  const int b = blockIdx.x;
  if (b < obs_blocks) {
    const int i = b * kBlock + threadIdx.x;
    if (i < o.n) {
      obs_update_one(o, i, branch, g.nx, g.ny, ntempread);
      if (d.mode == RBN_DEPOSIT_BILINEAR) bin_one(g, d, o, i);
    }
  } else if (b < obs_blocks + tail_blocks) {
    dep_tile(g, tail, half_tile(b - obs_blocks, dep_tiles(g)), L, true);
  } else {
    const int zb = tail_blocks + obs_blocks;
    zero_part(g.plane, d, b - zb, gridDim.x - zb);
  }
}

// An observation table on the device: the observer map of every record (map[r n + i] = the
// first entry of record r with observer i's id, INT_MAX when none) from the records' ids and
// a lookup of the observers' ids (dense range lo .. lo + range - 1, -1 where no observer).
__global__ __launch_bounds__(kBlock) void k_tab_fill(int *map, size_t count) {
  for (size_t e = (size_t)blockIdx.x * kBlock + threadIdx.x; e < count; e += (size_t)gridDim.x * kBlock)
    map[e] = INT_MAX;
}

__global__ __launch_bounds__(kBlock) void k_tab_map(const int *ids, size_t count, int n, const int *lookup, int lo,
                                                   int range, int *map) {
  for (size_t e = (size_t)blockIdx.x * kBlock + threadIdx.x; e < count; e += (size_t)gridDim.x * kBlock) {
    const long long k = (long long)ids[e] - lo;
    const int i = k >= 0 && k < range ? lookup[k] : -1;
    if (i < 0) continue;
    const size_t r = e / (size_t)n;
    atomicMin(&map[r * (size_t)n + i], (int)(e - r * (size_t)n));
  }
}

// Record blockIdx.x: the first observer its map misses (n when none).
__global__ __launch_bounds__(kBlock) void k_tab_check(const int *map, int n, int *first_missing) {
  __shared__ int s[kBlock];
  const int *m = map + (size_t)blockIdx.x * n;
  int first = n;
  for (int i = threadIdx.x; i < n; i += kBlock)
    if (m[i] == INT_MAX) {
      first = i;
      break;
    }
  s[threadIdx.x] = first;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] = min(s[threadIdx.x], s[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) first_missing[blockIdx.x] = s[0];
}

__global__ __launch_bounds__(kBlock) void k_dep_tiles(Geom g, DepPtrs d) {
  __shared__ DepTilePairLds L;
  dep_tile(g, d, half_tile(blockIdx.x, dep_tiles(g)), L, true);
}

// Per node: order the contributions by observer and fold them in that order, the
// reference's order: the running mean of the 7x7 deposit (N:1578-1579).  Lists are sorted in
// registers (odd-even
// transposition network, static indices) sized by the longest list of the wavefront (2, 4
// or 16 keys), longer ones in place.  The keys of a list are distinct (an observer reaches
// a node at most once), so every sort gives the same order.
constexpr int kSortRegs = 16;

struct Fold {
  const ObsPtrs &o;
  double t = 0.0;
  __device__ __forceinline__ void take(int a, int key) {
    if (key >= 0 && key < o.n) t = (t * (double)a + o.Tint[key]) / ((double)a + 1.0);
  }
};

// The n <= NS keys of a list sorted in registers and folded in that order.
template <int NS>
__device__ __forceinline__ void fold_sorted(Fold &f, const int *seg, int n) {
  int key[NS];
#pragma unroll
  for (int a = 0; a < NS; a++) key[a] = a < n ? seg[a] : 0x7fffffff;
#pragma unroll
  for (int pass = 0; pass < NS; pass++) {
#pragma unroll
    for (int a = pass & 1; a + 1 < NS; a += 2) {
      const int lo = min(key[a], key[a + 1]), hi = max(key[a], key[a + 1]);
      key[a] = lo;
      key[a + 1] = hi;
    }
  }
#pragma unroll
  for (int a = 0; a < NS; a++)
    if (a < n) f.take(a, key[a]);
}

// The longest list of the wavefront (every lane takes part).
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
  return v;
}

// 7x7 deposit at node c with n > 0 contributions; nmax: the wavefront's longest list.  Writes
// the mean and the packed deposit (weight 1).
__device__ __forceinline__ void fold_node(const int *off, const int *tile_sum, int *list, const ObsPtrs &o,
                                          const DepPtrs &d, long long cap, size_t c, int n, int nmax) {
  Fold f{o};
  const long long base = n > 0 ? seg_start(off, tile_sum, c) : 0;
  if (n > 0 && base + n <= cap) {
    int *seg = list + base;
    if (nmax <= 2) {
      fold_sorted<2>(f, seg, n);
    } else if (nmax <= 4) {
      fold_sorted<4>(f, seg, n);
    } else if (n <= kSortRegs) {
      fold_sorted<kSortRegs>(f, seg, n);
    } else {
      for (int a = 1; a < n; a++) {  // insertion sort in place
        const int key = seg[a];
        int b = a - 1;
        while (b >= 0 && seg[b] > key) {
          seg[b + 1] = seg[b];
          b--;
        }
        seg[b + 1] = key;
      }
      for (int a = 0; a < n; a++) f.take(a, seg[a]);
    }
  }
  d.Tbar[c] = f.t;
  tw_put(d, c, f.t, 1.0);
}

// The active nodes of a 7x7 deposit, block b of nb, grid-stride (every lane of a wavefront
// takes part in the wavefront's maximum).
__device__ __forceinline__ void fold_active(const ObsPtrs &o, const DepPtrs &d, int b, int nb) {
  const int na = (int)min((long long)*d.nactive, d.active_cap);
  for (int k0 = b * kBlock; k0 < na; k0 += nb * kBlock) {
    const int k = k0 + threadIdx.x;
    const size_t c = k < na ? (size_t)d.active[k] : 0;
    const int n = k < na ? d.cnt[c] : 0;
    const int nmax = wave_max(n);
    if (k < na) fold_node(d.off, d.blocksum, d.list, o, d, d.list_cap, c, n, nmax);
  }
}

// Row j of a dense 7x7 deposit, nodes i0 .. i0 + kBlock - 1: the nodes with contributions
// are folded.
__device__ __forceinline__ void fold_row(const Geom &g, const ObsPtrs &o, const DepPtrs &d, int i0, int j) {
  const int i = i0 + threadIdx.x;
  const size_t c = (size_t)j * g.pitch + i;
  const int n = i < g.nx ? d.cnt[c] : 0;
  const int nmax = wave_max(n);
  if (n > 0) fold_node(d.off, d.blocksum, d.list, o, d, d.list_cap, c, n, nmax);
}

__global__ __launch_bounds__(kBlock) void k_dep_fold(ObsPtrs o, DepPtrs d) { fold_active(o, d, blockIdx.x, gridDim.x); }

__global__ __launch_bounds__(kBlock) void k_dep_fold_rows(Geom g, ObsPtrs o, DepPtrs d) {
  fold_row(g, o, d, blockIdx.x * kBlock, blockIdx.y);
}

// ---------------------------------------------------------------- batched (ensembles)
//
// One launch per pass over the nudged members of an ensemble: the blocks of pass p are the
// members' blocks in member order.  A block map (NudgePlan::map, built on the host) gives
// each block its member, its index among the member's blocks of the pass and their number,
// in one load; each block runs the single-member body of its member with that index.

__device__ __forceinline__ int nudge_blocks(const Geom &g) { return (g.nx + kBlock - 1) / kBlock; }

// Batched zero passes give each block more work than the single-member ones, so that the
// descriptor loads of a block start are spread over more nodes.
constexpr int kZeroPerBlock = 16 * kBlock;  // counts zeroed per block

// Blocks of a fold over a deposit's active nodes (at most kFoldBlocks, grid-stride).
inline int fold_blocks(const DepPtrs &d, int max_blocks) {
  return (int)std::max<long long>(1, std::min<long long>((d.active_cap + kBlock - 1) / kBlock, max_blocks));
}

// Tile pass of entry k of a batched launch's tile map (two tiles per block, one per half, of
// members M; k < 0: none).  A half is whole wavefronts, so its entry and its member's
// geometry and deposit descriptors are wavefront-uniform: they are read once, by scalar
// loads, into registers, instead of field by field after each of dep_tile's barriers.
__device__ __forceinline__ void member_tile(const NudgeMember *M, const int4 *map, int k, DepTilePairLds &L,
                                            bool api) {
  const int ku = __builtin_amdgcn_readfirstlane(k);
  const int4 w = map[ku < 0 ? 0 : ku];
  const int m = __builtin_amdgcn_readfirstlane(w.x), t = __builtin_amdgcn_readfirstlane(w.y);
  const Geom g = M[m].g;
  const DepPtrs d = M[m].d;
  dep_tile(g, d, ku < 0 ? -1 : t, L, api);
}

// The head of a step's batched deposit: every member's observer update (bilinear members
// bin their contributions) and the 7x7 members' zero passes (nhead map entries `hmap`, pass
// NB_OBS or NB_ZERO) in the first blocks, then the tile passes of the bilinear members'
// previous deposit when it was deferred (ntail_tiles map entries `tmap` of members `tail`,
// two per block; they read only their own set's bins and write their own set).  The
// observer blocks go first: they wait on atomics, and the tile passes fill in around them.
__global__ __launch_bounds__(kBlock) void k_nb_head(const NudgeMember *M, const NudgeMember *tail, const int4 *tmap,
                                                    int ntail_tiles, const int4 *hmap, int nhead, int t) {
  __shared__ DepTilePairLds L;
  if ((int)blockIdx.x >= nhead) {
    member_tile(tail, tmap, half_tile(blockIdx.x - nhead, ntail_tiles), L, false);
    return;
  }
  const int4 w = hmap[blockIdx.x];
  const NudgeMember &d = M[w.x];
  if (w.w == NB_OBS) {
    // the step's place in the record interval as the single-member loop derives it
    head_one(d, w.y * kBlock + threadIdx.x, t);
  } else {
    zero_part(d.g.plane, d.d, w.y, w.z);
  }
}

// Counts of every 7x7 member's candidates.
__global__ __launch_bounds__(kBlock) void k_nb_count(const NudgeMember *M, const int4 *map) {
  const int4 w = map[blockIdx.x];
  const NudgeMember &d = M[w.x];
  count_one(d.g, d.right, d.o, d.d.cnt, (long long)w.y * kBlock + threadIdx.x);
}

__global__ __launch_bounds__(kBlock) void k_nb_scan_tiles(const NudgeMember *M, const int4 *map) {
  __shared__ int s[kBlock];
  const int4 w = map[blockIdx.x];
  const NudgeMember &d = M[w.x];
  scan_tile(d.g.plane, d.d.cnt, d.d.off, d.d.blocksum, w.y, s);
}

__global__ __launch_bounds__(kBlock) void k_nb_scan_sums(const NudgeMember *M) {
  __shared__ int s[kBlock];
  __shared__ int carry;
  const NudgeMember &d = M[blockIdx.x];
  if (d.d.mode == RBN_DEPOSIT_BILINEAR) return;  // uniform per block: no count scan
  const int ntiles = (int)((d.g.plane + kScanTile - 1) / kScanTile);
  scan_sums(ntiles, d.d.blocksum, d.d.off + d.g.plane, s, carry);
}

// The list fill of the 7x7 members.
__global__ __launch_bounds__(kBlock) void k_nb_fill(const NudgeMember *M, const int4 *map) {
  const int4 w = map[blockIdx.x];
  const NudgeMember &d = M[w.x];
  const DepPtrs &p = d.d;
  fill_one(d.g, d.right, d.o, p.off, p.blocksum, p.fill, p.list, p.list_cap, p.overflow, p.dense ? nullptr : p.active,
           p.nactive, p.active_cap, (long long)w.y * kBlock + threadIdx.x);
}

// The 7x7 members' folds.
__global__ __launch_bounds__(kBlock) void k_nb_fold(const NudgeMember *M, const int4 *map) {
  const int4 w = map[blockIdx.x];
  const NudgeMember &d = M[w.x];
  if (d.d.dense) {
    const int bx = nudge_blocks(d.g), j = w.y / bx, i0 = (w.y - j * bx) * kBlock;
    fold_row(d.g, d.o, d.d, i0, j);
  } else {
    fold_active(d.o, d.d, w.y, w.z);
  }
}

// The bilinear members' tile passes (api: as dep_tile's).
__global__ __launch_bounds__(kBlock) void k_nb_tiles(const NudgeMember *M, const int4 *map, int ntiles, int api) {
  __shared__ DepTilePairLds L;
  member_tile(M, map, half_tile(blockIdx.x, ntiles), L, api != 0);
}

}  // namespace

namespace {

// Job blockIdx.y, split over gridDim.x blocks.
__global__ __launch_bounds__(kBlock) void k_obs_scatter(const ObsJob *jobs, const unsigned char *staging) {
  const ObsJob j = jobs[blockIdx.y];
  const double *src = (const double *)(staging + j.src);
  const int k0 = blockIdx.x * kBlock + threadIdx.x, stride = gridDim.x * kBlock;
  for (int k = k0; k < 3 * j.n; k += stride) j.rx[k] = src[k];
  if (j.chk) {
    const int *map = (const int *)(src + 3 * (size_t)j.n);
    for (int k = k0; k < j.n; k += stride) j.chk[k] = map[k];
  }
}

}  // namespace

hipError_t launch_obs_scatter(hipStream_t s, const ObsJob *jobs, int njobs, const unsigned char *staging) {
  if (njobs <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_obs_scatter, dim3(kScatterSplit, njobs), dim3(kBlock), 0, s, jobs, staging);
  return hipGetLastError();
}

void nudge_batch_blocks(const NudgeMember &d, int blocks[kNudgeBatches]) {
  for (int p = 0; p < kNudgeBatches; p++) blocks[p] = 0;
  blocks[NB_OBS] = (d.o.n + kBlock - 1) / kBlock;
  if (d.d.mode == RBN_DEPOSIT_BILINEAR) {  // binned in the observer pass, then the tile pass
    blocks[NB_TILES] = dep_tiles(d.g);
    return;
  }
  const long long nc = (long long)d.o.n * kCells;
  blocks[NB_ZERO] = (int)((d.g.plane + kZeroPerBlock - 1) / kZeroPerBlock);
  blocks[NB_CAND] = (int)((nc + kBlock - 1) / kBlock);
  blocks[NB_SCAN] = (int)((d.g.plane + kScanTile - 1) / kScanTile);
  blocks[NB_FILL] = blocks[NB_CAND];
  blocks[NB_FOLD] = d.d.dense ? (d.g.nx + kBlock - 1) / kBlock * d.g.ny : fold_blocks(d.d, 256);
}

hipError_t launch_nudge_batch(hipStream_t s, const NudgeMember *members, const NudgePlan &plan, int t,
                              const NudgeMember *tail, bool defer, bool api) {
  if (plan.nm <= 0) return hipSuccess;
  const dim3 b(kBlock);
  const int *tot = plan.tot;
  auto at = [&](int p) { return plan.map + plan.off[p]; };
  const int ntail = tail ? tot[NB_TILES] : 0, nhead = tot[NB_OBS] + tot[NB_ZERO];
  if (ntail + nhead > 0)
    hipLaunchKernelGGL(k_nb_head, dim3((ntail + 1) / 2 + nhead), b, 0, s, members, tail ? tail : members, at(NB_TILES),
                       ntail, at(NB_OBS), nhead, t);
  if (tot[NB_CAND] > 0) hipLaunchKernelGGL(k_nb_count, dim3(tot[NB_CAND]), b, 0, s, members, at(NB_CAND));
  if (tot[NB_SCAN] > 0) {  // the 7x7 members' count scan and list fill
    hipLaunchKernelGGL(k_nb_scan_tiles, dim3(tot[NB_SCAN]), b, 0, s, members, at(NB_SCAN));
    hipLaunchKernelGGL(k_nb_scan_sums, dim3(plan.nm), b, 0, s, members);
  }
  if (tot[NB_FILL] > 0) hipLaunchKernelGGL(k_nb_fill, dim3(tot[NB_FILL]), b, 0, s, members, at(NB_FILL));
  if (tot[NB_FOLD] > 0) hipLaunchKernelGGL(k_nb_fold, dim3(tot[NB_FOLD]), b, 0, s, members, at(NB_FOLD));
  if (!defer && tot[NB_TILES] > 0)
    hipLaunchKernelGGL(k_nb_tiles, dim3((tot[NB_TILES] + 1) / 2), b, 0, s, members, at(NB_TILES), tot[NB_TILES],
                       api ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_nudge_tiles(hipStream_t s, const NudgeMember *members, const NudgePlan &plan, bool api) {
  if (plan.nm <= 0 || plan.tot[NB_TILES] <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_nb_tiles, dim3((plan.tot[NB_TILES] + 1) / 2), dim3(kBlock), 0, s, members,
                     plan.map + plan.off[NB_TILES], plan.tot[NB_TILES], api ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_obs_table(hipStream_t s, const int *ids, int nrec, int n, const int *lookup, int lo, int range,
                            int *map, int *first_missing) {
  if (nrec <= 0 || n <= 0) return hipSuccess;
  const size_t count = (size_t)nrec * (size_t)n;
  const unsigned grid = (unsigned)std::min<size_t>((count + kBlock - 1) / kBlock, 16384);
  hipLaunchKernelGGL(k_tab_fill, dim3(grid), dim3(kBlock), 0, s, map, count);
  hipLaunchKernelGGL(k_tab_map, dim3(grid), dim3(kBlock), 0, s, ids, count, n, lookup, lo, range, map);
  hipLaunchKernelGGL(k_tab_check, dim3((unsigned)nrec), dim3(kBlock), 0, s, map, n, first_missing);
  return hipGetLastError();
}

hipError_t launch_obs_deposit(hipStream_t s, int branch, int ntempread, const Geom &g, int right_limit,
                              const ObsPtrs &o, const DepPtrs &d, const DepPtrs *tail, bool defer) {
  const size_t n = g.plane;
  const bool bil = d.mode == RBN_DEPOSIT_BILINEAR;
  const int zb = bil ? 0 : (int)std::min<size_t>((n + kBlock - 1) / kBlock, 4096);
  const int obs_blocks = (o.n + kBlock - 1) / kBlock;
  // only a bilinear deposit has a tile pass to carry (a 7x7 set has no bins)
  const int tb = tail && tail->mode == RBN_DEPOSIT_BILINEAR && tail->tcnt && tail->bin ? (dep_tiles(g) + 1) / 2 : 0;
  if (tb + obs_blocks + zb > 0)
    hipLaunchKernelGGL(k_obs_zero, dim3(tb + obs_blocks + zb), dim3(kBlock), 0, s, g, o, branch, ntempread, tb,
                       obs_blocks, d, tail ? *tail : d);
  if (bil) {
    // the observers binned their contributions; the tile pass writes every node
    if (!defer) hipLaunchKernelGGL(k_dep_tiles, dim3((dep_tiles(g) + 1) / 2), dim3(kBlock), 0, s, g, d);
    return hipGetLastError();
  }
  // a 7x7 deposit completes here whatever defer says (it has no pass to defer)
  if (o.n <= 0) return hipGetLastError();
  const long long work = (long long)o.n * kCells;
  const dim3 ob((unsigned)((work + kBlock - 1) / kBlock));
  const dim3 fb((unsigned)fold_blocks(d, 1024)), rows((g.nx + kBlock - 1) / kBlock, g.ny);
  const int ntiles = (int)((n + kScanTile - 1) / kScanTile);
  hipLaunchKernelGGL(k_dep_count, ob, dim3(kBlock), 0, s, g, right_limit, o, d.cnt);
  hipLaunchKernelGGL(k_scan_tiles, dim3(ntiles), dim3(kBlock), 0, s, n, d.cnt, d.off, d.blocksum);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kBlock), 0, s, ntiles, d.blocksum, d.off + n);
  hipLaunchKernelGGL(k_dep_fill, ob, dim3(kBlock), 0, s, g, right_limit, o, d.off, d.blocksum, d);
  if (d.dense)
    hipLaunchKernelGGL(k_dep_fold_rows, rows, dim3(kBlock), 0, s, g, o, d);
  else
    hipLaunchKernelGGL(k_dep_fold, fb, dim3(kBlock), 0, s, o, d);
  return hipGetLastError();
}

}  // namespace rbn
