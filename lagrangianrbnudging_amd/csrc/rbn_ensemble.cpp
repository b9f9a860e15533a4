// rbn_ensemble.cpp — host side of the batched ensemble (rbn_ensemble_* in include/rbn.h):
// member validation, work lists, the captured cycle of kCycle steps, the autotuner.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "rbn_ensemble.h"
#include "rbn_host.h"
#include "rbn_internal.h"

struct rbn_ensemble {
  int dev = 0, prec = RBN_FP32;
  std::vector<rbn_ctx *> m;
  hipStream_t stream = nullptr, pstream = nullptr;
  hipEvent_t ev_lbm = nullptr, ev_part[2] = {nullptr, nullptr};
  void *d_members = nullptr;
  int2 *d_items = nullptr;  // one-step tiles, particle chunks, two-step tiles
  int n_lbm = 0, n_part = 0, n_lbm2 = 0;
  int *d_counters = nullptr;  // [2 sets][2 phases]
  int cus = 1;
  // nudged members: their observer/deposit descriptors (two rows: deposit set 0, set 1) and
  // per-pass block prefixes; the second deposit set of each (indexed by member); the events
  // between the deposits (on pstream) and the lattice launches (on stream)
  std::vector<int> nudged;  // member indices
  std::vector<rbn::DepPtrs> spare;
  hipEvent_t ev_dep[2] = {nullptr, nullptr}, ev_lat[2] = {nullptr, nullptr}, ev_seq = nullptr;
  // autotuner: when seg_a is set, a nudged call records seg_a / seg_b on stream around its
  // graph replays, or around its eager steps cyc_t0 .. cyc_t0 + 2L - 1 without replays
  // (seg_steps steps)
  hipEvent_t seg_a = nullptr, seg_b = nullptr;
  int seg_steps = 0;
  void *d_nmembers = nullptr;
  int4 *d_nmap = nullptr;
  size_t nmap_cap = 0;
  rbn::NudgePlan plan{};
  // their record uploads: pinned staging buffers with a device twin each, taken in turn by
  // the steps that load records, copied on cstream beside the running kernels (copied: the
  // copy ended; done: the device twin's reader ended)
  struct Stage {
    unsigned char *h = nullptr, *d = nullptr;
    size_t cap = 0;
    hipEvent_t done = nullptr, copied = nullptr;
  };
  static constexpr int kStages = 4;
  Stage stage[kStages];
  int stage_next = 0;
  hipStream_t cstream = nullptr;
  std::vector<int> loading, rid;
  // and their captured cycle: ncycle steps from call step cyc_t0 (the least common multiple of
  // the velocity ring and every record interval, and of 4 in the two-step mode; 0: no graph),
  // the record scatters inside it (call step t of the cycle, the members loading, their
  // staging region), two staging sets (uploaded whole before a replay) and two graph
  // instances alternating between replays
  int ncycle = 0, cyc_t0 = 1;
  // first call step of a replay's record window: cyc_t0, or cyc_t0 + 1 when two-step
  // launches carry the next pair's first observer update (its records load one launch early)
  int plan_t0 = 1;
  struct CycleLoad {
    int t;
    std::vector<int> ld;
    size_t off, bytes;
  };
  std::vector<CycleLoad> cyc;
  size_t cyc_bytes = 0;
  Stage cstage[2];
  hipGraph_t ngraph[2] = {nullptr, nullptr};
  hipGraphExec_t nexec[2] = {nullptr, nullptr};
  // tuning
  int tile_w = 256, tile_h = 1, wg_per_cu = 4, order = 0;
  bool built = false;
  // each member's state as the last successful call left it (step, ring position, planes):
  // a member stepped on its own since then makes the next call rebuild descriptors and graphs
  struct Sig {
    int istep, cur;
    const void *fA, *F, *nud;
  };
  std::vector<Sig> sig;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::string err;
};

namespace {

int efail(rbn_ensemble *e, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (e) e->err = buf;
  rbn::thread_fail(code, buf);
  return code;
}

#define EHIP(e, expr)                                                                          \
  do {                                                                                         \
    hipError_t e_ = (expr);                                                                    \
    if (e_ != hipSuccess) return efail((e), RBN_EDEVICE, "%s failed: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

void drop_graph(rbn_ensemble *e) {
  if (e->exec) (void)hipGraphExecDestroy(e->exec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  e->exec = nullptr;
  e->graph = nullptr;
  for (int k = 0; k < 2; k++) {
    if (e->nexec[k]) (void)hipGraphExecDestroy(e->nexec[k]);
    if (e->ngraph[k]) (void)hipGraphDestroy(e->ngraph[k]);
    e->nexec[k] = nullptr;
    e->ngraph[k] = nullptr;
  }
}

// The second deposit set of a nudged member: the sizes of its own, the overflow flag shared.
void free_spare(rbn::DepPtrs &s) {
  for (void *p : {(void *)s.cnt, (void *)s.fill, (void *)s.off, (void *)s.list, (void *)s.blocksum,
                  (void *)s.active, (void *)s.nactive, (void *)s.Tbar, (void *)s.wq, (void *)s.S, (void *)s.Wi,
                  s.tw, (void *)s.tcnt, (void *)s.bin, (void *)s.ocnt})
    if (p) (void)hipFree(p);
  s = rbn::DepPtrs{};
}

int alloc_spare(rbn_ensemble *e, const rbn_ctx *c, rbn::DepPtrs &s) {
  free_spare(s);
  rbn::DepPtrs x = c->dep;
  x.cnt = x.fill = x.off = x.list = x.blocksum = x.active = x.nactive = nullptr;
  x.Tbar = x.wq = nullptr;
  x.S = x.Wi = nullptr;
  x.tw = nullptr;
  x.tcnt = x.ocnt = nullptr;
  x.bin = nullptr;
  s = x;
  const size_t plane = c->geom.plane, ntiles = (plane + rbn::kDepScanTile - 1) / rbn::kDepScanTile;
  const bool bil = x.mode == RBN_DEPOSIT_BILINEAR;
  const size_t sums = bil ? plane * sizeof(long long) : 0;
  const size_t dtiles = (size_t)rbn::dep_tiles(c->geom);
  struct A {
    void **p;
    size_t b;
  } a[] = {{(void **)&s.S, sums}, {(void **)&s.Wi, sums}, {(void **)&s.cnt, plane * sizeof(int)},
           {&s.tw, plane * 2 * c->rs},
           {(void **)&s.tcnt, bil ? dtiles * rbn::kTcntStride * sizeof(int) : 0},
           {(void **)&s.ocnt, bil ? plane * sizeof(int) : 0},
           {(void **)&s.bin, bil ? dtiles * (size_t)x.bin_cap * 2 * sizeof(long long) : 0},
           {(void **)&s.fill, plane * sizeof(int)},
           {(void **)&s.off, (plane + 1) * sizeof(int)},
           {(void **)&s.blocksum, ntiles * sizeof(int)},
           {(void **)&s.list, (size_t)x.list_cap * sizeof(int)},
           {(void **)&s.active, (size_t)x.active_cap * sizeof(int)},
           {(void **)&s.nactive, sizeof(int)},
           {(void **)&s.Tbar, plane * sizeof(double)},
           {(void **)&s.wq, plane * sizeof(double)}};
  for (auto &y : a) {
    if (y.b == 0 && y.p != (void **)&s.cnt) continue;  // parts of the other deposit mode
    const size_t b = std::max<size_t>(y.b, sizeof(int));
    EHIP(e, hipMalloc(y.p, b));
    EHIP(e, hipMemset(*y.p, 0, b));
  }
  return RBN_OK;
}

// Second deposit sets for the nudged members whose own set they do not match.
int ensure_spares(rbn_ensemble *e) {
  e->spare.resize(e->m.size());
  for (int q : e->nudged) {
    const rbn::DepPtrs &d = e->m[(size_t)q]->dep;
    rbn::DepPtrs &s = e->spare[(size_t)q];
    if (s.cnt && s.list_cap == d.list_cap && s.active_cap == d.active_cap && s.mode == d.mode &&
        s.dense == d.dense && s.overflow == d.overflow && s.bin_cap == d.bin_cap)
      continue;
    int rc = alloc_spare(e, e->m[(size_t)q], s);
    if (rc) return rc;
  }
  for (hipEvent_t *ev : {&e->ev_dep[0], &e->ev_dep[1], &e->ev_lat[0], &e->ev_lat[1], &e->ev_seq})
    if (!*ev) EHIP(e, hipEventCreateWithFlags(ev, hipEventDisableTiming));
  return RBN_OK;
}

template <typename R>
void fill_member(const rbn_ctx *c, const rbn::DepPtrs *spare, int tile_w, int tile_h, rbn::EnsMember<R> &d) {
  memset(&d, 0, sizeof d);
  d.g = c->geom;
  d.c = *reinterpret_cast<const rbn::StepConst<R> *>(c->consts.data());
  d.pf[0] = (R *)c->fA;
  d.pf[1] = (R *)c->fB;
  d.pg[0] = (R *)c->gA;
  d.pg[1] = (R *)c->gB;
  d.F[0] = (R *)c->F;
  d.F[1] = (R *)c->F2;
  for (int k = 0; k < rbn::kCycle; k++) d.uv[k] = (R *)c->uv[k];
  d.rho = (R *)c->rho;
  d.T = (R *)c->T;
  d.cur0 = c->cur;
  d.tiles_x = (c->geom.nx + tile_w - 1) / tile_w;
  d.tiles_y = (c->geom.ny + tile_h - 1) / tile_h;
  const int th2 = sizeof(R) == 8 ? rbn::kEnsTH64 : rbn::kEnsTH32;
  d.t2x = (c->geom.nx + 63) / 64;
  d.t2y = (c->geom.ny + th2 - 1) / th2;
  d.part = c->d_part;
  d.np = (c->P.particle && c->d_part && c->P.variant != RBN_VARIANT_NUDGING) ? c->np : 0;
  d.err = c->d_err;
  d.bad = c->d_err + 1;
  d.variant = c->P.variant;
  if (c->P.variant == RBN_VARIANT_NUDGING) {
    const bool bil = c->dep.mode == RBN_DEPOSIT_BILINEAR;
    for (int p = 0; p < 2; p++) {
      const rbn::DepPtrs &s = p ? c->dep : *spare;
      d.tw[p] = s.tw;
      d.bin[p] = bil ? s.bin : nullptr;
      d.tcnt[p] = bil ? s.tcnt : nullptr;
      d.S[p] = bil ? s.S : nullptr;
      d.Wi[p] = bil ? s.Wi : nullptr;
      d.ocnt[p] = bil ? s.ocnt : nullptr;
    }
    d.bin_cap = c->dep.bin_cap;
    d.nud[0] = (R *)c->nud;
    d.nud[1] = (R *)c->nud2;
  }
}

size_t loading_members(const rbn_ensemble *e, int t, std::vector<int> &ld);
bool pairs(const rbn_ensemble *e);
bool rides(const rbn_ensemble *e);

// Nudged members' graph replays start kReplayLead steps after the first launch of a call:
// those steps run eagerly while the host fetches the first replay's records (a cycle's
// worth, about 120 MB for 64 members of 1e3..1e4 observers every 5 steps: 5-6 ms of host
// copying) and uploads them, which would otherwise leave the GPU idle before that replay.
// Even, and a multiple of the velocity ring.
constexpr int kReplayLead = 30;
static_assert(kReplayLead % 2 == 0 && kReplayLead % rbn::kCycle == 0, "the lead keeps the launch parities");

// Observer/deposit descriptors of the nudged members as of their current step, the block
// map of the batched passes, and the record uploads of the captured cycle.
int build_nudged(rbn_ensemble *e) {
  const int nm = (int)e->nudged.size();
  if (nm == 0) return RBN_OK;
  std::vector<rbn::NudgeMember> d(2 * (size_t)nm);
  std::vector<int> blocks((size_t)rbn::kNudgeBatches * nm, 0);
  for (int q = 0; q < nm; q++) {
    rbn_ctx *c = e->m[(size_t)e->nudged[(size_t)q]];
    rbn::NudgeMember &x = d[(size_t)q];
    x.g = c->geom;
    x.right = c->P.fix_q3 ? c->P.nx - 3 : c->P.nx - 2;
    x.ntempread = c->P.ntempread;
    x.istep0 = c->istep;
    x.o = rbn::host::obs_ptrs(c);
    x.d = c->dep;
    d[(size_t)nm + q] = x;  // set 1: the member's own deposit arrays
    x.d = e->spare[(size_t)e->nudged[(size_t)q]];
    rbn::nudge_batch_blocks(x, &blocks[(size_t)q * rbn::kNudgeBatches]);
  }
  std::vector<int4> map;
  for (int p = 0; p < rbn::kNudgeBatches; p++) {
    e->plan.off[p] = (int)map.size();
    for (int q = 0; q < nm; q++) {
      const int nb = blocks[(size_t)q * rbn::kNudgeBatches + p];
      for (int b = 0; b < nb; b++) map.push_back(make_int4(q, b, nb, p));
    }
    e->plan.tot[p] = (int)map.size() - e->plan.off[p];
  }
  e->plan.nm = nm;
  if (!e->d_nmembers) EHIP(e, hipMalloc(&e->d_nmembers, d.size() * sizeof(rbn::NudgeMember)));
  if (e->nmap_cap < map.size()) {
    if (e->d_nmap) (void)hipFree(e->d_nmap);
    e->d_nmap = nullptr;
    e->nmap_cap = 0;
    EHIP(e, hipMalloc((void **)&e->d_nmap, std::max<size_t>(map.size(), 1) * sizeof(int4)));
    e->nmap_cap = map.size();
  }
  e->plan.map = e->d_nmap;
  EHIP(e, hipMemcpy(e->d_nmembers, d.data(), d.size() * sizeof(rbn::NudgeMember), hipMemcpyHostToDevice));
  if (!map.empty()) EHIP(e, hipMemcpy(e->d_nmap, map.data(), map.size() * sizeof(int4), hipMemcpyHostToDevice));
  // A cycle of L steps replays the same kernel arguments when L is a multiple of the
  // velocity ring and of every record interval, in the two-step mode also of 4 (an even
  // number of launches, each flipping the A/B planes, an even number of them flipping the F
  // planes); longer cycles than kMaxNudgedCycle run eagerly.
  constexpr int kMaxNudgedCycle = 120;
  long long L = pairs(e) ? 2 * rbn::kCycle : rbn::kCycle;
  for (int q : e->nudged) {
    const long long a = e->m[(size_t)q]->P.ntempread;
    long long x = L, y = a;
    while (y) {
      const long long r = x % y;
      x = y;
      y = r;
    }
    L = L / x * a;
    if (L > kMaxNudgedCycle) break;
  }
  e->ncycle = L <= kMaxNudgedCycle ? (int)L : 0;
  e->cyc_t0 = (pairs(e) ? 2 : 1) + kReplayLead;
  e->plan_t0 = e->cyc_t0 + (rides(e) ? 1 : 0);
  e->cyc.clear();
  e->cyc_bytes = 0;
  for (int t = e->plan_t0; t < e->plan_t0 + e->ncycle; t++) {
    rbn_ensemble::CycleLoad cl;
    cl.t = t;
    cl.bytes = loading_members(e, t, cl.ld);
    if (!cl.bytes) continue;
    cl.off = e->cyc_bytes;
    e->cyc_bytes += cl.bytes;
    e->cyc.push_back(std::move(cl));
  }
  return RBN_OK;
}

// Descriptors and work lists for the current tuning and member state.
int build(rbn_ensemble *e) {
  drop_graph(e);
  int rc = ensure_spares(e);
  if (rc) return rc;
  const int n = (int)e->m.size();
  const size_t dsz = e->prec == RBN_FP64 ? sizeof(rbn::EnsMember<double>) : sizeof(rbn::EnsMember<float>);
  std::vector<unsigned char> desc(dsz * (size_t)n);
  std::vector<int> tiles(n), chunks(n), tiles2(n);
  for (int i = 0; i < n; i++) {
    if (e->prec == RBN_FP64) {
      auto *d = reinterpret_cast<rbn::EnsMember<double> *>(desc.data()) + i;
      fill_member(e->m[i], &e->spare[(size_t)i], e->tile_w, e->tile_h, *d);
      tiles[i] = d->tiles_x * d->tiles_y;
      tiles2[i] = d->t2x * d->t2y;
      chunks[i] = (d->np + 255) / 256;
    } else {
      auto *d = reinterpret_cast<rbn::EnsMember<float> *>(desc.data()) + i;
      fill_member(e->m[i], &e->spare[(size_t)i], e->tile_w, e->tile_h, *d);
      tiles[i] = d->tiles_x * d->tiles_y;
      tiles2[i] = d->t2x * d->t2y;
      chunks[i] = (d->np + 255) / 256;
    }
  }
  std::vector<int2> items;
  // order bit 0 clear: member after member; set: round-robin over members (small and large
  // members interleaved, largest first).  Bit 1: odd steps visit the lattice items in
  // reverse (applied at launch).
  auto emit = [&](const std::vector<int> &cnt) {
    if ((e->order & 1) == 0) {
      for (int i = 0; i < n; i++)
        for (int t = 0; t < cnt[i]; t++) items.push_back(make_int2(i, t));
    } else {
      std::vector<int> idx(n);
      for (int i = 0; i < n; i++) idx[i] = i;
      std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return cnt[a] > cnt[b]; });
      const int mx = n ? cnt[idx[0]] : 0;
      for (int t = 0; t < mx; t++)
        for (int i : idx)
          if (t < cnt[i]) items.push_back(make_int2(i, t));
    }
  };
  emit(tiles);
  e->n_lbm = (int)items.size();
  emit(chunks);
  e->n_part = (int)items.size() - e->n_lbm;
  emit(tiles2);
  e->n_lbm2 = (int)items.size() - e->n_lbm - e->n_part;
  if (e->d_members) (void)hipFree(e->d_members);
  if (e->d_items) (void)hipFree(e->d_items);
  e->d_members = nullptr;
  e->d_items = nullptr;
  EHIP(e, hipMalloc(&e->d_members, std::max<size_t>(desc.size(), 64)));
  EHIP(e, hipMalloc((void **)&e->d_items, std::max<size_t>(items.size() * sizeof(int2), 64)));
  EHIP(e, hipMemcpy(e->d_members, desc.data(), desc.size(), hipMemcpyHostToDevice));
  EHIP(e, hipMemcpy(e->d_items, items.data(), items.size() * sizeof(int2), hipMemcpyHostToDevice));
  if ((rc = build_nudged(e))) return rc;
  e->built = true;
  return RBN_OK;
}

int grid_size(const rbn_ensemble *e) { return e->wg_per_cu > 0 ? e->wg_per_cu * e->cus : 0; }

// order bit 2: separate lattice and particle launches (the particle launch on a second
// stream beside the next lattice launch) instead of one launch per step.  Not with nudged
// members (one launch per step).
bool separate(const rbn_ensemble *e) { return (e->order & 4) != 0 && e->nudged.empty(); }

// order bit 3 (without bit 2): two steps per launch on two-step tiles, the previous launch's
// particle updates carried; with nudged members, both steps' deposits are built before the
// launch.
bool pairs(const rbn_ensemble *e) { return (e->order & 8) != 0 && !separate(e); }

// Two steps per launch with every nudged member on the bilinear deposit: a two-step launch
// carries the observer update of the next pair's first step as extra work items (it writes
// that step's bins and the observers, which the launch does not read; a 7x7 member's update
// shares its launch with the zero pass of the deposit set the launch reads).
bool rides(const rbn_ensemble *e) {
  return pairs(e) && !e->nudged.empty() && e->plan.tot[rbn::NB_ZERO] == 0 && e->plan.tot[rbn::NB_CAND] == 0;
}

// order bit 4: no graphs, every launch eager (for debugging and tracing tools).
bool graphs(const rbn_ensemble *e) { return (e->order & 16) == 0; }

// Steps of one replay of the captured graph: the velocity ring's period, doubled in the
// two-step mode so that the A/B and F planes (flipped once per launch) also return.
int graph_steps(const rbn_ensemble *e) { return pairs(e) ? 2 * rbn::kCycle : rbn::kCycle; }

// Two-step mode: where a launch stands (steps done t, launches L, two-step launches P, particle
// updates waiting).
struct PairState {
  int t = 0, L = 0, P = 0, pending = 0;
  int head_done = -1;  // the step whose observer update an earlier launch carried
};

// 0: truth runs only, 1: nudged runs only, 2: both (the two-step kernels compile each apart).
int composition(const rbn_ensemble *e) {
  return e->nudged.empty() ? 0 : (e->nudged.size() == e->m.size() ? 1 : 2);
}

int launch_pairs(rbn_ensemble *e, const rbn::EnsStep &st, const rbn::EnsHead &head = rbn::EnsHead{}) {
  EHIP(e, rbn::launch_ensemble_pairs(e->stream, e->prec, e->d_members, e->d_items, e->n_lbm, e->n_part, e->n_lbm2,
                                     e->d_counters, st, e->tile_w, e->tile_h, grid_size(e), composition(e), head));
  return RBN_OK;
}

// The next launch of the two-step mode: two steps when two remain, carrying what is pending.
// bins2: the bilinear members' second-step deposits are summed in the tiles.
int enqueue_pair(rbn_ensemble *e, PairState &ps, int nsteps, bool moments, bool bins2 = false,
                 const rbn::EnsHead &head = rbn::EnsHead{}) {
  const bool two = ps.t + 1 < nsteps;
  rbn::EnsStep st{};
  st.k = ps.t % rbn::kCycle;
  st.lpar = ps.L & 1;
  st.fpar = ps.P & 1;
  st.ncarry = e->n_part > 0 ? ps.pending : 0;
  st.lat = two ? 2 : 1;
  st.moments = moments ? 1 : 0;
  st.reverse = (e->order & 2) && (ps.L & 1);
  st.cset = ps.L & 1;
  st.bins2 = two && bins2 ? 1 : 0;
  int rc = launch_pairs(e, st, head);
  if (rc) return rc;
  ps.pending = two ? 2 : 1;
  ps.t += two ? 2 : 1;
  ps.L++;
  if (two) ps.P++;
  return RBN_OK;
}

int launch(rbn_ensemble *e, hipStream_t s, int phase, int k, bool moments) {
  EHIP(e, rbn::launch_ensemble_phase(s, e->prec, phase, e->d_members, e->d_items, e->n_lbm, e->n_part,
                                     e->d_counters, k, e->tile_w, e->tile_h, moments ? 1 : 0,
                                     grid_size(e), phase != 1 && (e->order & 2) && (k & 1)));
  return RBN_OK;
}

// Separate phases: enqueue step k (0 .. kCycle-1) of a cycle on the two streams.  cross_waits: the
// lattice launch waits for the particle launch of step k-2 (it overwrites the velocity that
// one reads); inside a captured cycle that dependency exists only for k >= 2, across cycles
// the graph launches are ordered by the stream.
int enqueue_step(rbn_ensemble *e, int k, bool last, bool cross_waits) {
  int rc;
  if (cross_waits) EHIP(e, hipStreamWaitEvent(e->stream, e->ev_part[k & 1], 0));
  if ((rc = launch(e, e->stream, 0, k, last))) return rc;
  if (e->n_part > 0) {
    EHIP(e, hipEventRecord(e->ev_lbm, e->stream));
    EHIP(e, hipStreamWaitEvent(e->pstream, e->ev_lbm, 0));
    if ((rc = launch(e, e->pstream, 1, k, false))) return rc;
    EHIP(e, hipEventRecord(e->ev_part[k & 1], e->pstream));
  }
  return RBN_OK;
}

// One launch per step: step t (counted from the descriptors' state) carries the particle
// update of step t-1; step 0 has none to carry.
int enqueue_fused(rbn_ensemble *e, int t, bool last) {
  return launch(e, e->stream, t > 0 && e->n_part > 0 ? 2 : 0, t % rbn::kCycle, last);
}

// Captured cycle: separate phases, steps 0 .. kCycle-1 (moments on the last); one launch per
// step, steps 1 .. kCycle (no moments: the last step of a call runs eagerly and writes them).
int capture(rbn_ensemble *e) {
  drop_graph(e);
  EHIP(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
  int rc = RBN_OK;
  if (separate(e)) {
    for (int k = 0; k < rbn::kCycle && !rc; k++) rc = enqueue_step(e, k, k == rbn::kCycle - 1, k >= 2);
  } else if (pairs(e)) {
    // launches 1 .. kCycle (steps 2 .. 2 kCycle + 1), each of two steps carrying two updates
    PairState ps;
    ps.t = 2;
    ps.L = 1;
    ps.P = 1;
    ps.pending = 2;
    for (int l = 0; l < rbn::kCycle && !rc; l++) rc = enqueue_pair(e, ps, 1 << 30, false);
  } else {
    for (int t = 1; t <= rbn::kCycle && !rc; t++) rc = enqueue_fused(e, t, false);
  }
  if (!rc && separate(e) && e->n_part > 0) {
    // join the particle stream back into the origin stream
    if (hipStreamWaitEvent(e->stream, e->ev_part[1], 0) != hipSuccess)
      rc = efail(e, RBN_EDEVICE, "capture join failed");
  }
  hipGraph_t g = nullptr;
  hipError_t ee = hipStreamEndCapture(e->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (ee != hipSuccess) return efail(e, RBN_EDEVICE, "hipStreamEndCapture: %s", hipGetErrorString(ee));
  e->graph = g;
  EHIP(e, hipGraphInstantiate(&e->exec, g, nullptr, nullptr, 0));
  return RBN_OK;
}

// Advance every member's host-side state by n steps made of `launches` lattice launches, of
// which `pairs` advanced two steps (each flips the A/B planes once, each two-step launch
// the F planes once).
void advance_members(rbn_ensemble *e, int n, int launches, int pairs) {
  for (rbn_ctx *c : e->m) {
    c->istep += n;
    if (n > 0) c->at_init = false;
    if (launches & 1) {
      std::swap(c->fA, c->fB);
      std::swap(c->gA, c->gB);
    }
    if (pairs & 1) {
      std::swap(c->F, c->F2);
      std::swap(c->nud, c->nud2);
    }
    c->cur = (c->cur + n) % rbn::kCycle;
    c->moments_valid = true;
  }
}

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// The nudged members' descriptors of the deposit set of call step t.
const rbn::NudgeMember *nudge_row(const rbn_ensemble *e, int t) {
  return (const rbn::NudgeMember *)e->d_nmembers + (size_t)(t & 1) * e->nudged.size();
}

// The nudged members whose record interval starts at step t of a call (t counted from the
// call's first step, their step counts those at its start), and the staging bytes of their
// records: the jobs, then each member's 3n doubles and n-entry map, 256-aligned.
size_t loading_members(const rbn_ensemble *e, int t, std::vector<int> &ld) {
  ld.clear();
  size_t data = 0;
  for (int q : e->nudged) {
    const rbn_ctx *c = e->m[(size_t)q];
    if ((c->istep + t + 1) % c->P.ntempread == 1) {
      ld.push_back(q);
      data += align256((size_t)c->nobs * (3 * sizeof(double) + sizeof(int)));
    }
  }
  return ld.empty() ? 0 : align256(ld.size() * sizeof(rbn::ObsJob)) + data;
}

// Fetch the records of members ld for step t of the call (their providers, on the host)
// into the staging region h laid out as loading_members says; job offsets are relative to
// the region.
int fetch_records(rbn_ensemble *e, const std::vector<int> &ld, int t, unsigned char *h) {
  auto *jobs = (rbn::ObsJob *)h;
  size_t off = align256(ld.size() * sizeof(rbn::ObsJob));
  for (size_t k = 0; k < ld.size(); k++) {
    rbn_ctx *c = e->m[(size_t)ld[k]];
    const int n = c->nobs, ntr = c->P.ntempread, istep = c->istep + t + 1;
    double *rx = (double *)(h + off);
    int *chk = (int *)(rx + 3 * (size_t)n);
    e->rid.resize((size_t)std::max(n, 1));
    bool same = false;
    int rc = rbn::host::fetch_observation(c, istep + ntr - 1, rx, rx + n, rx + 2 * (size_t)n, e->rid.data(), chk, &same);
    if (rc) return efail(e, rc, "member %d: %s", ld[k], rbn_last_error(c));
    const rbn::ObsPtrs o = rbn::host::obs_ptrs(c);
    jobs[k] = {(double *)o.rx, same ? nullptr : (int *)o.chk, (long long)off, n};
    off += align256((size_t)n * (3 * sizeof(double) + sizeof(int)));
  }
  return RBN_OK;
}

// Pinned staging buffer of at least `need` bytes with a device twin.
int ensure_stage(rbn_ensemble *e, rbn_ensemble::Stage &st, size_t need) {
  if (!st.done) EHIP(e, hipEventCreateWithFlags(&st.done, hipEventDisableTiming));
  if (!st.copied) EHIP(e, hipEventCreateWithFlags(&st.copied, hipEventDisableTiming));
  EHIP(e, hipEventSynchronize(st.done));  // this buffer's previous upload has been read
  if (st.cap >= need) return RBN_OK;
  if (st.h) (void)hipHostFree(st.h);
  if (st.d) (void)hipFree(st.d);
  st.h = st.d = nullptr;
  st.cap = 0;
  EHIP(e, hipHostMalloc((void **)&st.h, need, hipHostMallocDefault));
  EHIP(e, hipMalloc((void **)&st.d, need));
  st.cap = need;
  // touch the pinned pages now: their first writes cost about 0.3 ms per MB (a 120 MB cycle
  // set took 40 ms to fill the first time, 5 ms after), which would otherwise land in a step
  std::memset(st.h, 0, need);
  return RBN_OK;
}

// Eager step t: fetch the records that start an interval at t, stage them in one pinned
// buffer and send them to the observers with one copy (on cstream, beside the kernels queued
// before it) and one scatter launch on s after it.
int load_records(rbn_ensemble *e, int t, hipStream_t s) {
  const size_t need = loading_members(e, t, e->loading);
  if (!need) return RBN_OK;
  rbn_ensemble::Stage &st = e->stage[e->stage_next];
  e->stage_next = (e->stage_next + 1) % rbn_ensemble::kStages;
  int rc = ensure_stage(e, st, need);
  if (rc || (rc = fetch_records(e, e->loading, t, st.h))) return rc;
  EHIP(e, hipMemcpyAsync(st.d, st.h, need, hipMemcpyHostToDevice, e->cstream));
  EHIP(e, hipEventRecord(st.copied, e->cstream));
  EHIP(e, hipStreamWaitEvent(s, st.copied, 0));
  EHIP(e, rbn::launch_obs_scatter(s, (const rbn::ObsJob *)st.d, (int)e->loading.size(), st.d));
  EHIP(e, hipEventRecord(st.done, s));
  return RBN_OK;
}

// Eager nudged steps run the record uploads, observer updates and deposits on pstream and
// the lattice launches on stream.  The deposit of step t goes to set t & 1 and the lattice
// launch of step t waits for it; the deposit of step t + 1 then runs beside that launch,
// after the lattice launch of step t - 1 (the last reader of its set) when that one is a
// step of this run of eager steps (from t0; earlier work is ordered before it).  last: the
// call's last step (its deposit also fills the arrays the getters read).
int enqueue_deposit(rbn_ensemble *e, int t, int t0, bool last) {
  if (t - 2 >= t0) EHIP(e, hipStreamWaitEvent(e->pstream, e->ev_lat[t & 1], 0));
  EHIP(e, rbn::launch_nudge_batch(e->pstream, nudge_row(e, t), e->plan, t, nullptr, false, last));
  EHIP(e, hipEventRecord(e->ev_dep[t & 1], e->pstream));
  return RBN_OK;
}

int enqueue_fused(rbn_ensemble *e, int t, bool last);

int enqueue_nudged_lattice(rbn_ensemble *e, int t, bool last) {
  EHIP(e, hipStreamWaitEvent(e->stream, e->ev_dep[t & 1], 0));
  int rc = enqueue_fused(e, t, last);
  if (rc) return rc;
  EHIP(e, hipEventRecord(e->ev_lat[t & 1], e->stream));
  return RBN_OK;
}

// The records of replay r of the captured cycle (call steps t0 + r L .. t0 + r L + L - 1)
// into staging set k.
int fill_cycle(rbn_ensemble *e, int k, int r) {
  for (const auto &cl : e->cyc) {
    int rc = fetch_records(e, cl.ld, cl.t + r * e->ncycle, e->cstage[k].h + cl.off);
    if (rc) return rc;
  }
  return RBN_OK;
}

// Two-step tiles of the ensemble (FP32) are deposit tiles: a launch can sum its bilinear
// members' second-step deposits from their bins.
bool tile_bins(const rbn_ensemble *e) { return e->prec == RBN_FP32; }
static_assert(rbn::kEnsTH32 == rbn::kDepTileH, "FP32 ensemble tiles are deposit tiles");

// Two-step mode with nudged members: the records and deposits of a launch's steps (step t
// into set t & 1, t + 1 into set (t + 1) & 1) run on stream before it, after the previous
// launch, which read both sets.  The first step's bilinear tile passes ride in the second
// step's first launch; the second step's are summed in the lattice launch's tiles, except in
// the call's last launch (its deposit fills the arrays the getters read).
// When the launch is followed by another two-step launch and rides(e), it carries that
// launch's first observer update (its records loaded before it); a step whose update an
// earlier launch carried runs only its tile pass.
int enqueue_nudged_pair(rbn_ensemble *e, PairState &ps, int nsteps, bool moments) {
  const int t0 = ps.t;
  const bool two = t0 + 1 < nsteps;
  const bool in_tile = two && tile_bins(e) && t0 + 2 < nsteps;
  const bool carry = rides(e) && two && t0 + 3 < nsteps;
  int rc;
  for (int t = t0; t < t0 + (two ? 2 : 1); t++) {
    const bool second = t > t0;
    if (t == ps.head_done) {
      // observers updated and binned by the previous launch: the tile pass, unless deferred
      if (!two) EHIP(e, rbn::launch_nudge_tiles(e->stream, nudge_row(e, t), e->plan, t == nsteps - 1));
      continue;
    }
    if ((rc = load_records(e, t, e->stream))) return rc;
    EHIP(e, rbn::launch_nudge_batch(e->stream, nudge_row(e, t), e->plan, t, second ? nudge_row(e, t - 1) : nullptr,
                                    second ? in_tile : two, t == nsteps - 1));
  }
  rbn::EnsHead head{};
  if (carry) {
    if ((rc = load_records(e, t0 + 2, e->stream))) return rc;
    head = {nudge_row(e, t0 + 2), e->plan.map + e->plan.off[rbn::NB_OBS], e->plan.tot[rbn::NB_OBS], t0 + 2};
  }
  if ((rc = enqueue_pair(e, ps, nsteps, moments, in_tile, head))) return rc;
  ps.head_done = carry ? t0 + 2 : -1;
  return RBN_OK;
}

// Capture the cycle of nudged steps t0 .. t0 + L - 1 of a call twice, the scatters reading
// the records from staging set 0 and set 1 (each uploaded whole, outside the graph, before
// its replay): per step the record scatter (where members start an interval) and the batched
// observer/deposit passes, and the lattice launches — one per step, or in the two-step mode
// one per two steps after both steps' deposits — all on stream (graph replays of the
// two-stream form of eager steps measured slower than those eager steps: 1.01 against
// 0.67 ms per step for 64 members).
int capture_nudged(rbn_ensemble *e) {
  for (int k = 0; k < 2; k++) {
    int rc = ensure_stage(e, e->cstage[k], std::max<size_t>(e->cyc_bytes, 256));
    if (rc) return rc;
  }
  for (int k = 0; k < 2; k++) {
    rbn_ensemble::Stage &st = e->cstage[k];
    EHIP(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    int rc = RBN_OK;
    size_t next = 0;
    // the record scatter of step t, where members start an interval
    auto scatter = [&](int t) -> int {
      if (next < e->cyc.size() && e->cyc[next].t == t) {
        const auto &cl = e->cyc[next++];
        if (rbn::launch_obs_scatter(e->stream, (const rbn::ObsJob *)(st.d + cl.off), (int)cl.ld.size(),
                                    st.d + cl.off) != hipSuccess)
          return efail(e, RBN_EDEVICE, "record scatter capture failed");
      }
      return RBN_OK;
    };
    // tail: the previous step's tile passes ride along (two-step mode, a pair's second step);
    // defer: this step's wait for the next (a pair's first step)
    auto deposit = [&](int t, bool tail, bool defer) -> int {
      int r = scatter(t);
      if (r) return r;
      if (rbn::launch_nudge_batch(e->stream, nudge_row(e, t), e->plan, t, tail ? nudge_row(e, t - 1) : nullptr,
                                  defer, false) != hipSuccess)
        return efail(e, RBN_EDEVICE, "observer/deposit capture failed");
      return RBN_OK;
    };
    if (pairs(e)) {
      // launches t0 / 2 + 1 .. t0 / 2 + L / 2 of a call (steps t0 .. t0 + L - 1, t0 = cyc_t0),
      // two steps each, carrying two updates; with rides(e) each also carries the next
      // launch's first observer update (the first launch's was carried by the launch before
      // the replay)
      PairState ps;
      ps.t = e->cyc_t0;
      ps.L = e->cyc_t0 / 2;
      ps.P = e->cyc_t0 / 2;
      ps.pending = 2;
      const bool ride = rides(e);
      for (int l = 0; l < e->ncycle / 2 && !rc; l++) {
        const int t0 = ps.t;
        if (!ride) rc = deposit(t0, false, true);
        if (!rc) rc = deposit(t0 + 1, true, tile_bins(e));
        rbn::EnsHead head{};
        if (!rc && ride) {
          rc = scatter(t0 + 2);
          head = {nudge_row(e, t0 + 2), e->plan.map + e->plan.off[rbn::NB_OBS], e->plan.tot[rbn::NB_OBS], t0 + 2};
        }
        if (!rc) rc = enqueue_pair(e, ps, 1 << 30, false, tile_bins(e), head);
      }
    } else {
      for (int t = e->cyc_t0; t < e->cyc_t0 + e->ncycle && !rc; t++)
        if (!(rc = deposit(t, false, false))) rc = enqueue_fused(e, t, false);
    }
    hipGraph_t g = nullptr;
    hipError_t ee = hipStreamEndCapture(e->stream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (ee != hipSuccess) return efail(e, RBN_EDEVICE, "hipStreamEndCapture: %s", hipGetErrorString(ee));
    e->ngraph[k] = g;
    EHIP(e, hipGraphInstantiate(&e->nexec[k], g, nullptr, nullptr, 0));
  }
  return RBN_OK;
}

rbn_ensemble::Sig member_sig(const rbn_ctx *c) { return {c->istep, c->cur, c->fA, c->F, c->nud}; }

void record_sigs(rbn_ensemble *e) {
  e->sig.clear();
  for (const rbn_ctx *c : e->m) e->sig.push_back(member_sig(c));
}

// Wait for the members' own streams.  A member whose latest velocity sits past the
// ensemble's ring (rbn_step's three-step launches use slots kCycle and up) has that slot's
// buffer swapped into the ring (pointers only: no step is pending between calls, and only
// the latest velocity is read again).  Either that or a member stepped on its own since the
// last call invalidates the descriptors and graphs.
int sync_members(rbn_ensemble *e) {
  bool same = e->sig.size() == e->m.size();
  for (size_t i = 0; i < e->m.size(); i++) {
    rbn_ctx *c = e->m[i];
    EHIP(e, hipStreamSynchronize(c->stream));
    EHIP(e, hipStreamSynchronize(c->pstream));
    if (c->cur >= rbn::kCycle) {
      const int k = c->cur % rbn::kCycle;
      std::swap(c->uv[c->cur], c->uv[k]);
      c->cur = k;
      same = false;
    }
    if (same) {
      const rbn_ensemble::Sig &a = e->sig[i];
      const rbn_ensemble::Sig b = member_sig(c);
      same = a.istep == b.istep && a.cur == b.cur && a.fA == b.fA && a.F == b.F && a.nud == b.nud;
    }
  }
  if (!same) {
    e->built = false;
    drop_graph(e);
  }
  return RBN_OK;
}

// Steps of an ensemble with nudged members.  A step is: the records of the members whose
// record interval starts (host providers, one upload), the batched observer update and
// deposit, then the lattice.  One launch per step: the lattice launch of step t carries the
// particle updates of step t - 1, and the first three run on pstream beside the previous
// step's lattice launch (enqueue_deposit).  Two-step mode: everything on stream, both
// steps' deposits before their launch (enqueue_nudged_pair).  Either way the first launch
// and the kReplayLead steps after it run eagerly, whole cycles of L steps from the graphs
// (their records fetched on the host before each replay, into the set the previous replay
// but one has read; the first replay's while the lead runs), the rest eagerly (the last
// writing the moments), then the last launch's particle updates on their own.
int run_nudged(rbn_ensemble *e, int nsteps) {
  int rc;
  const bool two = pairs(e);
  auto eager_deposit = [&](int t, int t0) {
    int r = load_records(e, t, e->pstream);
    if (r) return r;
    return enqueue_deposit(e, t, t0, t == nsteps - 1);
  };
  PairState ps;
  if (two) {
    if ((rc = enqueue_nudged_pair(e, ps, nsteps, nsteps <= 2))) return rc;
  } else {
    if ((rc = eager_deposit(0, 0)) || (rc = enqueue_nudged_lattice(e, 0, nsteps == 1))) return rc;
    ps.t = 1;
  }
  const int L = e->ncycle, t0 = e->cyc_t0;
  const int cycles = graphs(e) && L > 0 && nsteps > t0 ? (nsteps - t0 - 1) / L : 0;
  if (cycles > 0) {
    if (!e->nexec[0] && (rc = capture_nudged(e))) return rc;
    // the lead: the steps before the cycle's first, eagerly (the first replay's records are
    // fetched on the host meanwhile)
    if (two) {
      while (ps.t < t0)
        if ((rc = enqueue_nudged_pair(e, ps, nsteps, false))) return rc;
    } else {
      for (; ps.t < t0; ps.t++)
        if ((rc = eager_deposit(ps.t, 0)) || (rc = enqueue_nudged_lattice(e, ps.t, false))) return rc;
    }
    if (e->seg_a) {
      EHIP(e, hipEventRecord(e->seg_a, e->stream));
      e->seg_steps = cycles * L;
    }
    for (int r = 0; r < cycles; r++) {
      // replay r's records: filled on the host once the copy from this set two replays ago
      // has read it, copied on cstream once that replay has read the device twin, beside the
      // replay before
      rbn_ensemble::Stage &st = e->cstage[r & 1];
      EHIP(e, hipEventSynchronize(st.copied));
      if ((rc = fill_cycle(e, r & 1, r))) return rc;
      if (e->cyc_bytes) {
        EHIP(e, hipStreamWaitEvent(e->cstream, st.done, 0));
        EHIP(e, hipMemcpyAsync(st.d, st.h, e->cyc_bytes, hipMemcpyHostToDevice, e->cstream));
        EHIP(e, hipEventRecord(st.copied, e->cstream));
        EHIP(e, hipStreamWaitEvent(e->stream, st.copied, 0));
      }
      EHIP(e, hipGraphLaunch(e->nexec[r & 1], e->stream));
      EHIP(e, hipEventRecord(st.done, e->stream));
    }
    if (e->seg_a) EHIP(e, hipEventRecord(e->seg_b, e->stream));
    ps.t += cycles * L;
    if (two) {
      ps.L += cycles * L / 2;
      ps.P += cycles * L / 2;
      ps.head_done = rides(e) ? ps.t : -1;  // the last replayed launch carried it
    } else {
      // the eager steps after the replays: their deposits follow the replays' work, and the
      // lattice events they wait for are recorded again outside the captured graphs
      EHIP(e, hipEventRecord(e->ev_lat[0], e->stream));
      EHIP(e, hipEventRecord(e->ev_lat[1], e->stream));
      EHIP(e, hipEventRecord(e->ev_seq, e->stream));
      EHIP(e, hipStreamWaitEvent(e->pstream, e->ev_seq, 0));
    }
  }
  // the autotuner's eager form: steps t0 .. t0 + 2L - 1, when there were no replays
  const int seg_len = 2 * L;
  const bool seg = e->seg_a && L > 0 && cycles == 0 && nsteps > t0 + seg_len;
  if (seg) e->seg_steps = seg_len;
  const int tail = ps.t;
  if (two) {
    while (ps.t < nsteps) {
      if (seg && ps.t == t0) EHIP(e, hipEventRecord(e->seg_a, e->stream));
      if ((rc = enqueue_nudged_pair(e, ps, nsteps, ps.t + 2 >= nsteps))) return rc;
      if (seg && ps.t == t0 + seg_len) EHIP(e, hipEventRecord(e->seg_b, e->stream));
    }
    if (e->n_part > 0) {
      rbn::EnsStep st{};
      st.k = ps.t % rbn::kCycle;
      st.ncarry = ps.pending;
      st.lat = 0;
      st.cset = ps.L & 1;
      if ((rc = launch_pairs(e, st))) return rc;
    }
  } else {
    for (int t = tail; t < nsteps; t++) {
      if (seg && t == t0) EHIP(e, hipEventRecord(e->seg_a, e->stream));
      if ((rc = eager_deposit(t, t == tail ? tail : 0))) return rc;
      if ((rc = enqueue_nudged_lattice(e, t, t == nsteps - 1))) return rc;
      if (seg && t + 1 == t0 + seg_len) EHIP(e, hipEventRecord(e->seg_b, e->stream));
    }
    if (e->n_part > 0 && (rc = launch(e, e->stream, 1, (nsteps - 1) % rbn::kCycle, false))) return rc;
    ps.L = ps.t = nsteps;
  }
  // The last step's deposit is in set (nsteps - 1) & 1: set 0 (the second set) after an
  // odd number of steps, copied into the members' own arrays for their readers.
  if ((nsteps & 1) == 1) {
    for (int q : e->nudged) {
      rbn_ctx *c = e->m[(size_t)q];
      const rbn::DepPtrs &s = e->spare[(size_t)q];
      const size_t plane = c->geom.plane;
      EHIP(e, hipMemcpyAsync(c->dep.cnt, s.cnt, plane * sizeof(int), hipMemcpyDeviceToDevice, e->stream));
      EHIP(e, hipMemcpyAsync(c->dep.Tbar, s.Tbar, plane * sizeof(double), hipMemcpyDeviceToDevice, e->stream));
      if (c->dep.mode == RBN_DEPOSIT_BILINEAR)
        EHIP(e, hipMemcpyAsync(c->dep.wq, s.wq, plane * sizeof(double), hipMemcpyDeviceToDevice, e->stream));
    }
  }
  EHIP(e, hipStreamSynchronize(e->stream));
  EHIP(e, hipStreamSynchronize(e->pstream));
  advance_members(e, nsteps, ps.L, ps.P);
  record_sigs(e);
  // The observer descriptors and the cycle's record plan hold the members' step counts
  // modulo the cycle (and, in the two-step mode, an even number of launches and of two-step
  // launches): they (and the graphs) stay valid after a whole number of cycles.
  if (L == 0 || nsteps % L) {
    e->built = false;
    drop_graph(e);
  }
  return RBN_OK;
}

int run_steps(rbn_ensemble *e, int nsteps) {
  int rc = sync_members(e);
  if (rc) return rc;
  if (nsteps == 0) return RBN_OK;
  if (!e->built && (rc = build(e))) return rc;
  // Every launch leaves the counters of the next step clear, but eager steps that end
  // mid-cycle do not leave set 0 clear for a cycle that starts over: clear both here
  // (an ordinary stream operation, never part of the captured graph).
  EHIP(e, hipMemsetAsync(e->d_counters, 0, 4 * sizeof(int), e->stream));
  for (int q : e->nudged)
    if (!e->m[(size_t)q]->d_obs_i) return efail(e, RBN_ESTATE, "member %d: nudged run without observers", q);
  if (!e->nudged.empty()) return run_nudged(e, nsteps);
  int launches = nsteps, npairs = 0;
  if (pairs(e)) {
    // launch 0 eagerly, whole cycles of kCycle two-step launches from the graph, the rest
    // eagerly (at least one launch, the last writing the moments), then the particle
    // updates of the last launch on their own.
    PairState ps;
    if ((rc = enqueue_pair(e, ps, nsteps, nsteps <= 2))) return rc;
    const int span = 2 * rbn::kCycle;
    const int cycles = graphs(e) && nsteps - ps.t >= 1 ? (nsteps - ps.t - 1) / span : 0;
    if (cycles > 0) {
      if (!e->exec && (rc = capture(e))) return rc;
      for (int c = 0; c < cycles; c++) EHIP(e, hipGraphLaunch(e->exec, e->stream));
      ps.t += cycles * span;
      ps.L += cycles * rbn::kCycle;
      ps.P += cycles * rbn::kCycle;
    }
    while (ps.t < nsteps) {
      const bool last = ps.t + 2 >= nsteps;
      if ((rc = enqueue_pair(e, ps, nsteps, last))) return rc;
    }
    if (e->n_part > 0) {
      rbn::EnsStep st{};
      st.k = ps.t % rbn::kCycle;
      st.ncarry = ps.pending;
      st.lat = 0;
      st.cset = ps.L & 1;
      if ((rc = launch_pairs(e, st))) return rc;
    }
    launches = ps.L;
    npairs = ps.P;
  } else if (separate(e)) {
    const int cycles = graphs(e) ? nsteps / rbn::kCycle : 0, rem = nsteps - cycles * rbn::kCycle;
    if (cycles > 0) {
      if (!e->exec && (rc = capture(e))) return rc;
      for (int c = 0; c < cycles; c++) EHIP(e, hipGraphLaunch(e->exec, e->stream));
    }
    for (int k = 0; k < rem; k++)
      if ((rc = enqueue_step(e, k, k == rem - 1, k >= 2))) return rc;
    if (rem > 0 && e->n_part > 0) EHIP(e, hipStreamWaitEvent(e->stream, e->ev_part[(rem - 1) & 1], 0));
  } else {
    // step 0 eagerly, whole cycles of steps 1..kCycle*c from the graph, the rest (at least one
    // step when nsteps >= 2, the last writing the moments) eagerly, then the particle
    // update of the last step on its own.
    if ((rc = enqueue_fused(e, 0, nsteps == 1))) return rc;
    const int cycles = graphs(e) && nsteps >= 2 ? (nsteps - 2) / rbn::kCycle : 0;
    if (cycles > 0) {
      if (!e->exec && (rc = capture(e))) return rc;
      for (int c = 0; c < cycles; c++) EHIP(e, hipGraphLaunch(e->exec, e->stream));
    }
    for (int t = 1 + rbn::kCycle * cycles; t < nsteps; t++)
      if ((rc = enqueue_fused(e, t, t == nsteps - 1))) return rc;
    if (e->n_part > 0 && (rc = launch(e, e->stream, 1, (nsteps - 1) % rbn::kCycle, false))) return rc;
  }
  EHIP(e, hipStreamSynchronize(e->stream));
  EHIP(e, hipStreamSynchronize(e->pstream));
  advance_members(e, nsteps, launches, npairs);
  record_sigs(e);
  if (nsteps % rbn::kCycle || (launches & 1) || (npairs & 1)) {
    // member parity or ring position changed: rebuild descriptors and graph
    e->built = false;
    drop_graph(e);
  }
  return RBN_OK;
}

// Member state that the autotuner's trial steps change, saved on the device.
struct Snapshot {
  std::vector<void *> bufs;
  std::vector<std::pair<void *, size_t>> src;
  std::vector<int> istep, cur;
  std::vector<bool> moments_valid;
  std::vector<void *> fA, gA, F, nud;
};

void free_snapshot(Snapshot &s) {
  for (void *b : s.bufs) (void)hipFree(b);
  s.bufs.clear();
  s.src.clear();
}

int take_snapshot(rbn_ensemble *e, Snapshot &s) {
  for (rbn_ctx *c : e->m) {
    const size_t fb = c->geom.plane * c->rs;
    std::vector<std::pair<void *, size_t>> parts = {{c->fA, 9 * fb}, {c->gA, 9 * fb}, {c->F, fb},
                                                    {c->F2, fb},     {c->rho, fb},    {c->T, fb}};
    for (int k = 0; k < rbn_ctx::kVel; k++) parts.push_back({c->uv[k], 2 * fb});
    if (c->d_part && c->np > 0) parts.push_back({c->d_part, 10 * (size_t)c->np * sizeof(double)});
    if (c->P.variant == RBN_VARIANT_NUDGING) {
      // observers, nudging term and the latest deposit
      const size_t no = (size_t)c->nobs, plane = c->geom.plane;
      if (no > 0) {
        parts.push_back({c->d_obs_i, rbn::host::ONI * no * sizeof(int)});
        parts.push_back({c->d_obs_d, rbn::host::OND * no * sizeof(double)});
      }
      parts.push_back({c->nud, fb});
      parts.push_back({c->dep.cnt, plane * sizeof(int)});
      parts.push_back({c->dep.Tbar, plane * sizeof(double)});
      if (c->dep.wq) parts.push_back({c->dep.wq, plane * sizeof(double)});
    }
    for (auto &p : parts) {
      void *b = nullptr;
      hipError_t he = hipMalloc(&b, p.second);
      if (he == hipSuccess) {
        s.bufs.push_back(b);
        s.src.push_back(p);
        he = hipMemcpy(b, p.first, p.second, hipMemcpyDeviceToDevice);
      }
      if (he != hipSuccess) {
        free_snapshot(s);
        return efail(e, RBN_EDEVICE, "autotune snapshot failed: %s", hipGetErrorString(he));
      }
    }
    s.istep.push_back(c->istep);
    s.cur.push_back(c->cur);
    s.moments_valid.push_back(c->moments_valid);
    s.fA.push_back(c->fA);
    s.gA.push_back(c->gA);
    s.F.push_back(c->F);
    s.nud.push_back(c->nud);
  }
  return RBN_OK;
}

// Put the saved state back; keep: the snapshot stays for another restore.  Waits for the
// ensemble's streams first (a failed step may leave launches in flight that would overwrite
// the restored planes); every copy is attempted, and the first error is returned.
int restore_snapshot(rbn_ensemble *e, Snapshot &s, bool keep = false) {
  (void)hipStreamSynchronize(e->stream);
  (void)hipStreamSynchronize(e->pstream);
  (void)hipStreamSynchronize(e->cstream);
  for (size_t i = 0; i < s.istep.size() && i < e->m.size(); i++) {
    rbn_ctx *c = e->m[i];
    if (c->fA != s.fA[i]) {
      std::swap(c->fA, c->fB);
      std::swap(c->gA, c->gB);
    }
    if (c->F != s.F[i]) std::swap(c->F, c->F2);
    if (c->nud != s.nud[i]) std::swap(c->nud, c->nud2);
    // the observers' record map on the device goes back too: forget which record it is for
    c->prev_rid.clear();
    c->istep = s.istep[i];
    c->cur = s.cur[i];
    c->moments_valid = s.moments_valid[i];
  }
  hipError_t first = hipSuccess;
  for (size_t k = 0; k < s.bufs.size(); k++) {
    const hipError_t he = hipMemcpy(s.src[k].first, s.bufs[k], s.src[k].second, hipMemcpyDeviceToDevice);
    if (he != hipSuccess && first == hipSuccess) first = he;
  }
  // a failed trial can stop between an observer update and its tile pass: the bilinear tile
  // counters and overflow sums (zero between deposits) of both sets go back to zero
  for (int q : e->nudged) {
    const rbn_ctx *c = e->m[(size_t)q];
    hipError_t he = rbn::host::clear_dep_bins(c->dep, c->geom, e->stream);
    if (he == hipSuccess && (size_t)q < e->spare.size() && e->spare[(size_t)q].cnt)
      he = rbn::host::clear_dep_bins(e->spare[(size_t)q], c->geom, e->stream);
    if (he != hipSuccess && first == hipSuccess) first = he;
  }
  if (hipStreamSynchronize(e->stream) != hipSuccess && first == hipSuccess) first = hipErrorUnknown;
  if (!keep) free_snapshot(s);
  e->built = false;
  drop_graph(e);
  if (first != hipSuccess) return efail(e, RBN_EDEVICE, "autotune restore failed: %s", hipGetErrorString(first));
  return RBN_OK;
}

}  // namespace

extern "C" {

void rbn_ensemble_destroy(rbn_ensemble *e) {
  if (!e) return;
  (void)hipSetDevice(e->dev);
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  if (e->pstream) (void)hipStreamSynchronize(e->pstream);
  if (e->cstream) (void)hipStreamSynchronize(e->cstream);
  drop_graph(e);
  if (e->d_members) (void)hipFree(e->d_members);
  if (e->d_items) (void)hipFree(e->d_items);
  if (e->d_counters) (void)hipFree(e->d_counters);
  if (e->d_nmembers) (void)hipFree(e->d_nmembers);
  if (e->d_nmap) (void)hipFree(e->d_nmap);
  for (auto *set : {&e->stage[0], &e->stage[1], &e->stage[2], &e->stage[3], &e->cstage[0], &e->cstage[1]}) {
    if (set->h) (void)hipHostFree(set->h);
    if (set->d) (void)hipFree(set->d);
    if (set->done) (void)hipEventDestroy(set->done);
    if (set->copied) (void)hipEventDestroy(set->copied);
  }
  if (e->ev_lbm) (void)hipEventDestroy(e->ev_lbm);
  for (hipEvent_t ev : {e->ev_part[0], e->ev_part[1], e->ev_dep[0], e->ev_dep[1], e->ev_lat[0], e->ev_lat[1], e->ev_seq})
    if (ev) (void)hipEventDestroy(ev);
  for (rbn::DepPtrs &s : e->spare) free_spare(s);
  if (e->cstream) (void)hipStreamDestroy(e->cstream);
  if (e->pstream) (void)hipStreamDestroy(e->pstream);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
}

int rbn_ensemble_create(rbn_ctx *const *members, int n, rbn_ensemble **out) {
  if (!members || n < 1 || !out) return rbn::thread_fail(RBN_EINVAL, "rbn_ensemble_create: bad arguments");
  *out = nullptr;
  for (int i = 0; i < n; i++) {
    const rbn_ctx *c = members[i];
    if (!c) return rbn::thread_fail(RBN_EINVAL, "null member");
    if (!c->initialised) return rbn::thread_fail(RBN_ESTATE, "member not initialised");
    if (c->dev != members[0]->dev || c->prec != members[0]->prec)
      return rbn::thread_fail(RBN_EINVAL, "ensemble members must share device and precision");
    for (int j = 0; j < i; j++)
      if (members[j] == c) return rbn::thread_fail(RBN_EINVAL, "a context appears twice in the ensemble");
  }
  rbn_ensemble *e = new rbn_ensemble();
  e->dev = members[0]->dev;
  e->prec = members[0]->prec;
  e->m.assign(members, members + n);
  for (int i = 0; i < n; i++)
    if (members[i]->P.variant == RBN_VARIANT_NUDGING) e->nudged.push_back(i);
  if (hipSetDevice(e->dev) != hipSuccess ||
      hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&e->pstream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&e->cstream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&e->ev_lbm, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&e->ev_part[0], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&e->ev_part[1], hipEventDisableTiming) != hipSuccess ||
      hipMalloc((void **)&e->d_counters, 4 * sizeof(int)) != hipSuccess ||
      hipMemset(e->d_counters, 0, 4 * sizeof(int)) != hipSuccess ||
      hipDeviceGetAttribute(&e->cus, hipDeviceAttributeMultiprocessorCount, e->dev) != hipSuccess) {
    rbn_ensemble_destroy(e);
    return rbn::thread_fail(RBN_EDEVICE, "ensemble device set-up failed");
  }
  *out = e;
  return RBN_OK;
}

const char *rbn_ensemble_last_error(const rbn_ensemble *e) { return e ? e->err.c_str() : rbn::thread_error(); }

int rbn_ensemble_set_tuning(rbn_ensemble *e, int tile_w, int tile_h, int wg_per_cu, int order) {
  if (!e) return rbn::thread_fail(RBN_EINVAL, "null ensemble");
  if (!(tile_w == 64 || tile_w == 128 || tile_w == 256) || tile_h < 1 || tile_h > 64 ||
      (tile_w * tile_h) % 256 != 0 || wg_per_cu < 0 || wg_per_cu > 32 || order < 0 || order > 31)
    return efail(e, RBN_EINVAL, "tile must be 64/128/256 wide with a multiple of 256 nodes; wg_per_cu 0..32; order 0..31");
  e->tile_w = tile_w;
  e->tile_h = tile_h;
  e->wg_per_cu = wg_per_cu;
  e->order = order;
  e->built = false;
  drop_graph(e);
  return RBN_OK;
}

int rbn_ensemble_get_tuning(const rbn_ensemble *e, int *tile_w, int *tile_h, int *wg_per_cu, int *order) {
  if (!e) return rbn::thread_fail(RBN_EINVAL, "null ensemble");
  if (tile_w) *tile_w = e->tile_w;
  if (tile_h) *tile_h = e->tile_h;
  if (wg_per_cu) *wg_per_cu = e->wg_per_cu;
  if (order) *order = e->order;
  return RBN_OK;
}

int rbn_ensemble_step(rbn_ensemble *e, int nsteps) {
  if (!e || nsteps < 0) return rbn::thread_fail(RBN_EINVAL, "rbn_ensemble_step: bad arguments");
  if (hipSetDevice(e->dev) != hipSuccess) return efail(e, RBN_EDEVICE, "hipSetDevice failed");
  int rc = run_steps(e, nsteps);
  if (rc) return rc;
  for (size_t q = 0; q < e->m.size(); q++) {
    rbn_ctx *c = e->m[q];
    int h = 0, he[2] = {0, 0};
    EHIP(e, hipMemcpy(he, c->d_err, 2 * sizeof(int), hipMemcpyDeviceToHost));
    if (he[0] || he[1]) EHIP(e, hipMemset(c->d_err, 0, 2 * sizeof(int)));
    if (he[1]) return efail(e, RBN_ENUMERIC, "ensemble member %zu: non-finite moments (rho or T) by step %d", q, c->istep);
    if (he[0]) return efail(e, RBN_ERANGE, "a particle of ensemble member %zu left its lattice", q);
    if (c->P.variant == RBN_VARIANT_NUDGING && c->dep.overflow) {
      EHIP(e, hipMemcpy(&h, c->dep.overflow, sizeof(int), hipMemcpyDeviceToHost));
      if (h) return efail(e, RBN_ERANGE, "deposit list overflow in a nudged ensemble member");
    }
  }
  return RBN_OK;
}

int rbn_ensemble_autotune(rbn_ensemble *e, int trial_steps, double *best_ms_per_step) {
  if (!e) return rbn::thread_fail(RBN_EINVAL, "null ensemble");
  if (hipSetDevice(e->dev) != hipSuccess) return efail(e, RBN_EDEVICE, "hipSetDevice failed");
  struct Cand {
    int tw, th, wpc, order;
  };
  std::vector<Cand> cands;
  const int shapes[][2] = {{256, 1}, {256, 2}, {256, 4}, {128, 2}, {128, 4}, {64, 4}, {64, 8}};
  const int wpcs[] = {0, 2, 4, 8};
  // with nudged members the separate-launch orders (bit 2) do not apply
  const int norders = e->nudged.empty() ? 8 : 4;
  for (auto &sh : shapes)
    for (int w : wpcs)
      for (int o = 0; o < norders; o++) cands.push_back({sh[0], sh[1], w, o});
  // two steps per launch: the tile is fixed (64 x 8 in FP32, 64 x 4 in FP64)
  for (int w : wpcs)
    for (int o = 8; o < 12; o++) cands.push_back({256, 1, w, o});
  const int tile_w0 = e->tile_w, tile_h0 = e->tile_h, wpc0 = e->wg_per_cu, order0 = e->order;
  int rc = sync_members(e);
  if (rc) return rc;
  Snapshot snap;
  if ((rc = take_snapshot(e, snap))) return rc;
  hipEvent_t a = nullptr, b = nullptr;
  double best = 1e300;
  Cand bestc = cands[0];
  // Time one candidate: a warm-up call, then the timed steps, then back to the snapshot.
  auto trial_of = [&](const Cand &c, float *ms_per_step) -> int {
    int r = rbn_ensemble_set_tuning(e, c.tw, c.th, c.wpc, c.order);
    if (r) return r;
    const int gs = graph_steps(e);
    const int reps = std::max(1, (trial_steps + gs - 1) / gs), trial = reps * gs;
    float ms = 0.f;
    if (!e->nudged.empty()) {
      // eager steps (nudged members): time a call of `trial` steps after a warm-up call,
      // then restore, so that every candidate replays the same steps (and records)
      if ((r = run_steps(e, gs))) return r;
      EHIP(e, hipEventRecord(a, e->stream));
      if ((r = run_steps(e, trial))) return r;
      EHIP(e, hipEventRecord(b, e->stream));
      EHIP(e, hipEventSynchronize(b));
      if ((r = restore_snapshot(e, snap, true))) return r;
    } else {
      if ((r = run_steps(e, gs))) return r;  // build, warm (a whole cycle: descriptors stay valid)
      if (!e->built && (r = build(e))) return r;
      if (!e->exec && (r = capture(e))) return r;
      // The captured cycle starts past the first launch of a call (counter set 1), which the
      // eager launches above leave used: clear the counters before replaying.
      EHIP(e, hipMemsetAsync(e->d_counters, 0, 4 * sizeof(int), e->stream));
      EHIP(e, hipEventRecord(a, e->stream));
      for (int s = 0; s < reps; s++) EHIP(e, hipGraphLaunch(e->exec, e->stream));
      EHIP(e, hipEventRecord(b, e->stream));
      EHIP(e, hipEventSynchronize(b));
      EHIP(e, hipStreamSynchronize(e->pstream));
      if (pairs(e))
        advance_members(e, trial, trial / 2, trial / 2);
      else
        advance_members(e, trial, trial, 0);
    }
    EHIP(e, hipEventElapsedTime(&ms, a, b));
    *ms_per_step = ms / trial;
    return RBN_OK;
  };
  // Nudged members: eager steps overlap each step's deposit with the previous lattice launch
  // (one launch per step), captured cycles replay the steps in order.  Time both with the
  // best launch shape over steps t0 .. t0 + 2L - 1 of a call of 2L + t0 + 1 steps (t0: the
  // cycle's first step; the graph form: two replays).
  auto forms = [&]() -> int {
    int r = rbn_ensemble_set_tuning(e, bestc.tw, bestc.th, bestc.wpc, bestc.order & ~16);
    if (r || (!e->built && (r = build(e)))) return r;
    if (e->ncycle <= 0) return RBN_OK;
    const int T = 2 * e->ncycle + e->cyc_t0 + 1;
    double ms_form[2] = {1e300, 1e300};
    e->seg_a = a;
    e->seg_b = b;
    for (int form = 0; form < 2 && !r; form++) {
      if ((r = rbn_ensemble_set_tuning(e, bestc.tw, bestc.th, bestc.wpc, (bestc.order & ~16) | (form ? 16 : 0))))
        break;
      e->seg_steps = 0;
      if ((r = run_steps(e, T))) break;
      float ms = 0.f;
      if (e->seg_steps > 0 && hipEventElapsedTime(&ms, a, b) == hipSuccess) ms_form[form] = ms / e->seg_steps;
      r = restore_snapshot(e, snap, true);
    }
    e->seg_a = e->seg_b = nullptr;
    if (r) return r;
    if (ms_form[1] < ms_form[0]) bestc.order |= 16;
    if (std::min(ms_form[0], ms_form[1]) < 1e300) best = std::min(ms_form[0], ms_form[1]);
    return RBN_OK;
  };
  if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess)
    rc = efail(e, RBN_EDEVICE, "autotune: hipEventCreate failed");
  for (size_t k = 0; k < cands.size() && !rc; k++) {
    float ms = 0.f;
    if ((rc = trial_of(cands[k], &ms))) break;
    if (ms < best) {
      best = ms;
      bestc = cands[k];
    }
  }
  if (!rc && !e->nudged.empty() && best < 1e300) rc = forms();
  // Every exit path: the members go back to the snapshot (their state before the call).
  const int rc2 = restore_snapshot(e, snap);
  if (a) (void)hipEventDestroy(a);
  if (b) (void)hipEventDestroy(b);
  if (rc || rc2) {
    const std::string msg = e->err;
    (void)rbn_ensemble_set_tuning(e, tile_w0, tile_h0, wpc0, order0);
    e->err = msg;
    return rc ? rc : rc2;
  }
  if ((rc = rbn_ensemble_set_tuning(e, bestc.tw, bestc.th, bestc.wpc, bestc.order))) return rc;
  if (best_ms_per_step) *best_ms_per_step = best;
  return RBN_OK;
}

}  // extern "C"
