// rbn_api.cpp — the C-ABI of librbn (declared in include/rbn.h): context lifetime, initial
// state, the per-step launch sequence, records, observers and measurement.
//
// The launch sequence of rbn_step restates one iteration of the reference main loops:
//   truth run (Trajectories/code/temppart_dump.c:1674-1759):
//     K1 fused step  ->  K2 particles (+ record on istep % noutconfig == 0)  ->  energy
//   nudged run (Nudging/code/tempvel_nudge.c:2346-2411):
//     observer update + ordered deposit (read_prtcl)  ->  K1 fused step (nudged)  ->  energy
// Particles only read the step's velocity, so running K2 after the collision (the
// reference runs it between hydrovar and the collision) yields identical results.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <unordered_map>
#include <vector>

#include "rbn_host.h"
#include "rbn_internal.h"

using rbn::Geom;

namespace rbn {
namespace host {

int fail(rbn_ctx *c, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  rbn::thread_fail(code, buf);
  return code;
}

int dalloc(rbn_ctx *c, void **p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) bytes = 256;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(c, RBN_ENOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  }
  return RBN_OK;
}

void dfree(void *&p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

// The second deposit set of a nudged run (dep_overlap): the sizes of dep, the overflow flag
// shared.
void free_dep2(rbn_ctx *c) {
  rbn::DepPtrs &s = c->dep2;
  for (void *p : {(void *)s.cnt, (void *)s.fill, (void *)s.off, (void *)s.list, (void *)s.blocksum,
                  (void *)s.active, (void *)s.nactive, (void *)s.Tbar, (void *)s.wq, (void *)s.S, (void *)s.Wi,
                  s.tw, (void *)s.tcnt, (void *)s.bin, (void *)s.ocnt})
    if (p) (void)hipFree(p);
  s = rbn::DepPtrs{};
}

hipError_t clear_dep_bins(const rbn::DepPtrs &d, const rbn::Geom &g, hipStream_t s) {
  if (d.mode != RBN_DEPOSIT_BILINEAR) return hipSuccess;
  const size_t plane = g.plane, tiles = (size_t)rbn::dep_tiles(g);
  hipError_t e = hipSuccess;
  if (d.tcnt && e == hipSuccess) e = hipMemsetAsync(d.tcnt, 0, tiles * rbn::kTcntStride * sizeof(int), s);
  if (d.ocnt && e == hipSuccess) e = hipMemsetAsync(d.ocnt, 0, plane * sizeof(int), s);
  if (d.S && e == hipSuccess) e = hipMemsetAsync(d.S, 0, plane * sizeof(long long), s);
  if (d.Wi && e == hipSuccess) e = hipMemsetAsync(d.Wi, 0, plane * sizeof(long long), s);
  return e;
}

int ensure_dep2(rbn_ctx *c) {
  for (hipEvent_t *ev : {&c->ev_dep[0], &c->ev_dep[1], &c->ev_lat[0], &c->ev_lat[1]})
    if (!*ev) HIPCHK(c, hipEventCreateWithFlags(ev, hipEventDisableTiming));
  if (c->dep2.cnt) return RBN_OK;
  rbn::DepPtrs &s = c->dep2;
  s = c->dep;
  s.cnt = s.fill = s.off = s.list = s.blocksum = s.active = s.nactive = nullptr;
  s.Tbar = s.wq = nullptr;
  s.S = s.Wi = nullptr;
  s.tw = nullptr;
  s.tcnt = s.ocnt = nullptr;
  s.bin = nullptr;
  const size_t plane = c->geom.plane, ntiles = (plane + rbn::kDepScanTile - 1) / rbn::kDepScanTile;
  const bool bil = s.mode == RBN_DEPOSIT_BILINEAR;
  const size_t sums = bil ? plane * sizeof(long long) : 0;
  const size_t dtiles = (size_t)rbn::dep_tiles(c->geom);
  struct B {
    void **p;
    size_t b;
  } a[] = {{(void **)&s.S, sums}, {(void **)&s.Wi, sums},
           {&s.tw, plane * 2 * c->rs},
           {(void **)&s.tcnt, bil ? dtiles * rbn::kTcntStride * sizeof(int) : 0},
           {(void **)&s.ocnt, bil ? plane * sizeof(int) : 0},
           {(void **)&s.bin, bil ? dtiles * (size_t)s.bin_cap * 2 * sizeof(long long) : 0},
           {(void **)&s.off, (plane + 1) * sizeof(int)},
           {(void **)&s.fill, plane * sizeof(int)},
           {(void **)&s.blocksum, ntiles * sizeof(int)},
           {(void **)&s.list, (size_t)c->dep.list_cap * sizeof(int)},
           {(void **)&s.active, (size_t)c->dep.active_cap * sizeof(int)},
           {(void **)&s.nactive, sizeof(int)},
           {(void **)&s.Tbar, plane * sizeof(double)},
           {(void **)&s.wq, plane * sizeof(double)},
           {(void **)&s.cnt, plane * sizeof(int)}};  // last: a set without cnt is incomplete
  for (auto &x : a) {
    if (x.b == 0) continue;
    int rc = dalloc(c, x.p, x.b);
    if (rc) {
      free_dep2(c);
      return rc;
    }
    HIPCHK(c, hipMemsetAsync(*x.p, 0, x.b, c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));  // cleared before pstream writes it
  return RBN_OK;
}

template <typename R>
void fill_consts(const rbn_params &P, rbn::StepConst<R> &c) {
  // The FP64 values are computed with the reference's expressions (T:1318-1320, 752,
  // 896-921; N:1929-1931); the FP32 ones are those values rounded, with the plate rule
  // expressed for the shifted g.
  const double om = 1. / P.relax, omg = 1. / P.relaxg;
  c.om = (R)om;
  c.omg = (R)omg;
  if (sizeof(R) == 8) {
    c.one_m_om = (R)(1. - om);
    c.one_m_omg = (R)(1. - omg);
  } else {
    // FP32 relaxes the deviations from the reference level: the level's share,
    // w_k*level*((1-om) + om - 1), is dropped, which is exact only if the two rounded factors
    // sum to 1.  Rounding 1-om and om separately leaves a relative defect of up to an ulp,
    // which multiplies the deviation sum (T - T0) every step and drifts T linearly in time
    // (about 4e-8 per step at T - T0 = 0.5); 1 - (float)om is exact for om in [0.5, 2].
    c.one_m_om = (R)1 - c.om;
    c.one_m_omg = (R)1 - c.omg;
  }
  const double B = 1.0 - 0.5 * omg;
  for (int k = 0; k < 9; k++) {
    c.src[k] = (R)((1.0 - 0.5 * om) * rbn::kW[k]);
    c.bw[k] = (R)(B * rbn::kW[k]);
  }
  c.g0 = (R)(-P.gravity);
  c.alpha = (R)P.alpha;
  const bool shifted = sizeof(R) == 4;
  const double T0 = shifted ? P.T0_shift : 0.0;
  c.rho0 = (R)(shifted ? P.rho0_shift : 0.0);
  c.T0 = (R)T0;
  const int top_k[3] = {2, 5, 6}, bot_k[3] = {4, 7, 8};
  for (int q = 0; q < 3; q++) {
    c.wtop[q] = (R)(2.0 * rbn::kW[top_k[q]] * (shifted ? (P.Tup - T0) : P.Tup));
    c.wbot[q] = (R)(2.0 * rbn::kW[bot_k[q]] * (shifted ? (P.Tdown - T0) : P.Tdown));
  }
}

const char *const kKernelNames[KT_COUNT] = {"lbm", "particles", "deposit", "adjoint", "lbm_fused", "lbm2", "lbm3"};

int kernel_index(const char *name) {
  for (int k = 0; k < KT_COUNT; k++)
    if (name && strcmp(name, kKernelNames[k]) == 0) return k;
  return -1;
}

hipEvent_t take_event(rbn_ctx *c) {
  if (!c->event_pool.empty()) {
    hipEvent_t e = c->event_pool.back();
    c->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

int drain_timing(rbn_ctx *c) {
  for (auto &p : c->pending) {
    HIPCHK(c, hipEventSynchronize(p.b));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, p.a, p.b));
    c->k_launches[p.kernel]++;
    c->k_ms[p.kernel] += ms;
    c->event_pool.push_back(p.a);
    c->event_pool.push_back(p.b);
  }
  c->pending.clear();
  return RBN_OK;
}


// Host [nx][ny] (y fastest) <-> device [j][i] with pitch, converting to the real type.
template <typename R>
void to_device_layout(const rbn_ctx *c, const double *h, std::vector<unsigned char> &buf) {
  const Geom &g = c->geom;
  buf.assign(g.plane * sizeof(R), 0);
  R *d = reinterpret_cast<R *>(buf.data());
  for (int i = 0; i < g.nx; i++)
    for (int j = 0; j < g.ny; j++) d[(size_t)j * g.pitch + i] = (R)h[(size_t)i * g.ny + j];
}

int upload_field(rbn_ctx *c, void *dst, const double *h) {
  std::vector<unsigned char> buf;
  if (c->prec == RBN_FP64)
    to_device_layout<double>(c, h, buf);
  else
    to_device_layout<float>(c, h, buf);
  HIPCHK(c, hipMemcpyAsync(dst, buf.data(), buf.size(), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return RBN_OK;
}

int upload_uv(rbn_ctx *c, void *dst, const double *hu, const double *hv) {
  const Geom &g = c->geom;
  std::vector<unsigned char> buf(2 * g.plane * c->rs, 0);
  for (int i = 0; i < g.nx; i++)
    for (int j = 0; j < g.ny; j++) {
      const size_t o = 2 * ((size_t)j * g.pitch + i), h = (size_t)i * g.ny + j;
      if (c->prec == RBN_FP64) {
        reinterpret_cast<double *>(buf.data())[o] = hu[h];
        reinterpret_cast<double *>(buf.data())[o + 1] = hv[h];
      } else {
        reinterpret_cast<float *>(buf.data())[o] = (float)hu[h];
        reinterpret_cast<float *>(buf.data())[o + 1] = (float)hv[h];
      }
    }
  HIPCHK(c, hipMemcpyAsync(dst, buf.data(), buf.size(), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return RBN_OK;
}

int download_field(rbn_ctx *c, const void *src, double *h, double add, int cs) {
  const Geom &g = c->geom;
  // the last node's value is element (plane - 1) * cs of src: copy no further
  std::vector<unsigned char> buf(((g.plane - 1) * (size_t)cs + 1) * c->rs);
  HIPCHK(c, hipMemcpyAsync(buf.data(), src, buf.size(), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int i = 0; i < g.nx; i++)
    for (int j = 0; j < g.ny; j++) {
      size_t o = ((size_t)j * g.pitch + i) * cs;
      double v = c->prec == RBN_FP64 ? reinterpret_cast<const double *>(buf.data())[o]
                                     : (double)reinterpret_cast<const float *>(buf.data())[o];
      h[(size_t)i * g.ny + j] = v + add;
    }
  return RBN_OK;
}

int check_ctx(rbn_ctx *c) {
  if (!c) return rbn::thread_fail(RBN_EINVAL, "null context");
  if (hipSetDevice(c->dev) != hipSuccess) return fail(c, RBN_EDEVICE, "hipSetDevice(%d) failed", c->dev);
  return RBN_OK;
}

uint64_t splitmix64(uint64_t &s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

// ---------------------------------------------------------------- particle storage

enum { PX, PY, PU, PV, PXO, PYO, PUO, PVO, PINVTAU, PBETA, PNARR };

rbn::PartPtrs part_ptrs(rbn_ctx *c, double *rec) {
  rbn::PartPtrs pp{};
  const size_t n = (size_t)c->np;
  double *b = c->d_part;
  pp.n = c->np;
  pp.x = b + PX * n;
  pp.y = b + PY * n;
  pp.u = b + PU * n;
  pp.v = b + PV * n;
  pp.xo = b + PXO * n;
  pp.yo = b + PYO * n;
  pp.uo = b + PUO * n;
  pp.vo = b + PVO * n;
  pp.inv_tau = b + PINVTAU * n;
  pp.beta = b + PBETA * n;
  pp.rec = rec;
  pp.err = c->d_err;
  return pp;
}

rbn::ObsPtrs obs_view(int nobs, int *ints, double *d) {
  rbn::ObsPtrs o{};
  const size_t n = (size_t)nobs;
  o.n = nobs;
  o.chk = ints + OCHK * n;
  o.bcy = ints + OBCY * n;
  o.xidx = ints + OXIDX * n;
  o.yidx = ints + OYIDX * n;
  o.dx = d + ODX * n;
  o.dy = d + ODY * n;
  o.dT = d + ODT * n;
  o.xint = d + OXINT * n;
  o.yint = d + OYINT * n;
  o.Tint = d + OTINT * n;
  o.rx = d + ORX * n;
  o.ry = d + ORY * n;
  o.rT = d + ORT * n;
  return o;
}

rbn::ObsPtrs obs_ptrs(rbn_ctx *c) {
  rbn::ObsPtrs o = obs_view(c->nobs, c->d_obs_i, c->d_obs_d);
  if (c->tab_rec >= 0) {
    const size_t n = (size_t)c->nobs, r = (size_t)c->tab_rec;
    o.rx = c->d_tab + 3 * r * n;
    o.ry = o.rx + n;
    o.rT = o.rx + 2 * n;
    o.chk = c->d_tabmap + r * n;
  }
  return o;
}

int check_err_flag(rbn_ctx *c) {
  int e[2] = {0, 0}, h = 0;
  HIPCHK(c, hipMemcpyAsync(e, c->d_err, 2 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (e[0] || e[1]) HIPCHK(c, hipMemsetAsync(c->d_err, 0, 2 * sizeof(int), c->stream));
  // The reference stops on a non-finite state (Trajectories/code/temppart_dump.c:414-425);
  // here the step call reports it.
  if (e[1])
    return fail(c, RBN_ENUMERIC, "non-finite moments (rho or T) in a lattice step by step %d", c->istep);
  if (e[0])
    return fail(c, RBN_ERANGE, "a particle left the lattice (interpolation corner out of range) by step %d", c->istep);
  if (c->dep.overflow) {
    HIPCHK(c, hipMemcpyAsync(&h, c->dep.overflow, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (h) return fail(c, RBN_ERANGE, "deposit list overflow at step %d", c->istep);
  }
  return RBN_OK;
}

// Hand a copied record slot to the callback and free the slot.
int deliver_slot(rbn_ctx *c, rbn_ctx::RecSlot &s) {
  if (s.step < 0) return RBN_OK;
  const int step = s.step;
  s.step = -1;
  HIPCHK(c, hipEventSynchronize(s.copied));
  const size_t n = (size_t)c->np;
  const double *r = s.h;
  rbn_record rec{};
  rec.step = step;
  rec.n = c->np;
  rec.id = c->pid.data();
  const double **dst[9] = {&rec.x, &rec.y, &rec.u, &rec.v, &rec.xo, &rec.yo, &rec.uo, &rec.vo, &rec.temp};
  for (int q = 0; q < 9; q++) *dst[q] = r + q * n;
  if (c->P.prtcl_writefull) {
    const double **full[8] = {&rec.dtu, &rec.dtv, &rec.dxu, &rec.dxv, &rec.dyu, &rec.dyv, &rec.fu, &rec.fv};
    for (int q = 0; q < 8; q++) *full[q] = r + (9 + q) * n;
  }
  if (c->rec_fn && c->rec_fn(c->rec_user, &rec) != 0)
    return fail(c, RBN_ECALLBACK, "record callback failed at step %d", step);
  return RBN_OK;
}

// Deliver every pending record, oldest first.
int deliver_all(rbn_ctx *c) {
  for (int k = 0; k < rbn_ctx::kRecSlots; k++) {
    rbn_ctx::RecSlot &s = c->rec[(c->rec_next + k) % rbn_ctx::kRecSlots];
    int rc = deliver_slot(c, s);
    if (rc) return rc;
  }
  return RBN_OK;
}

// The slot K2 writes this record step: deliver its previous content first (the oldest
// pending record), which also guarantees its copy has finished before it is overwritten.
int claim_record_slot(rbn_ctx *c, rbn_ctx::RecSlot **out) {
  rbn_ctx::RecSlot &s = c->rec[c->rec_next];
  c->rec_next = (c->rec_next + 1) % rbn_ctx::kRecSlots;
  int rc = deliver_slot(c, s);
  if (rc) return rc;
  *out = &s;
  return RBN_OK;
}

// Queue the copy of a freshly written slot (written by K2 on stream ks) on the copy stream.
int queue_record_copy(rbn_ctx *c, rbn_ctx::RecSlot &s, int step, hipStream_t ks) {
  const size_t narr = c->P.prtcl_writefull ? 17 : 9;
  HIPCHK(c, hipEventRecord(s.written, ks));
  HIPCHK(c, hipStreamWaitEvent(c->copy_stream, s.written, 0));
  HIPCHK(c, hipMemcpyAsync(s.h, s.d, narr * (size_t)c->np * sizeof(double), hipMemcpyDeviceToHost, c->copy_stream));
  HIPCHK(c, hipEventRecord(s.copied, c->copy_stream));
  s.step = step;
  return RBN_OK;
}

void free_record_slots(rbn_ctx *c) {
  for (auto &s : c->rec) {
    if (s.d) (void)hipFree(s.d);
    if (s.h) (void)hipHostFree(s.h);
    s.d = nullptr;
    s.h = nullptr;
    s.step = -1;
  }
}

int run_energy(rbn_ctx *c, double out[2]) {
  const int nb = rbn::energy_blocks(c->geom);
  double *partial = c->d_energy, *res = c->d_energy + 2 * 1024;
  hipError_t e = rbn::launch_energy(c->stream, c->prec, c->geom, c->rho, c->uv[c->cur],
                                    c->T, partial, nb, res);
  if (e != hipSuccess) return fail(c, RBN_EDEVICE, "energy launch failed: %s", hipGetErrorString(e));
  HIPCHK(c, hipMemcpyAsync(out, res, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return RBN_OK;
}

// Fetch record `rstep` through the provider into rx, ry, rT (ids into rid) and map observer
// ids to record entries in chk (the first match, as the reference's search N:1480-1487).
int fetch_observation(rbn_ctx *c, int rstep, double *rx, double *ry, double *rT, int *rid, int *chk,
                      bool *same_map) {
  const int n = c->nobs;
  if (!c->obs_fn) return fail(c, RBN_ESTATE, "no observation provider set (record %d needed)", rstep);
  if (c->obs_fn(c->obs_user, rstep, n, rid, rx, ry, rT) != 0)
    return fail(c, RBN_ECALLBACK, "observation provider failed for record %d", rstep);
  // chk[i] = the first record entry with observer i's id.  When the ids are those of the
  // previous record, the map on the device is still current.
  *same_map = false;
  if (c->prev_rid.size() == (size_t)n && std::equal(rid, rid + n, c->prev_rid.begin())) {
    *same_map = true;
    return RBN_OK;
  }
  int lo = 0, hi = -1;
  if (n > 0) {
    lo = *std::min_element(rid, rid + n);
    hi = *std::max_element(rid, rid + n);
  }
  const long long range = (long long)hi - (long long)lo + 1;
  if (n > 0 && range <= 4LL * n + 1024) {  // dense ids: a direct table
    std::vector<int> first((size_t)range, -1);
    for (int j = n - 1; j >= 0; j--) first[(size_t)(rid[j] - lo)] = j;
    for (int i = 0; i < n; i++) {
      const long long id = c->oid[(size_t)i];
      const int j = (id >= lo && id <= hi) ? first[(size_t)(id - lo)] : -1;
      if (j < 0) return fail(c, RBN_ENOTFOUND, "observer id %d missing from record %d", c->oid[(size_t)i], rstep);
      chk[i] = j;
    }
  } else {
    std::unordered_map<int, int> first;
    first.reserve((size_t)n * 2);
    for (int j = 0; j < n; j++) first.emplace(rid[j], j);
    for (int i = 0; i < n; i++) {
      auto it = first.find(c->oid[(size_t)i]);
      if (it == first.end()) return fail(c, RBN_ENOTFOUND, "observer id %d missing from record %d", c->oid[(size_t)i], rstep);
      chk[i] = it->second;
    }
  }
  c->prev_rid.assign(rid, rid + n);
  return RBN_OK;
}

// Fetch record `rstep` and upload it on stream s_up.
int load_observation(rbn_ctx *c, int rstep, hipStream_t s_up) {
  const int n = c->nobs;
  if (!c->obs_fn) return fail(c, RBN_ESTATE, "no observation provider set (record %d needed)", rstep);
  rbn_ctx::ObsSlot &s = c->obs_slot[c->obs_next];
  c->obs_next ^= 1;
  const size_t N = (size_t)(n > 0 ? n : 1);
  const size_t bytes = 3 * N * sizeof(double) + 2 * N * sizeof(int);
  if (!s.copied) HIPCHK(c, hipEventCreateWithFlags(&s.copied, hipEventDisableTiming));
  HIPCHK(c, hipEventSynchronize(s.copied));  // the slot's previous upload has finished
  if (s.cap < N) {
    if (s.h) HIPCHK(c, hipHostFree(s.h));
    s.h = nullptr;
    s.cap = 0;
    if (hipHostMalloc((void **)&s.h, bytes, hipHostMallocDefault) != hipSuccess) {
      s.h = nullptr;
      return fail(c, RBN_ENOMEM, "pinned observation buffer allocation failed");
    }
    s.cap = N;
  }
  double *rx = (double *)s.h, *ry = rx + N, *rT = ry + N;
  int *rid = (int *)(rT + N), *chk = rid + N;
  bool same_map = false;
  int rc = fetch_observation(c, rstep, rx, ry, rT, rid, chk, &same_map);
  if (rc) return rc;
  rbn::ObsPtrs o = obs_ptrs(c);
  // rx, ry, rT are consecutive arrays on the device and in the slot
  HIPCHK(c, hipMemcpyAsync((void *)o.rx, rx, 3 * (size_t)n * sizeof(double), hipMemcpyHostToDevice, s_up));
  if (!same_map) HIPCHK(c, hipMemcpyAsync((void *)o.chk, chk, (size_t)n * sizeof(int), hipMemcpyHostToDevice, s_up));
  HIPCHK(c, hipEventRecord(s.copied, s_up));
  return RBN_OK;
}

int load_observation(rbn_ctx *c, int rstep) { return load_observation(c, rstep, c->stream); }

// ---------------------------------------------------------------- observation table on the device
// rbn_step over an observation table (obs_device): the caller's table is copied to the
// device at the first step after it or the observers change, and the observer map of every
// record is computed there, so that a record becomes current by pointer (obs_ptrs): no host
// copy, no upload and no map per record (each cost about 28 us of host time at 1e4
// observers, about 5 us per step at records every 5 steps).  Tables over kTabCap bytes,
// whose entries per record differ from the observers' count, or whose observers have
// sparse or repeated ids are served record by record as other providers are.

constexpr size_t kTabCap = 8ull << 30;

void free_obs_table(rbn_ctx *c) {
  dfree_t(c->d_tab);
  dfree_t(c->d_tabmap);
  c->tab_miss.clear();
  c->tab_rec = -1;
  c->tab_tried = false;
}

// Build the device table when it suits; false (and nothing built) when it does not.
int build_obs_table(rbn_ctx *c, bool *built) {
  *built = false;
  c->tab_tried = true;
  const rbn_ctx::ObsTable &t = c->obs_table;
  const int n = c->nobs;
  if (n <= 0 || t.n != n || t.nrec <= 0) return RBN_OK;
  const size_t N = (size_t)n, R = (size_t)t.nrec;
  if (R * N * (3 * sizeof(double) + sizeof(int)) > kTabCap) return RBN_OK;
  const int lo = *std::min_element(c->oid.begin(), c->oid.end());
  const int hi = *std::max_element(c->oid.begin(), c->oid.end());
  const long long range = (long long)hi - lo + 1;
  if (range > 4LL * n + 1024) return RBN_OK;
  std::vector<int> lookup((size_t)range, -1);
  for (int i = 0; i < n; i++) {
    int &l = lookup[(size_t)(c->oid[(size_t)i] - lo)];
    if (l >= 0) return RBN_OK;  // two observers share an id
    l = i;
  }
  void *p = nullptr;
  int rc;
  if ((rc = dalloc(c, &p, R * 3 * N * sizeof(double)))) return rc;
  c->d_tab = (double *)p;
  if ((rc = dalloc(c, &p, R * N * sizeof(int)))) {
    free_obs_table(c);
    c->tab_tried = true;
    return rc;
  }
  c->d_tabmap = (int *)p;
  int *ids = nullptr, *dl = nullptr, *miss = nullptr;
  hipError_t e = hipMalloc(&ids, R * N * sizeof(int));
  if (e == hipSuccess) e = hipMalloc(&dl, (size_t)range * sizeof(int));
  if (e == hipSuccess) e = hipMalloc(&miss, R * sizeof(int));
  const size_t w = N * sizeof(double);
  if (e == hipSuccess) e = hipMemcpy2DAsync(c->d_tab, 3 * w, t.x, w, w, R, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipMemcpy2DAsync(c->d_tab + N, 3 * w, t.y, w, w, R, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipMemcpy2DAsync(c->d_tab + 2 * N, 3 * w, t.T, w, w, R, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(ids, t.id, R * N * sizeof(int), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(dl, lookup.data(), (size_t)range * sizeof(int), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = rbn::launch_obs_table(c->stream, ids, t.nrec, n, dl, lo, (int)range, c->d_tabmap, miss);
  c->tab_miss.assign(R, 0);
  if (e == hipSuccess) e = hipMemcpyAsync(c->tab_miss.data(), miss, R * sizeof(int), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  for (void *q : {(void *)ids, (void *)dl, (void *)miss})
    if (q) (void)hipFree(q);
  if (e != hipSuccess) {
    free_obs_table(c);
    c->tab_tried = true;
    return fail(c, RBN_EDEVICE, "observation table upload failed: %s", hipGetErrorString(e));
  }
  *built = true;
  return RBN_OK;
}

// Make record rstep current for rbn_step: from the device table when it holds it, else as
// load_observation does (which also reports a record the table lacks).
int load_observation_dev(rbn_ctx *c, int rstep, hipStream_t s_up) {
  if (c->d_tab) {
    const rbn_ctx::ObsTable &t = c->obs_table;
    const int *it = std::lower_bound(t.steps, t.steps + t.nrec, rstep);
    if (it != t.steps + t.nrec && *it == rstep) {
      const int r = (int)(it - t.steps);
      if (c->tab_miss[(size_t)r] < c->nobs)
        return fail(c, RBN_ENOTFOUND, "observer id %d missing from record %d", c->oid[(size_t)c->tab_miss[(size_t)r]],
                    rstep);
      c->tab_rec = r;
      return RBN_OK;
    }
  }
  c->tab_rec = -1;
  return load_observation(c, rstep, s_up);
}

// End of rbn_step: the current record back in the observer arrays' own buffer, where the
// adjoint, the ensemble and record-by-record loads expect it.
int settle_records(rbn_ctx *c) {
  if (c->tab_rec < 0) return RBN_OK;
  const size_t n = (size_t)c->nobs, r = (size_t)c->tab_rec;
  const rbn::ObsPtrs o = obs_view(c->nobs, c->d_obs_i, c->d_obs_d);
  HIPCHK(c, hipMemcpyAsync((void *)o.rx, c->d_tab + 3 * r * n, 3 * n * sizeof(double), hipMemcpyDeviceToDevice,
                           c->stream));
  HIPCHK(c, hipMemcpyAsync((void *)o.chk, c->d_tabmap + r * n, n * sizeof(int), hipMemcpyDeviceToDevice, c->stream));
  c->tab_rec = -1;
  c->prev_rid.clear();  // the map in the observer arrays is no longer that of the last fetched record
  return RBN_OK;
}

int check_sample_args(rbn_ctx *c, const char *who, int field, int n, const double *x, const double *y,
                      const double *out) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (field < 0 || field > 3 || n < 0 || (n > 0 && (!x || !y || !out))) return fail(c, RBN_EINVAL, "%s: bad arguments", who);
  if (!c->initialised) return fail(c, RBN_ESTATE, "not initialised");
  if ((field == 0 || field == 1) && !c->moments_valid) return fail(c, RBN_ESTATE, "rho/T of the latest step not available");
  return RBN_OK;
}

int sample_staged(rbn_ctx *c, int field, int n, const double *x, const double *y, double *out, double *buf) {
  const void *src = field == 0   ? c->rho
                    : field == 1 ? c->T
                    : field == 2 ? c->uv[c->cur]
                                 : (const void *)((const char *)c->uv[c->cur] + c->rs);
  const int cs = field >= 2 ? 2 : 1;
  double *dx = buf, *dy = dx + n, *dout = dy + n;
  if (hipMemcpyAsync(dx, x, (size_t)n * sizeof(double), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
      hipMemcpyAsync(dy, y, (size_t)n * sizeof(double), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
      rbn::launch_sample(c->stream, c->prec, c->geom, src, cs, n, dx, dy, dout, c->d_err) != hipSuccess ||
      hipMemcpyAsync(out, dout, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    return fail(c, RBN_EDEVICE, "bilinear sampling: device operation failed");
  return check_err_flag(c);
}

int init_common_after_fields(rbn_ctx *c) {
  c->istep = c->P.nstart;
  c->at_init = false;
  c->initialised = true;
  c->moments_valid = true;
  if (c->nud) HIPCHK(c, hipMemsetAsync(c->nud, 0, c->geom.plane * c->rs, c->stream));
  if (c->dep.cnt) HIPCHK(c, hipMemsetAsync(c->dep.cnt, 0, c->geom.plane * sizeof(int), c->stream));
  if (c->dep.tw) HIPCHK(c, hipMemsetAsync(c->dep.tw, 0, c->geom.plane * 2 * c->rs, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return RBN_OK;
}

}  // namespace host
}  // namespace rbn

using namespace rbn::host;

extern "C" {

const char *rbn_version(void) { return "librbn 0.1 (gfx950)"; }
int rbn_api_version(void) { return RBN_API_VERSION; }

const char *rbn_last_error(const rbn_ctx *ctx) {
  if (ctx) return ctx->err.c_str();
  return rbn::thread_error();
}

int rbn_device_count(int *n) {
  if (!n) return rbn::thread_fail(RBN_EINVAL, "null pointer");
  int k = 0;
  if (hipGetDeviceCount(&k) != hipSuccess) k = 0;
  *n = k;
  return RBN_OK;
}

void rbn_destroy(rbn_ctx *c) {
  if (!c) return;
  (void)hipSetDevice(c->dev);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->pstream) (void)hipStreamSynchronize(c->pstream);
  rbn::adj_free(c);
  void *ptrs[] = {c->fA, c->fB, c->gA, c->gB, c->F, c->F2, c->rho, c->T, c->nud, c->nud2};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  for (int k = 0; k < rbn_ctx::kVel; k++)
    if (c->uv[k]) (void)hipFree(c->uv[k]);
  if (c->ev_lbm) (void)hipEventDestroy(c->ev_lbm);
  for (hipEvent_t e : c->ev_part)
    if (e) (void)hipEventDestroy(e);
  if (c->pstream) (void)hipStreamDestroy(c->pstream);
  if (c->copy_stream) (void)hipStreamSynchronize(c->copy_stream);
  for (auto &s : c->obs_slot) {
    if (s.h) (void)hipHostFree(s.h);
    if (s.copied) (void)hipEventDestroy(s.copied);
  }
  dfree_t(c->d_part);
  free_record_slots(c);
  for (auto &s : c->rec) {
    if (s.written) (void)hipEventDestroy(s.written);
    if (s.copied) (void)hipEventDestroy(s.copied);
  }
  dfree_t(c->d_err);
  dfree_t(c->d_obs_i);
  dfree_t(c->d_obs_d);
  free_obs_table(c);
  dfree_t(c->d_energy);
  dfree_t(c->dep.cnt);
  dfree_t(c->dep.fill);
  dfree_t(c->dep.off);
  dfree_t(c->dep.list);
  dfree_t(c->dep.blocksum);
  dfree_t(c->dep.Tbar);
  dfree_t(c->dep.wq);
  dfree_t(c->dep.overflow);
  dfree_t(c->dep.active);
  dfree_t(c->dep.nactive);
  dfree_t(c->dep.S);
  dfree_t(c->dep.Wi);
  dfree_t(c->dep.tw);
  dfree_t(c->dep.tcnt);
  dfree_t(c->dep.bin);
  dfree_t(c->dep.ocnt);
  free_dep2(c);
  for (hipEvent_t ev : {c->ev_dep[0], c->ev_dep[1], c->ev_lat[0], c->ev_lat[1]})
    if (ev) (void)hipEventDestroy(ev);
  for (auto &p : c->pending) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  for (hipEvent_t e : c->event_pool) (void)hipEventDestroy(e);
  if (c->copy_stream) (void)hipStreamDestroy(c->copy_stream);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int rbn_create(const rbn_params *p, int device, rbn_ctx **out) {
  if (!p || !out) return rbn::thread_fail(RBN_EINVAL, "rbn_create: null argument");
  *out = nullptr;
  int rc = rbn_params_check(p);
  if (rc) return rc;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
    return rbn::thread_fail(RBN_EDEVICE, "rbn_create: no GPU visible (librbn has no CPU fallback)");
  if (device < 0 || device >= ndev) return rbn::thread_fail(RBN_EINVAL, "rbn_create: device index out of range");
  rbn_ctx *c = new rbn_ctx();
  c->P = *p;
  c->dev = device;
  c->prec = p->precision;
  c->rs = c->prec == RBN_FP64 ? 8 : 4;
  if (hipSetDevice(device) != hipSuccess) {
    rbn_destroy(c);
    return rbn::thread_fail(RBN_EDEVICE, "hipSetDevice failed");
  }
  // All streams keep the default priority.
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&c->pstream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_lbm, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_part[0], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_part[1], hipEventDisableTiming) != hipSuccess) {
    rbn_destroy(c);
    return rbn::thread_fail(RBN_EDEVICE, "hipStreamCreate failed");
  }
  for (auto &s : c->rec) {
    if (hipEventCreateWithFlags(&s.written, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&s.copied, hipEventDisableTiming) != hipSuccess) {
      rbn_destroy(c);
      return rbn::thread_fail(RBN_EDEVICE, "hipEventCreate failed");
    }
  }
  Geom &g = c->geom;
  g.nx = p->nx;
  g.ny = p->ny;
  g.pitch = (p->nx + 63) / 64 * 64;
  g.plane = (size_t)g.pitch * (size_t)g.ny;
  // Lattices whose working set (two population plane sets, the force and velocity planes:
  // about 52 reals per node) fits well inside the 256 MB MALL keep the step's outputs in the
  // caches for the next step: cached stores and a band of rows per XCD (measured at 1024 x
  // 512 FP32 nudged: the lattice step 0.018 ms against 0.021 ms).  Larger ones stream past
  // the caches (non-temporal stores, hardware block order: the fastest at 4096 x 2048).
  if ((double)g.plane * 52.0 * (double)c->rs <= 160e6) {
    c->tune.nt = 0;
    c->tune.xcd = 1;
  }
  // FP32 truth runs of 2^21 nodes and more advance three steps per launch by row sweeps
  // (measured with particles: 2048 x 1024 0.0465 against 0.0489 ms per step, 4096 x 2048
  // 0.156 against 0.175, 8192 x 4096 0.663 against 0.700); smaller lattices give the sweep
  // too few workgroups (1024 x 512: 0.0212 against 0.0138).  FP64 keeps two: its sweep
  // (128-thread strips, 251 VGPRs, 1.5 waves per SIMD) carries three particle updates at a
  // cost that cancels its gain (with particles 4096 x 2048 0.3024 against 0.3019 ms per step,
  // 2048 x 1024 0.0791 against 0.0832, 8192 x 4096 1.340 against 1.314; without particles
  // 4096 x 2048 0.2589 against 0.2789).
  c->tune.steps3 = c->prec == RBN_FP32 && p->variant == RBN_VARIANT_TRAJECTORIES &&
                   (long long)p->nx * p->ny >= (1LL << 21);
  const size_t fb = g.plane * c->rs;
  struct A {
    void **p;
    size_t b;
  };
  std::vector<A> allocs = {{&c->fA, 9 * fb}, {&c->fB, 9 * fb}, {&c->gA, 9 * fb}, {&c->gB, 9 * fb},
                           {&c->F, fb},      {&c->F2, fb},      {&c->rho, fb},     {&c->T, fb}};
  for (int k = 0; k < rbn_ctx::kVel; k++) allocs.push_back({&c->uv[k], 2 * fb});
  for (auto &a : allocs) {
    if ((rc = dalloc(c, a.p, a.b))) {
      rbn::thread_fail(rc, c->err.c_str());
      rbn_destroy(c);
      return rc;
    }
    (void)hipMemsetAsync(*a.p, 0, a.b, c->stream);
  }
  void *tmp = nullptr;
  // d_err[0]: a particle left the lattice; d_err[1]: a step met non-finite moments
  if ((rc = dalloc(c, &tmp, 2 * sizeof(int)))) goto fail_out;
  c->d_err = (int *)tmp;
  (void)hipMemsetAsync(c->d_err, 0, 2 * sizeof(int), c->stream);
  if ((rc = dalloc(c, &tmp, (2 * 1024 + 2) * sizeof(double)))) goto fail_out;
  c->d_energy = (double *)tmp;
  if (p->variant == RBN_VARIANT_NUDGING) {
    if ((rc = dalloc(c, &c->nud, fb)) || (rc = dalloc(c, &c->nud2, fb))) goto fail_out;
    (void)hipMemsetAsync(c->nud, 0, fb, c->stream);
    (void)hipMemsetAsync(c->nud2, 0, fb, c->stream);
    const size_t ntiles = (g.plane + rbn::kDepScanTile - 1) / rbn::kDepScanTile;
    const bool bil = p->deposit == RBN_DEPOSIT_BILINEAR;
    struct B {
      void **p;
      size_t b;
    } dallocs[] = {{(void **)&c->dep.cnt, g.plane * sizeof(int)},
                   {(void **)&c->dep.fill, g.plane * sizeof(int)},
                   {(void **)&c->dep.off, (g.plane + 1) * sizeof(int)},
                   {(void **)&c->dep.blocksum, ntiles * sizeof(int)},
                   {(void **)&c->dep.Tbar, g.plane * sizeof(double)},
                   {(void **)&c->dep.wq, g.plane * sizeof(double)},
                   {(void **)&c->dep.overflow, sizeof(int)},
                   {(void **)&c->dep.nactive, sizeof(int)},
                   {(void **)&c->dep.S, bil ? g.plane * sizeof(long long) : 0},
                   {(void **)&c->dep.Wi, bil ? g.plane * sizeof(long long) : 0},
                   {&c->dep.tw, g.plane * 2 * c->rs},
                   {(void **)&c->dep.tcnt, bil ? (size_t)rbn::dep_tiles(g) * rbn::kTcntStride * sizeof(int) : 0},
                   {(void **)&c->dep.ocnt, bil ? g.plane * sizeof(int) : 0}};
    for (auto &a : dallocs) {
      if (a.b == 0) continue;
      if ((rc = dalloc(c, a.p, a.b))) goto fail_out;
      (void)hipMemsetAsync(*a.p, 0, a.b, c->stream);
    }
    c->dep.mode = p->deposit;
    c->dep.tw64 = c->prec == RBN_FP64;
  }
  c->consts.resize(std::max(sizeof(rbn::StepConst<double>), sizeof(rbn::StepConst<float>)));
  if (c->prec == RBN_FP64)
    fill_consts<double>(*p, *reinterpret_cast<rbn::StepConst<double> *>(c->consts.data()));
  else
    fill_consts<float>(*p, *reinterpret_cast<rbn::StepConst<float> *>(c->consts.data()));
  if (hipStreamSynchronize(c->stream) != hipSuccess) {
    rc = fail(c, RBN_EDEVICE, "device initialisation failed");
    goto fail_out;
  }
  c->istep = p->nstart;
  *out = c;
  return RBN_OK;
fail_out:
  rbn::thread_fail(rc, c->err.c_str());
  rbn_destroy(c);
  return rc;
}

int rbn_get_params(const rbn_ctx *ctx, rbn_params *out) {
  if (!ctx || !out) return rbn::thread_fail(RBN_EINVAL, "null argument");
  *out = ctx->P;
  return RBN_OK;
}

int rbn_init_fields(rbn_ctx *c, const double *rho, const double *T, const double *u, const double *v) {
  int rc = check_ctx(c);
  if (rc) return rc;
  const Geom &g = c->geom;
  const size_t nn = (size_t)g.nx * g.ny;
  std::vector<double> hr(nn), hT(nn), hu(nn, 0.0), hv(nn, 0.0);
  for (size_t k = 0; k < nn; k++) hr[k] = rho ? rho[k] : c->P.rhom1;
  if (u) std::copy(u, u + nn, hu.begin());
  if (v) std::copy(v, v + nn, hv.begin());
  if (T) {
    std::copy(T, T + nn, hT.begin());
  } else {
    // inithydro scratch branch (Trajectories/code/temppart_dump.c:686-700) on rank 0: the
    // per-node rand() stream visits halo nodes too, in (i, j) order over (nx+2)(ny+2).
    const double randnorm = 1.0 / RAND_MAX, epstan = 10.;
    srand(0);
    for (int i = 0; i < g.nx + 2; i++)
      for (int j = 0; j < g.ny + 2; j++) {
        double yyytemp = ((double)(j - 1 + (g.ny) * 0));
        double y_c = (double)c->P.ny * (1) / 2. + (randnorm * rand() - 0.5) * 2. * epstan;
        double t = c->P.Tdown + (c->P.Tup - c->P.Tdown) * (tanh((yyytemp - y_c) / epstan) + 1.0) / 2.0;
        if (i >= 1 && i <= g.nx && j >= 1 && j <= g.ny) hT[(size_t)(i - 1) * g.ny + (j - 1)] = t;
      }
  }
  c->cur = 0;
  if ((rc = upload_field(c, c->rho, hr.data()))) return rc;
  if ((rc = upload_field(c, c->T, hT.data()))) return rc;
  if ((rc = upload_uv(c, c->uv[0], hu.data(), hv.data()))) return rc;
  hipError_t e = rbn::launch_init_eq(c->stream, c->prec, g, c->rho, c->T, c->uv[0], c->fA,
                                     c->gA, c->F, c->consts.data());
  if (e != hipSuccess) return fail(c, RBN_EDEVICE, "init launch failed: %s", hipGetErrorString(e));
  if ((rc = init_common_after_fields(c))) return rc;
  c->at_init = true;
  return RBN_OK;
}

int rbn_init_conduction(rbn_ctx *c, double noise_amp, unsigned long long seed) {
  int rc = check_ctx(c);
  if (rc) return rc;
  const Geom &g = c->geom;
  std::vector<double> T((size_t)g.nx * g.ny);
  uint64_t s = seed;
  for (int i = 0; i < g.nx; i++)
    for (int j = 0; j < g.ny; j++) {
      // Tdown + (Tup-Tdown)*(j-1)/(ny-1), the commented-out profile of T:697.
      double lin = c->P.Tdown + (c->P.Tup - c->P.Tdown) * ((double)j) / ((double)(c->P.ny - 1));
      double r = (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0);  // [0,1)
      T[(size_t)i * g.ny + j] = lin + noise_amp * (2.0 * r - 1.0);
    }
  return rbn_init_fields(c, nullptr, T.data(), nullptr, nullptr);
}

int rbn_set_populations(rbn_ctx *c, const double *f, const double *gpop, const double *rho_prev,
                        const double *T_prev, const double *u_prev, const double *v_prev) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!f || !gpop || !rho_prev || !T_prev) return fail(c, RBN_EINVAL, "rbn_set_populations: null array");
  const Geom &g = c->geom;
  const size_t nn = (size_t)g.nx * g.ny;
  const bool shifted = c->prec == RBN_FP32;
  for (int which = 0; which < 2; which++) {
    const double *src = which ? gpop : f;
    void *dst = which ? c->gA : c->fA;
    const double lvl = shifted ? (which ? c->P.T0_shift : c->P.rho0_shift) : 0.0;
    std::vector<unsigned char> buf(9 * g.plane * c->rs, 0);
    for (int k = 0; k < 9; k++)
      for (int i = 0; i < g.nx; i++)
        for (int j = 0; j < g.ny; j++) {
          size_t o = (size_t)k * g.plane + (size_t)j * g.pitch + i;
          double val = src[((size_t)i * g.ny + j) * 9 + k];
          if (shifted)
            reinterpret_cast<float *>(buf.data())[o] = (float)(val - rbn::kW[k] * lvl);
          else
            reinterpret_cast<double *>(buf.data())[o] = val;
        }
    HIPCHK(c, hipMemcpyAsync(dst, buf.data(), buf.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  // Force of the first step from the given moments (forcingconstructWW, T:752).
  std::vector<double> F(nn), zero(nn, 0.0);
  for (size_t k = 0; k < nn; k++) F[k] = -c->P.gravity * rho_prev[k] * (T_prev[k] - 0.0);
  c->cur = 0;
  if ((rc = upload_field(c, c->F, F.data()))) return rc;
  if ((rc = upload_field(c, c->rho, rho_prev))) return rc;
  if ((rc = upload_field(c, c->T, T_prev))) return rc;
  if ((rc = upload_uv(c, c->uv[0], u_prev ? u_prev : zero.data(), v_prev ? v_prev : zero.data()))) return rc;
  return init_common_after_fields(c);
}

int rbn_get_populations(rbn_ctx *c, double *f, double *gpop) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!c->initialised) return fail(c, RBN_ESTATE, "not initialised");
  const Geom &g = c->geom;
  const bool shifted = c->prec == RBN_FP32;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int which = 0; which < 2; which++) {
    double *dst = which ? gpop : f;
    if (!dst) continue;
    const void *src = which ? c->gA : c->fA;
    const double lvl = shifted ? (which ? c->P.T0_shift : c->P.rho0_shift) : 0.0;
    std::vector<unsigned char> buf(9 * g.plane * c->rs);
    HIPCHK(c, hipMemcpy(buf.data(), src, buf.size(), hipMemcpyDeviceToHost));
    for (int k = 0; k < 9; k++)
      for (int i = 0; i < g.nx; i++)
        for (int j = 0; j < g.ny; j++) {
          size_t o = (size_t)k * g.plane + (size_t)j * g.pitch + i;
          double val = shifted ? (double)reinterpret_cast<const float *>(buf.data())[o] + rbn::kW[k] * lvl
                               : reinterpret_cast<const double *>(buf.data())[o];
          dst[((size_t)i * g.ny + j) * 9 + k] = val;
        }
  }
  return RBN_OK;
}

int rbn_get_fields(rbn_ctx *c, double *rho, double *T, double *u, double *v) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!c->initialised) return fail(c, RBN_ESTATE, "not initialised");
  if ((rho || T) && !c->moments_valid) return fail(c, RBN_ESTATE, "rho/T of the latest step not available");
  if (rho && (rc = download_field(c, c->rho, rho))) return rc;
  if (T && (rc = download_field(c, c->T, T))) return rc;
  if (u && (rc = download_field(c, c->uv[c->cur], u, 0.0, 2))) return rc;
  if (v && (rc = download_field(c, (const char *)c->uv[c->cur] + c->rs, v, 0.0, 2))) return rc;
  return RBN_OK;
}

int rbn_energy(rbn_ctx *c, double *mean_energy, double *mean_T) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!c->initialised || !c->moments_valid) return fail(c, RBN_ESTATE, "moments not available");
  double r[2];
  if ((rc = run_energy(c, r))) return rc;
  if (mean_energy) *mean_energy = r[0];
  if (mean_T) *mean_T = r[1];
  return RBN_OK;
}

int rbn_profiles(rbn_ctx *c, double *out) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!out) return fail(c, RBN_EINVAL, "null output");
  if (!c->initialised || !c->moments_valid) return fail(c, RBN_ESTATE, "moments of the latest step not available");
  const size_t bytes = (size_t)c->geom.ny * 9 * sizeof(double);
  void *buf = nullptr;
  if ((rc = dalloc(c, &buf, bytes))) return rc;
  rc = RBN_OK;
  if (rbn::launch_profiles(c->stream, c->prec, c->geom, c->rho, c->T, c->uv[c->cur], (double *)buf) != hipSuccess ||
      hipMemcpyAsync(out, buf, bytes, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    rc = fail(c, RBN_EDEVICE, "rbn_profiles: device operation failed");
  (void)hipFree(buf);
  return rc;
}

int rbn_sample(rbn_ctx *c, int field, int n, const double *x, const double *y, double *out) {
  int rc = check_sample_args(c, "rbn_sample", field, n, x, y, out);
  if (rc || n == 0) return rc;
  void *buf = nullptr;
  if ((rc = dalloc(c, &buf, 3 * (size_t)n * sizeof(double)))) return rc;
  rc = sample_staged(c, field, n, x, y, out, (double *)buf);
  (void)hipFree(buf);
  return rc;
}

int rbn_init_particles(rbn_ctx *c, int n, const int *id, const double *x, const double *y,
                       const double *tau_p, const double *beta_p, const double *up, const double *vp) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (n < 0 || (n > 0 && (!id || !x || !y))) return fail(c, RBN_EINVAL, "rbn_init_particles: bad arguments");
  if (!c->initialised) return fail(c, RBN_ESTATE, "initialise the fields before the particles");
  if ((up == nullptr) != (vp == nullptr)) return fail(c, RBN_EINVAL, "give both up and vp or neither");
  if (n > (1 << 28)) return fail(c, RBN_ERANGE, "too many particles");
  HIPCHK(c, hipStreamSynchronize(c->copy_stream));
  dfree_t(c->d_part);
  free_record_slots(c);
  c->np = n;
  c->pid.clear();
  if (n == 0) return RBN_OK;
  std::vector<int> order((size_t)n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return id[a] < id[b]; });
  const size_t N = (size_t)n;
  std::vector<double> h(PNARR * N, 0.0);
  c->pid.resize(N);
  for (size_t q = 0; q < N; q++) {
    const int p = order[q];
    c->pid[q] = id[p];
    h[PX * N + q] = x[p];
    h[PY * N + q] = y[p];
    h[PXO * N + q] = x[p];
    h[PYO * N + q] = y[p];
    double tp = tau_p ? tau_p[p] : (c->P.heavyprtcl && !c->P.tracer ? c->P.taulag : 0.0);
    h[PINVTAU * N + q] = tp > 0.0 ? (1 / tp) : 0.0;
    h[PBETA * N + q] = beta_p ? beta_p[p] : c->P.beta;
    if (up) {
      h[PU * N + q] = up[p];
      h[PV * N + q] = vp[p];
      h[PUO * N + q] = up[p];
      h[PVO * N + q] = vp[p];
    }
  }
  void *tmp = nullptr;
  if ((rc = dalloc(c, &tmp, PNARR * N * sizeof(double)))) return rc;
  c->d_part = (double *)tmp;
  for (auto &s : c->rec) {
    if ((rc = dalloc(c, &tmp, 17 * N * sizeof(double)))) return rc;
    s.d = (double *)tmp;
    HIPCHK(c, hipMemsetAsync(s.d, 0, 17 * N * sizeof(double), c->stream));
    if (hipHostMalloc((void **)&s.h, 17 * N * sizeof(double), hipHostMallocDefault) != hipSuccess) {
      s.h = nullptr;
      return fail(c, RBN_ENOMEM, "pinned record buffer allocation failed");
    }
  }
  HIPCHK(c, hipMemcpyAsync(c->d_part, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  if (!up) {
    hipError_t e = rbn::launch_particle_init(c->stream, c->prec, c->geom, c->uv[c->cur],
                                             part_ptrs(c, nullptr), c->P.heavy_init_factor);
    if (e != hipSuccess) return fail(c, RBN_EDEVICE, "particle init launch failed: %s", hipGetErrorString(e));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return check_err_flag(c);
}

int rbn_num_particles(const rbn_ctx *c, int *n) {
  if (!c || !n) return rbn::thread_fail(RBN_EINVAL, "null argument");
  *n = c->np;
  return RBN_OK;
}

int rbn_get_particles(rbn_ctx *c, rbn_particles *out) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!out) return fail(c, RBN_EINVAL, "null output");
  if (out->n < c->np) return fail(c, RBN_EINVAL, "output arrays hold %d particles, need %d", out->n, c->np);
  out->n = c->np;
  if (c->np == 0) return RBN_OK;
  const size_t N = (size_t)c->np;
  std::vector<double> h(8 * N);
  HIPCHK(c, hipMemcpyAsync(h.data(), c->d_part, 8 * N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  double *dst[8] = {out->x, out->y, out->u, out->v, out->xo, out->yo, out->uo, out->vo};
  for (int q = 0; q < 8; q++)
    if (dst[q]) std::copy(h.begin() + q * N, h.begin() + (q + 1) * N, dst[q]);
  if (out->id) std::copy(c->pid.begin(), c->pid.end(), out->id);
  return RBN_OK;
}

int rbn_set_record_callback(rbn_ctx *c, rbn_record_fn fn, void *user) {
  if (!c) return rbn::thread_fail(RBN_EINVAL, "null context");
  c->rec_fn = fn;
  c->rec_user = user;
  return RBN_OK;
}

int rbn_set_diag_callback(rbn_ctx *c, rbn_diag_fn fn, void *user) {
  if (!c) return rbn::thread_fail(RBN_EINVAL, "null context");
  c->diag_fn = fn;
  c->diag_user = user;
  return RBN_OK;
}

int rbn_init_observers(rbn_ctx *c, int n, const int *id, const double *x, const double *y, const double *T) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (c->P.variant != RBN_VARIANT_NUDGING) return fail(c, RBN_ESTATE, "observers belong to the nudged variant");
  if (n < 0 || (n > 0 && (!id || !x || !y || !T))) return fail(c, RBN_EINVAL, "rbn_init_observers: bad arguments");
  if (n > (1 << 24)) return fail(c, RBN_ERANGE, "too many observers");
  dfree_t(c->d_obs_i);
  dfree_t(c->d_obs_d);
  free_obs_table(c);
  dfree_t(c->dep.list);
  dfree_t(c->dep.active);
  dfree_t(c->dep.bin);
  free_dep2(c);  // sized by the observers: allocated again on first use
  c->nobs = n;
  c->oid.assign(id, id + n);
  c->prev_rid.clear();
  const size_t N = (size_t)(n ? n : 1);
  void *tmp = nullptr;
  if ((rc = dalloc(c, &tmp, ONI * N * sizeof(int)))) return rc;
  c->d_obs_i = (int *)tmp;
  if ((rc = dalloc(c, &tmp, OND * N * sizeof(double)))) return rc;
  c->d_obs_d = (double *)tmp;
  // Each observer reaches at most 49 nodes of its box plus 28 of one wrap window.
  c->dep.list_cap = 77LL * (long long)N;
  if ((rc = dalloc(c, &tmp, (size_t)c->dep.list_cap * sizeof(int)))) return rc;
  c->dep.list = (int *)tmp;
  // at most one active node per contribution; a deposit with at least one candidate
  // contribution per four nodes folds every node in row order instead (an active list would
  // serialise on its counter)
  c->dep.active_cap = std::min<long long>((long long)c->geom.plane, c->dep.list_cap);
  const long long ncand = (long long)n * (c->dep.mode == RBN_DEPOSIT_BILINEAR ? 4 : 77);
  c->dep.dense = 4 * ncand >= (long long)c->geom.plane;
  if ((rc = dalloc(c, &tmp, (size_t)c->dep.active_cap * sizeof(int)))) return rc;
  c->dep.active = (int *)tmp;
  if (c->dep.mode == RBN_DEPOSIT_BILINEAR) {
    // the bilinear deposit's tile slots, sized by the observers' mean density
    c->dep.bin_cap = rbn::dep_bin_cap(c->geom, n);
    const size_t bb = (size_t)rbn::dep_tiles(c->geom) * (size_t)c->dep.bin_cap * 2 * sizeof(long long);
    if ((rc = dalloc(c, &tmp, bb))) return rc;
    c->dep.bin = (long long *)tmp;
  }
  HIPCHK(c, hipMemsetAsync(c->d_obs_i, 0, ONI * N * sizeof(int), c->stream));
  std::vector<double> h(OND * N, 0.0);
  for (int i = 0; i < n; i++) {
    h[OXINT * N + i] = h[ORX * N + i] = x[i];
    h[OYINT * N + i] = h[ORY * N + i] = y[i];
    h[OTINT * N + i] = h[ORT * N + i] = T[i];
  }
  HIPCHK(c, hipMemcpyAsync(c->d_obs_d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  // set_nudge_vol zeroes nud_term and nudge_vol (N:2051-2056).
  if (c->nud) HIPCHK(c, hipMemsetAsync(c->nud, 0, c->geom.plane * c->rs, c->stream));
  HIPCHK(c, hipMemsetAsync(c->dep.cnt, 0, c->geom.plane * sizeof(int), c->stream));
  HIPCHK(c, hipMemsetAsync(c->dep.tw, 0, c->geom.plane * 2 * c->rs, c->stream));  // empty deposit
  HIPCHK(c, hipMemsetAsync(c->dep.overflow, 0, sizeof(int), c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return RBN_OK;
}

int rbn_set_obs_provider(rbn_ctx *c, rbn_obs_fn fn, void *user) {
  if (!c) return rbn::thread_fail(RBN_EINVAL, "null context");
  free_obs_table(c);
  c->obs_fn = fn;
  c->obs_user = user;
  return RBN_OK;
}

namespace {

// The provider over an observation table: record `step` by binary search of the ascending
// steps, copied out as stored.
int table_provider(void *user, int step, int n, int *id, double *x, double *y, double *T) {
  const rbn_ctx::ObsTable &t = *(const rbn_ctx::ObsTable *)user;
  if (n != t.n) return 1;
  const int *end = t.steps + t.nrec, *it = std::lower_bound(t.steps, end, step);
  if (it == end || *it != step) return 1;
  const size_t r = (size_t)(it - t.steps) * (size_t)n;
  std::copy(t.id + r, t.id + r + n, id);
  std::copy(t.x + r, t.x + r + n, x);
  std::copy(t.y + r, t.y + r + n, y);
  std::copy(t.T + r, t.T + r + n, T);
  return 0;
}

}  // namespace

int rbn_set_obs_table(rbn_ctx *c, int nrec, int n, const int *steps, const int *id, const double *x,
                      const double *y, const double *T) {
  if (!c) return rbn::thread_fail(RBN_EINVAL, "null context");
  if (nrec < 0 || n < 0 || (nrec > 0 && (!steps || (n > 0 && (!id || !x || !y || !T)))))
    return fail(c, RBN_EINVAL, "rbn_set_obs_table: bad arguments");
  for (int r = 1; r < nrec; r++)
    if (steps[r] <= steps[r - 1]) return fail(c, RBN_EINVAL, "rbn_set_obs_table: record steps must ascend");
  free_obs_table(c);  // copied to the device again at the next step
  c->obs_table = {nrec, n, steps, id, x, y, T};
  c->obs_fn = table_provider;
  c->obs_user = &c->obs_table;
  return RBN_OK;
}

int rbn_get_nudging_weights(rbn_ctx *c, double *weight) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (c->P.variant != RBN_VARIANT_NUDGING) return fail(c, RBN_ESTATE, "not a nudged run");
  if (!weight) return fail(c, RBN_EINVAL, "null output");
  const Geom &g = c->geom;
  std::vector<double> h(g.plane, 0.0);
  std::vector<int> n(g.plane, 0);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (c->dep.mode == RBN_DEPOSIT_BILINEAR) {
    HIPCHK(c, hipMemcpy(h.data(), c->dep.wq, g.plane * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(n.data(), c->dep.cnt, g.plane * sizeof(int), hipMemcpyDeviceToHost));
  }
  // weights are written only where the deposit has contributions
  for (int i = 0; i < g.nx; i++)
    for (int j = 0; j < g.ny; j++) {
      const size_t o = (size_t)j * g.pitch + i;
      weight[(size_t)i * g.ny + j] = n[o] > 0 ? h[o] : 0.0;
    }
  return RBN_OK;
}

int rbn_get_nudging(rbn_ctx *c, int *nudge_vol, double *tempreadint, double *nud_term) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (c->P.variant != RBN_VARIANT_NUDGING) return fail(c, RBN_ESTATE, "not a nudged run");
  const Geom &g = c->geom;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  std::vector<int> cnt(g.plane);
  HIPCHK(c, hipMemcpy(cnt.data(), c->dep.cnt, g.plane * sizeof(int), hipMemcpyDeviceToHost));
  if (nudge_vol)
    for (int i = 0; i < g.nx; i++)
      for (int j = 0; j < g.ny; j++) nudge_vol[(size_t)i * g.ny + j] = cnt[(size_t)j * g.pitch + i];
  if (tempreadint) {
    // the means are written only where the deposit has contributions (tempreadint is zero
    // elsewhere in the reference)
    std::vector<double> h(g.plane);
    HIPCHK(c, hipMemcpy(h.data(), c->dep.Tbar, g.plane * sizeof(double), hipMemcpyDeviceToHost));
    for (int i = 0; i < g.nx; i++)
      for (int j = 0; j < g.ny; j++) {
        const size_t o = (size_t)j * g.pitch + i;
        tempreadint[(size_t)i * g.ny + j] = cnt[o] > 0 ? h[o] : 0.0;
      }
  }
  if (nud_term && (rc = download_field(c, c->nud, nud_term))) return rc;
  return RBN_OK;
}

namespace {

// ---------------------------------------------------------------- pending particle updates
// With particle_fuse the particle update of a step waits for the next lattice launch and
// runs inside it; after a two- or three-step launch the updates of all its steps wait.  None
// is pending when rbn_step returns.

void add_pending(rbn_ctx *c, int istep, int cur, int prv, bool record, bool full, rbn_ctx::RecSlot *slot) {
  rbn_ctx::PendingParticles &j = c->pend[c->npend++];
  j.istep = istep;
  j.cur = cur;
  j.prv = prv;
  j.record = record;
  j.full = full;
  j.slot = slot;
}

bool pending_reads_T(const rbn_ctx *c) {
  for (int u = 0; u < c->npend; u++)
    if (c->pend[u].record) return true;
  return false;
}

rbn::PartJob pending_job(rbn_ctx *c) {
  rbn::PartJob job;
  job.n = c->npend;
  for (int u = 0; u < c->npend; u++) {
    const rbn_ctx::PendingParticles &j = c->pend[u];
    job.uv[u] = c->uv[j.cur];
    job.uvp[u] = c->uv[j.prv];
    job.rec[u] = j.slot ? j.slot->d : nullptr;
    job.record[u] = j.record;
    job.full[u] = j.full;
  }
  job.T = c->T;
  job.pp = part_ptrs(c, nullptr);
  return job;
}

// The pending particle updates have been launched (on the lattice stream): queue their
// records, in step order.
int pending_launched(rbn_ctx *c) {
  const int n = c->npend;
  c->npend = 0;
  for (int u = 0; u < n; u++) {
    const rbn_ctx::PendingParticles &j = c->pend[u];
    if (j.slot && c->rec_fn) {
      int rc = queue_record_copy(c, *j.slot, j.istep, c->stream);
      if (rc) return rc;
    }
  }
  return RBN_OK;
}

// Run the pending particle updates as launches of their own, in step order.
int flush_particles(rbn_ctx *c) {
  for (int u = 0; u < c->npend; u++) {
    const rbn_ctx::PendingParticles &j = c->pend[u];
    const rbn::PartPtrs pp = part_ptrs(c, j.slot ? j.slot->d : nullptr);
    int rc = timed(c, KT_PARTICLES, c->stream, [&] {
      return rbn::launch_particles(c->stream, c->prec, c->geom, c->uv[j.cur], c->uv[j.prv], c->T, pp, j.record,
                                   j.full);
    });
    if (rc) {
      c->npend = 0;
      return rc;
    }
  }
  return pending_launched(c);
}

int report_energy(rbn_ctx *c, int istep) {
  double r[2];
  int rc = run_energy(c, r);
  if (rc) return rc;
  if (c->diag_fn(c->diag_user, istep, r[0], r[1]) != 0)
    return fail(c, RBN_ECALLBACK, "diagnostics callback failed at step %d", istep);
  return RBN_OK;
}

// Observer update and deposit of step istep into d on stream s (with the record upload of an
// interval's first step).
// tail, defer: as launch_obs_deposit's.
int nudged_deposit(rbn_ctx *c, int istep, hipStream_t s, const rbn::DepPtrs &d, const rbn::DepPtrs *tail = nullptr,
                   bool defer = false) {
  const rbn_params &P = c->P;
  const int right = P.fix_q3 ? P.nx - 3 : P.nx - 2;
  const int r = istep % P.ntempread;
  int rc;
  if (r == 1 && (rc = load_observation_dev(c, istep + P.ntempread - 1, s))) return rc;
  const int branch = r == 1 ? 1 : (r == 0 ? 0 : 2);
  const rbn::ObsPtrs o = obs_ptrs(c);
  return timed(c, KT_DEPOSIT, s,
               [&] { return rbn::launch_obs_deposit(s, branch, P.ntempread, c->geom, right, o, d, tail, defer); });
}

int step_loop(rbn_ctx *c, int nsteps) {
  int rc = RBN_OK;
  const rbn_params &P = c->P;
  const bool nudged = P.variant == RBN_VARIANT_NUDGING;
  const bool parts = !nudged && P.particle && c->np > 0;
  const int nd = P.ndiag > 0 ? P.ndiag : (nudged ? 5000 : P.nout);
  const int right = P.fix_q3 ? P.nx - 3 : P.nx - 2;
  // Particle schedules.  particle_fuse (default): K2 of a step rides in the grid of the next
  // lattice launch, unless it is a record step's (it reads T) and that launch rewrites the
  // moments.  Otherwise K2 runs on the lattice stream after K1, ordered without events (a
  // cross-stream event wait costs about 20 us per hand-over), or with particle_overlap on
  // its own stream beside the next K1 (both kernels are bandwidth bound, so concurrency only
  // moves the contention); with overlap, K1 of step s waits for the particle update of step
  // s-2 (the latest event of the same parity).
  const bool overlap = c->tune.overlap != 0;
  const bool fuse = parts && !overlap && c->tune.fuse != 0;
  // lbm_steps 2: two truth-run steps per lattice launch, the first of which writes no
  // moments; with particles only when their updates ride in lattice launches.
  const bool pairs = c->tune.steps2 && (!parts || fuse);
  // lbm_steps 3: three truth-run steps per launch (row sweeps) wherever at most one of them
  // writes moments and neither of the first two is a diagnostics step; a launch of two or one
  // step elsewhere.
  const bool triples = pairs && c->tune.steps3 && !nudged;
  hipStream_t ks = overlap ? c->pstream : c->stream;
  // dep_overlap: the observer update and deposit of step s+1 run on pstream (unused by nudged
  // runs, which have no particles) beside the lattice step s.  Off by default: measured at
  // 1024x512 FP32, 1e4 observers 54 against 50 us per step, 1e5 observers 99 against 81 us,
  // and at 512x256 51 against 39 us (the hand-over between the streams costs more than the
  // 16 us lattice step it hides).
  const bool dovl = nudged && c->tune.dep_overlap && nsteps > 1;
  if (dovl && (rc = ensure_dep2(c))) return rc;
  // Nudged pairs (lbm_steps 2, without dep_overlap): the deposit of the pair's first step goes
  // to dep2 and that of its second step to dep, the set the getters and a following one-step
  // launch read.
  const bool npairs = nudged && pairs && !dovl && nsteps > 1;
  if (npairs && (rc = ensure_dep2(c))) return rc;
  if (nsteps > 0) c->at_init = false;
  if (nudged && nsteps > 0 && c->tune.obs_device && c->obs_user == &c->obs_table && !c->tab_tried) {
    bool built = false;
    if ((rc = build_obs_table(c, &built))) return rc;
  }
  // nudged pairs, bilinear deposit: the step whose observer update (and binning into dep2) the
  // previous two-step launch carried, when the next launch is a two-step one too
  int head_done = -1;
  for (int n = 0; n < nsteps; n++) {
    const int istep = ++c->istep;
    const bool record = parts && (istep % P.noutconfig) == 0;
    const bool diag = nd > 0 && (istep % nd) == 0;
    const bool moments = record || diag || n == nsteps - 1;
    // a three-step launch: at most one of its steps writes moments (the sweep writes those of
    // that step's level), so a record step may be any of the three; a diagnostics step (the
    // energy is taken from the latest velocity) or the call's last step only the third
    bool tri = false;
    if (triples && n + 2 < nsteps) {
      const int i2 = istep + 1, i3 = istep + 2;
      const bool d2 = nd > 0 && (i2 % nd) == 0;
      const bool m2 = (parts && (i2 % P.noutconfig) == 0) || d2;
      const bool m3 = (parts && (i3 % P.noutconfig) == 0) || (nd > 0 && (i3 % nd) == 0) || n + 2 == nsteps - 1;
      tri = !diag && !d2 && (int)moments + (int)m2 + (int)m3 <= 1;
    }
    const bool pair = tri || (pairs && (!nudged || npairs) && !moments && n + 1 < nsteps);
    int istep2 = 0;
    bool record2 = false, diag2 = false, moments2 = false;
    if (pair) {
      istep2 = ++c->istep;
      n++;
      record2 = parts && (istep2 % P.noutconfig) == 0;
      diag2 = nd > 0 && (istep2 % nd) == 0;
      moments2 = record2 || diag2 || n == nsteps - 1;
    }
    const bool triple = tri;
    int istep3 = 0;
    bool record3 = false, diag3 = false, moments3 = false;
    if (triple) {
      istep3 = ++c->istep;
      n++;
      record3 = parts && (istep3 % P.noutconfig) == 0;
      diag3 = nd > 0 && (istep3 % nd) == 0;
      moments3 = record3 || diag3 || n == nsteps - 1;
    }
    // the level of a three-step launch whose moments it writes (-1: none)
    const int mlev = !triple ? -1 : (moments ? 0 : (moments2 ? 1 : (moments3 ? 2 : -1)));
    const bool writes_moments = triple ? mlev >= 0 : (pair ? moments2 : moments);
    // bilinear deposit: the second step's deposit summed inside the two-step launch from its
    // bins, except in the call's last launch (its deposit fills the arrays the getters read)
    const bool in_tile = nudged && pair && c->dep.mode == RBN_DEPOSIT_BILINEAR && c->tune.dep_in_tile &&
                         rbn::lbm2_tile_rows(c->prec, c->tune.tile_h) == rbn::kDepTileH && n < nsteps - 1;
    // bilinear deposit: the two-step launch also carries the next two-step launch's first
    // observer update (into dep2, whose bins the tile pass of this pair's first step has
    // consumed; the launch reads dep2's packed deposit and dep's bins only)
    bool carry_head = false;
    int hstep = 0;
    if (nudged && pair && c->dep.mode == RBN_DEPOSIT_BILINEAR) {
      hstep = istep2 + 1;
      const bool hmoments = (nd > 0 && (hstep % nd) == 0) || n + 1 == nsteps - 1;
      carry_head = !hmoments && n + 2 < nsteps;
    }
    if (nudged && pair) {
      // the first step's bilinear tile pass rides in the second step's first launch
      const bool bil = c->dep.mode == RBN_DEPOSIT_BILINEAR;
      if (head_done != istep && (rc = nudged_deposit(c, istep, c->stream, c->dep2, nullptr, bil))) return rc;
      if ((rc = nudged_deposit(c, istep2, c->stream, c->dep, bil ? &c->dep2 : nullptr, in_tile))) return rc;
      if (carry_head && hstep % P.ntempread == 1 &&
          (rc = load_observation_dev(c, hstep + P.ntempread - 1, c->stream)))
        return rc;
    } else if (nudged && !dovl) {
      if (head_done == istep) return fail(c, RBN_ESTATE, "internal: a carried observer update before a single step");
      if ((rc = nudged_deposit(c, istep, c->stream, c->dep))) return rc;
    } else if (nudged) {
      // step n's deposit is in set n & 1 (dep, dep2), built on pstream: the first here, the
      // others beside the previous lattice launch
      if (n == 0 && (rc = nudged_deposit(c, istep, c->pstream, c->dep))) return rc;
      if (n == 0) HIPCHK(c, hipEventRecord(c->ev_dep[0], c->pstream));
      HIPCHK(c, hipStreamWaitEvent(c->stream, c->ev_dep[n & 1], 0));
    }
    const rbn::DepPtrs &dset = (dovl && (n & 1)) || (nudged && pair) ? c->dep2 : c->dep;
    // the observers as the carried observer update reads them: after its record's load
    const rbn::ObsPtrs o3 = nudged ? obs_ptrs(c) : rbn::ObsPtrs{};
    // A launch writes the velocity slots after cur: K1 of step s overwrites the slot of step
    // s-8, a two-step launch also that of s-7, a three-step one those of s-7 and s-6; pending
    // updates read slots s-4 .. s-1.
    bool carry = false;
    if (c->npend > 0) {
      if (pending_reads_T(c) && writes_moments) {
        if ((rc = flush_particles(c))) return rc;
      } else {
        carry = true;
      }
    }
    if (parts && overlap) {
      HIPCHK(c, hipStreamWaitEvent(c->stream, c->ev_part[istep % 2], 0));
      if (moments && c->part_record_pending)
        HIPCHK(c, hipStreamWaitEvent(c->stream, c->ev_part[(istep + 1) % 2], 0));
    }
    const int s1 = (c->cur + 1) % rbn_ctx::kVel, s2 = (c->cur + 2) % rbn_ctx::kVel,
              s3 = (c->cur + 3) % rbn_ctx::kVel;
    rbn::LbmPtrs lp{};
    lp.fA = c->fA;
    lp.gA = c->gA;
    lp.fB = c->fB;
    lp.gB = c->gB;
    lp.F = c->F;
    lp.uv = c->uv[s1];
    lp.rho = !pair && moments ? c->rho : nullptr;
    lp.T = !pair && moments ? c->T : nullptr;
    if (nudged) {
      lp.tw = dset.tw;
      lp.nud = c->nud;
    }
    lp.bad = c->d_err + 1;
    const int kclass = triple ? KT_LBM3 : (pair ? KT_LBM2 : (carry ? KT_LBM_FUSED : KT_LBM));
    if ((rc = timed(c, kclass, c->stream, [&] {
           rbn::LbmTuning t = c->tune;
           if (triple) {
             const rbn::PartJob job = carry ? pending_job(c) : rbn::PartJob{};
             return rbn::launch_lbm3(c->stream, c->prec, c->geom, lp, c->F2, c->uv[s2], c->uv[s3],
                                     mlev >= 0 ? c->rho : nullptr, mlev >= 0 ? c->T : nullptr, mlev, c->consts.data(),
                                     t.seg3, t.xcd2, t.reverse && ((istep / 3) & 1), t.nt3,
                                     carry ? &job : nullptr);
           }
           if (pair) {
             const rbn::PartJob job = carry ? pending_job(c) : rbn::PartJob{};
             rbn::NudgePair np{};
             if (nudged) {
               np.tw2 = c->dep.tw;
               np.nud_out = c->nud2;
               np.bins2 = in_tile ? &c->dep : nullptr;
               if (carry_head) {
                 np.head_obs = &o3;
                 np.head_dep = &c->dep2;
                 const int r3 = hstep % P.ntempread;
                 np.head_branch = r3 == 1 ? 1 : (r3 == 0 ? 0 : 2);
                 np.head_ntempread = P.ntempread;
               }
             }
             return rbn::launch_lbm2(c->stream, c->prec, c->geom, lp, c->F2, c->uv[s2], moments2 ? c->rho : nullptr,
                                     moments2 ? c->T : nullptr, c->consts.data(), t.tile_h, t.seg, t.xcd2, t.variant,
                                     t.reverse && ((istep >> 1) & 1), t.nt, carry ? &job : nullptr,
                                     nudged ? &np : nullptr);
           }
           t.flip = t.reverse && (istep & 1);
           if (carry)
             return rbn::launch_lbm_fused(c->stream, c->prec, c->geom, lp, c->consts.data(), t, pending_job(c));
           return rbn::launch_lbm(c->stream, c->prec, P.variant, c->geom, lp, c->consts.data(), t);
         }))) {
      c->npend = 0;
      return rc;
    }
    head_done = carry_head ? hstep : -1;
    if (carry && (rc = pending_launched(c))) return rc;
    if (nudged && dovl) {
      HIPCHK(c, hipEventRecord(c->ev_lat[n & 1], c->stream));
      if (n + 1 < nsteps) {
        // the next step's deposit, into the set the lattice launch before this one read
        if (n >= 1) HIPCHK(c, hipStreamWaitEvent(c->pstream, c->ev_lat[(n + 1) & 1], 0));
        if ((rc = nudged_deposit(c, istep + 1, c->pstream, (n & 1) ? c->dep : c->dep2))) return rc;
        HIPCHK(c, hipEventRecord(c->ev_dep[(n + 1) & 1], c->pstream));
      } else if (n & 1) {
        // the last deposit is in dep2: its counts, means and weights to dep for the readers
        const size_t plane = c->geom.plane;
        HIPCHK(c, hipMemcpyAsync(c->dep.cnt, c->dep2.cnt, plane * sizeof(int), hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->dep.Tbar, c->dep2.Tbar, plane * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        if (c->dep.mode == RBN_DEPOSIT_BILINEAR)
          HIPCHK(c, hipMemcpyAsync(c->dep.wq, c->dep2.wq, plane * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
      }
    }
    std::swap(c->fA, c->fB);
    std::swap(c->gA, c->gB);
    if (pair) std::swap(c->F, c->F2);
    if (pair && nudged) std::swap(c->nud, c->nud2);
    const int prv = c->cur;
    c->cur = triple ? s3 : (pair ? s2 : s1);
    // rho and T hold the latest step's moments (a three-step launch may write an earlier step's)
    c->moments_valid = triple ? mlev == 2 : writes_moments;
    if (fuse) {
      // this launch's particle updates wait for the next launch; a record step (the last step
      // of a one- or two-step launch, any step of a three-step one) writes its record then,
      // interpolating T of that step (no later step of the launch writes T)
      auto pend = [&](int st, int cur_slot, int prv_slot, bool rec) -> int {
        rbn_ctx::RecSlot *slot = nullptr;
        if (rec && (rc = claim_record_slot(c, &slot))) return rc;
        add_pending(c, st, cur_slot, prv_slot, rec, rec && P.prtcl_writefull, slot);
        return RBN_OK;
      };
      if (pair && (rc = pend(istep, s1, prv, record))) return rc;
      if (triple && (rc = pend(istep2, s2, s1, record2))) return rc;
      const int rstep = triple ? istep3 : (pair ? istep2 : istep);
      const bool rec = triple ? record3 : (pair ? record2 : record);
      if ((rc = pend(rstep, c->cur, triple ? s2 : (pair ? s1 : prv), rec))) return rc;
    } else if (parts) {
      // K2 of step s; with overlap, on the particle stream alongside K1 of step s+1.
      if (overlap) {
        HIPCHK(c, hipEventRecord(c->ev_lbm, c->stream));
        HIPCHK(c, hipStreamWaitEvent(c->pstream, c->ev_lbm, 0));
      }
      const int full = record && P.prtcl_writefull;
      rbn_ctx::RecSlot *slot = nullptr;
      if (record && (rc = claim_record_slot(c, &slot))) return rc;
      rbn::PartPtrs pp = part_ptrs(c, slot ? slot->d : nullptr);
      if ((rc = timed(c, KT_PARTICLES, ks, [&] {
             return rbn::launch_particles(ks, c->prec, c->geom, c->uv[c->cur], c->uv[prv], c->T, pp, record,
                                          full);
           })))
        return rc;
      if (overlap) HIPCHK(c, hipEventRecord(c->ev_part[istep % 2], c->pstream));
      c->part_record_pending = overlap && record;
      if (slot && c->rec_fn && (rc = queue_record_copy(c, *slot, istep, ks))) return rc;
    }
    const bool dstep = triple ? diag3 : (pair ? diag2 : diag);
    if (dstep && c->diag_fn && (rc = report_energy(c, triple ? istep3 : (pair ? istep2 : istep)))) return rc;
  }
  return RBN_OK;
}

}  // namespace

int rbn_step(rbn_ctx *c, int nsteps) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!c->initialised) return fail(c, RBN_ESTATE, "rbn_step before initialisation");
  if (nsteps < 0) return fail(c, RBN_EINVAL, "negative step count");
  const rbn_params &P = c->P;
  const bool nudged = P.variant == RBN_VARIANT_NUDGING;
  if (nudged && c->nobs == 0 && !c->d_obs_i) return fail(c, RBN_ESTATE, "nudged run without observers");
  const bool parts = !nudged && P.particle && c->np > 0;
  rc = step_loop(c, nsteps);
  if (nudged) {
    const int rs = settle_records(c);
    if (!rc) rc = rs;
  }
  // The last step's particle update (particle_fuse) has no lattice launch to ride in.
  const int rf = flush_particles(c);
  if (rc) {
    // a failure between an observer update and its tile pass (a provider error) leaves
    // bilinear tile counters set: clear them, so that later deposits start clean
    if (nudged) {
      (void)hipStreamSynchronize(c->stream);
      (void)hipStreamSynchronize(c->pstream);
      (void)clear_dep_bins(c->dep, c->geom, c->stream);
      if (c->dep2.cnt) (void)clear_dep_bins(c->dep2, c->geom, c->stream);
      (void)hipStreamSynchronize(c->stream);
    }
    return rc;
  }
  if (rf) return rf;
  if (parts) HIPCHK(c, hipStreamSynchronize(c->pstream));
  c->part_record_pending = false;
  if (c->timing && (rc = drain_timing(c))) return rc;
  // Non-finite moments, or a particle that left the lattice, invalidate the records of this
  // call: report them before delivering the records.
  if (nsteps > 0 && (rc = check_err_flag(c))) return rc;
  return deliver_all(c);
}

int rbn_get_step(const rbn_ctx *c, int *istep) {
  if (!c || !istep) return rbn::thread_fail(RBN_EINVAL, "null argument");
  *istep = c->istep;
  return RBN_OK;
}

int rbn_synchronize(rbn_ctx *c) {
  int rc = check_ctx(c);
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return RBN_OK;
}

int rbn_set_tuning(rbn_ctx *c, const char *key, int value) {
  if (!c || !key) return rbn::thread_fail(RBN_EINVAL, "null argument");
  if (!strcmp(key, "lbm_block")) {
    if (value != 64 && value != 128 && value != 256 && value != 512)
      return fail(c, RBN_EINVAL, "lbm_block must be 64, 128, 256 or 512");
    c->tune.block = value;
  } else if (!strcmp(key, "lbm_nt")) {
    c->tune.nt = value != 0;
  } else if (!strcmp(key, "lbm_npt")) {
    if (value != 1 && value != 2 && value != 4 && value != 8)
      return fail(c, RBN_EINVAL, "lbm_npt must be 1, 2, 4 or 8");
    c->tune.npt = value;
  } else if (!strcmp(key, "lbm_ntload")) {
    c->tune.ntload = value != 0;
  } else if (!strcmp(key, "lbm_uv_keep")) {
    c->tune.uv_keep = value != 0;
  } else if (!strcmp(key, "lbm_xcd")) {
    if (value < 0 || value > 2) return fail(c, RBN_EINVAL, "lbm_xcd must be 0, 1 or 2");
    c->tune.xcd = value;
  } else if (!strcmp(key, "lbm_reverse")) {
    c->tune.reverse = value != 0;
  } else if (!strcmp(key, "lbm_steps")) {
    if (value < 1 || value > 3) return fail(c, RBN_EINVAL, "lbm_steps must be 1, 2 or 3");
    c->tune.steps2 = value >= 2;
    c->tune.steps3 = value == 3;
  } else if (!strcmp(key, "lbm3_seg")) {
    if (value != 0 && (value < 4 || value > 4096)) return fail(c, RBN_EINVAL, "lbm3_seg must be 0 or 4..4096");
    c->tune.seg3 = value;
  } else if (!strcmp(key, "lbm3_nt")) {
    c->tune.nt3 = value != 0;
  } else if (!strcmp(key, "lbm2_tile_h")) {
    if (value != 0 && value != 4 && value != 8 && value != 16)
      return fail(c, RBN_EINVAL, "lbm2_tile_h must be 0, 4, 8 or 16");
    c->tune.tile_h = value;
  } else if (!strcmp(key, "lbm2_seg")) {
    if (value != 0 && (value < 8 || value > 64)) return fail(c, RBN_EINVAL, "lbm2_seg must be 0 or 8..64");
    c->tune.seg = value;
  } else if (!strcmp(key, "lbm2_xcd")) {
    c->tune.xcd2 = value != 0;
  } else if (!strcmp(key, "lbm2_variant")) {
    if (value < 0 || value > 1) return fail(c, RBN_EINVAL, "lbm2_variant must be 0 or 1");
    c->tune.variant = value;
  } else if (!strcmp(key, "dep_overlap")) {
    c->tune.dep_overlap = value != 0;
  } else if (!strcmp(key, "dep_in_tile")) {
    c->tune.dep_in_tile = value != 0;
  } else if (!strcmp(key, "obs_device")) {
    c->tune.obs_device = value != 0;
    if (!value) free_obs_table(c);
  } else if (!strcmp(key, "particle_overlap")) {
    c->tune.overlap = value != 0;
  } else if (!strcmp(key, "particle_fuse")) {
    c->tune.fuse = value != 0;
  } else if (!strcmp(key, "adj_store_deposit")) {
    c->adj_store_deposit = value != 0;
  } else {
    return fail(c, RBN_EINVAL, "unknown tuning key %s", key);
  }
  return RBN_OK;
}

int rbn_get_tuning(const rbn_ctx *c, const char *key, int *value) {
  if (!c || !key || !value) return rbn::thread_fail(RBN_EINVAL, "null argument");
  if (!strcmp(key, "lbm_block")) *value = c->tune.block;
  else if (!strcmp(key, "lbm_nt")) *value = c->tune.nt;
  else if (!strcmp(key, "lbm_npt")) *value = c->tune.npt;
  else if (!strcmp(key, "lbm_ntload")) *value = c->tune.ntload;
  else if (!strcmp(key, "lbm_uv_keep")) *value = c->tune.uv_keep;
  else if (!strcmp(key, "lbm_xcd")) *value = c->tune.xcd;
  else if (!strcmp(key, "lbm_reverse")) *value = c->tune.reverse;
  else if (!strcmp(key, "lbm_steps")) *value = c->tune.steps3 ? 3 : (c->tune.steps2 ? 2 : 1);
  else if (!strcmp(key, "lbm3_seg")) *value = c->tune.seg3;
  else if (!strcmp(key, "lbm3_nt")) *value = c->tune.nt3;
  else if (!strcmp(key, "lbm2_tile_h")) *value = c->tune.tile_h;
  else if (!strcmp(key, "lbm2_seg")) *value = c->tune.seg;
  else if (!strcmp(key, "lbm2_xcd")) *value = c->tune.xcd2;
  else if (!strcmp(key, "lbm2_variant")) *value = c->tune.variant;
  else if (!strcmp(key, "dep_overlap")) *value = c->tune.dep_overlap;
  else if (!strcmp(key, "dep_in_tile")) *value = c->tune.dep_in_tile;
  else if (!strcmp(key, "obs_device")) *value = c->tune.obs_device;
  else if (!strcmp(key, "particle_overlap")) *value = c->tune.overlap;
  else if (!strcmp(key, "particle_fuse")) *value = c->tune.fuse;
  else if (!strcmp(key, "adj_store_deposit")) *value = c->adj_store_deposit;
  else return rbn::thread_fail(RBN_EINVAL, "unknown tuning key");
  return RBN_OK;
}

int rbn_set_kernel_timing(rbn_ctx *c, int every) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (every < 0) return fail(c, RBN_EINVAL, "negative timing interval");
  if ((rc = drain_timing(c))) return rc;
  c->timing = every != 0;
  c->timing_every = every > 0 ? every : 1;
  for (long long &n : c->k_seen) n = 0;
  return RBN_OK;
}

int rbn_kernel_stats(rbn_ctx *c, const char *name, int reset, long long *launches, double *total_ms) {
  int rc = check_ctx(c);
  if (rc) return rc;
  const int k = kernel_index(name);
  if (k < 0) return fail(c, RBN_EINVAL, "unknown kernel name");
  if ((rc = drain_timing(c))) return rc;
  if (launches) *launches = c->k_launches[k];
  if (total_ms) *total_ms = c->k_ms[k];
  if (reset) {
    c->k_launches[k] = 0;
    c->k_ms[k] = 0.0;
  }
  return RBN_OK;
}

int rbn_time_steps(rbn_ctx *c, int nsteps, double *elapsed_ms) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!elapsed_ms) return fail(c, RBN_EINVAL, "null output");
  hipEvent_t a = nullptr, b = nullptr;
  HIPCHK(c, hipEventCreate(&a));
  HIPCHK(c, hipEventCreate(&b));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipEventRecord(a, c->stream));
  rc = rbn_step(c, nsteps);
  if (rc == RBN_OK) {
    HIPCHK(c, hipEventRecord(b, c->stream));
    HIPCHK(c, hipEventSynchronize(b));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, a, b));
    *elapsed_ms = ms;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return rc;
}

}  // extern "C"
