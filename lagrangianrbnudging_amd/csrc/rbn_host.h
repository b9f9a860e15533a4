// rbn_host.h — host-side helpers shared by the C-ABI translation units (rbn_api.cpp,
// rbn_adjoint.cpp): error reporting, device allocation, per-launch timing, observer array
// views and the observation-record upload.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

#include "rbn_internal.h"

namespace rbn {
namespace host {

// Record an error on the context (and the calling thread) and return code.
int fail(rbn_ctx *c, int code, const char *fmt, ...);

#define HIPCHK(ctx, expr)                                                                   \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return ::rbn::host::fail((ctx), RBN_EDEVICE, "%s failed: %s (%s:%d)", #expr,          \
                               hipGetErrorString(e_), __FILE__, __LINE__);                  \
  } while (0)

int dalloc(rbn_ctx *c, void **p, size_t bytes);
void dfree(void *&p);

template <typename T>
void dfree_t(T *&p) {
  void *q = p;
  dfree(q);
  p = nullptr;
}

int check_ctx(rbn_ctx *c);

// Kernel classes of the per-launch timing (rbn_set_kernel_timing / rbn_kernel_stats).
// KT_LBM_FUSED: a lattice launch carrying earlier steps' particle updates; KT_LBM2 / KT_LBM3:
// a two-step / three-step lattice launch (with or without carried particle updates).
enum { KT_LBM, KT_PARTICLES, KT_DEPOSIT, KT_ADJOINT, KT_LBM_FUSED, KT_LBM2, KT_LBM3, KT_COUNT };
static_assert(KT_COUNT == rbn_ctx::kKernelClasses, "kernel classes");
extern const char *const kKernelNames[KT_COUNT];
hipEvent_t take_event(rbn_ctx *c);
int drain_timing(rbn_ctx *c);

// Launch through fn() on stream s, bracketed with events on s when kernel timing is on
// (every timing_every-th launch of the class).
template <typename Fn>
int timed(rbn_ctx *c, int kernel, hipStream_t s, Fn &&fn) {
  hipEvent_t a = nullptr, b = nullptr;
  const bool sample = c->timing && (c->k_seen[kernel]++ % c->timing_every) == 0;
  if (sample) {
    a = take_event(c);
    b = take_event(c);
    if (!a || !b) return fail(c, RBN_EDEVICE, "hipEventCreate failed");
    HIPCHK(c, hipEventRecord(a, s));
  }
  hipError_t e = fn();
  if (e != hipSuccess) return fail(c, RBN_EDEVICE, "%s launch failed: %s", kKernelNames[kernel], hipGetErrorString(e));
  if (sample) {
    HIPCHK(c, hipEventRecord(b, s));
    c->pending.push_back({kernel, a, b});
    if (c->pending.size() > 8192) return drain_timing(c);
  }
  return RBN_OK;
}

// Device field <-> host array in the reference's [nx][ny] order (y fastest).
int upload_field(rbn_ctx *c, void *dst, const double *h);
// (u, v) interleaved into one velocity slot.
int upload_uv(rbn_ctx *c, void *dst, const double *hu, const double *hv);
// Node o of the source is src[o * cs] (cs = 2 for one component of a velocity slot).
int download_field(rbn_ctx *c, const void *src, double *h, double add = 0.0, int cs = 1);

// Observer arrays: ints chk, bcy, xidx, yidx; doubles dx, dy, dT, xint, yint, Tint and the
// current record rx, ry, rT (record order).  Each array holds nobs entries.
enum { OCHK, OBCY, OXIDX, OYIDX, ONI };
enum { ODX, ODY, ODT, OXINT, OYINT, OTINT, ORX, ORY, ORT, OND };
ObsPtrs obs_view(int n, int *ints, double *dbls);
ObsPtrs obs_ptrs(rbn_ctx *c);

// Fetch record rstep from the observation provider, map observers to record entries by id
// and upload it as the current record (on the context's stream, or on s).
int load_observation(rbn_ctx *c, int rstep);
int load_observation(rbn_ctx *c, int rstep, hipStream_t s);
// The host half: record rstep into rx, ry, rT (n = nobs each; ids into rid) and the observer
// map into chk; *same_map: the map is the one already on the device (chk not written).
int fetch_observation(rbn_ctx *c, int rstep, double *rx, double *ry, double *rT, int *rid, int *chk,
                      bool *same_map);

// Report (and clear) a particle / observer range error or a deposit overflow.
int check_err_flag(rbn_ctx *c);

// rbn_sample's argument and state checks, and its work given a device buffer of 3n doubles
// (positions in, samples out): copies, the sampling launch, the copy back, the range check.
int check_sample_args(rbn_ctx *c, const char *who, int field, int n, const double *x, const double *y,
                      const double *out);
int sample_staged(rbn_ctx *c, int field, int n, const double *x, const double *y, double *out, double *buf);
// Clear a bilinear deposit set's tile counters and overflow sums (zero between deposits) on
// stream s: a call that failed between an observer update and its tile pass leaves them set.
hipError_t clear_dep_bins(const rbn::DepPtrs &d, const rbn::Geom &g, hipStream_t s);

}  // namespace host
}  // namespace rbn
