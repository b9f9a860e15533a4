// rbn_noalloc.hip — rbn_sample_noalloc: bilinear sampling (rbn_sample's kernel) with the
// positions and samples staged in a fixed device array instead of a buffer allocated per
// call, for callers that cannot allocate device memory (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc).
//
// The array `scratch` is a file-scope __device__ array of kSampleNoallocMax positions
// (3 doubles each: x, y and the sample), 6 MiB that every GPU the library runs on holds
// until the process ends.  Its name is a host-side stand-in for the symbol; the device
// address is resolved once per device with hipGetSymbolAddress (with that device current)
// and kept, and only that address reaches the copies and the kernel.  All contexts of one
// GPU share the array: calls that overlap on one GPU race (one call's positions or samples
// overwritten by another's); sequential calls do not, because each call waits for its copy
// back before returning.  Nothing in the array needs to start at zero: every call writes the
// positions it reads and the samples it returns.
#include <hip/hip_runtime.h>

#include <mutex>

#include "rbn_host.h"

using namespace rbn::host;

namespace {

constexpr int kSampleNoallocMax = RBN_SAMPLE_NOALLOC_MAX;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[3 * (size_t)kSampleNoallocMax];

std::mutex g_scratch_mu;
double *g_scratch[kMaxDevices] = {};

// The device address of `scratch` on the context's device (current after check_ctx).
int scratch_of(rbn_ctx *c, double **out) {
  if (c->dev < 0 || c->dev >= kMaxDevices)
    return fail(c, RBN_ERANGE, "rbn_sample_noalloc: device %d beyond the %d devices it serves", c->dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[c->dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(c, RBN_EDEVICE, "rbn_sample_noalloc: hipGetSymbolAddress(scratch) on device %d failed: %s", c->dev,
                  hipGetErrorString(e));
    g_scratch[c->dev] = (double *)p;
  }
  *out = g_scratch[c->dev];
  return RBN_OK;
}

}  // namespace

extern "C" int rbn_sample_noalloc(rbn_ctx *c, int field, int n, const double *x, const double *y, double *out) {
  int rc = check_sample_args(c, "rbn_sample_noalloc", field, n, x, y, out);
  if (rc) return rc;
  if (n > kSampleNoallocMax)
    return fail(c, RBN_ERANGE, "rbn_sample_noalloc: %d positions exceed its maximum of %d (use rbn_sample)", n,
                kSampleNoallocMax);
  if (n == 0) return RBN_OK;
  double *buf = nullptr;
  if ((rc = scratch_of(c, &buf))) return rc;
  return sample_staged(c, field, n, x, y, out, buf);
}
