// copy_probe.hip — a plain device-to-device copy of known size, for two measurements:
//   1. the HBM bandwidth a streaming kernel reaches on this GPU (read + write bytes / time);
//   2. the scale of rocprofv3's FETCH_SIZE / WRITE_SIZE counters: profiling this program
//      gives the counter values for exactly `bytes` read and `bytes` written per launch.
// usage: copy_probe [MiB=1024] [repeats=20] [nt=0]   (nt=1: non-temporal loads and stores)
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define CHECK(x)                                                                        \
  do {                                                                                  \
    hipError_t e_ = (x);                                                                \
    if (e_ != hipSuccess) {                                                             \
      fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                          \
      return 1;                                                                         \
    }                                                                                   \
  } while (0)

typedef float v4f __attribute__((ext_vector_type(4)));

template <int NT>
__global__ __launch_bounds__(256) void k_copy(const v4f *__restrict__ a, v4f *__restrict__ b,
                                              size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    if (NT) {
      v4f v = __builtin_nontemporal_load(&a[i]);
      __builtin_nontemporal_store(v, &b[i]);
    } else {
      b[i] = a[i];
    }
  }
}

int main(int argc, char **argv) {
  const size_t mib = argc > 1 ? strtoull(argv[1], nullptr, 10) : 1024;
  const int reps = argc > 2 ? atoi(argv[2]) : 20;
  const int nt = argc > 3 ? atoi(argv[3]) : 0;
  const size_t bytes = mib << 20, n = bytes / sizeof(v4f);
  v4f *a = nullptr, *b = nullptr;
  CHECK(hipMalloc(&a, bytes));
  CHECK(hipMalloc(&b, bytes));
  CHECK(hipMemset(a, 1, bytes));
  CHECK(hipMemset(b, 0, bytes));
  int dev = 0, cus = 0;
  CHECK(hipGetDevice(&dev));
  CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const unsigned grid = (unsigned)(cus * 16);
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  for (int w = 0; w < 3; w++) {
    if (nt) hipLaunchKernelGGL(k_copy<1>, dim3(grid), dim3(256), 0, 0, a, b, n);
    else hipLaunchKernelGGL(k_copy<0>, dim3(grid), dim3(256), 0, 0, a, b, n);
  }
  CHECK(hipDeviceSynchronize());
  CHECK(hipEventRecord(e0, 0));
  for (int r = 0; r < reps; r++) {
    if (nt) hipLaunchKernelGGL(k_copy<1>, dim3(grid), dim3(256), 0, 0, a, b, n);
    else hipLaunchKernelGGL(k_copy<0>, dim3(grid), dim3(256), 0, 0, a, b, n);
  }
  CHECK(hipEventRecord(e1, 0));
  CHECK(hipEventSynchronize(e1));
  float ms = 0.f;
  CHECK(hipEventElapsedTime(&ms, e0, e1));
  const double per = ms / reps;
  printf("{\"probe\": \"copy\", \"bytes_read\": %zu, \"bytes_written\": %zu, \"avg_ms\": %.5f, "
         "\"GBps\": %.1f, \"nt\": %d, \"cus\": %d}\n",
         bytes, bytes, per, 2.0 * bytes / (per * 1e-3) / 1e9, nt, cus);
  CHECK(hipFree(a));
  CHECK(hipFree(b));
  return 0;
}
