// k_stream.hip -- a plain streaming copy, the measured HBM ceiling that the
// iteration kernels' bandwidth is compared with (the reference has no
// counterpart; it is measurement only).  One 16-byte element per thread and
// 256-thread workgroups: on MI355X this form moves more bytes per second
// than unrolled or non-temporal variants and than the runtime's own copy.
#include <hip/hip_runtime.h>

#include <string>

#include "rism3d_amd.h"
#include "rism_host.h"

namespace rism {

__global__ __launch_bounds__(256) void k_stream_copy(const float4 *__restrict__ src, float4 *__restrict__ dst,
                                                     size_t n) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

}  // namespace rism

extern "C" int rism_stream_copy(int device, size_t bytes, int reps, double *gbps, double *us) {
  using rism::set_error;
  if (bytes < 4096 || reps < 1 || (bytes & 15)) {
    set_error("rism_stream_copy: bytes must be a multiple of 16 and >= 4096, reps >= 1");
    return RISM_ERR_ARG;
  }
  float4 *a = nullptr, *b = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  std::string err;
  auto ok = [&](hipError_t e, const char *what) {
    if (e != hipSuccess && err.empty()) err = std::string(what) + ": " + hipGetErrorString(e);
    return e == hipSuccess;
  };
  const size_t n = bytes / 16;
  const unsigned grid = (unsigned)((n + 255) / 256);
  float ms = 0.f;
  if (ok(hipSetDevice(device), "hipSetDevice") && ok(hipMalloc((void **)&a, bytes), "hipMalloc") &&
      ok(hipMalloc((void **)&b, bytes), "hipMalloc") && ok(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "stream") &&
      ok(hipEventCreate(&e0), "event") && ok(hipEventCreate(&e1), "event") &&
      ok(hipMemsetAsync(a, 0, bytes, s), "memset")) {
    for (int w = 0; w < 2; ++w) rism::k_stream_copy<<<grid, 256, 0, s>>>(a, b, n);
    ok(hipGetLastError(), "launch");
    ok(hipEventRecord(e0, s), "record");
    for (int r = 0; r < reps; ++r) rism::k_stream_copy<<<grid, 256, 0, s>>>(a, b, n);
    ok(hipGetLastError(), "launch");
    ok(hipEventRecord(e1, s), "record");
    ok(hipEventSynchronize(e1), "sync");
    ok(hipEventElapsedTime(&ms, e0, e1), "elapsed");
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s) (void)hipStreamDestroy(s);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  if (!err.empty()) {
    set_error("rism_stream_copy: " + err);
    return RISM_ERR_HIP;
  }
  const double t = (double)ms / 1e3 / reps;
  if (gbps) *gbps = 2.0 * (double)bytes / t / 1e9;
  if (us) *us = t * 1e6;
  return RISM_OK;
}
