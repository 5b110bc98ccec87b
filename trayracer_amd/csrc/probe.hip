// Measurement hook: the vector-memory ceiling of the traversal kernels' loads.
//
// Extend and Shadow read 48-byte records (a node: three 16-byte rows; a triangle: three
// 16-byte rows) at data-dependent addresses, one record per lane, and their rate is set
// by how many such gathers the CUs' vector-memory path (address and data units, L1, L2)
// completes, not by HBM bandwidth: the working set is mostly served by the L2 and the
// Infinity Cache. This kernel issues the same access pattern with nothing else in the
// way: each lane reads 48-byte records at pseudo-random record indices of a buffer, the
// loads independent of each other so that latency hides behind occupancy. Over a buffer
// that fits the L2 it measures the highest rate such gathers reach, the ceiling that
// the traversal's loaded bytes are compared with.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "../../include/trt.h"

namespace {

constexpr int kProbeBlock = 256;

__global__ __launch_bounds__(kProbeBlock) void k_gather48(const float4* __restrict__ buf, uint32_t n_rec,
                                                          uint32_t iters, uint32_t seed, float* sink) {
  uint32_t h = (blockIdx.x * kProbeBlock + threadIdx.x) * 2654435761u ^ seed;
  float acc = 0.f;
  for (uint32_t k = 0; k < iters; ++k) {
    h = h * 1664525u + 1013904223u;
    const uint32_t r = __umulhi(h, n_rec);
    const float4* p = buf + static_cast<size_t>(r) * 3u;
    const float4 a = p[0], b = p[1], c = p[2];
    // every loaded component is used, so the loads stay full 16-byte loads
    acc += ((a.x + a.y) + (a.z + a.w)) + ((b.x + b.y) + (b.z + b.w)) + ((c.x + c.y) + (c.z + c.w));
  }
  if (acc == -1.25f) sink[0] = acc;  // never true for the zero-filled buffer: keeps the loads live
}

}  // namespace

extern "C" trt_status trt_measure_gather(int device, uint64_t buffer_bytes, uint32_t iters, double* gbs) {
  if (!gbs || buffer_bytes < 48 || iters == 0 || buffer_bytes / 48 > 0xffffffffull) return TRT_INVALID_ARG;
  if (hipSetDevice(device) != hipSuccess) return TRT_DEVICE;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return TRT_DEVICE;
  const uint32_t n_rec = static_cast<uint32_t>(buffer_bytes / 48);
  void* buf = nullptr;
  float* sink = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  trt_status st = TRT_OK;
  auto check = [&](hipError_t e, const char*) {
    if (e != hipSuccess && st == TRT_OK) {
      (void)hipGetLastError();
      st = e == hipErrorOutOfMemory ? TRT_OUT_OF_MEMORY : TRT_DEVICE;
    }
    return st == TRT_OK;
  };
  if (check(hipMalloc(&buf, size_t(n_rec) * 48), "alloc probe buffer") && check(hipMalloc(&sink, 16), "alloc") &&
      check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "stream") && check(hipEventCreate(&e0), "event") &&
      check(hipEventCreate(&e1), "event") && check(hipMemsetAsync(buf, 0, size_t(n_rec) * 48, s), "clear")) {
    // eight 256-thread blocks per CU (32 waves, the most a CU holds at this size)
    const unsigned grid = static_cast<unsigned>(prop.multiProcessorCount) * 8u;
    double best = 0.0;
    for (int rep = 0; rep < 4 && st == TRT_OK; ++rep) {
      check(hipEventRecord(e0, s), "event");
      hipLaunchKernelGGL(k_gather48, dim3(grid), dim3(kProbeBlock), 0, s, static_cast<const float4*>(buf), n_rec,
                         iters, 0x9e3779b9u * (rep + 1), sink);
      check(hipGetLastError(), "probe launch");
      check(hipEventRecord(e1, s), "event");
      check(hipEventSynchronize(e1), "probe");
      float ms = 0.f;
      check(hipEventElapsedTime(&ms, e0, e1), "event time");
      if (rep > 0 && ms > 0.f)  // the first launch warms up
        best = std::max(best, double(grid) * kProbeBlock * iters * 48.0 / (ms * 1e-3) / 1e9);
    }
    *gbs = best;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s) (void)hipStreamDestroy(s);
  if (sink) (void)hipFree(sink);
  if (buf) (void)hipFree(buf);
  return st;
}
