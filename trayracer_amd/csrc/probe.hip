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
// the traversal's loaded bytes are compared with. A second entry times the same loads
// per wave-level instruction with fewer active lanes, or with every lane of a wave at
// one record, to show what a load instruction costs the load path when lanes idle.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "../../include/trt.h"

namespace {

constexpr int kProbeBlock = 256;

// Lanes below `active` read records; per_wave = 0: each lane its own pseudo-random
// record, 1: every active lane of a wave the same (pseudo-random) record.
__global__ __launch_bounds__(kProbeBlock) void k_gather48(const float4* __restrict__ buf, uint32_t n_rec,
                                                          uint32_t iters, uint32_t seed, uint32_t active,
                                                          uint32_t per_wave, float* sink) {
  const uint32_t tid = blockIdx.x * kProbeBlock + threadIdx.x;
  uint32_t h = (per_wave ? (tid >> 6) : tid) * 2654435761u ^ seed;
  float acc = 0.f;
  if ((threadIdx.x & 63u) < active) {
    for (uint32_t k = 0; k < iters; ++k) {
      h = h * 1664525u + 1013904223u;
      const uint32_t r = __umulhi(h, n_rec);
      const float4* p = buf + static_cast<size_t>(r) * 3u;
      const float4 a = p[0], b = p[1], c = p[2];
      // every loaded component is used, so the loads stay full 16-byte loads
      acc += ((a.x + a.y) + (a.z + a.w)) + ((b.x + b.y) + (b.z + b.w)) + ((c.x + c.y) + (c.z + c.w));
    }
  }
  if (acc == -1.25f) sink[0] = acc;  // never true for the zero-filled buffer: keeps the loads live
}

// Best of three timed launches (after one warm-up) of k_gather48 over a zero-filled
// buffer of n_rec records; ms = launch time.
trt_status time_gather(int device, uint32_t n_rec, uint32_t iters, uint32_t active, uint32_t per_wave, unsigned& grid,
                       float& best_ms) {
  if (hipSetDevice(device) != hipSuccess) return TRT_DEVICE;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return TRT_DEVICE;
  void* buf = nullptr;
  float* sink = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  trt_status st = TRT_OK;
  auto check = [&](hipError_t e) {
    if (e != hipSuccess && st == TRT_OK) {
      (void)hipGetLastError();
      st = e == hipErrorOutOfMemory ? TRT_OUT_OF_MEMORY : TRT_DEVICE;
    }
    return st == TRT_OK;
  };
  // eight 256-thread blocks per CU (32 waves, the most a CU holds at this size)
  grid = static_cast<unsigned>(prop.multiProcessorCount) * 8u;
  best_ms = 0.f;
  if (check(hipMalloc(&buf, size_t(n_rec) * 48)) && check(hipMalloc(&sink, 16)) &&
      check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking)) && check(hipEventCreate(&e0)) &&
      check(hipEventCreate(&e1)) && check(hipMemsetAsync(buf, 0, size_t(n_rec) * 48, s))) {
    for (int rep = 0; rep < 4 && st == TRT_OK; ++rep) {
      check(hipEventRecord(e0, s));
      hipLaunchKernelGGL(k_gather48, dim3(grid), dim3(kProbeBlock), 0, s, static_cast<const float4*>(buf), n_rec,
                         iters, 0x9e3779b9u * (rep + 1), active, per_wave, sink);
      check(hipGetLastError());
      check(hipEventRecord(e1, s));
      check(hipEventSynchronize(e1));
      float ms = 0.f;
      check(hipEventElapsedTime(&ms, e0, e1));
      if (rep > 0 && ms > 0.f && (best_ms == 0.f || ms < best_ms)) best_ms = ms;  // the first launch warms up
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s) (void)hipStreamDestroy(s);
  if (sink) (void)hipFree(sink);
  if (buf) (void)hipFree(buf);
  return st;
}

}  // namespace

extern "C" trt_status trt_measure_gather(int device, uint64_t buffer_bytes, uint32_t iters, double* gbs) {
  if (!gbs || buffer_bytes < 48 || iters == 0 || buffer_bytes / 48 > 0xffffffffull) return TRT_INVALID_ARG;
  unsigned grid = 0;
  float ms = 0.f;
  const trt_status st = time_gather(device, static_cast<uint32_t>(buffer_bytes / 48), iters, 64, 0, grid, ms);
  if (st != TRT_OK) return st;
  *gbs = ms > 0.f ? double(grid) * kProbeBlock * iters * 48.0 / (ms * 1e-3) / 1e9 : 0.0;
  return TRT_OK;
}

extern "C" trt_status trt_measure_gather_lanes(int device, uint64_t buffer_bytes, uint32_t iters, uint32_t active_lanes,
                                               uint32_t same_record, double* ns_per_load) {
  if (!ns_per_load || buffer_bytes < 48 || iters == 0 || buffer_bytes / 48 > 0xffffffffull || active_lanes == 0 ||
      active_lanes > 64 || same_record > 1)
    return TRT_INVALID_ARG;
  hipDeviceProp_t prop;
  if (hipSetDevice(device) != hipSuccess || hipGetDeviceProperties(&prop, device) != hipSuccess) return TRT_DEVICE;
  unsigned grid = 0;
  float ms = 0.f;
  const trt_status st =
      time_gather(device, static_cast<uint32_t>(buffer_bytes / 48), iters, active_lanes, same_record, grid, ms);
  if (st != TRT_OK) return st;
  // wave-level 16-byte load instructions per CU: grid x 4 waves x iters x 3 over the CUs
  const double per_cu = double(grid) * (kProbeBlock / 64) * iters * 3.0 / prop.multiProcessorCount;
  *ns_per_load = ms * 1e6 / per_cu;
  return TRT_OK;
}
