// Device-side data structures of the wavefront renderer (shared by kernels.hip and
// the host context in api.cpp).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bvh_build.h"
#include "trt/integrator.h"
#include "trt/render_modes.h"
#include "trt/tonemap.h"

namespace trt {

// Structure-of-arrays queue of path segments (one slot per live ray). Each field is
// its own array so that a wave reads/writes 64 consecutive 16-byte records.
struct RayQueue {
  float4* o_tmin;   // origin, tmin
  float4* d_tmax;   // direction, tmax
  float4* t_pdf;    // throughput rgb, pdf of the sampled direction
  uint32_t* meta;   // path << 2 | flags
};

// Closest-hit results, indexed like the ray queue slot that produced them.
struct HitBuf {
  float4* thit;     // t, b1, b2, prim (bits)
  uint32_t* inst;   // instance index (two-level scenes)
};

// Shadow-ray queue, and beside it the emission list of the same bounce: radiance that
// Shade found (emitter hits, sky) waiting to be added to its path.
struct ShadowQueue {
  float4* o_tmin;
  float4* d_tmax;
  float4* contrib;  // rgb, path (bits)
  float4* emit;     // emission list: rgb, path (bits); null where no Shade writes one
};

// Acceleration structure views.
struct DevAccel {
  const DevNode48* nodes;   // single-level BVH, or all BLAS nodes for two-level scenes
  const float4* tris;       // reordered traversal triangles: 3 float4 per triangle (w of row 0 = prim id)
  const DevNode48* tlas;    // top level over instances (nullptr for single-level)
  const Instance* inst;     // instances (two-level)
  int stack_cap;            // stack entries per lane (tree depth bound)
  int two_level;
  uint32_t* spill;          // single-level kernels: per-lane stack spill area, stack_cap entries
                            // per thread of the traversal grid (blockIdx.x * 256 + threadIdx.x)
  uint32_t* touch_nodes;    // counting kernels only (else null): one bit per (bottom-level) node
  uint32_t* touch_tris;     // and per triangle record, set when the launch reads it
};

// Queues are split into kSegs segments of `cap` slots each (slot = segment * cap + i).
// Every segment has its own append and work-fetch counters, each on its own 128-byte
// line: returning atomics on one line serialise at about 88 per microsecond on MI355X,
// so a single shared counter caps a queue at ~5.6 G rays/s (one atomic per 64 rays);
// spread over 32 lines they stop mattering. Producer waves append to their home
// segment (wave index mod kSegs) and move on to the next segment when it is full;
// consumer waves fetch 64-ray batches from their home segment first, then from the
// others. Capacity is 1.5x the path count, so appends always find room.
constexpr uint32_t kSegs = 32;
constexpr uint32_t kLine = 32;  // uint32 words per 128-byte line

// Counters in device memory; zeroed at the start of every frame by a kernel.
// Layout (uint32 indices):
//   [kCtrExt + d]    total rays in the extension queue of depth d (d = 0: primary rays),
//   [kCtrShadow + d] total shadow rays emitted at depth d; both written at the end of a
//                    frame by k_sum_counts from the per-segment counters
//   [kCtrOverflow]   bit 0: traversal stack overflow; bit 1: queue append failed
//   [kCtrStats + k]  traversal statistics (node visits, triangle tests, stack entries moved
//                    from the LDS ring to the spill area: 64-bit pairs)
//   [seg_ctr(kind, d) + s * kLine] per-segment counters, kinds below
constexpr int kMaxDepth = 32;
constexpr int kCtrExt = 0;
constexpr int kCtrShadow = 32;
constexpr int kCtrOverflow = 161;
constexpr int kCtrStats = 164;  // 10 x uint64: closest nodes, closest tris, any nodes, any tris,
                                // closest stack spills, any stack spills, then the distinct
                                // nodes and triangle records read (summed over launches) by
                                // closest and any traversals
constexpr int kCtrSegBase = 256;
enum SegKind : int { kSegExt = 0, kSegShadow = 1, kSegWorkExt = 2, kSegWorkShade = 3, kSegWorkShadow = 4, kSegEmit = 5 };
constexpr int kNumSegKinds = 6;
constexpr int kSegDepths = kMaxDepth + 1;  // the last shade may address depth kMaxDepth
constexpr int kNumCounters = kCtrSegBase + kNumSegKinds * kSegDepths * int(kSegs * kLine);
constexpr int seg_ctr(int kind, int depth) { return kCtrSegBase + (kind * kSegDepths + depth) * int(kSegs * kLine); }

// Slots per segment for n paths (a multiple of 64, kSegs * cap >= 1.5 n), and the
// primary-ray layout: path p sits in segment p / prim_per_seg(n) (prim_per_seg <= cap).
inline uint32_t seg_cap(uint64_t n) {
  const uint64_t c = (3 * n + 2 * kSegs - 1) / (2 * kSegs);
  return static_cast<uint32_t>((c + 63) / 64 * 64);
}
inline uint32_t prim_per_seg(uint64_t n) {
  const uint64_t c = (n + kSegs - 1) / kSegs;
  return static_cast<uint32_t>((c + 63) / 64 * 64);
}

// A segmented queue's counters as seen by a kernel.
struct SegIn {
  const uint32_t* cnt;  // kSegs lines: items appended per segment (may exceed cap)
  uint32_t* work;       // kSegs lines: fetch cursors
  uint32_t cap;
};

struct FrameBuffers {
  float4* radiance;   // per path (rgb)
  float4* accum;      // per pixel: rgb sum, sample count
  float4* color;      // per pixel mean
  float4* albedo;     // per pixel mean over the accumulated samples of the first-hit albedo
  float4* normal;     // per pixel mean of the first-hit shading normal (xyz) and view depth (w)
  float4* path_aov_a; // per path: first-hit albedo (rgb), view depth (w)
  float4* path_aov_n; // per path: first-hit shading normal (xyz)
  float4* aov_sum_a;  // per pixel sums of the above over the accumulated samples
  float4* aov_sum_n;
};

// Adjoint (path replay) state. Records: one per (path, depth) slot r = path*depth + d,
// key = material id of the vertex (0xffffffff = none), rec_g[k * n_records + r] =
// gradient of parameter k (SoA so a wave's loads coalesce).
struct AdjBuffers {
  const float* dldi;   // dLoss/dImage, 3 floats per pixel
  float4* lrem;        // per path: radiance not yet attributed to earlier vertices
  float4* pend;        // per path: 4 float4 = (d f / d theta)/f of the pending continuation + material
  uint32_t* rec_key;
  float* rec_g;
  float4* s_ga;        // per shadow slot: NEE gradient, params 0..3
  float* s_gb;         // per shadow slot: NEE gradient, param 4
  size_t n_records;
  uint32_t depth;
};
constexpr int kGradBlocks = 1024;       // fixed: the reduction order is device independent
constexpr uint32_t kGradMatsPerPass = 512;

// Resident 256-thread blocks per CU of the traversal kernels for this accel's LDS stack.
int trace_blocks_per_cu(const DevAccel& acc);
bool traversal_lds_check();  // false if a queue traversal kernel has static LDS

// Kernel launchers (kernels.hip). All run on `stream`. `cap` is the per-segment
// capacity of every queue; out_cnt / shadow_cnt point at kSegs counter lines.
void launch_zero_counters(uint32_t* ctr, hipStream_t stream);
void launch_raygen(const RenderParams& rp, uint32_t n_paths, uint32_t cap, RayQueue q, const FrameBuffers& fb,
                   uint32_t* ctr, hipStream_t stream);
void launch_extend(const DevAccel& acc, RayQueue q, HitBuf h, SegIn in, uint32_t* ctr, int grid, bool stats,
                   hipStream_t stream);
// Shade writes no path radiance: what a vertex emits goes to the emission list sq.emit
// (segmented, counters emit_cnt), which launch_emit adds to the paths.
void launch_shade(const SceneView& sv, const Material* sampling_mats, const RenderParams& rp, RayQueue q,
                  HitBuf h, RayQueue qn, ShadowQueue sq, FrameBuffers fb, SegIn in, uint32_t* next_cnt,
                  uint32_t* shadow_cnt, uint32_t* emit_cnt, uint32_t* ctr, uint32_t depth, int grid,
                  hipStream_t stream);
// Adds an emission list (kSegs segments of `cap` slots, counters emit_cnt) to the path
// radiance; every path is listed at most once, so entries need no atomics.
void launch_emit(const float4* emit, const uint32_t* emit_cnt, uint32_t cap, float4* radiance, hipStream_t stream);
void launch_shadow(const DevAccel& acc, ShadowQueue sq, float4* radiance, SegIn in, uint32_t* ctr, int grid,
                   bool stats, hipStream_t stream);
void launch_sum_counts(uint32_t* ctr, uint32_t cap, hipStream_t stream);
// Adds the set bits of a touch bitmap (`words` 32-bit words) to *out and clears it.
void launch_touch_count(uint32_t* bits, uint32_t words, unsigned long long* out, hipStream_t stream);
void launch_tonemap(const float4* src, uint32_t n, const TonemapParams& tp, uchar4* dst, hipStream_t stream);
// reset: restart the accumulator; aov_reset: restart the AOV sums; aov_count: samples in
// the AOV sums after this frame.
void launch_finalize(const RenderParams& rp, const float4* radiance, FrameBuffers fb, uint32_t reset,
                     uint32_t aov_reset, uint32_t aov_count, hipStream_t stream);
void launch_raygen_adj(const RenderParams& rp, uint32_t n_paths, uint32_t cap, RayQueue q, const float4* radiance,
                       const AdjBuffers& ab, uint32_t* ctr, hipStream_t stream);
void launch_shade_adj(const SceneView& sv, const Material* sampling_mats, const RenderParams& rp, RayQueue q,
                      HitBuf h, RayQueue qn, ShadowQueue sq, const AdjBuffers& ab, SegIn in, uint32_t* next_cnt,
                      uint32_t* shadow_cnt, uint32_t* ctr, uint32_t depth, int grid, hipStream_t stream);
void launch_shadow_adj(const DevAccel& acc, ShadowQueue sq, const AdjBuffers& ab, SegIn in, uint32_t* ctr,
                       uint32_t depth, int grid, hipStream_t stream);
// partial: kGradBlocks * n_mat * 5 floats; out: n_mat * 5 floats (device)
void launch_grad_reduce(const AdjBuffers& ab, uint32_t mat_lo, uint32_t n_mat, float* partial, float* out,
                        hipStream_t stream);

}  // namespace trt
