// Wavefront path-tracing kernels for gfx950 (CDNA4, wave64).
//
// Replaces the reference's device programs: __raygen__ Primary/Secondary/Shadow with
// the OptiX traversal behind optixTrace (Optix/GPU/optix.cu:225-343), the shading
// kernel PathTracingKernel (CUDA/GPU/Shade/PathTracing.h:8-175, launched from
// CUDA/GPU/cuda.cu:22-84), InitCudaCountersKernel and FimalizeFrameKernel
// (CUDA/GPU/CudaUtilityKernels.h:6-44).
//
// Design:
//  * Extend / Shadow: persistent waves refill idle lanes from a queue of
//    structure-of-arrays ray records (one atomic per refill); each lane walks the BVH2
//    (speculative while-while loop, leaf triangles dealt out over the wave) with a
//    two-tier stack: a 16-entry LDS ring per lane (lane-interleaved) over a global spill
//    area.
//  * Shade: the shared integrator step (trt/integrator.h) per ray; extension and
//    shadow rays are appended with one atomic per wave (ballot + mbcnt prefix) rather
//    than one atomic per ray as in the reference (PathTracing.h:135-141, 168-173).
//  * Queues are segmented (device.h: kSegs segments, each with its own counters on
//    separate cache lines) so that fetch and append atomics spread over 32 lines.
//  * RayGen writes thread t's ray to a fixed slot (no atomics; the primary queue is
//    dense), the paths of each sample in 8x8 pixel tiles.
//  * No host synchronisation between bounces: queue sizes stay in device counters.
//  * Radiance is accumulated per path (one writer per path per kernel) and reduced
//    per pixel in a fixed sample order in Finalize, so images are deterministic. Shade
//    lists emission rather than adding it; the emission and shadow launches of the
//    shadow stream add every term in vertex order, and Shade runs beside the previous
//    bounce's shadow launch.
#include <hip/hip_runtime.h>

#include "device.h"
#include "trt/tonemap.h"

namespace trt {

namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / 64;

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint32_t prefix_count(uint64_t mask) {
  // number of set bits of mask below this lane
  const uint32_t lo = static_cast<uint32_t>(mask), hi = static_cast<uint32_t>(mask >> 32);
  return __builtin_amdgcn_mbcnt_hi(hi, __builtin_amdgcn_mbcnt_lo(lo, 0u));
}

__device__ __forceinline__ void wave_add_stat(unsigned long long* dst, uint32_t v) {
  // sum over the wave, one atomic
  uint32_t s = v;
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if (lane_id() == 0 && s) atomicAdd(dst, static_cast<unsigned long long>(s));
}

// Fetches the next batch of up to `want` (1..64) items of a segmented queue for the
// wave: one atomic on the fetch cursor of the first segment, from `home` on, not yet
// found exhausted. `n_seg` holds, in lane s < kSegs, the item count of segment s
// (constant during the kernel); `done` collects exhausted segments (cursors only grow,
// so they stay exhausted; callers start it with seg_empty(), which also closes the
// segments outside the wave's fetch window). Every wave makes at most kFetchWindow
// failed attempts (kSegs in grids of fewer than kSegs waves), so the loop always ends. The cursors are not pre-read: coherent
// reads of a hot line serialise like the atomics themselves (measured: a pre-read of all
// 32 cursors per batch cost as much as the whole Shade kernel's work).
__device__ __forceinline__ bool seg_fetch_k(const SegIn& in, uint32_t n_seg, uint32_t home, uint64_t& done,
                                            uint32_t want, uint32_t& slot0, uint32_t& n_got) {
  const uint32_t lane = lane_id();
  const uint64_t all = (1ull << kSegs) - 1ull;
  for (;;) {
    const uint64_t m = ~done & all;
    if (m == 0ull) return false;
    const uint64_t rot = ((m >> home) | (m << (kSegs - home))) & all;
    uint32_t sgm = home + static_cast<uint32_t>(__builtin_ctzll(rot));
    if (sgm >= kSegs) sgm -= kSegs;
    uint32_t b = 0;
    if (lane == 0) b = atomicAdd(&in.work[sgm * kLine], want);
    b = __shfl(b, 0);
    const uint32_t ns = __shfl(n_seg, static_cast<int>(sgm));
    if (b < ns) {
      slot0 = sgm * in.cap + b;
      n_got = ns - b < want ? ns - b : want;
      return true;
    }
    done |= 1ull << sgm;
  }
}

__device__ __forceinline__ bool seg_fetch(const SegIn& in, uint32_t n_seg, uint32_t home, uint64_t& done,
                                          uint32_t& slot0, uint32_t& n_valid) {
  return seg_fetch_k(in, n_seg, home, done, 64u, slot0, n_valid);
}

// Home segment of the calling wave.
__device__ __forceinline__ uint32_t wave_home() { return (blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6)) % kSegs; }

// Initial `done` mask of seg_fetch (uniform per wave): segments without items, and every
// segment outside the wave's window {home, home + 1}. A wave thus only ever touches two
// fetch cursors; once the queue runs dry each cursor line takes (waves / kSegs) x 2
// failed attempts instead of one from every wave (the tail of small launches was spent
// in those serialised atomics). Every segment lies in the window of the waves whose
// home it is, so when every segment is some wave's home (at least kSegs waves in the
// grid with a home per wave, kSegs blocks with a home per block: every_segment_homed)
// every item is fetched; otherwise all segments stay open.
constexpr uint32_t kFetchWindow = 2;
__device__ __forceinline__ uint64_t seg_empty(uint32_t n_seg, uint32_t home, bool every_segment_homed) {
  const uint64_t all = (1ull << kSegs) - 1ull;
  uint64_t closed = ~__ballot(n_seg > 0u);
  if (every_segment_homed) {
    const uint64_t w = (1ull << kFetchWindow) - 1ull;
    closed |= ~(((w << home) | (w >> (kSegs - home))) & all);
  }
  return closed & all;
}
// seg_empty for a home per wave (wave_home).
__device__ __forceinline__ uint64_t seg_empty(uint32_t n_seg) {
  return seg_empty(n_seg, wave_home(), gridDim.x * kWavesPerBlock >= kSegs);
}

// Home segment of a wave in the single-level traversal and in Shade: one per block, so
// that the four waves of a block fetch neighbouring batches of the same part of the
// queue (neighbouring rays, which share nodes, triangles and materials in the CU's L1),
// and Shade's waves of a block append their outputs next to each other: 4 % faster Extend
// and 3.5 % faster Shade on the 1M-triangle soup than a home per wave. (The two-level
// traversal keeps a home per wave: there the block home measured 1.5 % slower, and a
// home per XCD, segment (b % 8) x 4 + (b / 8) % 4 for block b, slower on both scenes.)
// Every segment is some block's home once the grid has kSegs blocks.
__device__ __forceinline__ uint32_t block_home() { return blockIdx.x % kSegs; }

// Per-lane item count of segment `lane` (lanes >= kSegs: 0).
__device__ __forceinline__ uint32_t seg_counts(const SegIn& in) {
  const uint32_t lane = lane_id();
  if (lane >= kSegs) return 0u;
  const uint32_t c = in.cnt[lane * kLine];
  return c < in.cap ? c : in.cap;
}


// Appends the lanes in `mask` to a segmented queue; returns this lane's slot (kNoHit
// for lanes not in the mask). Starts at segment `home` and moves on while segments are
// full; a partly full segment takes what fits. Total capacity (kSegs * cap) is at
// least the number of items ever appended per queue, so the loop places everything
// within kSegs attempts; otherwise bit 1 of *overflow is set.
__device__ __forceinline__ uint32_t seg_append(uint32_t* cnt, uint32_t cap, uint32_t home, uint64_t mask,
                                               uint32_t* overflow) {
  const uint32_t lane = lane_id();
  const uint32_t need = static_cast<uint32_t>(__popcll(mask));
  const uint32_t rank = prefix_count(mask);
  const bool mine = (mask >> lane) & 1ull;
  uint32_t slot = kNoHit, done = 0, sgm = home;
  for (uint32_t t = 0; t < kSegs && done < need; ++t) {
    uint32_t b = 0;
    if (lane == 0) b = atomicAdd(&cnt[sgm * kLine], need - done);
    b = __shfl(b, 0);
    const uint32_t fit = b < cap ? (need - done < cap - b ? need - done : cap - b) : 0u;
    if (mine && rank >= done && rank < done + fit) slot = sgm * cap + b + (rank - done);
    done += fit;
    sgm = sgm + 1 == kSegs ? 0u : sgm + 1;
  }
  if (done < need && lane == 0) atomicOr(overflow, 2u);
  return slot;
}

struct TraceOut {
  float t, b1, b2;
  uint32_t prim, inst;
  uint32_t nodes, tests, spills;  // statistics: node visits, triangle tests, stack entries spilled
};

// Counting kernels: marks item i of a touch bitmap (null: not collected).
__device__ __forceinline__ void touch(uint32_t* bits, uint32_t i) {
  if (bits) atomicOr(&bits[i >> 5], 1u << (i & 31u));
}

// Ray/box slab test of the device traversals (after Ize 2013's conservative form), on
// boxes stored as centre c and half extent h per axis. With x = c - o and s = sign(inv)
// (per ray, +-1), a slab's entry and exit are (x - s h) * inv and (x + s h) * inv, the
// entry never above the exit; each is within a few ulp of itself plus u |x inv|. The
// stored half extent exceeds the exact box by 2^-10 of itself, which covers the error
// relative to h, and the far distance is widened by kBoxSlack, which covers the rest, so
// no box the ray touches is culled and traversal equals a brute-force loop. inv = 1/d is
// the IEEE reciprocal: for a zero direction component it is +-inf and the products give
// the exact slab of a ray parallel to the axis: (-inf, +inf) when the origin lies inside
// it, both +inf or both -inf (an empty interval) when outside, and NaN for an end whose
// plane the origin lies on; the axes are combined with max / min that ignore NaN, so that
// end limits nothing. Only FP32 subtracts, multiplies and adds, then three-operand max /
// min: on gfx950 FP32 arithmetic issues at twice the rate of min / max, so this form is
// cheaper than ordering each axis's two plane distances with min / max.
struct RayBox {
  V3 o;
  V3 inv;  // 1 / d, infinite for a zero component
  V3 sgn;  // +-1, the sign of inv
};

__device__ __forceinline__ RayBox ray_box_setup(V3 o, V3 d) {
  RayBox b;
  b.o = o;
  b.inv = v3(1.f / d.x, 1.f / d.y, 1.f / d.z);
  b.sgn = v3(__builtin_copysignf(1.f, b.inv.x), __builtin_copysignf(1.f, b.inv.y), __builtin_copysignf(1.f, b.inv.z));
  return b;
}

__device__ __forceinline__ bool slab_ch(float cx, float hx, float cy, float hy, float cz, float hz,
                                        const RayBox& b, float tmin, float tmax, float& tnear) {
  const float xx = cx - b.o.x, sx = hx * b.sgn.x;
  const float xy = cy - b.o.y, sy = hy * b.sgn.y;
  const float xz = cz - b.o.z, sz = hz * b.sgn.z;
  const float n = __builtin_fmaxf(__builtin_fmaxf((xx - sx) * b.inv.x, (xy - sy) * b.inv.y),
                                  __builtin_fmaxf((xz - sz) * b.inv.z, tmin));
  const float f = __builtin_fminf(__builtin_fminf((xx + sx) * b.inv.x, (xy + sy) * b.inv.y),
                                  __builtin_fminf((xz + sz) * b.inv.z, tmax));
  tnear = n;
  return n <= f * kBoxSlack;
}

// Child reference from the low bytes of one node row (v_perm_b32 gathers two bytes per
// instruction).
__device__ __forceinline__ int32_t row_child_ref(float4 r) {
  const uint32_t lo = __builtin_amdgcn_perm(__float_as_uint(r.y), __float_as_uint(r.x), 0x0c0c0400u);
  const uint32_t hi = __builtin_amdgcn_perm(__float_as_uint(r.w), __float_as_uint(r.z), 0x04000c0cu);
  return static_cast<int32_t>(lo | hi);
}

// One node at byte offset i (node references are byte offsets): three 16-byte loads,
// child references from the x / y values' low bytes.
__device__ __forceinline__ void load_node(const DevNode48* __restrict__ nodes, int32_t i, float4& r0, float4& r1,
                                          float4& r2, int4& c) {
  // a 32-bit offset from the (wave-uniform) array base: one add per visit and a load with
  // a scalar base, instead of 64-bit address arithmetic
  const float4* p = reinterpret_cast<const float4*>(reinterpret_cast<const char*>(nodes) + static_cast<uint32_t>(i));
  r0 = p[0];
  r1 = p[1];
  r2 = p[2];
  c = make_int4(row_child_ref(r0), row_child_ref(r1), 0, 0);
}

// BVH2 traversal of one bottom-level tree (or the single-level scene) for an
// (object-space) ray. Stack lives in LDS: stk[k * 64] for entry k. Returns true on the
// first hit when ANY. Updates tbest/pbest/ibest with the (t, inst, prim) order.
template <bool ANY, bool STATS>
__device__ __forceinline__ bool traverse_bvh(const DevNode48* __restrict__ nodes, const float4* __restrict__ tris,
                                             int32_t root, V3 o, V3 d, float tmin, uint32_t inst, uint32_t* stk,
                                             int sp0, int cap, TraceOut& r, uint32_t* overflow) {
  const RayTri rt = ray_tri_setup(o, d);
  const RayBox rb = ray_box_setup(o, d);
  int sp = sp0;
  int32_t node = root;
  for (;;) {
    if (node >= 0) {
      float4 r0, r1, r2;
      int4 c;
      load_node(nodes, node, r0, r1, r2, c);
      if (STATS) r.nodes++;
      float t0, t1;
      const bool h0 = slab_ch(r0.x, r0.y, r0.z, r0.w, r2.x, r2.y, rb, tmin, r.t, t0);
      const bool h1 = slab_ch(r1.x, r1.y, r1.z, r1.w, r2.z, r2.w, rb, tmin, r.t, t1);
      if (h0 && h1) {
        int32_t nearc = c.x, farc = c.y;
        if (t1 < t0) {
          nearc = c.y;
          farc = c.x;
        }
        if (sp < cap) {
          stk[sp * 64] = static_cast<uint32_t>(farc);
          ++sp;
        } else {
          atomicOr(overflow, 1u);
        }
        node = nearc;
        continue;
      }
      if (h0) {
        node = c.x;
        continue;
      }
      if (h1) {
        node = c.y;
        continue;
      }
    } else {
      const uint32_t code = ~static_cast<uint32_t>(node);
      const uint32_t first = code >> 3, cnt = code & 7u;
      for (uint32_t k = 0; k < cnt; ++k) {
        const float4* tp = tris + static_cast<size_t>(first + k) * 3u;
        const float4 a = tp[0], b = tp[1], cc = tp[2];
        if (STATS) r.tests++;
        TriHit th;
        if (intersect_triangle(rt, v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), v3(cc.x, cc.y, cc.z), tmin, r.t, th)) {
          const uint32_t prim = __float_as_uint(a.w);
          if (ANY) {
            r.t = th.t;
            r.prim = prim;
            r.inst = inst;
            return true;
          }
          if (th.t < r.t || inst < r.inst || (inst == r.inst && prim < r.prim)) {
            r.t = th.t;
            r.b1 = th.b1;
            r.b2 = th.b2;
            r.prim = prim;
            r.inst = inst;
          }
        }
      }
    }
    if (sp == sp0) break;
    --sp;
    node = static_cast<int32_t>(stk[sp * 64]);
  }
  return false;
}

// Sentinel of an exhausted traversal (never a leaf code: leaf first indices are < 2^28).
constexpr int32_t kTravDone = static_cast<int32_t>(0x80000000u);

// The while-while node loops end once at most kSearchBreak of the lanes still walking
// have no leaf parked, instead of none: the last few searching lanes would otherwise
// hold the wave for a long tail of iterations in which most lanes idle. They keep their
// state and go on walking in the next round. With per-lane leaf loops (traverse_ww), 4
// to 8 lanes traced 9% more rays per second than 0 on the 1M-triangle soup (16 slightly
// worse); the queue loops, whose leaf phase is spread over the wave and so costs less
// per round, measured best at 16 (soup: 8 -1.5%, 32 -6%; instanced: 8 -2%).
constexpr uint32_t kSearchBreak = 8;
constexpr uint32_t kSearchBreakSpread = 16;

// End of a queue traversal's node loop, evaluated by all lanes after each iteration as a
// wave-uniform branch (two compares and scalar work): no lane has an inner node left
// (walks), or at most kSearchBreakSpread of the lanes that walked in this iteration
// (walked) hold no parked leaf.
__device__ __forceinline__ bool walk_done(bool walked, bool walks, int32_t leaf) {
  const uint64_t searching = __ballot(walked) & __ballot(leaf == 0);
  // two 32-bit counts: a 64-bit compare of the count would be a vector instruction
  const uint32_t n = static_cast<uint32_t>(__builtin_popcount(static_cast<uint32_t>(searching))) +
                     static_cast<uint32_t>(__builtin_popcount(static_cast<uint32_t>(searching >> 32)));
  return __ballot(walks) == 0 || n <= kSearchBreakSpread;
}

// Lanes of the wave (among the active ones) that hold no parked leaf.
__device__ __forceinline__ uint32_t searching_lanes(int32_t leaf) {
  const uint64_t m = __ballot(leaf == 0);
  // two 32-bit counts: a 64-bit compare of the count would be a vector instruction
  return static_cast<uint32_t>(__builtin_popcount(static_cast<uint32_t>(m))) +
         static_cast<uint32_t>(__builtin_popcount(static_cast<uint32_t>(m >> 32)));
}

// Single-level BVH2 traversal in the "speculative while-while" form of Aila & Laine
// (2009, "Understanding the efficiency of ray traversal on GPUs"): lanes walk inner
// nodes until (nearly) every lane of the wave holds a leaf (a lane that meets its first
// leaf parks it and keeps walking), then the wave intersects leaves together. Triangle
// tests are the costly step; this keeps them from running for a handful of lanes at
// every iteration. The result does not depend on the visiting order (closest hit under
// the (t, inst, prim) order, or any hit), so it equals traverse_bvh's.
template <bool ANY, bool STATS>
__device__ __forceinline__ bool traverse_ww(const DevNode48* __restrict__ nodes, const float4* __restrict__ tris, V3 o,
                                            V3 d, float tmin, uint32_t* stk, int cap, TraceOut& r,
                                            uint32_t* overflow) {
  const RayTri rt = ray_tri_setup(o, d);
  const RayBox rb = ray_box_setup(o, d);
  int sp = 0;
  int32_t node = 0;  // next node: inner index, leaf code, or kTravDone
  int32_t leaf = 0;  // parked leaf code (0 = none)
  auto pop = [&]() -> int32_t {
    if (sp == 0) return kTravDone;
    --sp;
    return static_cast<int32_t>(stk[sp * 64]);
  };
  while (node != kTravDone || leaf != 0) {
    // inner nodes
    while (node >= 0) {
      float4 r0, r1, r2;
      int4 c;
      load_node(nodes, node, r0, r1, r2, c);
      if (STATS) r.nodes++;
      float t0, t1;
      const bool h0 = slab_ch(r0.x, r0.y, r0.z, r0.w, r2.x, r2.y, rb, tmin, r.t, t0);
      const bool h1 = slab_ch(r1.x, r1.y, r1.z, r1.w, r2.z, r2.w, rb, tmin, r.t, t1);
      if (h0 && h1) {
        int32_t nearc = c.x, farc = c.y;
        if (t1 < t0) {
          nearc = c.y;
          farc = c.x;
        }
        if (sp < cap) {
          stk[sp * 64] = static_cast<uint32_t>(farc);
          ++sp;
        } else {
          atomicOr(overflow, 1u);
        }
        node = nearc;
      } else if (h0) {
        node = c.x;
      } else if (h1) {
        node = c.y;
      } else {
        node = pop();
      }
      if (node < 0 && node != kTravDone && leaf == 0) {  // park the first leaf, keep walking
        leaf = node;
        node = pop();
      }
      if (searching_lanes(leaf) <= kSearchBreak) break;
    }
    // leaves
    while (leaf != 0) {
      const uint32_t code = ~static_cast<uint32_t>(leaf);
      const uint32_t first = code >> 3, cnt = code & 7u;
      for (uint32_t k = 0; k < cnt; ++k) {
        const float4* tp = tris + static_cast<size_t>(first + k) * 3u;
        const float4 a = tp[0], b = tp[1], cc = tp[2];
        if (STATS) r.tests++;
        TriHit th;
        if (intersect_triangle(rt, v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), v3(cc.x, cc.y, cc.z), tmin, r.t, th)) {
          const uint32_t prim = __float_as_uint(a.w);
          if (ANY) {
            r.t = th.t;
            r.prim = prim;
            r.inst = kNoInstance;
            return true;
          }
          if (th.t < r.t || prim < r.prim) {
            r.t = th.t;
            r.b1 = th.b1;
            r.b2 = th.b2;
            r.prim = prim;
            r.inst = kNoInstance;
          }
        }
      }
      if (node < 0 && node != kTravDone) {
        leaf = node;
        node = pop();
      } else {
        leaf = 0;
      }
    }
  }
  return false;
}

// Persistent single-level traversal of a segmented ray queue with lane refill
// ("dynamic fetch", Aila & Laine 2009): the wave runs rounds of the speculative
// while-while loop above, and before each round lanes whose ray has finished take new
// rays from the queue once at least kRefill of them are idle, so a few long rays no
// longer hold the whole wave (40 of 64 measured best on MI355X with the early end of
// the node loop and the spread leaf phase, 24 and 48 about 1-3% slower; 64, i.e. no
// refill, was 18% slower with per-lane leaf loops).
// load(slot, o, d, tmin, tmax) reads a ray;
// finish(slot, r, hit) consumes a result (hit: a hit was found).
constexpr uint32_t kRefill = 40;

// Traversal stack of the queue traversals: the top kRingEntries entries live in an
// LDS ring (entry k in slot k % kRingEntries, lane-interleaved), older ones are moved
// to the lane's area of a global spill buffer when the ring is full and brought back,
// a ring at a time, when it runs empty. Deep traversals are rare, so the ring keeps
// LDS small (4 KiB per wave, beside the 1.6 KiB leaf-task area) and occupancy bound by
// registers, not by the tree depth.
constexpr int kRingEntries = 16;

// LDS word at a byte address (ds_read / ds_write with a 32-bit address formed by integer
// arithmetic, so that a ring slot's address is one and-or of the lane's ring base and
// the slot offset).
typedef __attribute__((address_space(3))) uint32_t lds_u32;

__device__ __forceinline__ uint32_t lds_address(uint32_t* p) {
  return static_cast<uint32_t>(reinterpret_cast<uintptr_t>((lds_u32*)p));  // address-space cast
}
__device__ __forceinline__ uint32_t lds_load(uint32_t a) { return *reinterpret_cast<lds_u32*>(static_cast<uintptr_t>(a)); }
__device__ __forceinline__ void lds_store(uint32_t a, uint32_t v) {
  *reinterpret_cast<lds_u32*>(static_cast<uintptr_t>(a)) = v;
}

// Entry counts are kept in units of kSlotBytes (the ring's slot stride: 64 lanes x 4 bytes),
// so that entry k's ring slot is at ring | (k * kSlotBytes & kRingMask). The wave's ring is
// 4 KiB aligned in LDS (the traversal kernels have no static LDS: dynamic LDS starts at 0,
// checked at start-up by traversal_lds_check), so the slot bits never meet the lane's base.
constexpr uint32_t kSlotBytes = 64 * 4;
constexpr uint32_t kRingMask = (kRingEntries - 1) * kSlotBytes;

struct TwoTierStack {
  uint32_t ring;         // LDS byte address of the lane's first ring slot
  uint32_t* spill_wave;  // the wave's spill area; lane l's `cap` entries start at l * cap
  uint32_t cap;          // capacity (entries)
  uint32_t sp;           // entries on the stack, x kSlotBytes
  uint32_t lo;           // entries [0, lo) are in the spill area, [lo, sp) in the ring (x kSlotBytes)
  uint32_t lim;          // min(lo + kRingEntries, cap) x kSlotBytes: a push at sp == lim either moves
                         // the ring's oldest entry out or, at the capacity, replaces the top entry
};

__device__ __forceinline__ void stack_reset(TwoTierStack& s) {
  s.sp = 0;
  s.lo = 0;
  s.lim = min(static_cast<uint32_t>(kRingEntries), s.cap) * kSlotBytes;
}

// The wave's spill area of the traversal grid (stack_cap entries per thread), taken
// wave-uniform so that it stays in scalar registers; a lane's own area is formed only
// on the rare ring overflow or underflow.
__device__ __forceinline__ uint32_t* wave_spill(uint32_t* spill, int cap) {
  const uint32_t wave_first = __builtin_amdgcn_readfirstlane(threadIdx.x) & ~63u;
  return spill + static_cast<size_t>(blockIdx.x * kBlock + wave_first) * static_cast<uint32_t>(cap);
}

__device__ __forceinline__ uint32_t* lane_spill(const TwoTierStack& s) { return s.spill_wave + lane_id() * s.cap; }

// Pushes v. At the capacity (a tree deeper than the stack was sized for) v replaces the
// top entry, which keeps every access in bounds, and on_full() reports the inexact
// traversal.
template <bool STATS, class OnFull>
__device__ __forceinline__ void stack_push(TwoTierStack& s, uint32_t v, TraceOut& stats, OnFull on_full) {
  if (s.sp == s.lim) {
    if (s.sp == s.cap * kSlotBytes) {
      on_full();
      s.sp -= kSlotBytes;
    } else {  // ring full: move its oldest entry out
      lane_spill(s)[s.lo / kSlotBytes] = lds_load(s.ring | (s.lo & kRingMask));
      s.lo += kSlotBytes;
      s.lim = min(s.lo + kRingEntries * kSlotBytes, s.cap * kSlotBytes);
      if (STATS) stats.spills++;
    }
  }
  lds_store(s.ring | (s.sp & kRingMask), v);
  s.sp += kSlotBytes;
}

__device__ __forceinline__ int32_t stack_pop(TwoTierStack& s) {
  if (s.sp == s.lo) {
    if (s.sp == 0) return kTravDone;
    // ring empty: bring back up to a ring of spilled entries
    const uint32_t lo = s.lo / kSlotBytes;
    const uint32_t n = lo < kRingEntries ? lo : kRingEntries;
    const uint32_t* sp = lane_spill(s);
    for (uint32_t k = lo - n; k < lo; ++k) lds_store(s.ring | ((k * kSlotBytes) & kRingMask), sp[k]);
    s.lo -= n * kSlotBytes;
    s.lim = min(s.lo + kRingEntries * kSlotBytes, s.cap * kSlotBytes);
  }
  s.sp -= kSlotBytes;
  return static_cast<int32_t>(lds_load(s.ring | (s.sp & kRingMask)));
}

// Leaf phase of the single-level loop, spread over the whole wave: every (ray, triangle)
// pair of the leaves the lanes hold is a task, and the wave deals the tasks out 64 at a
// time, each lane testing one with the owning lane's ray (read across lanes), instead of
// each lane looping over its own leaves while most lanes wait (a round's leaves hold
// ~120 triangles, which the per-lane loop tested at ~16 active lanes per iteration).
// An owner's closest hit is the minimum of a (t, prim) key over its tasks (LDS atomic
// min; the lexicographic order of the closest-hit search), and the task that holds the
// minimum writes the barycentrics. Per wave, in LDS: one window of kTaskWin task
// triangles and owners per round, and a key and barycentrics per lane.
constexpr uint32_t kTaskWin = 128;
constexpr uint32_t kTaskWords = kTaskWin + kTaskWin / 4 + 64 * 2 + 64 * 2;

struct LeafTasks {
  uint32_t* tri;            // [kTaskWin] triangle index
  uint8_t* own;             // [kTaskWin] owning lane
  unsigned long long* key;  // [64] hit key per lane
  float2* bary;             // [64] barycentrics of the key's hit
};

__device__ __forceinline__ LeafTasks leaf_tasks(uint32_t* base) {
  LeafTasks t;
  t.tri = base;
  t.own = reinterpret_cast<uint8_t*>(base + kTaskWin);
  t.key = reinterpret_cast<unsigned long long*>(base + kTaskWin + kTaskWin / 4);
  t.bary = reinterpret_cast<float2*>(base + kTaskWin + kTaskWin / 4 + 128);
  return t;
}

// Order-preserving (t, prim) key: t's bits mapped so that unsigned order is float
// order (-0 taken as +0), prim below.
__device__ __forceinline__ unsigned long long hit_key(float t, uint32_t prim) {
  uint32_t u = __float_as_uint(t + 0.f);
  u ^= (u >> 31) ? 0xffffffffu : 0x80000000u;
  return (static_cast<unsigned long long>(u) << 32) | prim;
}

__device__ __forceinline__ float key_t(unsigned long long k) {
  uint32_t u = static_cast<uint32_t>(k >> 32);
  u ^= (u >> 31) ? 0x80000000u : 0xffffffffu;
  return __uint_as_float(u);
}

// One leaf phase: lane l contributes the tasks of ca triangles from fa and cb from fb,
// tested with its ray (object-space origin o, ray_tri_setup's packed axes and shear,
// tmin, tmax). A task hit counts if it beats the lane's best (t = tmax, prim0, and for
// TWO, instance inst0, with the lane's current instance cur). Returns whether the
// lane's tasks were in this round's window (false: it keeps its leaves for the next
// round); `better`: a task hit counted, the smallest (t, prim) key of those is kf, and
// its barycentrics are in tk.bary[lane].
template <bool STATS, bool TWO>
__device__ __forceinline__ bool leaf_phase(const LeafTasks& tk, const float4* __restrict__ tris, uint32_t lane,
                                           uint32_t fa, uint32_t ca, uint32_t fb, uint32_t cb, V3 o, int32_t kxyz,
                                           V3 shear, float tmin, float tmax, uint32_t prim0, uint32_t inst0,
                                           uint32_t cur, bool& better, unsigned long long& kf, TraceOut& stats,
                                           uint32_t* touch_tris) {
  const uint32_t n_tasks = ca + cb;  // at most 14
  uint32_t off = 0, total = 0;       // exclusive prefix and sum of n_tasks over the wave
  for (uint32_t bit = 0; bit < 4; ++bit) {
    const uint64_t m = __ballot((n_tasks >> bit) & 1u);
    off += prefix_count(m) << bit;
    total += static_cast<uint32_t>(__popcll(m)) << bit;
  }
  better = false;
  kf = 0;
  if (total == 0) return true;
  // one window of kTaskWin tasks per round: lanes whose tasks do not fit keep their
  // leaves for the next round (the first lane with tasks always fits)
  const bool fits = off + n_tasks <= kTaskWin;
  const uint64_t late = __ballot(!fits);
  const uint32_t m = late != 0 ? __shfl(off, static_cast<int>(__builtin_ctzll(late))) : total;
  // TWO: every hit up to tmax is a candidate, the owner's order is checked per hit
  const unsigned long long key0 = hit_key(tmax, TWO ? kNoHit : prim0);
  tk.key[lane] = key0;
  if (fits) {
    for (uint32_t k = 0; k < n_tasks; ++k) {
      tk.tri[off + k] = k < ca ? fa + k : fb + (k - ca);
      tk.own[off + k] = static_cast<uint8_t>(lane);
    }
  }
  for (uint32_t s0 = 0; s0 < m; s0 += 64) {
    const uint32_t j = s0 + lane;
    const bool act = j < m;
    const uint32_t ti = act ? tk.tri[j] : 0u;
    const int own = act ? static_cast<int>(tk.own[j]) : static_cast<int>(lane);
    RayTri ort;
    ort.o = v3(__shfl(o.x, own), __shfl(o.y, own), __shfl(o.z, own));
    const int kk = __shfl(kxyz, own);
    ort.kx = kk & 3;
    ort.ky = (kk >> 2) & 3;
    ort.kz = kk >> 4;
    ort.sx = __shfl(shear.x, own);
    ort.sy = __shfl(shear.y, own);
    ort.sz = __shfl(shear.z, own);
    const float otmin = __shfl(tmin, own), otmax = __shfl(tmax, own);
    uint32_t oprim = 0, oinst = 0, ocur = 0;
    if (TWO) {
      oprim = __shfl(prim0, own);
      oinst = __shfl(inst0, own);
      ocur = __shfl(cur, own);
    }
    if (act) {
      const float4* tp = tris + static_cast<size_t>(ti) * 3u;
      const float4 a = tp[0], b = tp[1], cc = tp[2];
      if (STATS) {
        stats.tests++;
        touch(touch_tris, ti);
      }
      TriHit th;
      const uint32_t prim = __float_as_uint(a.w);
      if (intersect_triangle(ort, v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), v3(cc.x, cc.y, cc.z), otmin, otmax, th) &&
          (!TWO || th.t < otmax || ocur < oinst || (ocur == oinst && prim < oprim))) {
        const unsigned long long key = hit_key(th.t, prim);
        atomicMin(&tk.key[own], key);
        if (tk.key[own] == key) tk.bary[own] = make_float2(th.b1, th.b2);
      }
    }
  }
  if (fits && n_tasks != 0) {
    kf = tk.key[lane];
    better = kf < key0;
  }
  return fits;
}

__device__ __forceinline__ int32_t pack_axes(const RayTri& t) { return t.kx | (t.ky << 2) | (t.kz << 4); }

template <bool ANY, bool STATS, class Load, class Finish>
__device__ __forceinline__ void trace_queue(const DevAccel& acc, const SegIn& in, uint32_t* stk, uint32_t* tsk,
                                            uint32_t* overflow, TraceOut& stats, Load load, Finish finish) {
  const uint32_t lane = lane_id();
  const uint32_t n_seg = seg_counts(in), home = block_home();
  const int cap = acc.stack_cap;
  const DevNode48* __restrict__ nodes = acc.nodes;
  const float4* __restrict__ tris = acc.tris;
  uint64_t seg_done = seg_empty(n_seg, home, gridDim.x >= kSegs);
  bool drained = false;
  uint32_t slot = kNoHit;
  V3 o = v3(0.f), d = v3(0.f);
  RayBox rb{};
  float tmin = 0.f;
  int32_t kxyz = 0;        // ray_tri_setup's axis permutation, 2 bits per axis
  V3 shear = v3(0.f);      // and its shear constants
  TraceOut r;
  r.t = 0.f;
  r.b1 = r.b2 = 0.f;
  r.prim = kNoHit;
  r.inst = kNoInstance;
  TwoTierStack st;
  st.ring = lds_address(stk);
  st.spill_wave = wave_spill(acc.spill, cap);
  st.cap = static_cast<uint32_t>(cap);
  stack_reset(st);
  int32_t node = kTravDone, leaf = 0;
  auto pop = [&]() -> int32_t { return stack_pop(st); };
  const LeafTasks tk = leaf_tasks(tsk);
  for (;;) {
    const uint64_t idle = __ballot(slot == kNoHit);
    const uint32_t n_idle = static_cast<uint32_t>(__popcll(idle));
    if (!drained && n_idle >= kRefill) {
      uint32_t base = 0, got = 0;
      if (seg_fetch_k(in, n_seg, home, seg_done, n_idle, base, got)) {
        const uint32_t rank = prefix_count(idle);
        if (slot == kNoHit && rank < got) {
          slot = base + rank;
          float tmax;
          load(slot, o, d, tmin, tmax);
          const RayTri rt = ray_tri_setup(o, d);
          kxyz = pack_axes(rt);
          shear = v3(rt.sx, rt.sy, rt.sz);
          rb = ray_box_setup(o, d);
          r.t = tmax;
          r.prim = kNoHit;
          r.inst = kNoInstance;
          stack_reset(st);
          node = 0;
          leaf = 0;
        }
      } else {
        drained = true;
      }
    }
    if (!__any(slot != kNoHit)) break;  // drained and every lane idle
    // inner nodes: a wave-uniform loop in which lanes without an inner node idle
    for (bool walks = node >= 0;;) {
      if (walks) {
        float4 r0, r1, r2;
        int4 c;
        load_node(nodes, node, r0, r1, r2, c);
        if (STATS) {
          stats.nodes++;
          touch(acc.touch_nodes, static_cast<uint32_t>(node) / kNodeBytes);
        }
        float t0, t1;
        const bool h0 = slab_ch(r0.x, r0.y, r0.z, r0.w, r2.x, r2.y, rb, tmin, r.t, t0);
        const bool h1 = slab_ch(r1.x, r1.y, r1.z, r1.w, r2.z, r2.w, rb, tmin, r.t, t1);
        if (h0 && h1) {
          int32_t nearc = c.x, farc = c.y;
          if (t1 < t0) {
            nearc = c.y;
            farc = c.x;
          }
          stack_push<STATS>(st, static_cast<uint32_t>(farc), stats, [&] { atomicOr(overflow, 1u); });
          node = nearc;
        } else if (h0) {
          node = c.x;
        } else if (h1) {
          node = c.y;
        } else {
          node = pop();
        }
        if (node < 0 && node != kTravDone && leaf == 0) {
          leaf = node;
          node = pop();
        }
      }
      const bool walked = walks;
      walks = node >= 0;
      if (walk_done(walked, walks, leaf)) break;
    }
    // leaves (the parked one, and `node` when it is a leaf), spread over the wave
    bool hit = false;
    const bool node_leaf = node < 0 && node != kTravDone;
    const uint32_t ca = leaf != 0 ? (~static_cast<uint32_t>(leaf) & 7u) : 0u;
    const uint32_t cb = node_leaf ? (~static_cast<uint32_t>(node) & 7u) : 0u;
    bool better;
    unsigned long long kf;
    if (leaf_phase<STATS, false>(tk, tris, lane, ~static_cast<uint32_t>(leaf) >> 3, ca, ~static_cast<uint32_t>(node) >> 3,
                                 cb, rb.o, kxyz, shear, tmin, r.t, r.prim, kNoInstance, kNoInstance, better, kf,
                                 stats, acc.touch_tris) &&
        ca + cb != 0) {
      if (better) {  // the barycentrics stay in tk.bary[lane] until the ray finishes
        r.t = key_t(kf);
        r.prim = static_cast<uint32_t>(kf);
        hit = ANY;
      }
      leaf = 0;
      if (node_leaf) node = pop();
    }
    if (slot != kNoHit && (hit || (node == kTravDone && leaf == 0))) {
      r.b1 = r.b2 = 0.f;
      if (r.prim != kNoHit) {
        const float2 bb = tk.bary[lane];
        r.b1 = bb.x;
        r.b2 = bb.y;
      }
      finish(slot, r, hit || r.prim != kNoHit);
      slot = kNoHit;
      node = kTravDone;
      leaf = 0;
    }
  }
}

// Two-level counterpart of trace_queue (instanced scenes): the same speculative
// while-while loop with lane refill, two-tier stack and wave-wide leaf phase, walking the
// top-level tree over instance boxes and, inside an instance, its mesh's tree with the
// ray in object space (leaf tasks use the owner's object-space ray).
// Entering an instance pushes what is left of the top-level leaf and an exit marker;
// popping the marker returns the ray to world space. A lane with a parked (not yet
// intersected) leaf never leaves its instance: it waits at the marker until the leaf
// is done, since the leaf's triangles are tested in that instance's space.
constexpr int32_t kTravExit = 0x7fffffff;                               // leave the instance (stack marker)
constexpr int32_t kTravBlocked = static_cast<int32_t>(0x80000001u);     // waits at the marker with a parked leaf

template <bool ANY, bool STATS, class Load, class Finish>
__device__ __forceinline__ void trace_queue2(const DevAccel& acc, const SegIn& in, uint32_t* stk, uint32_t* tsk,
                                             uint32_t* overflow, TraceOut& stats, Load load, Finish finish) {
  const uint32_t lane = lane_id();
  const uint32_t n_seg = seg_counts(in), home = wave_home();
  const int cap = acc.stack_cap;
  const DevNode48* __restrict__ blas = acc.nodes;
  const DevNode48* __restrict__ tlas = acc.tlas;
  const float4* __restrict__ tris = acc.tris;
  uint64_t seg_done = seg_empty(n_seg);
  bool drained = false;
  uint32_t slot = kNoHit;
  V3 ow = v3(0.f), dw = v3(0.f);
  RayBox rb{};  // in the current space (world outside instances)
  int32_t kxyz = 0;    // ray_tri_setup's axis permutation in the current instance, 2 bits per axis
  V3 shear = v3(0.f);  // and its shear constants
  float tmin = 0.f;
  TraceOut r;
  r.t = 0.f;
  r.b1 = r.b2 = 0.f;
  r.prim = kNoHit;
  r.inst = kNoInstance;
  uint32_t cur_id = kNoInstance;
  bool in_blas = false;
  TwoTierStack st;
  st.ring = lds_address(stk);
  st.spill_wave = wave_spill(acc.spill, cap);
  st.cap = static_cast<uint32_t>(cap);
  stack_reset(st);
  int32_t node = kTravDone, leaf = 0;
  bool ok = true;  // false after a stack overflow (reported once per lane)
  auto push = [&](uint32_t v) {
    stack_push<STATS>(st, v, stats, [&] {
      if (ok) {
        ok = false;
        atomicOr(overflow, 1u);
      }
    });
  };
  // Pops the next node; leaves the instance at its marker unless a leaf is parked.
  auto pop = [&]() -> int32_t {
    for (;;) {
      const int32_t e = stack_pop(st);
      if (e != kTravExit) return e;
      if (leaf != 0) {
        stack_push<STATS>(st, static_cast<uint32_t>(kTravExit), stats, [] {});  // just popped: never full
        return kTravBlocked;
      }
      in_blas = false;
      rb = ray_box_setup(ow, dw);  // back in world space (recomputed: fewer live registers)
    }
  };
  // Top-level leaf (instances [first, first + count)): enter the first instance.
  auto enter = [&](int32_t code) -> int32_t {
    const uint32_t c = ~static_cast<uint32_t>(code);
    const uint32_t first = c >> 3, cnt = c & 7u;
    if (cnt > 1u) push(~(((first + 1u) << 3) | (cnt - 1u)));  // the rest of the leaf
    push(static_cast<uint32_t>(kTravExit));
    const Instance& ins = acc.inst[first];
    const V3 o = xform_point(ins.w2o, ow);
    const V3 d = xform_vector(ins.w2o, dw);
    rb = ray_box_setup(o, d);
    const RayTri rt = ray_tri_setup(o, d);
    kxyz = pack_axes(rt);
    shear = v3(rt.sx, rt.sy, rt.sz);
    cur_id = ins.id;
    in_blas = true;
    if (STATS) stats.nodes++;
    return static_cast<int32_t>(ins.blas_root);
  };
  // Resolves top-level leaves (entering instances) until node is an inner node, a
  // mesh leaf, kTravBlocked or kTravDone.
  auto settle = [&](int32_t n) -> int32_t {
    while (!in_blas && n < 0 && n != kTravDone && n != kTravBlocked) n = enter(n);
    return n;
  };
  const LeafTasks tk = leaf_tasks(tsk);
  for (;;) {
    const uint64_t idle = __ballot(slot == kNoHit);
    const uint32_t n_idle = static_cast<uint32_t>(__popcll(idle));
    if (!drained && n_idle >= kRefill) {
      uint32_t base = 0, got = 0;
      if (seg_fetch_k(in, n_seg, home, seg_done, n_idle, base, got)) {
        const uint32_t rank = prefix_count(idle);
        if (slot == kNoHit && rank < got) {
          slot = base + rank;
          float tmax;
          load(slot, ow, dw, tmin, tmax);
          rb = ray_box_setup(ow, dw);
          in_blas = false;
          r.t = tmax;
          r.prim = kNoHit;
          r.inst = kNoInstance;
          stack_reset(st);
          node = 0;
          leaf = 0;
        }
      } else {
        drained = true;
      }
    }
    if (!__any(slot != kNoHit)) break;
    if (slot != kNoHit) {
      // inner nodes of either level (a while loop: the uniform form of the single-level
      // loop costs this kernel register spills)
      while (node >= 0) {
        float4 r0, r1, r2;
        int4 c;
        load_node(in_blas ? blas : tlas, node, r0, r1, r2, c);
        if (STATS) {
          stats.nodes++;
          if (in_blas) touch(acc.touch_nodes, static_cast<uint32_t>(node) / kNodeBytes);
        }
        float t0, t1;
        const bool h0 = slab_ch(r0.x, r0.y, r0.z, r0.w, r2.x, r2.y, rb, tmin, r.t, t0);
        const bool h1 = slab_ch(r1.x, r1.y, r1.z, r1.w, r2.z, r2.w, rb, tmin, r.t, t1);
        if (h0 && h1) {
          int32_t nearc = c.x, farc = c.y;
          if (t1 < t0) {
            nearc = c.y;
            farc = c.x;
          }
          push(static_cast<uint32_t>(farc));
          node = nearc;
        } else if (h0) {
          node = c.x;
        } else if (h1) {
          node = c.y;
        } else {
          node = pop();
        }
        node = settle(node);
        if (in_blas && node < 0 && node != kTravDone && node != kTravBlocked && leaf == 0) {
          leaf = node;
          node = settle(pop());
        }
        if (searching_lanes(leaf) <= kSearchBreakSpread) break;
      }
    }
    // mesh leaves (the parked one, and `node` when it is one; both in the current
    // instance, which a lane with a parked leaf does not leave), spread over the wave.
    // Task hits that beat the ray's best under the (t, inst, prim) order compete; the
    // phase's (t, prim) minimum of those becomes the best.
    bool hit = false;
    const bool node_leaf = in_blas && node < 0 && node != kTravDone && node != kTravBlocked;
    const uint32_t ca = leaf != 0 ? (~static_cast<uint32_t>(leaf) & 7u) : 0u;
    const uint32_t cb = node_leaf ? (~static_cast<uint32_t>(node) & 7u) : 0u;
    bool better;
    unsigned long long kf;
    if (leaf_phase<STATS, true>(tk, tris, lane, ~static_cast<uint32_t>(leaf) >> 3, ca, ~static_cast<uint32_t>(node) >> 3,
                                cb, rb.o, kxyz, shear, tmin, r.t, r.prim, r.inst, cur_id, better, kf, stats,
                                acc.touch_tris) &&
        ca + cb != 0) {
      if (better) {  // the barycentrics stay in tk.bary[lane] until the ray finishes
        r.t = key_t(kf);
        r.prim = static_cast<uint32_t>(kf);
        r.inst = cur_id;
        hit = ANY;
      }
      leaf = 0;
      if (node_leaf) node = settle(pop());
      if (node == kTravBlocked) node = settle(pop());  // the leaf is done: leave the instance
    }
    if (slot != kNoHit && (hit || (node == kTravDone && leaf == 0))) {
      r.b1 = r.b2 = 0.f;
      if (r.prim != kNoHit) {
        const float2 bb = tk.bary[lane];
        r.b1 = bb.x;
        r.b2 = bb.y;
      }
      finish(slot, r, hit || r.prim != kNoHit);
      slot = kNoHit;
      node = kTravDone;
      leaf = 0;
    }
  }
}

// Full trace (single or two-level). r.t must hold tmax on entry.
template <bool ANY, bool STATS>
__device__ __forceinline__ bool trace_ray(const DevAccel& acc, V3 o, V3 d, float tmin, uint32_t* stk,
                                          TraceOut& r, uint32_t* overflow) {
  if (!acc.two_level) return traverse_ww<ANY, STATS>(acc.nodes, acc.tris, o, d, tmin, stk, acc.stack_cap, r, overflow);
  // top level: leaf ranges index the instance array in TLAS build order
  const RayBox rb = ray_box_setup(o, d);
  int sp = 0;
  int32_t node = 0;
  const int cap = acc.stack_cap;
  for (;;) {
    if (node >= 0) {
      float4 r0, r1, r2;
      int4 c;
      load_node(acc.tlas, node, r0, r1, r2, c);
      if (STATS) r.nodes++;
      float t0, t1;
      const bool h0 = slab_ch(r0.x, r0.y, r0.z, r0.w, r2.x, r2.y, rb, tmin, r.t, t0);
      const bool h1 = slab_ch(r1.x, r1.y, r1.z, r1.w, r2.z, r2.w, rb, tmin, r.t, t1);
      if (h0 && h1) {
        int32_t nearc = c.x, farc = c.y;
        if (t1 < t0) {
          nearc = c.y;
          farc = c.x;
        }
        if (sp < cap) {
          stk[sp * 64] = static_cast<uint32_t>(farc);
          ++sp;
        } else {
          atomicOr(overflow, 1u);
        }
        node = nearc;
        continue;
      }
      if (h0) {
        node = c.x;
        continue;
      }
      if (h1) {
        node = c.y;
        continue;
      }
    } else {
      const uint32_t code = ~static_cast<uint32_t>(node);
      const uint32_t first = code >> 3, cnt = code & 7u;
      for (uint32_t k = 0; k < cnt; ++k) {
        const Instance& in = acc.inst[first + k];
        const V3 oo = xform_point(in.w2o, o);
        const V3 od = xform_vector(in.w2o, d);
        if (traverse_bvh<ANY, STATS>(acc.nodes, acc.tris, static_cast<int32_t>(in.blas_root), oo, od, tmin, in.id,
                                     stk, sp, cap, r, overflow))
          return true;
      }
    }
    if (sp == 0) break;
    --sp;
    node = static_cast<int32_t>(stk[sp * 64]);
  }
  return false;
}

// ------------------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_zero_counters(uint32_t* ctr) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i < static_cast<uint32_t>(kNumCounters)) ctr[i] = 0u;
}

// Totals per depth of the extension and shadow queues (lane = depth, two waves).
__global__ __launch_bounds__(64) void k_sum_counts(uint32_t* ctr, uint32_t cap) {
  const uint32_t t = threadIdx.x;
  if (t >= 2u * kMaxDepth) return;
  const uint32_t kind = t < kMaxDepth ? kSegExt : kSegShadow, d = t % kMaxDepth;
  const uint32_t* c = ctr + seg_ctr(static_cast<int>(kind), static_cast<int>(d));
  uint32_t total = 0;
  for (uint32_t sgm = 0; sgm < kSegs; ++sgm) {
    const uint32_t v = c[sgm * kLine];
    total += v < cap ? v : cap;
  }
  ctr[(kind == kSegExt ? kCtrExt : kCtrShadow) + d] = total;
}

// Distinct items a launch read: set bits of its touch bitmap, which is cleared for the
// next launch.
__global__ __launch_bounds__(kBlock) void k_touch_count(uint32_t* bits, uint32_t words, unsigned long long* out) {
  uint32_t n = 0;
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < words; i += gridDim.x * kBlock) {
    n += static_cast<uint32_t>(__popc(bits[i]));
    bits[i] = 0u;
  }
  wave_add_stat(out, n);
}

// Primary-ray slot of path p: segment p / per, in order (no atomics).
__device__ __forceinline__ uint32_t primary_slot(uint32_t p, uint32_t per, uint32_t cap) {
  return (p / per) * cap + p % per;
}

// Writes the primary-queue counts (one lane per segment).
__device__ __forceinline__ void primary_counts(uint32_t n_paths, uint32_t per, uint32_t* cnt) {
  if (blockIdx.x == 0 && threadIdx.x < kSegs) {
    const uint32_t lo = threadIdx.x * per;
    cnt[threadIdx.x * kLine] = n_paths > lo ? (n_paths - lo < per ? n_paths - lo : per) : 0u;
  }
}

// Stores a primary ray; a camera sample with zero weight is stored with tmax = -1, so
// the traversal rejects it at the root and Shade skips it.
__device__ __forceinline__ void store_primary(const RayQueue& q, uint32_t s, uint32_t p, const PathState& ps,
                                              bool alive, float eps) {
  q.o_tmin[s] = make_float4(ps.o.x, ps.o.y, ps.o.z, eps);
  q.d_tmax[s] = make_float4(ps.d.x, ps.d.y, ps.d.z, alive ? 1e30f : -1.f);
  q.t_pdf[s] = make_float4(ps.T.x, ps.T.y, ps.T.z, ps.pdf);
  q.meta[s] = (p << 2) | ps.flags;
}

// Path whose primary ray RayGen thread t writes to slot t's place in the queue: the
// frame's paths sample by sample, each sample's pixels in 8x8 tiles (row-major tiles over
// the largest region with sides a multiple of 8, then the remaining bottom rows and right
// columns row by row; a bijection on [0, W*H)), so that the 64 rays of a fetched batch
// are a square of neighbouring pixels rather than 64 pixels of a row: the lanes of a wave
// share more nodes on every load. Results are per path, whatever the queue order.
__device__ __forceinline__ uint32_t tiled_path(uint32_t t, uint32_t W, uint32_t H) {
  const uint32_t npix = W * H;
  const uint32_t smp = t / npix, j = t - smp * npix;
  const uint32_t W8 = W & ~7u, H8 = H & ~7u, core = W8 * H8;
  uint32_t x, y;
  if (j < core) {
    const uint32_t tile = j >> 6, k = j & 63u, tiles_x = W8 >> 3;
    x = (tile % tiles_x) * 8u + (k & 7u);
    y = (tile / tiles_x) * 8u + (k >> 3);
  } else {
    uint32_t r = j - core;
    const uint32_t bottom = (H - H8) * W;
    if (r < bottom) {
      y = H8 + r / W;
      x = r % W;
    } else {  // reached only when W8 < W
      r -= bottom;
      const uint32_t wr = W - W8;
      y = r / wr;
      x = W8 + r % wr;
    }
  }
  return smp * npix + y * W + x;
}

__global__ __launch_bounds__(kBlock) void k_raygen(RenderParams rp, uint32_t n_paths, uint32_t per, uint32_t cap,
                                                    RayQueue q, FrameBuffers fb, uint32_t* ctr) {
  const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
  primary_counts(n_paths, per, ctr + seg_ctr(kSegExt, 0));
  if (t < n_paths) {
    const uint32_t p = tiled_path(t, rp.width, rp.height);
    PathState ps;
    const bool alive = raygen(rp, p, ps);
    fb.radiance[p] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (!alive) {  // Shade never sees this path: its first-hit AOVs are zero
      fb.path_aov_a[p] = make_float4(0.f, 0.f, 0.f, 0.f);
      fb.path_aov_n[p] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    store_primary(q, primary_slot(t, per, cap), p, ps, alive, rp.eps);
  }
}


// Single-level scenes: Extend and Shadow through trace_queue. Both are held to 72
// VGPRs, 7 waves per SIMD (the allocator otherwise settles just above, at 6 waves,
// which measured 5% slower).
template <bool STATS>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(7))) void k_extend_sl(DevAccel acc, RayQueue q, HitBuf h, SegIn in, uint32_t* ctr) {
  extern __shared__ uint32_t lds[];
  uint32_t* stk = lds + (threadIdx.x >> 6) * kRingEntries * 64 + lane_id();
  uint32_t* tsk = lds + kWavesPerBlock * kRingEntries * 64 + (threadIdx.x >> 6) * kTaskWords;
  TraceOut st;
  st.nodes = 0;
  st.tests = 0;
  st.spills = 0;
  trace_queue<false, STATS>(
      acc, in, stk, tsk, &ctr[kCtrOverflow], st,
      [&](uint32_t slot, V3& o, V3& d, float& tmin, float& tmax) {
        const float4 a = q.o_tmin[slot];
        const float4 b = q.d_tmax[slot];
        o = v3(a.x, a.y, a.z);
        d = v3(b.x, b.y, b.z);
        tmin = a.w;
        tmax = b.w;
      },
      [&](uint32_t slot, const TraceOut& r, bool) {
        h.thit[slot] = make_float4(r.t, r.b1, r.b2, __uint_as_float(r.prim));
      });
  if (STATS) {
    wave_add_stat(reinterpret_cast<unsigned long long*>(&ctr[kCtrStats + 0]), st.nodes);
    wave_add_stat(reinterpret_cast<unsigned long long*>(&ctr[kCtrStats + 2]), st.tests);
    wave_add_stat(reinterpret_cast<unsigned long long*>(&ctr[kCtrStats + 8]), st.spills);
  }
}

template <bool STATS>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(7))) void k_shadow_sl(DevAccel acc, ShadowQueue sq, float4* radiance, SegIn in,
                                                       uint32_t* ctr) {
  extern __shared__ uint32_t lds[];
  uint32_t* stk = lds + (threadIdx.x >> 6) * kRingEntries * 64 + lane_id();
  uint32_t* tsk = lds + kWavesPerBlock * kRingEntries * 64 + (threadIdx.x >> 6) * kTaskWords;
  TraceOut st;
  st.nodes = 0;
  st.tests = 0;
  st.spills = 0;
  trace_queue<true, STATS>(
      acc, in, stk, tsk, &ctr[kCtrOverflow], st,
      [&](uint32_t slot, V3& o, V3& d, float& tmin, float& tmax) {
        const float4 a = sq.o_tmin[slot];
        const float4 b = sq.d_tmax[slot];
        o = v3(a.x, a.y, a.z);
        d = v3(b.x, b.y, b.z);
        tmin = a.w;
        tmax = b.w;
      },
      [&](uint32_t slot, const TraceOut&, bool occluded) {
        if (!occluded) {
          const float4 c = sq.contrib[slot];
          const uint32_t path = __float_as_uint(c.w);
          float4 L = radiance[path];
          L.x += c.x;
          L.y += c.y;
          L.z += c.z;
          radiance[path] = L;
        }
      });
  if (STATS) {
    wave_add_stat(reinterpret_cast<unsigned long long*>(&ctr[kCtrStats + 4]), st.nodes);
    wave_add_stat(reinterpret_cast<unsigned long long*>(&ctr[kCtrStats + 6]), st.tests);
    wave_add_stat(reinterpret_cast<unsigned long long*>(&ctr[kCtrStats + 10]), st.spills);
  }
}

// Two-level scenes: Extend and Shadow through trace_queue2, held to 80 VGPRs, 6 waves
// per SIMD (a few spilled dwords; 5 waves at the allocator's 86 measured 5% slower).
template <bool STATS>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(6))) void k_extend_tl(DevAccel acc, RayQueue q, HitBuf h, SegIn in, uint32_t* ctr) {
  extern __shared__ uint32_t lds[];
  uint32_t* stk = lds + (threadIdx.x >> 6) * kRingEntries * 64 + lane_id();
  uint32_t* tsk = lds + kWavesPerBlock * kRingEntries * 64 + (threadIdx.x >> 6) * kTaskWords;
  TraceOut st;
  st.nodes = 0;
  st.tests = 0;
  st.spills = 0;
  trace_queue2<false, STATS>(
      acc, in, stk, tsk, &ctr[kCtrOverflow], st,
      [&](uint32_t slot, V3& o, V3& d, float& tmin, float& tmax) {
        const float4 a = q.o_tmin[slot];
        const float4 b = q.d_tmax[slot];
        o = v3(a.x, a.y, a.z);
        d = v3(b.x, b.y, b.z);
        tmin = a.w;
        tmax = b.w;
      },
      [&](uint32_t slot, const TraceOut& r, bool) {
        h.thit[slot] = make_float4(r.t, r.b1, r.b2, __uint_as_float(r.prim));
        h.inst[slot] = r.inst;
      });
  if (STATS) {
    wave_add_stat(reinterpret_cast<unsigned long long*>(&ctr[kCtrStats + 0]), st.nodes);
    wave_add_stat(reinterpret_cast<unsigned long long*>(&ctr[kCtrStats + 2]), st.tests);
    wave_add_stat(reinterpret_cast<unsigned long long*>(&ctr[kCtrStats + 8]), st.spills);
  }
}

template <bool STATS>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(6))) void k_shadow_tl(DevAccel acc, ShadowQueue sq, float4* radiance, SegIn in,
                                                       uint32_t* ctr) {
  extern __shared__ uint32_t lds[];
  uint32_t* stk = lds + (threadIdx.x >> 6) * kRingEntries * 64 + lane_id();
  uint32_t* tsk = lds + kWavesPerBlock * kRingEntries * 64 + (threadIdx.x >> 6) * kTaskWords;
  TraceOut st;
  st.nodes = 0;
  st.tests = 0;
  st.spills = 0;
  trace_queue2<true, STATS>(
      acc, in, stk, tsk, &ctr[kCtrOverflow], st,
      [&](uint32_t slot, V3& o, V3& d, float& tmin, float& tmax) {
        const float4 a = sq.o_tmin[slot];
        const float4 b = sq.d_tmax[slot];
        o = v3(a.x, a.y, a.z);
        d = v3(b.x, b.y, b.z);
        tmin = a.w;
        tmax = b.w;
      },
      [&](uint32_t slot, const TraceOut&, bool occluded) {
        if (!occluded) {
          const float4 c = sq.contrib[slot];
          const uint32_t path = __float_as_uint(c.w);
          float4 L = radiance[path];
          L.x += c.x;
          L.y += c.y;
          L.z += c.z;
          radiance[path] = L;
        }
      });
  if (STATS) {
    wave_add_stat(reinterpret_cast<unsigned long long*>(&ctr[kCtrStats + 4]), st.nodes);
    wave_add_stat(reinterpret_cast<unsigned long long*>(&ctr[kCtrStats + 6]), st.tests);
    wave_add_stat(reinterpret_cast<unsigned long long*>(&ctr[kCtrStats + 10]), st.spills);
  }
}


// Per-wave LDS list of Shade's pending work: queue slots whose ray needs shading, with
// the ray's (direction, tmax) and hit records already read when the slot was fetched.
constexpr uint32_t kShadeList = 128;
struct ShadeList {
  uint32_t slot[kShadeList];
  float4 dir[kShadeList];
  float4 hit[kShadeList];
};

// DEBUG: the render modes of render_modes.h instead of the path-tracing step.
// Slots are fetched 64 at a time; a ray that left the scene adds nothing when the sky
// is off (path tracing), so only the slots that hit are listed (a miss at depth 0 still
// writes its pixel's zero albedo and normal), and the wave shades 64 listed slots at a
// time: in the 1M soup 18% of the rays miss (35% at the second bounce), lanes that the
// per-slot loop left idle through the whole shading step. Held to 3 waves per SIMD (150
// VGPRs with one material set; the allocator could otherwise take more than 170, 2 waves;
// 4 waves spill 20 bytes per lane and measured 1% slower).
// SEP: the sampling materials are a set of their own (1) or the evaluation materials (0),
// fixed per launch from RenderParams::separate_sampling.
template <bool DEBUG, int SEP>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3))) void k_shade(SceneView sv, const Material* sampling_mats, RenderParams rp,
                                                   RayQueue q, HitBuf h, RayQueue qn, ShadowQueue sq,
                                                   FrameBuffers fb, SegIn in, uint32_t* next_cnt,
                                                   uint32_t* shadow_cnt, uint32_t* emit_cnt, uint32_t* ctr,
                                                   uint32_t depth, uint32_t two_level,
                                                   const SkyParams* __restrict__ sky) {
  // sv.sky again, as a restrict kernel argument: nothing the kernel writes can alias it, so
  // its fields are read with scalar loads where they are used
  sv.sky = sky;
  __shared__ ShadeList lists[kWavesPerBlock];
  ShadeList& L = lists[threadIdx.x >> 6];
  const uint32_t lane = lane_id();
  const uint32_t npix = rp.width * rp.height;
  const uint32_t n_seg = seg_counts(in), home = block_home();
  const bool keep_misses = DEBUG || sv.sky->enabled != 0u;
  uint64_t done = seg_empty(n_seg, home, gridDim.x >= kSegs);
  uint32_t pend = 0;  // listed slots (wave-uniform)
  bool more = true;
  for (;;) {
    while (more && pend < 64u) {
      uint32_t slot0, n_valid;
      more = seg_fetch(in, n_seg, home, done, slot0, n_valid);
      if (!more) break;
      const uint32_t slot = slot0 + lane;
      float4 b = make_float4(0.f, 0.f, 0.f, -1.f), hh = make_float4(0.f, 0.f, 0.f, 0.f);
      if (lane < n_valid) {
        b = q.d_tmax[slot];
        hh = h.thit[slot];
      }
      // a dead primary ray has tmax < 0 and is skipped
      const bool live = b.w >= 0.f;
      const bool work = live && (keep_misses || __float_as_uint(hh.w) != kNoHit);
      if (live && !work && depth == 0) {
        const uint32_t path = q.meta[slot] >> 2;
        fb.path_aov_a[path] = make_float4(0.f, 0.f, 0.f, 0.f);
        fb.path_aov_n[path] = make_float4(0.f, 0.f, 0.f, 0.f);
      }
      const uint64_t m = __ballot(work);
      if (work) {
        const uint32_t k = pend + prefix_count(m);
        L.slot[k] = slot;
        L.dir[k] = b;
        L.hit[k] = hh;
      }
      pend += static_cast<uint32_t>(__popcll(m));
    }
    if (pend == 0u) break;  // drained
    const uint32_t take = pend < 64u ? pend : 64u;
    ShadeOut out;
    out.has_ext = 0;
    out.has_shadow = 0;
    uint32_t path = 0;
    bool emits = false;
    if (lane < take) {
      const uint32_t slot = L.slot[lane];
      const float4 a = q.o_tmin[slot];
      const float4 b = L.dir[lane];
      const float4 tp = q.t_pdf[slot];
      const uint32_t meta = q.meta[slot];
      const float4 hh = L.hit[lane];
      const uint32_t hinst = two_level ? h.inst[slot] : kNoInstance;
      PathState ps;
      ps.o = v3(a.x, a.y, a.z);
      ps.d = v3(b.x, b.y, b.z);
      ps.T = v3(tp.x, tp.y, tp.z);
      ps.pdf = tp.w;
      ps.flags = meta & 3u;
      ps.path = meta >> 2;
      path = ps.path;
      Hit hit;
      hit.t = hh.x;
      hit.b1 = hh.y;
      hit.b2 = hh.z;
      hit.prim = __float_as_uint(hh.w);
      hit.inst = hinst;
      if (DEBUG) shade_debug(sv, rp, ps, hit, depth, out);
      else shade<SEP>(sv, sampling_mats, rp, ps, hit, depth, out);
      emits = !is_black(out.L);
      if (depth == 0) {
        fb.path_aov_a[path] = make_float4(out.albedo.x, out.albedo.y, out.albedo.z, out.vdepth);
        fb.path_aov_n[path] = make_float4(out.normal.x, out.normal.y, out.normal.z, 0.f);
      }
    }
    // emission (emitter hits, the sky) to the emission list: the path radiance belongs to
    // the previous bounce's shadow launch, which may still run
    const uint64_t mL = __ballot(emits);
    if (mL != 0ull) {
      const uint32_t sl = seg_append(emit_cnt, in.cap, home, mL, &ctr[kCtrOverflow]);
      if (emits && sl != kNoHit) sq.emit[sl] = make_float4(out.L.x, out.L.y, out.L.z, __uint_as_float(path));
    }
    // move the unprocessed tail (at most 63 slots, all at or above `take`) to the front
    if (take == 64u && lane + 64u < pend) {
      L.slot[lane] = L.slot[lane + 64u];
      L.dir[lane] = L.dir[lane + 64u];
      L.hit[lane] = L.hit[lane + 64u];
    }
    pend -= take;
    const bool ext = out.has_ext != 0;
    const bool shw = out.has_shadow != 0;
    const uint64_t me = __ballot(ext);
    const uint64_t ms = __ballot(shw);
    const uint32_t se = seg_append(next_cnt, in.cap, home, me, &ctr[kCtrOverflow]);
    const uint32_t ss = seg_append(shadow_cnt, in.cap, home, ms, &ctr[kCtrOverflow]);
    if (ext && se != kNoHit) {
      const uint32_t s = se;
      qn.o_tmin[s] = make_float4(out.next.o.x, out.next.o.y, out.next.o.z, rp.eps);
      qn.d_tmax[s] = make_float4(out.next.d.x, out.next.d.y, out.next.d.z, 1e30f);
      qn.t_pdf[s] = make_float4(out.next.T.x, out.next.T.y, out.next.T.z, out.next.pdf);
      qn.meta[s] = (path << 2) | out.next.flags;
    }
    if (shw && ss != kNoHit) {
      const uint32_t s = ss;
      sq.o_tmin[s] = make_float4(out.s_o.x, out.s_o.y, out.s_o.z, rp.eps);
      sq.d_tmax[s] = make_float4(out.s_d.x, out.s_d.y, out.s_d.z, out.s_tmax);
      sq.contrib[s] = make_float4(out.s_contrib.x, out.s_contrib.y, out.s_contrib.z, __uint_as_float(path));
    }
  }
}

// A bounce's emission list into the path radiance. It runs on the shadow stream after
// the previous bounce's shadow rays and before this bounce's, so every path's radiance
// is summed in vertex order: emission at vertex d, then the next-event estimate of d.
// Block (x, y) takes every gridDim.x-th group of kBlock entries of segment y.
__global__ __launch_bounds__(kBlock) void k_emit(const float4* __restrict__ emit, const uint32_t* __restrict__ cnt,
                                                 uint32_t cap, float4* radiance) {
  const uint32_t sgm = blockIdx.y;
  const uint32_t c = cnt[sgm * kLine];
  const uint32_t n = c < cap ? c : cap;
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const float4 e = emit[size_t(sgm) * cap + i];
    const uint32_t path = __float_as_uint(e.w);
    float4 L = radiance[path];
    L.x += e.x;
    L.y += e.y;
    L.z += e.z;
    radiance[path] = L;
  }
}

__device__ __forceinline__ float4 add4(float4 a, float4 b) {
  return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
}

// Per pixel, in sample order: the frame's radiance into the accumulator (mean colour),
// and its first-hit AOVs into their sums (means over the accumulated samples: the
// denoiser's guide images, anti-aliased like the colour).
__global__ __launch_bounds__(kBlock) void k_finalize(RenderParams rp, const float4* radiance, FrameBuffers fb,
                                                      uint32_t reset, uint32_t aov_reset, uint32_t aov_count) {
  const uint32_t pix = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t npix = rp.width * rp.height;
  if (pix >= npix) return;
  float sx = 0.f, sy = 0.f, sz = 0.f;
  float4 sa = make_float4(0.f, 0.f, 0.f, 0.f), sn = make_float4(0.f, 0.f, 0.f, 0.f);
  for (uint32_t s = 0; s < rp.spp; ++s) {
    const size_t i = static_cast<size_t>(s) * npix + pix;
    const float4 L = radiance[i];
    sx += L.x;
    sy += L.y;
    sz += L.z;
    sa = add4(sa, fb.path_aov_a[i]);
    sn = add4(sn, fb.path_aov_n[i]);
  }
  float4 acc = reset ? make_float4(0.f, 0.f, 0.f, 0.f) : fb.accum[pix];
  acc.x += sx;
  acc.y += sy;
  acc.z += sz;
  acc.w += static_cast<float>(rp.spp);
  fb.accum[pix] = acc;
  const float inv = acc.w > 0.f ? 1.f / acc.w : 0.f;
  fb.color[pix] = make_float4(acc.x * inv, acc.y * inv, acc.z * inv, 1.f);
  if (!aov_reset) {
    sa = add4(fb.aov_sum_a[pix], sa);
    sn = add4(fb.aov_sum_n[pix], sn);
  }
  fb.aov_sum_a[pix] = sa;
  fb.aov_sum_n[pix] = sn;
  const float ia = 1.f / static_cast<float>(aov_count);
  fb.albedo[pix] = make_float4(sa.x * ia, sa.y * ia, sa.z * ia, 1.f);
  fb.normal[pix] = make_float4(sn.x * ia, sn.y * ia, sn.z * ia, sa.w * ia);
}

__global__ __launch_bounds__(kBlock) void k_tonemap(const float4* src, uint32_t n, TonemapParams tp,
                                                     uchar4* dst) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float4 c = src[i];
  dst[i] = make_uchar4(static_cast<unsigned char>(tonemap_u8(tp, c.x)), static_cast<unsigned char>(tonemap_u8(tp, c.y)),
                       static_cast<unsigned char>(tonemap_u8(tp, c.z)), 255);
}

// ------------------------------------------------------------------------------------
// Adjoint (path-replay backpropagation). Gradient records live in fixed slots
// r = path * D + depth, so their values never depend on queue order; the reduction
// below sums them in slot order.
__global__ __launch_bounds__(kBlock) void k_raygen_adj(RenderParams rp, uint32_t n_paths, uint32_t per, uint32_t cap,
                                                        RayQueue q, const float4* radiance, AdjBuffers ab,
                                                        uint32_t* ctr) {
  const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
  primary_counts(n_paths, per, ctr + seg_ctr(kSegExt, 0));
  if (p < n_paths) {
    PathState ps;
    const bool alive = raygen(rp, p, ps);
    ab.lrem[p] = radiance[p];
    for (uint32_t d = 0; d < ab.depth; ++d) ab.rec_key[static_cast<size_t>(p) * ab.depth + d] = kNoHit;
    ab.pend[static_cast<size_t>(p) * 4 + 3] = make_float4(0.f, 0.f, 0.f, __uint_as_float(kNoHit));
    store_primary(q, primary_slot(p, per, cap), p, ps, alive, rp.eps);
  }
}

__global__ __launch_bounds__(kBlock) void k_shade_adj(SceneView sv, const Material* sampling_mats, RenderParams rp,
                                                       RayQueue q, HitBuf h, RayQueue qn, ShadowQueue sq,
                                                       AdjBuffers ab, SegIn in, uint32_t* next_cnt,
                                                       uint32_t* shadow_cnt, uint32_t* ctr, uint32_t depth,
                                                       uint32_t two_level, const SkyParams* __restrict__ sky) {
  sv.sky = sky;  // scalar loads of the sky's fields, as in k_shade
  const uint32_t lane = lane_id();
  const uint32_t npix = rp.width * rp.height;
  const float inv_spp = 1.f / static_cast<float>(rp.spp);
  const size_t R = ab.n_records;
  const uint32_t n_seg = seg_counts(in), home = wave_home();
  uint64_t done = seg_empty(n_seg);
  uint32_t slot0, n_valid;
  while (seg_fetch(in, n_seg, home, done, slot0, n_valid)) {
    const uint32_t slot = slot0 + lane;
    ShadeOut out;
    AdjOut adj;
    out.has_ext = 0;
    out.has_shadow = 0;
    uint32_t path = 0;
    const float4 b = lane < n_valid ? q.d_tmax[slot] : make_float4(0.f, 0.f, 0.f, -1.f);
    if (b.w >= 0.f) {
      const float4 a = q.o_tmin[slot];
      const float4 tp = q.t_pdf[slot];
      const uint32_t meta = q.meta[slot];
      const float4 hh = h.thit[slot];
      PathState ps;
      ps.o = v3(a.x, a.y, a.z);
      ps.d = v3(b.x, b.y, b.z);
      ps.T = v3(tp.x, tp.y, tp.z);
      ps.pdf = tp.w;
      ps.flags = meta & 3u;
      ps.path = meta >> 2;
      path = ps.path;
      const uint32_t pix = path % npix;
      const V3 delta = v3(ab.dldi[pix * 3 + 0] * inv_spp, ab.dldi[pix * 3 + 1] * inv_spp, ab.dldi[pix * 3 + 2] * inv_spp);
      float4 lr = ab.lrem[path];
      // continuation gradient of the previous vertex: the radiance still unaccounted
      // for arrived through its sampled direction
      if (depth > 0) {
        const float4* pp = ab.pend + static_cast<size_t>(path) * 4;
        const float4 p0 = pp[0], p1 = pp[1], p2 = pp[2], p3 = pp[3];
        if (__float_as_uint(p3.w) != kNoHit) {
          AdjOut prev;
          prev.r[0][0] = p0.x; prev.r[0][1] = p0.y; prev.r[0][2] = p0.z;
          prev.r[1][0] = p0.w; prev.r[1][1] = p1.x; prev.r[1][2] = p1.y;
          prev.r[2][0] = p1.z; prev.r[2][1] = p1.w; prev.r[2][2] = p2.x;
          prev.r[3][0] = p2.y; prev.r[3][1] = p2.z; prev.r[3][2] = p2.w;
          prev.r[4][0] = p3.x; prev.r[4][1] = p3.y; prev.r[4][2] = p3.z;
          float g[kNumParams];
          continuation_grad(prev, delta, v3(lr.x, lr.y, lr.z), g);
          const size_t r = static_cast<size_t>(path) * ab.depth + (depth - 1);
          for (int k = 0; k < kNumParams; ++k) ab.rec_g[k * R + r] += g[k];
        }
      }
      Hit hit;
      hit.t = hh.x;
      hit.b1 = hh.y;
      hit.b2 = hh.z;
      hit.prim = __float_as_uint(hh.w);
      hit.inst = two_level ? h.inst[slot] : kNoInstance;
      shade_adjoint(sv, sampling_mats, rp, ps, hit, depth, delta, out, adj);
      if (!is_black(out.L)) {
        lr.x -= out.L.x;
        lr.y -= out.L.y;
        lr.z -= out.L.z;
        ab.lrem[path] = lr;
      }
      const size_t r = static_cast<size_t>(path) * ab.depth + depth;
      if (adj.mat != kNoHit) {
        ab.rec_key[r] = adj.mat;
        for (int k = 0; k < kNumParams; ++k) ab.rec_g[k * R + r] = 0.f;
      }
      float4* pp = ab.pend + static_cast<size_t>(path) * 4;
      if (out.has_ext) {
        pp[0] = make_float4(adj.r[0][0], adj.r[0][1], adj.r[0][2], adj.r[1][0]);
        pp[1] = make_float4(adj.r[1][1], adj.r[1][2], adj.r[2][0], adj.r[2][1]);
        pp[2] = make_float4(adj.r[2][2], adj.r[3][0], adj.r[3][1], adj.r[3][2]);
        pp[3] = make_float4(adj.r[4][0], adj.r[4][1], adj.r[4][2], __uint_as_float(adj.mat));
      } else {
        pp[3] = make_float4(0.f, 0.f, 0.f, __uint_as_float(kNoHit));
      }
    }
    const bool ext = out.has_ext != 0;
    const bool shw = out.has_shadow != 0;
    const uint64_t me = __ballot(ext);
    const uint64_t ms = __ballot(shw);
    const uint32_t se = seg_append(next_cnt, in.cap, home, me, &ctr[kCtrOverflow]);
    const uint32_t ss = seg_append(shadow_cnt, in.cap, home, ms, &ctr[kCtrOverflow]);
    if (ext && se != kNoHit) {
      const uint32_t s = se;
      qn.o_tmin[s] = make_float4(out.next.o.x, out.next.o.y, out.next.o.z, rp.eps);
      qn.d_tmax[s] = make_float4(out.next.d.x, out.next.d.y, out.next.d.z, 1e30f);
      qn.t_pdf[s] = make_float4(out.next.T.x, out.next.T.y, out.next.T.z, out.next.pdf);
      qn.meta[s] = (path << 2) | out.next.flags;
    }
    if (shw && ss != kNoHit) {
      const uint32_t s = ss;
      sq.o_tmin[s] = make_float4(out.s_o.x, out.s_o.y, out.s_o.z, rp.eps);
      sq.d_tmax[s] = make_float4(out.s_d.x, out.s_d.y, out.s_d.z, out.s_tmax);
      sq.contrib[s] = make_float4(out.s_contrib.x, out.s_contrib.y, out.s_contrib.z, __uint_as_float(path));
      ab.s_ga[s] = make_float4(adj.nee_g[0], adj.nee_g[1], adj.nee_g[2], adj.nee_g[3]);
      ab.s_gb[s] = adj.nee_g[4];
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_shadow_adj(DevAccel acc, ShadowQueue sq, AdjBuffers ab, SegIn in,
                                                        uint32_t* ctr, uint32_t depth) {
  extern __shared__ uint32_t lds[];
  const uint32_t lane = lane_id();
  uint32_t* stk = lds + (threadIdx.x >> 6) * acc.stack_cap * 64 + lane;
  const size_t R = ab.n_records;
  const uint32_t n_seg = seg_counts(in), home = wave_home();
  uint64_t done = seg_empty(n_seg);
  uint32_t slot0, n_valid;
  while (seg_fetch(in, n_seg, home, done, slot0, n_valid)) {
    const uint32_t slot = slot0 + lane;
    if (lane < n_valid) {
      const float4 a = sq.o_tmin[slot];
      const float4 b = sq.d_tmax[slot];
      TraceOut r;
      r.t = b.w;
      r.prim = kNoHit;
      r.inst = kNoInstance;
      r.nodes = 0;
      r.tests = 0;
      const bool occ =
          trace_ray<true, false>(acc, v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), a.w, stk, r, &ctr[kCtrOverflow]);
      if (!occ) {
        const float4 c = sq.contrib[slot];
        const uint32_t path = __float_as_uint(c.w);
        float4 lr = ab.lrem[path];
        lr.x -= c.x;
        lr.y -= c.y;
        lr.z -= c.z;
        ab.lrem[path] = lr;
        const float4 ga = ab.s_ga[slot];
        const float gb = ab.s_gb[slot];
        const size_t rr = static_cast<size_t>(path) * ab.depth + depth;
        ab.rec_g[0 * R + rr] += ga.x;
        ab.rec_g[1 * R + rr] += ga.y;
        ab.rec_g[2 * R + rr] += ga.z;
        ab.rec_g[3 * R + rr] += ga.w;
        ab.rec_g[4 * R + rr] += gb;
      }
    }
  }
}

// Deterministic reduction of records into per-material gradients, materials
// [mat_lo, mat_lo + n_mat). Block b owns records [b R/B, (b+1) R/B); each of its waves
// a contiguous quarter; each wave adds its records in slot order into a private LDS
// accumulator (one lane per parameter), then the block adds its four accumulators in
// wave order. No float atomics anywhere: results are bitwise reproducible.
__global__ __launch_bounds__(kBlock) void k_grad_reduce(const uint32_t* rec_key, const float* rec_g, uint64_t R,
                                                         uint32_t mat_lo, uint32_t n_mat, float* partial) {
  extern __shared__ float acc_lds[];
  const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
  const uint32_t P = n_mat * kNumParams;
  float* acc = acc_lds + wave * P;
  for (uint32_t i = lane; i < P; i += 64) acc[i] = 0.f;
  const uint64_t b0 = R * blockIdx.x / gridDim.x, b1 = R * (blockIdx.x + 1) / gridDim.x;
  const uint64_t w0 = b0 + (b1 - b0) * wave / kWavesPerBlock, w1 = b0 + (b1 - b0) * (wave + 1) / kWavesPerBlock;
  for (uint64_t base = w0; base < w1; base += 64) {
    const uint64_t r = base + lane;
    uint32_t key = kNoHit;
    float g0 = 0.f, g1 = 0.f, g2 = 0.f, g3 = 0.f, g4 = 0.f;
    if (r < w1) {
      key = rec_key[r];
      if (key >= mat_lo && key - mat_lo < n_mat) {
        g0 = rec_g[0 * R + r];
        g1 = rec_g[1 * R + r];
        g2 = rec_g[2 * R + r];
        g3 = rec_g[3 * R + r];
        g4 = rec_g[4 * R + r];
      } else {
        key = kNoHit;
      }
    }
    uint64_t mask = __ballot(key != kNoHit);
    while (mask) {
      const int j = __builtin_ctzll(mask);
      mask &= mask - 1;
      const uint32_t kj = __shfl(key, j);
      const float v0 = __shfl(g0, j), v1 = __shfl(g1, j), v2 = __shfl(g2, j), v3v = __shfl(g3, j), v4 = __shfl(g4, j);
      if (lane < kNumParams) {
        const float v = lane == 0 ? v0 : lane == 1 ? v1 : lane == 2 ? v2 : lane == 3 ? v3v : v4;
        acc[(kj - mat_lo) * kNumParams + lane] += v;
      }
    }
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < P; i += kBlock) {
    float t = acc_lds[i];
    for (int w = 1; w < kWavesPerBlock; ++w) t += acc_lds[w * P + i];
    partial[static_cast<size_t>(blockIdx.x) * P + i] = t;
  }
}

__global__ __launch_bounds__(kBlock) void k_grad_final(const float* partial, uint32_t n_blocks, uint32_t P,
                                                        float* out) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= P) return;
  float t = 0.f;
  for (uint32_t b = 0; b < n_blocks; ++b) t += partial[static_cast<size_t>(b) * P + i];
  out[i] = t;
}

size_t stack_lds_bytes(const DevAccel& acc) { return static_cast<size_t>(acc.stack_cap) * 64u * 4u * kWavesPerBlock; }
size_t ring_lds_bytes() { return static_cast<size_t>(kRingEntries) * 64u * 4u * kWavesPerBlock; }
// queue traversals: rings, then each wave's leaf-task area
size_t trace_lds_bytes() { return ring_lds_bytes() + static_cast<size_t>(kTaskWords) * 4u * kWavesPerBlock; }

}  // namespace

int trace_blocks_per_cu(const DevAccel& acc) {
  int a = 0, b = 0;
  hipError_t ea, eb;
  if (acc.two_level) {
    ea = hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, k_extend_tl<false>, kBlock, trace_lds_bytes());
    eb = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_shadow_tl<false>, kBlock, trace_lds_bytes());
  } else {
    ea = hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, k_extend_sl<false>, kBlock, trace_lds_bytes());
    eb = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_shadow_sl<false>, kBlock, trace_lds_bytes());
  }
  if (ea != hipSuccess) a = 4;
  if (eb != hipSuccess) b = 4;
  const int m = a < b ? a : b;
  return m > 0 ? m : 1;
}

bool traversal_lds_check() {
  // The queue traversals address their LDS rings as 4 KiB-aligned blocks from LDS
  // address 0: they must have no static LDS, which would move dynamic LDS up.
  const void* ks[] = {reinterpret_cast<const void*>(k_extend_sl<false>), reinterpret_cast<const void*>(k_extend_sl<true>),
                      reinterpret_cast<const void*>(k_shadow_sl<false>), reinterpret_cast<const void*>(k_shadow_sl<true>),
                      reinterpret_cast<const void*>(k_extend_tl<false>), reinterpret_cast<const void*>(k_extend_tl<true>),
                      reinterpret_cast<const void*>(k_shadow_tl<false>), reinterpret_cast<const void*>(k_shadow_tl<true>)};
  for (const void* k : ks) {
    hipFuncAttributes a{};
    if (hipFuncGetAttributes(&a, k) != hipSuccess || a.sharedSizeBytes != 0) return false;
  }
  return true;
}

void launch_zero_counters(uint32_t* ctr, hipStream_t stream) {
  hipLaunchKernelGGL(k_zero_counters, dim3((kNumCounters + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, ctr);
}

void launch_tonemap(const float4* src, uint32_t n, const TonemapParams& tp, uchar4* dst, hipStream_t stream) {
  hipLaunchKernelGGL(k_tonemap, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, src, n, tp, dst);
}

void launch_sum_counts(uint32_t* ctr, uint32_t cap, hipStream_t stream) {
  hipLaunchKernelGGL(k_sum_counts, dim3(1), dim3(64), 0, stream, ctr, cap);
}

void launch_touch_count(uint32_t* bits, uint32_t words, unsigned long long* out, hipStream_t stream) {
  const uint32_t blocks = (words + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(k_touch_count, dim3(blocks < 1024 ? (blocks ? blocks : 1) : 1024), dim3(kBlock), 0, stream, bits,
                     words, out);
}

void launch_raygen(const RenderParams& rp, uint32_t n_paths, uint32_t cap, RayQueue q, const FrameBuffers& fb,
                   uint32_t* ctr, hipStream_t stream) {
  const uint32_t blocks = (n_paths + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(k_raygen, dim3(blocks), dim3(kBlock), 0, stream, rp, n_paths, prim_per_seg(n_paths), cap, q,
                     fb, ctr);
}

void launch_extend(const DevAccel& acc, RayQueue q, HitBuf h, SegIn in, uint32_t* ctr, int grid, bool stats,
                   hipStream_t stream) {
  if (!acc.two_level) {
    if (stats)
      hipLaunchKernelGGL(k_extend_sl<true>, dim3(grid), dim3(kBlock), trace_lds_bytes(), stream, acc, q, h, in, ctr);
    else
      hipLaunchKernelGGL(k_extend_sl<false>, dim3(grid), dim3(kBlock), trace_lds_bytes(), stream, acc, q, h, in,
                         ctr);
    return;
  }
  if (stats)
    hipLaunchKernelGGL(k_extend_tl<true>, dim3(grid), dim3(kBlock), trace_lds_bytes(), stream, acc, q, h, in, ctr);
  else
    hipLaunchKernelGGL(k_extend_tl<false>, dim3(grid), dim3(kBlock), trace_lds_bytes(), stream, acc, q, h, in, ctr);
}

void launch_shade(const SceneView& sv, const Material* sampling_mats, const RenderParams& rp, RayQueue q,
                  HitBuf h, RayQueue qn, ShadowQueue sq, FrameBuffers fb, SegIn in, uint32_t* next_cnt,
                  uint32_t* shadow_cnt, uint32_t* emit_cnt, uint32_t* ctr, uint32_t depth, int grid,
                  hipStream_t stream) {
  if (rp.mode != kModePathTracing)
    hipLaunchKernelGGL((k_shade<true, -1>), dim3(grid), dim3(kBlock), 0, stream, sv, sampling_mats, rp, q, h, qn, sq,
                       fb, in, next_cnt, shadow_cnt, emit_cnt, ctr, depth, sv.instances ? 1u : 0u, sv.sky);
  else if (rp.separate_sampling)
    hipLaunchKernelGGL((k_shade<false, 1>), dim3(grid), dim3(kBlock), 0, stream, sv, sampling_mats, rp, q, h, qn, sq,
                       fb, in, next_cnt, shadow_cnt, emit_cnt, ctr, depth, sv.instances ? 1u : 0u, sv.sky);
  else
    hipLaunchKernelGGL((k_shade<false, 0>), dim3(grid), dim3(kBlock), 0, stream, sv, sampling_mats, rp, q, h, qn, sq,
                       fb, in, next_cnt, shadow_cnt, emit_cnt, ctr, depth, sv.instances ? 1u : 0u, sv.sky);
}

void launch_emit(const float4* emit, const uint32_t* emit_cnt, uint32_t cap, float4* radiance, hipStream_t stream) {
  // 16 blocks per segment: most lists are short (emitter hits); sky scenes list a third
  // of the paths
  hipLaunchKernelGGL(k_emit, dim3(16, kSegs), dim3(kBlock), 0, stream, emit, emit_cnt, cap, radiance);
}

void launch_shadow(const DevAccel& acc, ShadowQueue sq, float4* radiance, SegIn in, uint32_t* ctr, int grid,
                   bool stats, hipStream_t stream) {
  if (!acc.two_level) {
    if (stats)
      hipLaunchKernelGGL(k_shadow_sl<true>, dim3(grid), dim3(kBlock), trace_lds_bytes(), stream, acc, sq, radiance,
                         in, ctr);
    else
      hipLaunchKernelGGL(k_shadow_sl<false>, dim3(grid), dim3(kBlock), trace_lds_bytes(), stream, acc, sq, radiance,
                         in, ctr);
    return;
  }
  if (stats)
    hipLaunchKernelGGL(k_shadow_tl<true>, dim3(grid), dim3(kBlock), trace_lds_bytes(), stream, acc, sq, radiance, in,
                       ctr);
  else
    hipLaunchKernelGGL(k_shadow_tl<false>, dim3(grid), dim3(kBlock), trace_lds_bytes(), stream, acc, sq, radiance, in,
                       ctr);
}

void launch_finalize(const RenderParams& rp, const float4* radiance, FrameBuffers fb, uint32_t reset,
                     uint32_t aov_reset, uint32_t aov_count, hipStream_t stream) {
  const uint32_t npix = rp.width * rp.height;
  hipLaunchKernelGGL(k_finalize, dim3((npix + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, rp, radiance, fb,
                     reset, aov_reset, aov_count);
}


void launch_raygen_adj(const RenderParams& rp, uint32_t n_paths, uint32_t cap, RayQueue q, const float4* radiance,
                       const AdjBuffers& ab, uint32_t* ctr, hipStream_t stream) {
  const uint32_t blocks = (n_paths + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(k_raygen_adj, dim3(blocks), dim3(kBlock), 0, stream, rp, n_paths, prim_per_seg(n_paths), cap, q,
                     radiance, ab, ctr);
}

void launch_shade_adj(const SceneView& sv, const Material* sampling_mats, const RenderParams& rp, RayQueue q,
                      HitBuf h, RayQueue qn, ShadowQueue sq, const AdjBuffers& ab, SegIn in, uint32_t* next_cnt,
                      uint32_t* shadow_cnt, uint32_t* ctr, uint32_t depth, int grid, hipStream_t stream) {
  hipLaunchKernelGGL(k_shade_adj, dim3(grid), dim3(kBlock), 0, stream, sv, sampling_mats, rp, q, h, qn, sq, ab, in,
                     next_cnt, shadow_cnt, ctr, depth, sv.instances ? 1u : 0u, sv.sky);
}

void launch_shadow_adj(const DevAccel& acc, ShadowQueue sq, const AdjBuffers& ab, SegIn in, uint32_t* ctr,
                       uint32_t depth, int grid, hipStream_t stream) {
  hipLaunchKernelGGL(k_shadow_adj, dim3(grid), dim3(kBlock), stack_lds_bytes(acc), stream, acc, sq, ab, in, ctr,
                     depth);
}

void launch_grad_reduce(const AdjBuffers& ab, uint32_t mat_lo, uint32_t n_mat, float* partial, float* out,
                        hipStream_t stream) {
  const uint32_t P = n_mat * kNumParams;
  hipLaunchKernelGGL(k_grad_reduce, dim3(kGradBlocks), dim3(kBlock), P * 4 * kWavesPerBlock, stream, ab.rec_key,
                     ab.rec_g, static_cast<uint64_t>(ab.n_records), mat_lo, n_mat, partial);
  hipLaunchKernelGGL(k_grad_final, dim3((P + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, partial,
                     static_cast<uint32_t>(kGradBlocks), P, out);
}

}  // namespace trt
