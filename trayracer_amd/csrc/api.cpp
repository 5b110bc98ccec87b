// C-ABI implementation: context, scene upload, BVH build, frame orchestration.
//
// Replaces the reference's Renderer / DeviceRenderer / OptixRenderer orchestration
// (Tracer/Renderer/Renderer.cpp:79-190, 442-504; Tracer/Renderer/DeviceRenderer.cpp:97-274;
// Tracer/Optix/OptixRenderer.cpp:90-131). Differences by design: all device state is
// owned by the context (no __constant__ pointer globals, CudaGlobals.h:17-38), one HIP
// stream per context, no host synchronisation between bounces (queue sizes stay on the
// device; the reference downloads counters every bounce, DeviceRenderer.cpp:256), and
// errors are returned as trt_status instead of exit() (Tracer/Utility/Errors.cpp:15-37).
#include <hip/hip_runtime.h>
#include <sched.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <mutex>
#include <random>
#include <chrono>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/trt.h"
#include "api_internal.h"
#include "bvh_build.h"
#include "denoiser.h"
#include "device.h"
#include "trt/lights_build.h"

using namespace trt;

namespace {

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  hipError_t ensure(size_t n) {
    if (n <= bytes && p) return hipSuccess;
    release();
    if (n == 0) return hipSuccess;
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) {
      (void)hipGetLastError();  // reported here: later launch checks must not see it again
      p = nullptr;
      return e;
    }
    bytes = n;
    return hipSuccess;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

struct Mesh {
  uint32_t tri_offset, n_tris;
};

}  // namespace

struct trt_context {
  int device = 0;
  hipStream_t stream = nullptr;
  int num_cus = 0;
  std::string err;

  // host scene
  std::vector<float> pos, nrm;
  std::vector<uint8_t> has_nrm;  // per mesh
  std::vector<uint32_t> mat_ids;
  std::vector<Mesh> meshes;
  std::vector<Material> materials;
  std::vector<trt_instance> instances_in;
  bool dirty = true;
  bool built = false;

  // device scene
  DevBuf d_pos, d_nrm, d_mat_ids, d_materials, d_inst, d_inst_tlas, d_lights, d_alias_p, d_alias_i;
  DevBuf d_nodes, d_tris, d_tlas;
  DevBuf d_spill;  // traversal stack spill areas
  DevBuf d_touch;  // counting frames: touch bitmaps of Extend (nodes, triangles), then Shadow's
  uint32_t touch_node_words = 0, touch_tri_words = 0;
  uint32_t n_lights = 0;
  float total_power = 0.f;
  float scene_radius = 0.f;  // half-diagonal of the world bounds (the sun's selection weight)
  bool use_normals = false;
  bool two_level = false;
  int stack_cap = 1;
  trt_bvh_info info{};
  SkyParams sky{};  // disabled

  // camera / film
  bool camera_set = false;
  CameraParams cam{};
  uint32_t width = 0, height = 0;
  uint32_t sample_count = 0;
  uint32_t aov_count = 0;  // samples in the AOV sums (restarted with the accumulator and by trt_write_accum)
  uint64_t last_paths = 0;

  // frame buffers
  DevBuf q_o[2], q_d[2], q_t[2], q_m[2];
  DevBuf h_t, h_i;
  // shadow queues and emission lists, two of each: bounce d uses set d % 2 while the
  // previous bounce's shadow launch still reads the other
  DevBuf s_o[2], s_d[2], s_c[2], s_e[2];
  DevBuf radiance;
  DevBuf accum, color, albedo, normal;
  DevBuf aov_sum_a, aov_sum_n;    // per pixel AOV sums
  DevBuf path_aov_a, path_aov_n;  // per path first-hit AOVs of the current frame
  DevBuf tonemapped;  // RGBA8 display image
  DevBuf counters;
  DevBuf scratch_rays, scratch_hits, scratch_inst, scratch_occ;
  uint64_t path_capacity = 0;
  uint64_t queue_slots = 0;  // kSegs * seg_cap slots allocated per queue
  uint32_t seg_cap = 0;      // slots per queue segment for the current frame size

  hipEvent_t ev[4 + 4 * kMaxDepth] = {};
  // Shadow rays of bounce d run on a second stream, concurrently with Extend of bounce
  // d + 1 (they share no buffer: Shadow adds to path radiance, Extend writes hits);
  // Shade of bounce d + 1 waits for them, which keeps every path's radiance additions
  // in the serial order. ev2: per bounce, Shadow start and end on that stream.
  hipStream_t stream2 = nullptr;
  hipEvent_t ev2[2 * kMaxDepth] = {};
  int grid_trace = 0, grid_shade = 0;
  size_t spill_half = 0;  // d_spill holds two areas: Extend's and the concurrent Shadow's

  // differentiable rendering
  std::vector<Material> sampling_materials;  // empty = same as materials
  DevBuf d_mat_sampling;
  DevBuf d_sky;  // `sky` as the kernels read it (SceneView::sky)
  bool has_frame = false;      // the last frame's per-path radiance matches the current scene
  bool phase_timing = true;    // record the per-phase timing events (trt_set_phase_timing)
  uint64_t frame_id = 0;       // frames rendered by this context
  trt_render_opts last_opts{};
  uint32_t last_sample_offset = 0;
  DevBuf a_lrem, a_pend, a_key, a_g, a_sga, a_sgb, a_dldi, a_partial, a_out;

  // denoiser
  std::unique_ptr<Unet> unet;
  DevBuf dn_feat, dn_denoised, dn_ds_feat, dn_ds_tgt, dn_batch_feat, dn_batch_tgt, dn_crops, dn_io;
  uint32_t dn_ds_n = 0, dn_ds_w = 0, dn_ds_h = 0;
  std::vector<uint8_t> dn_ds_have;  // per slot: bit0 input, bit1 target
  // denoising policy of the reference's Renderer::DenoiseFrame (Renderer.cpp:546-569)
  uint32_t dn_enabled = 0, dn_threshold = 0, dn_last_spp = 0;
};

namespace {

// Scene, camera or material changes make the last frame's per-path radiance stale for
// the adjoint, which replays that frame against the current state.
void invalidate_frame(trt_context* c) { c->has_frame = false; }

// Copies ctx->sky to the device copy the kernels read; called whenever it changes (no
// frame is in flight then: every render call ends synchronised).
hipError_t upload_sky(trt_context* c) {
  hipError_t e = hipMemcpyAsync(c->d_sky.p, &c->sky, sizeof(SkyParams), hipMemcpyHostToDevice, c->stream);
  return e == hipSuccess ? hipStreamSynchronize(c->stream) : e;
}

// Restarts progressive accumulation (Renderer::Reset, Tracer/Renderer/Renderer.cpp:186-190),
// and with it the denoiser's record of the sample count it last ran at (Denoiser::Reset).
void reset_accumulation(trt_context* c) {
  c->sample_count = 0;
  c->aov_count = 0;
  c->dn_last_spp = 0;
}

trt_status fail(trt_context* c, trt_status s, const std::string& msg) {
  if (c) c->err = msg;
  return s;
}

trt_status hip_fail(trt_context* c, hipError_t e, const char* what) {
  const trt_status s = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? TRT_OUT_OF_MEMORY : TRT_DEVICE;
  return fail(c, s, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIPC(expr, what)                           \
  do {                                             \
    hipError_t e_ = (expr);                        \
    if (e_ != hipSuccess) return hip_fail(ctx, e_, what); \
  } while (0)

hipError_t upload(DevBuf& b, const void* src, size_t bytes, hipStream_t s) {
  hipError_t e = b.ensure(std::max<size_t>(bytes, 16));
  if (e != hipSuccess) return e;
  if (bytes) return hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, s);
  return hipSuccess;
}

Material to_material(const trt_material& m) {
  Material r{};
  for (int k = 0; k < 3; ++k) {
    r.base[k] = m.base_color[k];
    r.emission[k] = m.emission[k];
  }
  r.roughness = m.roughness;
  r.metallic = m.metallic;
  r.specular = m.specular;
  r.specular_tint = m.specular_tint;
  r.sheen = m.sheen;
  r.sheen_tint = m.sheen_tint;
  r.clearcoat = m.clearcoat;
  r.clearcoat_gloss = m.clearcoat_gloss;
  r.subsurface = m.subsurface;
  r.anisotropic = m.anisotropic;
  r.model = m.model;
  return r;
}

bool finite_all(const float* p, size_t n) {
  for (size_t i = 0; i < n; ++i)
    if (!(p[i] == p[i]) || p[i] == __builtin_huge_valf() || p[i] == -__builtin_huge_valf()) return false;
  return true;
}

void tri_boxes(const float* pos, uint32_t first, uint32_t n, std::vector<float>& box) {
  box.resize(size_t(n) * 6);
  for (uint32_t i = 0; i < n; ++i) {
    const float* p = pos + size_t(first + i) * 9;
    for (int k = 0; k < 3; ++k) {
      box[size_t(i) * 6 + k] = std::min(p[k], std::min(p[3 + k], p[6 + k]));
      box[size_t(i) * 6 + 3 + k] = std::max(p[k], std::max(p[3 + k], p[6 + k]));
    }
  }
}

// Builds the BVH over a mesh's triangles pos[first .. first+n), pre-split when
// presplit > 0; ref_prim receives the mesh-local triangle of each build primitive.
BuildResult build_mesh(const float* pos, uint32_t first, uint32_t n, const trt_build_opts& o, int threads,
                       std::vector<uint32_t>& ref_prim) {
  std::vector<float> box;
  if (o.presplit > 0.f) {
    presplit_triangles(pos + size_t(first) * 9, n, o.presplit, box, ref_prim);
  } else {
    tri_boxes(pos, first, n, box);
    ref_prim.resize(n);
    for (uint32_t i = 0; i < n; ++i) ref_prim[i] = i;
  }
  return build_binned_sah(box.data(), uint32_t(ref_prim.size()), o.bins, o.max_leaf, threads, kTravCostTriangles);
}

// Appends a mesh BVH in device layout: node indices offset by node_base, leaf
// triangle ranges offset by tri_base; traversal triangles get mesh-local prim ids
// (ref_prim maps build primitives to triangles; a pre-split triangle is stored once
// per leaf that references it).
void append_blas(const float* pos, uint32_t tri_offset, const BuildResult& b, const std::vector<uint32_t>& ref_prim,
                 std::vector<DevNode2>& nodes, std::vector<float>& tris) {
  std::vector<DevNode2> dn = to_device_layout(b);
  const int32_t node_base = int32_t(nodes.size());
  const uint32_t tri_base = uint32_t(tris.size() / 12);
  for (DevNode2& n : dn) {
    for (int s = 0; s < 2; ++s) {
      if (n.c[s] >= 0) {
        n.c[s] += node_base;
      } else {
        const uint32_t code = ~uint32_t(n.c[s]);
        const uint32_t first = (code >> 3) + tri_base, cnt = code & 7u;
        n.c[s] = leaf_code(first, cnt);
      }
    }
    nodes.push_back(n);
  }
  for (uint32_t k = 0; k < b.prim_order.size(); ++k) {
    const uint32_t local = ref_prim[b.prim_order[k]];
    const float* p = pos + size_t(tri_offset + local) * 9;
    float rec[12] = {p[0], p[1], p[2], 0.f, p[3], p[4], p[5], 0.f, p[6], p[7], p[8], 0.f};
    std::memcpy(&rec[3], &local, 4);
    tris.insert(tris.end(), rec, rec + 12);
  }
}

trt_status ensure_frame(trt_context* ctx, uint32_t W, uint32_t H, uint32_t spp) {
  const uint64_t n = uint64_t(W) * H * spp;
  if (n == 0 || n >= (1ull << 30)) return fail(ctx, TRT_INVALID_ARG, "path count must be in [1, 2^30)");
  const uint64_t npix = uint64_t(W) * H;
  if (W != ctx->width || H != ctx->height) {
    HIPC(ctx->accum.ensure(npix * 16), "alloc accum");
    HIPC(ctx->color.ensure(npix * 16), "alloc color");
    HIPC(ctx->albedo.ensure(npix * 16), "alloc albedo");
    HIPC(ctx->normal.ensure(npix * 16), "alloc normal");
    HIPC(ctx->aov_sum_a.ensure(npix * 16), "alloc AOV sums");
    HIPC(ctx->aov_sum_n.ensure(npix * 16), "alloc AOV sums");
    HIPC(hipMemsetAsync(ctx->albedo.p, 0, npix * 16, ctx->stream), "clear albedo");
    HIPC(hipMemsetAsync(ctx->normal.p, 0, npix * 16, ctx->stream), "clear normal");
    ctx->width = W;
    ctx->height = H;
    reset_accumulation(ctx);
  }
  // queues: kSegs segments of seg_cap(n) slots (slack for uneven segment fill)
  ctx->seg_cap = seg_cap(n);
  const uint64_t slots = uint64_t(kSegs) * ctx->seg_cap;
  if (slots > ctx->queue_slots) {
    for (int i = 0; i < 2; ++i) {
      HIPC(ctx->q_o[i].ensure(slots * 16), "alloc queue");
      HIPC(ctx->q_d[i].ensure(slots * 16), "alloc queue");
      HIPC(ctx->q_t[i].ensure(slots * 16), "alloc queue");
      HIPC(ctx->q_m[i].ensure(slots * 4), "alloc queue");
    }
    HIPC(ctx->h_t.ensure(slots * 16), "alloc hits");
    HIPC(ctx->h_i.ensure(slots * 4), "alloc hits");
    for (int i = 0; i < 2; ++i) {
      HIPC(ctx->s_o[i].ensure(slots * 16), "alloc shadow queue");
      HIPC(ctx->s_d[i].ensure(slots * 16), "alloc shadow queue");
      HIPC(ctx->s_c[i].ensure(slots * 16), "alloc shadow queue");
      HIPC(ctx->s_e[i].ensure(slots * 16), "alloc emission list");
    }
    ctx->queue_slots = slots;
  }
  if (n > ctx->path_capacity) {
    HIPC(ctx->radiance.ensure(n * 16), "alloc radiance");
    HIPC(ctx->path_aov_a.ensure(n * 16), "alloc path AOVs");
    HIPC(ctx->path_aov_n.ensure(n * 16), "alloc path AOVs");
    ctx->path_capacity = n;
  }
  return TRT_OK;
}

DevAccel accel_of(const trt_context* ctx) {
  DevAccel a;
  a.nodes = ctx->d_nodes.as<DevNode48>();
  a.tris = ctx->d_tris.as<float4>();
  a.tlas = ctx->two_level ? ctx->d_tlas.as<DevNode48>() : nullptr;
  a.inst = ctx->two_level ? ctx->d_inst_tlas.as<Instance>() : nullptr;
  a.stack_cap = ctx->stack_cap;
  a.two_level = ctx->two_level ? 1 : 0;
  a.spill = ctx->d_spill.as<uint32_t>();
  a.touch_nodes = nullptr;
  a.touch_tris = nullptr;
  return a;
}

SceneView view_of(const trt_context* ctx) {
  SceneView s{};
  s.tri_pos = ctx->d_pos.as<float>();
  s.tri_nrm = ctx->use_normals ? ctx->d_nrm.as<float>() : nullptr;
  s.tri_mat = ctx->d_mat_ids.as<uint32_t>();
  s.materials = ctx->d_materials.as<Material>();
  s.instances = ctx->two_level ? ctx->d_inst.as<Instance>() : nullptr;
  s.lights = ctx->d_lights.as<LightTri>();
  s.alias_prob = ctx->d_alias_p.as<float>();
  s.alias_idx = ctx->d_alias_i.as<uint32_t>();
  s.n_lights = ctx->n_lights;
  s.total_power = ctx->total_power;
  s.n_tris = uint32_t(ctx->mat_ids.size());
  s.n_materials = uint32_t(ctx->materials.size());
  s.sky = ctx->d_sky.as<SkyParams>();
  return s;
}

const Material* sampling_mats(const trt_context* ctx) {
  return ctx->sampling_materials.empty() ? ctx->d_materials.as<Material>() : ctx->d_mat_sampling.as<Material>();
}

RenderParams params_of_opts(const trt_context* ctx, const trt_render_opts& o) {
  RenderParams rp{};
  rp.cam = ctx->cam;
  rp.width = o.width;
  rp.height = o.height;
  rp.spp = o.spp;
  rp.max_depth = o.max_depth;
  rp.rr_depth = o.rr_depth;
  rp.seed = o.seed;
  rp.sample_offset = ctx->sample_count;
  rp.eps = o.ray_epsilon;
  rp.clamp = o.clamp;
  rp.mode = o.mode;
  rp.mode_flags = o.mode_flags;
  rp.ao_dist = o.ao_dist;
  rp.zdepth_max = o.zdepth_max;
  rp.separate_sampling = ctx->sampling_materials.empty() ? 0u : 1u;
  return rp;
}

// Launches one frame on ctx->stream. stats = counting traversal kernels (no timing use).
trt_status run_frame(trt_context* ctx, const trt_render_opts& o, bool stats, bool accumulate, trt_stats* st) {
  trt_status s = ensure_frame(ctx, o.width, o.height, o.spp);
  if (s != TRT_OK) return s;
  const RenderParams rp = params_of_opts(ctx, o);
  const uint32_t n = o.width * o.height * o.spp;
  const DevAccel acc0 = accel_of(ctx);
  const SceneView sv = view_of(ctx);
  uint32_t* ctr = ctx->counters.as<uint32_t>();
  FrameBuffers fb;
  fb.radiance = ctx->radiance.as<float4>();
  fb.accum = ctx->accum.as<float4>();
  fb.color = ctx->color.as<float4>();
  fb.albedo = ctx->albedo.as<float4>();
  fb.normal = ctx->normal.as<float4>();
  fb.path_aov_a = ctx->path_aov_a.as<float4>();
  fb.path_aov_n = ctx->path_aov_n.as<float4>();
  fb.aov_sum_a = ctx->aov_sum_a.as<float4>();
  fb.aov_sum_n = ctx->aov_sum_n.as<float4>();
  RayQueue q[2];
  for (int i = 0; i < 2; ++i) {
    q[i].o_tmin = ctx->q_o[i].as<float4>();
    q[i].d_tmax = ctx->q_d[i].as<float4>();
    q[i].t_pdf = ctx->q_t[i].as<float4>();
    q[i].meta = ctx->q_m[i].as<uint32_t>();
  }
  HitBuf hb{ctx->h_t.as<float4>(), ctx->h_i.as<uint32_t>()};
  ShadowQueue sq[2];
  for (int i = 0; i < 2; ++i)
    sq[i] = ShadowQueue{ctx->s_o[i].as<float4>(), ctx->s_d[i].as<float4>(), ctx->s_c[i].as<float4>(),
                        ctx->s_e[i].as<float4>()};
  hipStream_t stm = ctx->stream;
  hipEvent_t* ev = ctx->ev;
  // debug modes: one bounce (two for ambient occlusion)
  const uint32_t D = o.mode == kModePathTracing ? std::min<uint32_t>(o.max_depth, kMaxDepth)
                     : std::min<uint32_t>(o.max_depth, (o.mode == kModeAmbientOcclusion ||
                                                        o.mode == kModeAmbientOcclusionShading) ? 2u : 1u);

  const uint32_t cap = ctx->seg_cap;
  // the concurrent Shadow launches use the second spill area
  DevAccel acc = acc0;
  DevAccel acc_sh = acc;
  acc_sh.spill = acc.spill + ctx->spill_half;
  // counting frames: every launch marks the nodes and triangle records it reads
  const uint32_t tw = ctx->touch_node_words + ctx->touch_tri_words;
  unsigned long long* ctr64 = reinterpret_cast<unsigned long long*>(ctr + kCtrStats);
  if (stats) {
    HIPC(ctx->d_touch.ensure(size_t(2) * tw * 4), "alloc touch bitmaps");
    HIPC(hipMemsetAsync(ctx->d_touch.p, 0, size_t(2) * tw * 4, stm), "clear touch bitmaps");
    uint32_t* t = ctx->d_touch.as<uint32_t>();
    acc.touch_nodes = t;
    acc.touch_tris = t + ctx->touch_node_words;
    acc_sh.touch_nodes = t + tw;
    acc_sh.touch_tris = t + tw + ctx->touch_node_words;
  }
  hipStream_t stm2 = ctx->stream2;
  hipEvent_t* ev2 = ctx->ev2;

  // Events: ev[4 + 4d + 1] (Shade done) and ev2[2d + 1] (Shadow done) order the two
  // streams, ev[0] and ev[2] bracket the frame; the others only time the phases. Every
  // event recorded between two kernels of a stream delays the second by a few
  // microseconds, so those are left out when phase timing is off.
  const bool timing = ctx->phase_timing;
  launch_zero_counters(ctr, stm);
  HIPC(hipEventRecord(ev[0], stm), "event");
  launch_raygen(rp, n, cap, q[0], fb, ctr, stm);
  if (timing) HIPC(hipEventRecord(ev[1], stm), "event");
  for (uint32_t d = 0; d < D; ++d) {
    const int cur = int(d & 1u), nxt = cur ^ 1;
    const int di = int(d);
    launch_extend(acc, q[cur], hb, SegIn{ctr + seg_ctr(kSegExt, di), ctr + seg_ctr(kSegWorkExt, di), cap}, ctr,
                  ctx->grid_trace, stats, stm);
    if (stats) {
      launch_touch_count(acc.touch_nodes, ctx->touch_node_words, ctr64 + 6, stm);
      launch_touch_count(acc.touch_tris, ctx->touch_tri_words, ctr64 + 7, stm);
    }
    if (timing) HIPC(hipEventRecord(ev[4 + 4 * d + 0], stm), "event");
    // Shade writes shadow queue and emission list d % 2, which the shadow launch of
    // bounce d - 2 read: wait for it. It writes no path radiance, so it runs beside the
    // previous bounce's shadow launch.
    if (d > 1) HIPC(hipStreamWaitEvent(stm, ev2[2 * (d - 2) + 1], 0), "wait");
    if (timing) HIPC(hipEventRecord(ev[4 + 4 * d + 3], stm), "event");
    launch_shade(sv, sampling_mats(ctx), rp, q[cur], hb, q[nxt], sq[cur], fb,
                 SegIn{ctr + seg_ctr(kSegExt, di), ctr + seg_ctr(kSegWorkShade, di), cap},
                 ctr + seg_ctr(kSegExt, di + 1), ctr + seg_ctr(kSegShadow, di), ctr + seg_ctr(kSegEmit, di), ctr, d,
                 ctx->grid_shade, stm);
    HIPC(hipEventRecord(ev[4 + 4 * d + 1], stm), "event");
    HIPC(hipStreamWaitEvent(stm2, ev[4 + 4 * d + 1], 0), "wait");
    if (timing) HIPC(hipEventRecord(ev2[2 * d], stm2), "event");
    // all path radiance is added on this stream, in vertex order: the emission Shade
    // found at bounce d after the shadow rays of bounce d - 1, before those of bounce d
    launch_emit(sq[cur].emit, ctr + seg_ctr(kSegEmit, di), cap, fb.radiance, stm2);
    launch_shadow(acc_sh, sq[cur], fb.radiance,
                  SegIn{ctr + seg_ctr(kSegShadow, di), ctr + seg_ctr(kSegWorkShadow, di), cap}, ctr, ctx->grid_trace,
                  stats, stm2);
    if (stats) {
      launch_touch_count(acc_sh.touch_nodes, ctx->touch_node_words, ctr64 + 8, stm2);
      launch_touch_count(acc_sh.touch_tris, ctx->touch_tri_words, ctr64 + 9, stm2);
    }
    HIPC(hipEventRecord(ev2[2 * d + 1], stm2), "event");
  }
  if (D > 0) HIPC(hipStreamWaitEvent(stm, ev2[2 * (D - 1) + 1], 0), "wait");
  if (timing) HIPC(hipEventRecord(ev[3], stm), "event");
  if (accumulate)
    launch_finalize(rp, fb.radiance, fb, ctx->sample_count == 0 ? 1u : 0u, ctx->aov_count == 0 ? 1u : 0u,
                    ctx->aov_count + o.spp, stm);
  HIPC(hipEventRecord(ev[2], stm), "event");
  launch_sum_counts(ctr, cap, stm);
  HIPC(hipGetLastError(), "kernel launch");
  HIPC(hipStreamSynchronize(stm), "frame");

  uint32_t hc[kCtrSegBase];
  HIPC(hipMemcpy(hc, ctr, sizeof(hc), hipMemcpyDeviceToHost), "read counters");
  ctx->last_paths = n;
  ctx->has_frame = true;
  ++ctx->frame_id;
  ctx->last_opts = o;
  ctx->last_sample_offset = rp.sample_offset;
  if (accumulate) {
    ctx->sample_count += o.spp;
    ctx->aov_count += o.spp;
  }
  if (st) {
    std::memset(st, 0, sizeof(*st));
    st->paths = n;
    st->primary_rays = hc[kCtrExt + 0];
    for (uint32_t d = 0; d < D; ++d) {
      st->extend_rays += hc[kCtrExt + d];
      st->shadow_rays += hc[kCtrShadow + d];
    }
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ev[0], ev[2]);
    st->frame_ms = ms;
    if (timing) {
      (void)hipEventElapsedTime(&ms, ev[0], ev[1]);
      st->raygen_ms = ms;
      // per-kernel intervals; Extend and Shadow overlap, so the phases add up to more
      // than the frame
      hipEvent_t prev = ev[1];
      for (uint32_t d = 0; d < D; ++d) {
        (void)hipEventElapsedTime(&ms, prev, ev[4 + 4 * d + 0]);
        st->extend_ms += ms;
        (void)hipEventElapsedTime(&ms, ev[4 + 4 * d + 3], ev[4 + 4 * d + 1]);
        st->shade_ms += ms;
        (void)hipEventElapsedTime(&ms, ev2[2 * d], ev2[2 * d + 1]);
        st->shadow_ms += ms;
        prev = ev[4 + 4 * d + 1];
      }
      (void)hipEventElapsedTime(&ms, D > 0 ? ev[3] : ev[1], ev[2]);
      st->finalize_ms = ms;
      st->trace_ms = st->frame_ms - st->raygen_ms - st->shade_ms - st->finalize_ms;
    }
    st->extend_launches = D;
    st->shadow_launches = D;
    st->sample_count = ctx->sample_count;
    st->overflow = hc[kCtrOverflow];
  }
  if (hc[kCtrOverflow] & 1u) return fail(ctx, TRT_DEVICE, "traversal stack overflow");
  if (hc[kCtrOverflow] & 2u) return fail(ctx, TRT_DEVICE, "ray queue overflow");
  (void)stats;
  return TRT_OK;
}

}  // namespace

trt_status trt_internal_fail(trt_context* ctx, trt_status s, const std::string& msg) { return fail(ctx, s, msg); }

extern "C" {

const char* trt_version(void) { return "trt 0.1 (gfx950 wavefront path tracer)"; }

trt_status trt_set_phase_timing(trt_context* ctx, int enabled) {
  if (!ctx) return TRT_INVALID_ARG;
  ctx->phase_timing = enabled != 0;
  return TRT_OK;
}

static trt_status init_context(trt_context* ctx) {
  HIPC(hipSetDevice(ctx->device), "hipSetDevice");
  HIPC(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking), "stream");
  HIPC(hipStreamCreateWithFlags(&ctx->stream2, hipStreamNonBlocking), "stream");
  hipDeviceProp_t prop;
  HIPC(hipGetDeviceProperties(&prop, ctx->device), "device properties");
  ctx->num_cus = prop.multiProcessorCount;
  for (auto& e : ctx->ev) HIPC(hipEventCreate(&e), "event");
  for (auto& e : ctx->ev2) HIPC(hipEventCreate(&e), "event");
  HIPC(ctx->counters.ensure(kNumCounters * 4), "alloc counters");
  HIPC(hipMemset(ctx->counters.p, 0, kNumCounters * 4), "clear counters");
  HIPC(ctx->d_sky.ensure(sizeof(SkyParams)), "alloc sky");
  HIPC(upload_sky(ctx), "upload sky");
  return TRT_OK;
}

trt_status trt_create(int device, trt_context** out) {
  if (!out) return TRT_INVALID_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return TRT_DEVICE;
  if (device < 0 || device >= ndev) return TRT_INVALID_ARG;
  trt_context* ctx = new trt_context();
  ctx->device = device;
  const trt_status s = init_context(ctx);
  if (s != TRT_OK) {
    trt_destroy(ctx);
    return s;
  }
  *out = ctx;
  return TRT_OK;
}

void trt_destroy(trt_context* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->stream2) (void)hipStreamSynchronize(ctx->stream2);
  for (auto& e : ctx->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : ctx->ev2)
    if (e) (void)hipEventDestroy(e);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  if (ctx->stream2) (void)hipStreamDestroy(ctx->stream2);
  delete ctx;
}

const char* trt_last_error(const trt_context* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

trt_status trt_clear_scene(trt_context* ctx) {
  if (!ctx) return TRT_INVALID_ARG;
  ctx->pos.clear();
  ctx->nrm.clear();
  ctx->has_nrm.clear();
  ctx->mat_ids.clear();
  ctx->meshes.clear();
  ctx->instances_in.clear();
  ctx->sky = SkyParams{};
  HIPC(upload_sky(ctx), "upload sky");
  ctx->built = false;
  ctx->dirty = true;
  invalidate_frame(ctx);
  return TRT_OK;
}

trt_status trt_set_sky(trt_context* ctx, const trt_sky* sky) {
  if (!ctx || !sky) return TRT_INVALID_ARG;
  if (!finite_all(sky->sun_dir, 3) || !(sky->angular_diameter >= 0.f) || !(sky->angular_diameter < 10800.f) ||
      !(sky->intensity >= 0.f) || !(sky->intensity < 1e30f) || !(sky->turbidity >= 0.f) || !(sky->turbidity < 1e6f))
    return fail(ctx, TRT_INVALID_ARG, "invalid sky parameters");
  ctx->sky = make_sky(sky->enabled != 0, sky->draw_sun != 0, sky->sun_dir, sky->angular_diameter, sky->intensity,
                      sky->turbidity, sky->sun_nee != 0);
  ctx->sky.sun_power = sun_selection_power(ctx->sky, ctx->scene_radius);
  HIPC(upload_sky(ctx), "upload sky");
  reset_accumulation(ctx);
  invalidate_frame(ctx);
  return TRT_OK;
}

trt_status trt_set_materials(trt_context* ctx, const trt_material* mats, uint32_t n) {
  if (!ctx) return TRT_INVALID_ARG;
  if (n && !mats) return fail(ctx, TRT_INVALID_ARG, "materials is null");
  std::vector<Material> m(n);
  for (uint32_t i = 0; i < n; ++i) {
    if (!finite_all(mats[i].base_color, 3) || !finite_all(mats[i].emission, 3) ||
        !finite_all(&mats[i].roughness, 1) || !finite_all(&mats[i].metallic, 1))
      return fail(ctx, TRT_INVALID_ARG, "material " + std::to_string(i) + " has a non-finite value");
    if (mats[i].model > 1) return fail(ctx, TRT_INVALID_ARG, "unknown material model");
    m[i] = to_material(mats[i]);
  }
  ctx->materials = std::move(m);
  ctx->sampling_materials.clear();
  ctx->built = false;
  invalidate_frame(ctx);
  return TRT_OK;
}

trt_status trt_add_mesh(trt_context* ctx, const float* positions, const float* normals,
                        const uint32_t* material_ids, uint32_t n_tris, uint32_t* mesh_id) {
  if (!ctx) return TRT_INVALID_ARG;
  if (n_tris && (!positions || !material_ids)) return fail(ctx, TRT_INVALID_ARG, "positions/material_ids is null");
  if (!finite_all(positions, size_t(n_tris) * 9)) return fail(ctx, TRT_INVALID_ARG, "non-finite vertex position");
  if (normals && !finite_all(normals, size_t(n_tris) * 9)) return fail(ctx, TRT_INVALID_ARG, "non-finite normal");
  const uint64_t total = uint64_t(ctx->mat_ids.size()) + n_tris;
  if (total >= (1ull << 28)) return fail(ctx, TRT_INVALID_ARG, "more than 2^28 triangles");
  Mesh m{uint32_t(ctx->mat_ids.size()), n_tris};
  ctx->pos.insert(ctx->pos.end(), positions, positions + size_t(n_tris) * 9);
  if (normals) {
    ctx->nrm.resize(ctx->pos.size(), 0.f);
    std::memcpy(&ctx->nrm[size_t(m.tri_offset) * 9], normals, size_t(n_tris) * 9 * 4);
  } else {
    ctx->nrm.resize(ctx->pos.size(), 0.f);
  }
  ctx->has_nrm.push_back(normals ? 1 : 0);
  ctx->mat_ids.insert(ctx->mat_ids.end(), material_ids, material_ids + n_tris);
  ctx->meshes.push_back(m);
  if (mesh_id) *mesh_id = uint32_t(ctx->meshes.size() - 1);
  ctx->built = false;
  invalidate_frame(ctx);
  return TRT_OK;
}

trt_status trt_set_instances(trt_context* ctx, const trt_instance* inst, uint32_t n) {
  if (!ctx) return TRT_INVALID_ARG;
  if (n && !inst) return fail(ctx, TRT_INVALID_ARG, "instances is null");
  for (uint32_t i = 0; i < n; ++i)
    if (!finite_all(inst[i].transform, 12)) return fail(ctx, TRT_INVALID_ARG, "non-finite instance transform");
  ctx->instances_in.assign(inst, inst + n);
  ctx->built = false;
  invalidate_frame(ctx);
  return TRT_OK;
}

// Host threads for the BVH build when the caller gives none: OMP_NUM_THREADS when set
// (the usual per-process thread budget on shared machines, where the hardware thread
// count can be far above the CPUs this process may use), else the CPUs in this
// process's affinity mask, else the hardware thread count.
static int default_build_threads() {
  if (const char* e = std::getenv("OMP_NUM_THREADS")) {
    const int v = std::atoi(e);
    if (v > 0) return v;
  }
  cpu_set_t set;
  CPU_ZERO(&set);
  if (sched_getaffinity(0, sizeof(set), &set) == 0 && CPU_COUNT(&set) > 0) return CPU_COUNT(&set);
  return int(std::thread::hardware_concurrency());
}

trt_status trt_build(trt_context* ctx, const trt_build_opts* opts) {
  if (!ctx) return TRT_INVALID_ARG;
  trt_build_opts o = opts ? *opts : trt_build_opts{32, 4, 0, 0.f};
  if (o.bins < 2 || o.bins > 256 || o.max_leaf < 1 || o.max_leaf > 7)
    return fail(ctx, TRT_INVALID_ARG, "bins must be in [2,256] and max_leaf in [1,7]");
  if (!(o.presplit >= 0.f && o.presplit <= 8.f))
    return fail(ctx, TRT_INVALID_ARG, "presplit must be in [0, 8]");
  int threads = o.threads > 0 ? o.threads : default_build_threads();
  if (threads <= 0) threads = 1;
  const uint32_t n_tris = uint32_t(ctx->mat_ids.size());
  const uint32_t n_mats = uint32_t(ctx->materials.size());
  if (n_mats == 0) return fail(ctx, TRT_INVALID_ARG, "no materials");
  if (n_tris == 0) return fail(ctx, TRT_INVALID_ARG, "scene has no triangles");
  for (uint32_t g = 0; g < n_tris; ++g)
    if (ctx->mat_ids[g] >= n_mats && ctx->instances_in.empty())
      return fail(ctx, TRT_INVALID_ARG, "material id out of range at triangle " + std::to_string(g));
  HIPC(hipSetDevice(ctx->device), "hipSetDevice");
  const auto t0 = std::chrono::steady_clock::now();

  // normals: use them when any mesh provides them; fill the others with face normals
  ctx->use_normals = false;
  for (uint8_t h : ctx->has_nrm) ctx->use_normals = ctx->use_normals || h;
  if (ctx->use_normals) {
    for (size_t mi = 0; mi < ctx->meshes.size(); ++mi) {
      if (ctx->has_nrm[mi]) continue;
      const Mesh& m = ctx->meshes[mi];
      for (uint32_t k = 0; k < m.n_tris; ++k) {
        const float* p = &ctx->pos[size_t(m.tri_offset + k) * 9];
        const V3 n = normalize(cross(v3(p[3] - p[0], p[4] - p[1], p[5] - p[2]), v3(p[6] - p[0], p[7] - p[1], p[8] - p[2])));
        for (int v = 0; v < 3; ++v) {
          ctx->nrm[size_t(m.tri_offset + k) * 9 + v * 3 + 0] = n.x;
          ctx->nrm[size_t(m.tri_offset + k) * 9 + v * 3 + 1] = n.y;
          ctx->nrm[size_t(m.tri_offset + k) * 9 + v * 3 + 2] = n.z;
        }
      }
    }
  }

  std::vector<DevNode2> nodes;
  std::vector<float> tris;
  std::vector<Instance> inst_orig, inst_tlas;
  std::vector<DevNode2> tlas_nodes;
  ctx->two_level = !ctx->instances_in.empty();
  uint64_t n_nodes = 0, n_leaves = 0, n_refs = 0;
  double sah = 0.0;
  int depth = 0, tlas_depth = 0;
  if (!ctx->two_level) {
    std::vector<uint32_t> ref_prim;
    BuildResult b = build_mesh(ctx->pos.data(), 0, n_tris, o, threads, ref_prim);
    append_blas(ctx->pos.data(), 0, b, ref_prim, nodes, tris);
    n_refs = ref_prim.size();
    // single-level: prim ids are global triangle indices (tri_offset 0)
    n_nodes = b.nodes.size();
    n_leaves = b.n_leaves;
    sah = b.sah_cost;
    depth = b.depth;
  } else {
    std::vector<uint32_t> blas_root(ctx->meshes.size());
    std::vector<float> mesh_lo(ctx->meshes.size() * 3), mesh_hi(ctx->meshes.size() * 3);
    for (size_t mi = 0; mi < ctx->meshes.size(); ++mi) {
      const Mesh& m = ctx->meshes[mi];
      std::vector<uint32_t> ref_prim;
      BuildResult b = build_mesh(ctx->pos.data(), m.tri_offset, m.n_tris, o, threads, ref_prim);
      blas_root[mi] = uint32_t(nodes.size());
      append_blas(ctx->pos.data(), m.tri_offset, b, ref_prim, nodes, tris);
      n_refs += ref_prim.size();
      n_nodes += b.nodes.size();
      n_leaves += b.n_leaves;
      depth = std::max(depth, b.depth);
      for (int k = 0; k < 3; ++k) {
        mesh_lo[mi * 3 + k] = b.nodes[0].lo[k];
        mesh_hi[mi * 3 + k] = b.nodes[0].hi[k];
      }
    }
    const uint32_t ni = uint32_t(ctx->instances_in.size());
    inst_orig.resize(ni);
    std::vector<float> ibox(size_t(ni) * 6);
    for (uint32_t i = 0; i < ni; ++i) {
      const trt_instance& ti = ctx->instances_in[i];
      if (ti.mesh >= ctx->meshes.size()) return fail(ctx, TRT_INVALID_ARG, "instance mesh out of range");
      Instance& in = inst_orig[i];
      std::memset(&in, 0, sizeof(in));
      std::memcpy(in.o2w, ti.transform, 48);
      invert_affine(in.o2w, in.w2o);
      in.mesh = ti.mesh;
      in.tri_offset = ctx->meshes[ti.mesh].tri_offset;
      in.mat_offset = ti.material_offset;
      in.blas_root = blas_root[ti.mesh] * kNodeBytes;  // byte offset, like every node reference
      in.id = i;
      const Mesh& m = ctx->meshes[ti.mesh];
      for (uint32_t k = 0; k < m.n_tris; ++k)
        if (ctx->mat_ids[m.tri_offset + k] + ti.material_offset >= n_mats)
          return fail(ctx, TRT_INVALID_ARG, "instance material id out of range");
      // world box of the transformed mesh box corners
      float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
      for (int cc = 0; cc < 8; ++cc) {
        const V3 p = v3((cc & 1) ? mesh_hi[ti.mesh * 3 + 0] : mesh_lo[ti.mesh * 3 + 0],
                        (cc & 2) ? mesh_hi[ti.mesh * 3 + 1] : mesh_lo[ti.mesh * 3 + 1],
                        (cc & 4) ? mesh_hi[ti.mesh * 3 + 2] : mesh_lo[ti.mesh * 3 + 2]);
        const V3 w = xform_point(in.o2w, p);
        const float wc[3] = {w.x, w.y, w.z};
        for (int k = 0; k < 3; ++k) {
          // widen by a relative ulp margin: the transform is rounded
          lo[k] = std::min(lo[k], wc[k]);
          hi[k] = std::max(hi[k], wc[k]);
        }
      }
      for (int k = 0; k < 3; ++k) {
        const float pad = 1e-5f * std::max(1.f, std::max(std::fabs(lo[k]), std::fabs(hi[k])));
        ibox[size_t(i) * 6 + k] = lo[k] - pad;
        ibox[size_t(i) * 6 + 3 + k] = hi[k] + pad;
      }
    }
    BuildResult tb = build_binned_sah(ibox.data(), ni, o.bins, o.max_leaf, threads, kTravCostInstances);
    tlas_nodes = to_device_layout(tb);
    inst_tlas.resize(ni);
    for (uint32_t k = 0; k < ni; ++k) inst_tlas[k] = inst_orig[tb.prim_order[k]];
    tlas_depth = tb.depth;
    n_nodes += tb.nodes.size();
  }
  // entries per lane: one per tree level, plus a leaf remainder and the exit marker per instance level
  const int cap = depth + tlas_depth + (ctx->two_level ? 4 : 2);
  if (cap > 64) return fail(ctx, TRT_UNSUPPORTED, "BVH too deep for the LDS stack (" + std::to_string(cap) + ")");
  ctx->stack_cap = cap;

  // lights
  std::vector<uint32_t> mesh_cnt(ctx->meshes.size());
  for (size_t i = 0; i < ctx->meshes.size(); ++i) mesh_cnt[i] = ctx->meshes[i].n_tris;
  LightTable lt = gather_lights(ctx->pos.data(), ctx->mat_ids.data(), n_tris, ctx->materials.data(), n_mats,
                                ctx->two_level ? inst_orig.data() : nullptr, uint32_t(inst_orig.size()),
                                mesh_cnt.data());
  ctx->n_lights = uint32_t(lt.lights.size());
  ctx->total_power = lt.total_power;
  {
    std::vector<uint32_t> mesh_off(ctx->meshes.size());
    for (size_t i = 0; i < ctx->meshes.size(); ++i) mesh_off[i] = ctx->meshes[i].tri_offset;
    ctx->scene_radius = scene_radius(ctx->pos.data(), n_tris, ctx->two_level ? inst_orig.data() : nullptr,
                                     uint32_t(inst_orig.size()), mesh_off.data(), mesh_cnt.data(),
                                     uint32_t(ctx->meshes.size()));
    ctx->sky.sun_power = sun_selection_power(ctx->sky, ctx->scene_radius);
    HIPC(upload_sky(ctx), "upload sky");
  }

  // child references are byte offsets below 2^31
  if (nodes.size() >= kMaxDeviceNodes || tlas_nodes.size() >= kMaxDeviceNodes)
    return fail(ctx, TRT_UNSUPPORTED, "BVH too large: more than 44M nodes in one tree array");
  std::vector<DevNode48> nodes48, tlas48;
  if (!pack_nodes(nodes, nodes48) || !pack_nodes(tlas_nodes, tlas48))
    return fail(ctx, TRT_UNSUPPORTED, "scene bounds too large for the node layout (infinite coordinates)");

  hipStream_t s = ctx->stream;
  HIPC(upload(ctx->d_pos, ctx->pos.data(), ctx->pos.size() * 4, s), "upload positions");
  if (ctx->use_normals) HIPC(upload(ctx->d_nrm, ctx->nrm.data(), ctx->nrm.size() * 4, s), "upload normals");
  HIPC(upload(ctx->d_mat_ids, ctx->mat_ids.data(), ctx->mat_ids.size() * 4, s), "upload material ids");
  HIPC(upload(ctx->d_materials, ctx->materials.data(), ctx->materials.size() * sizeof(Material), s), "upload materials");
  HIPC(upload(ctx->d_nodes, nodes48.data(), nodes48.size() * sizeof(DevNode48), s), "upload nodes");
  HIPC(upload(ctx->d_tris, tris.data(), tris.size() * 4, s), "upload triangles");
  HIPC(upload(ctx->d_lights, lt.lights.data(), lt.lights.size() * sizeof(LightTri), s), "upload lights");
  HIPC(upload(ctx->d_alias_p, lt.prob.data(), lt.prob.size() * 4, s), "upload alias");
  HIPC(upload(ctx->d_alias_i, lt.alias.data(), lt.alias.size() * 4, s), "upload alias");
  if (ctx->two_level) {
    HIPC(upload(ctx->d_inst, inst_orig.data(), inst_orig.size() * sizeof(Instance), s), "upload instances");
    HIPC(upload(ctx->d_inst_tlas, inst_tlas.data(), inst_tlas.size() * sizeof(Instance), s), "upload instances");
    HIPC(upload(ctx->d_tlas, tlas48.data(), tlas48.size() * sizeof(DevNode48), s), "upload tlas");
  }
  HIPC(hipStreamSynchronize(s), "upload");

  if (!traversal_lds_check()) return fail(ctx, TRT_DEVICE, "traversal kernels have static LDS: the stack rings need dynamic LDS at 0");
  // persistent grid sizes from occupancy (every resident slot gets a wave)
  ctx->grid_trace = ctx->num_cus * trace_blocks_per_cu(accel_of(ctx));
  ctx->grid_shade = ctx->num_cus * 8;
  // stack spill areas of the traversals: stack_cap entries per grid thread, twice (Extend
  // and the Shadow launch that runs beside it)
  ctx->spill_half = size_t(ctx->grid_trace) * 256 * size_t(ctx->stack_cap);
  ctx->touch_node_words = uint32_t((nodes48.size() + 31) / 32);
  ctx->touch_tri_words = uint32_t((tris.size() / 12 + 31) / 32);
  HIPC(ctx->d_spill.ensure(2 * ctx->spill_half * 4), "alloc stack spill");

  const auto t1 = std::chrono::steady_clock::now();
  ctx->info = trt_bvh_info{};
  ctx->info.nodes = n_nodes;
  ctx->info.leaves = n_leaves;
  ctx->info.depth = uint32_t(depth);
  ctx->info.stack_entries = uint32_t(cap);
  ctx->info.sah_cost = sah;
  ctx->info.build_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  ctx->info.device_bytes = nodes48.size() * sizeof(DevNode48) + tris.size() * 4 + tlas48.size() * sizeof(DevNode48);
  ctx->info.two_level = ctx->two_level ? 1 : 0;
  ctx->info.tlas_depth = uint32_t(tlas_depth);
  ctx->info.references = n_refs;
  ctx->built = true;
  reset_accumulation(ctx);
  invalidate_frame(ctx);
  return TRT_OK;
}

trt_status trt_bvh_info_get(trt_context* ctx, trt_bvh_info* info) {
  if (!ctx || !info) return TRT_INVALID_ARG;
  if (!ctx->built) return fail(ctx, TRT_NOT_BUILT, "scene not built");
  *info = ctx->info;
  return TRT_OK;
}

trt_status trt_set_camera(trt_context* ctx, const trt_camera* c) {
  if (!ctx || !c) return TRT_INVALID_ARG;
  if (!finite_all(c->position, 3) || !finite_all(c->target, 3) || !finite_all(c->up, 3) || !(c->fov > 0.f) ||
      !(c->fov < 3.14159f))
    return fail(ctx, TRT_INVALID_ARG, "invalid camera");
  const V3 p = v3(c->position[0], c->position[1], c->position[2]);
  const V3 t = v3(c->target[0], c->target[1], c->target[2]);
  const V3 u = v3(c->up[0], c->up[1], c->up[2]);
  if (length(t - p) == 0.f || length(cross(normalize(t - p), u)) == 0.f)
    return fail(ctx, TRT_INVALID_ARG, "degenerate camera basis");
  ctx->cam = make_camera(p, t, u, c->fov, c->aperture, c->focal_dist, c->distortion, c->bokeh_sides, c->bokeh_rotation);
  ctx->camera_set = true;
  reset_accumulation(ctx);  // the reference resets accumulation on camera change
  invalidate_frame(ctx);
  return TRT_OK;
}

static trt_status check_render(trt_context* ctx, const trt_render_opts* o) {
  if (!ctx || !o) return TRT_INVALID_ARG;
  if (!ctx->built) return fail(ctx, TRT_NOT_BUILT, "scene not built");
  if (!ctx->camera_set) return fail(ctx, TRT_INVALID_ARG, "camera not set");
  if (o->width == 0 || o->height == 0 || o->spp == 0 || o->max_depth == 0)
    return fail(ctx, TRT_INVALID_ARG, "width, height, spp and max_depth must be positive");
  if (o->max_depth > uint32_t(kMaxDepth - 1)) return fail(ctx, TRT_INVALID_ARG, "max_depth must be < 32");
  if (!(o->ray_epsilon >= 0.f) || !(o->clamp >= 0.f)) return fail(ctx, TRT_INVALID_ARG, "invalid epsilon/clamp");
  if (uint64_t(o->width) * o->height * o->spp >= (1ull << 30))
    return fail(ctx, TRT_INVALID_ARG, "width*height*spp must be below 2^30");
  if (o->mode >= kNumRenderModes) return fail(ctx, TRT_INVALID_ARG, "unknown render mode");
  if ((o->mode == kModeAmbientOcclusion || o->mode == kModeAmbientOcclusionShading) && !(o->ao_dist > 0.f))
    return fail(ctx, TRT_INVALID_ARG, "ambient occlusion needs ao_dist > 0");
  if (o->mode == kModeZDepth && !(o->zdepth_max > 0.f))
    return fail(ctx, TRT_INVALID_ARG, "z-depth mode needs zdepth_max > 0");
  if (o->mode == kModeMaterialProperty && o->mode_flags >= kNumMaterialProperties)
    return fail(ctx, TRT_INVALID_ARG, "unknown material property");
  return TRT_OK;
}

trt_status trt_render(trt_context* ctx, const trt_render_opts* o, trt_stats* stats) {
  trt_status s = check_render(ctx, o);
  if (s != TRT_OK) return s;
  HIPC(hipSetDevice(ctx->device), "hipSetDevice");
  return run_frame(ctx, *o, false, true, stats);
}

trt_status trt_measure_traversal(trt_context* ctx, const trt_render_opts* o, trt_trace_counts* out) {
  trt_status s = check_render(ctx, o);
  if (s != TRT_OK) return s;
  if (!out) return TRT_INVALID_ARG;
  HIPC(hipSetDevice(ctx->device), "hipSetDevice");
  trt_stats st;
  s = run_frame(ctx, *o, true, false, &st);
  if (s != TRT_OK) return s;
  unsigned long long c[10];
  HIPC(hipMemcpy(c, ctx->counters.as<uint32_t>() + kCtrStats, sizeof(c), hipMemcpyDeviceToHost), "read stats");
  std::memset(out, 0, sizeof(*out));
  out->closest_rays = st.extend_rays;
  out->any_rays = st.shadow_rays;
  out->closest_nodes_per_ray = st.extend_rays ? double(c[0]) / st.extend_rays : 0.0;
  out->closest_tris_per_ray = st.extend_rays ? double(c[1]) / st.extend_rays : 0.0;
  out->any_nodes_per_ray = st.shadow_rays ? double(c[2]) / st.shadow_rays : 0.0;
  out->any_tris_per_ray = st.shadow_rays ? double(c[3]) / st.shadow_rays : 0.0;
  out->node_bytes = sizeof(DevNode48);
  out->tri_bytes = 48;
  out->closest_spills = c[4];
  out->any_spills = c[5];
  out->closest_distinct_nodes = c[6];
  out->closest_distinct_tris = c[7];
  out->any_distinct_nodes = c[8];
  out->any_distinct_tris = c[9];
  out->launches = st.extend_launches;
  return TRT_OK;
}

trt_status trt_reset(trt_context* ctx) {
  if (!ctx) return TRT_INVALID_ARG;
  reset_accumulation(ctx);
  return TRT_OK;
}

uint32_t trt_sample_count(const trt_context* ctx) { return ctx ? ctx->sample_count : 0u; }

uint64_t trt_frame_id(const trt_context* ctx) { return ctx ? ctx->frame_id : 0u; }

static trt_status read_pixels(trt_context* ctx, DevBuf& b, float* out) {
  if (!ctx || !out) return TRT_INVALID_ARG;
  if (ctx->width == 0) return fail(ctx, TRT_NOT_BUILT, "nothing rendered yet");
  HIPC(hipSetDevice(ctx->device), "hipSetDevice");
  HIPC(hipMemcpy(out, b.p, size_t(ctx->width) * ctx->height * 16, hipMemcpyDeviceToHost), "read pixels");
  return TRT_OK;
}
trt_status trt_read_accum(trt_context* ctx, float* rgba) { return ctx ? read_pixels(ctx, ctx->accum, rgba) : TRT_INVALID_ARG; }
trt_status trt_read_color(trt_context* ctx, float* rgba) { return ctx ? read_pixels(ctx, ctx->color, rgba) : TRT_INVALID_ARG; }
trt_status trt_read_albedo(trt_context* ctx, float* rgba) { return ctx ? read_pixels(ctx, ctx->albedo, rgba) : TRT_INVALID_ARG; }
trt_status trt_read_normal(trt_context* ctx, float* rgba) { return ctx ? read_pixels(ctx, ctx->normal, rgba) : TRT_INVALID_ARG; }

// Restores a progressive render from an accumulator read by trt_read_accum: rgb sums and
// a sample count (alpha, the same whole number in every pixel). The next trt_render of
// the same film size continues it with the following samples (their keys start at the
// restored count), so a restored render adds up to the uninterrupted one.
trt_status trt_write_accum(trt_context* ctx, uint32_t width, uint32_t height, const float* rgba) {
  if (!ctx || !rgba || width == 0 || height == 0) return TRT_INVALID_ARG;
  const uint64_t npix = uint64_t(width) * height;
  if (npix >= (1ull << 30)) return fail(ctx, TRT_INVALID_ARG, "film too large");
  const float n = rgba[3];
  if (!(n >= 1.f && n < 4294967296.f) || n != float(uint32_t(n)))
    return fail(ctx, TRT_INVALID_ARG, "the accumulator's sample count (alpha) must be a positive whole number");
  for (uint64_t i = 0; i < npix; ++i) {
    if (rgba[i * 4 + 3] != n) return fail(ctx, TRT_INVALID_ARG, "every pixel must hold the same sample count");
    if (!finite_all(rgba + i * 4, 3)) return fail(ctx, TRT_INVALID_ARG, "non-finite accumulator value");
  }
  HIPC(hipSetDevice(ctx->device), "hipSetDevice");
  if (width != ctx->width || height != ctx->height) {
    HIPC(ctx->accum.ensure(npix * 16), "alloc accum");
    HIPC(ctx->color.ensure(npix * 16), "alloc color");
    HIPC(ctx->albedo.ensure(npix * 16), "alloc albedo");
    HIPC(ctx->normal.ensure(npix * 16), "alloc normal");
    HIPC(ctx->aov_sum_a.ensure(npix * 16), "alloc AOV sums");
    HIPC(ctx->aov_sum_n.ensure(npix * 16), "alloc AOV sums");
    HIPC(hipMemsetAsync(ctx->albedo.p, 0, npix * 16, ctx->stream), "clear albedo");
    HIPC(hipMemsetAsync(ctx->normal.p, 0, npix * 16, ctx->stream), "clear normal");
    ctx->width = width;
    ctx->height = height;
  }
  std::vector<float> col(npix * 4);
  const float inv = 1.f / n;
  for (uint64_t i = 0; i < npix; ++i) {
    for (int k = 0; k < 3; ++k) col[i * 4 + k] = rgba[i * 4 + k] * inv;
    col[i * 4 + 3] = 1.f;
  }
  HIPC(hipMemcpyAsync(ctx->accum.p, rgba, npix * 16, hipMemcpyHostToDevice, ctx->stream), "upload accum");
  HIPC(hipMemcpyAsync(ctx->color.p, col.data(), npix * 16, hipMemcpyHostToDevice, ctx->stream), "upload color");
  HIPC(hipStreamSynchronize(ctx->stream), "upload accum");
  ctx->sample_count = uint32_t(n);
  ctx->aov_count = 0;  // the AOV sums restart with the next frame
  ctx->dn_last_spp = 0;
  invalidate_frame(ctx);
  return TRT_OK;
}

trt_status trt_read_path_radiance(trt_context* ctx, float* rgb, uint64_t n_paths) {
  if (!ctx || !rgb) return TRT_INVALID_ARG;
  if (n_paths != ctx->last_paths || n_paths == 0) return fail(ctx, TRT_INVALID_ARG, "n_paths must equal the last frame's path count");
  HIPC(hipSetDevice(ctx->device), "hipSetDevice");
  std::vector<float> tmp(n_paths * 4);
  HIPC(hipMemcpy(tmp.data(), ctx->radiance.p, n_paths * 16, hipMemcpyDeviceToHost), "read radiance");
  for (uint64_t i = 0; i < n_paths; ++i) {
    rgb[i * 3 + 0] = tmp[i * 4 + 0];
    rgb[i * 3 + 1] = tmp[i * 4 + 1];
    rgb[i * 3 + 2] = tmp[i * 4 + 2];
  }
  return TRT_OK;
}

// Test hook of the traversal: the rays go through the render's own Extend (closest) or
// Shadow (any) kernels, as one depth-0 queue whose slot i is ray i (kSegs segments of
// ceil(n / kSegs) slots). Occlusion comes back through Shadow's own result path: each
// ray carries a unit contribution to its own radiance slot, added when unoccluded.
static trt_status trace_common(trt_context* ctx, const float* rays, uint32_t n, float* hits, uint32_t* inst,
                               uint8_t* occ, bool any) {
  if (!ctx || (n && !rays)) return TRT_INVALID_ARG;
  if (!ctx->built) return fail(ctx, TRT_NOT_BUILT, "scene not built");
  if (n == 0) return TRT_OK;
  HIPC(hipSetDevice(ctx->device), "hipSetDevice");
  HIPC(ctx->scratch_rays.ensure(size_t(n) * 32), "alloc rays");
  HIPC(ctx->scratch_hits.ensure(size_t(n) * 16), "alloc hits");
  HIPC(ctx->scratch_inst.ensure(size_t(n) * 4), "alloc hits");
  HIPC(ctx->scratch_occ.ensure(size_t(n) * 16), "alloc hits");
  // (o, tmin) rows first, then (d, tmax) rows
  std::vector<float> soa(size_t(n) * 8);
  for (uint32_t i = 0; i < n; ++i) {
    const float* r = &rays[size_t(i) * 8];
    // origins and directions finite, interval bounds not NaN (tmax may be +inf), tmin >= 0
    // (the traversal's box test widens far distances relatively, which is conservative
    // only for boxes at non-negative distances)
    if (!finite_all(r, 3) || !finite_all(r + 4, 3) || r[3] != r[3] || r[7] != r[7])
      return fail(ctx, TRT_INVALID_ARG, "ray " + std::to_string(i) + " has a non-finite origin or direction or a NaN bound");
    if (!(r[3] >= 0.f)) return fail(ctx, TRT_INVALID_ARG, "ray " + std::to_string(i) + " has a negative tmin");
    std::memcpy(&soa[size_t(i) * 4], &rays[size_t(i) * 8], 16);
    std::memcpy(&soa[(size_t(n) + i) * 4], &rays[size_t(i) * 8 + 4], 16);
  }
  float4* o_tmin = ctx->scratch_rays.as<float4>();
  float4* d_tmax = o_tmin + n;
  HIPC(hipMemcpyAsync(o_tmin, soa.data(), size_t(n) * 32, hipMemcpyHostToDevice, ctx->stream), "upload rays");
  uint32_t* ctr = ctx->counters.as<uint32_t>();
  launch_zero_counters(ctr, ctx->stream);
  const uint32_t per = (n + kSegs - 1) / kSegs;
  std::vector<uint32_t> cnt(size_t(kSegs) * kLine, 0u);
  for (uint32_t sgm = 0; sgm < kSegs; ++sgm) {
    const uint64_t first = uint64_t(sgm) * per;
    cnt[size_t(sgm) * kLine] = first >= n ? 0u : uint32_t(std::min<uint64_t>(per, n - first));
  }
  const int kind = any ? kSegShadow : kSegExt, work = any ? kSegWorkShadow : kSegWorkExt;
  HIPC(hipMemcpyAsync(ctr + seg_ctr(kind, 0), cnt.data(), cnt.size() * 4, hipMemcpyHostToDevice, ctx->stream),
       "upload queue counts");
  const SegIn in{ctr + seg_ctr(kind, 0), ctr + seg_ctr(work, 0), per};
  const DevAccel acc = accel_of(ctx);
  if (any) {
    std::vector<float> unit(size_t(n) * 4);
    for (uint32_t i = 0; i < n; ++i) {
      unit[size_t(i) * 4 + 0] = unit[size_t(i) * 4 + 1] = unit[size_t(i) * 4 + 2] = 1.f;
      std::memcpy(&unit[size_t(i) * 4 + 3], &i, 4);
    }
    float4* contrib = ctx->scratch_hits.as<float4>();
    float4* result = ctx->scratch_occ.as<float4>();
    HIPC(hipMemcpyAsync(contrib, unit.data(), size_t(n) * 16, hipMemcpyHostToDevice, ctx->stream), "upload contrib");
    HIPC(hipMemsetAsync(result, 0, size_t(n) * 16, ctx->stream), "clear result");
    launch_shadow(acc, ShadowQueue{o_tmin, d_tmax, contrib}, result, in, ctr, ctx->grid_trace, false, ctx->stream);
  } else {
    launch_extend(acc, RayQueue{o_tmin, d_tmax, nullptr, nullptr},
                  HitBuf{ctx->scratch_hits.as<float4>(), ctx->scratch_inst.as<uint32_t>()}, in, ctr, ctx->grid_trace,
                  false, ctx->stream);
  }
  HIPC(hipGetLastError(), "trace launch");
  HIPC(hipStreamSynchronize(ctx->stream), "trace");
  uint32_t overflow = 0;
  HIPC(hipMemcpy(&overflow, ctr + kCtrOverflow, 4, hipMemcpyDeviceToHost), "read overflow");
  if (overflow) return fail(ctx, TRT_DEVICE, "traversal stack overflow");
  if (any) {
    std::vector<float> r(size_t(n) * 4);
    HIPC(hipMemcpy(r.data(), ctx->scratch_occ.p, size_t(n) * 16, hipMemcpyDeviceToHost), "read occlusion");
    for (uint32_t i = 0; i < n; ++i) occ[i] = r[size_t(i) * 4] == 0.f ? 1u : 0u;
  } else {
    HIPC(hipMemcpy(hits, ctx->scratch_hits.p, size_t(n) * 16, hipMemcpyDeviceToHost), "read hits");
    if (inst) {
      if (ctx->two_level) {
        HIPC(hipMemcpy(inst, ctx->scratch_inst.p, size_t(n) * 4, hipMemcpyDeviceToHost), "read hits");
      } else {
        for (uint32_t i = 0; i < n; ++i) inst[i] = kNoInstance;
      }
    }
  }
  return TRT_OK;
}

trt_status trt_trace_closest(trt_context* ctx, const float* rays, uint32_t n, float* hits, uint32_t* inst) {
  if (!hits) return TRT_INVALID_ARG;
  return trace_common(ctx, rays, n, hits, inst, nullptr, false);
}

}  // extern "C"

// ---------------------------------------------------------------------------------
// Allocation-free closest-hit queries (trt_trace_closest_noalloc): the per-call staging
// of trt_trace_closest (ray records, their structure-of-arrays form, hit records) lives
// in one file-scope device array per GPU instead of context buffers grown on demand.
// ---------------------------------------------------------------------------------
namespace {

constexpr uint32_t kNoAllocMaxRays = TRT_NOALLOC_MAX_RAYS;
// per ray: 32 B caller record, 16 + 16 B origin/direction rows, 16 B hit, 4 B instance
constexpr size_t kNoAllocBytes = size_t(kNoAllocMaxRays) * 96;
static __device__ __align__(256) uint8_t scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch[kMaxDevices] = {};

// The device address of `scratch` on the current device (`device`), looked up once per
// device (each GPU holds its own copy of the array).
hipError_t scratch_ptr(int device, uint8_t** out) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch[device] = p;
  }
  *out = static_cast<uint8_t*>(g_scratch[device]);
  return hipSuccess;
}

// Ray records (8 floats: origin, tmin, direction, tmax) -> the (origin, tmin) and
// (direction, tmax) rows of a depth-0 queue whose slot i is ray i, and the queue's
// per-segment counts (kSegs segments of `per` slots).
__global__ void k_noalloc_prepare(const float4* __restrict__ rec, uint32_t n, uint32_t per, float4* o_tmin,
                                  float4* d_tmax, uint32_t* seg_cnt) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) {
    o_tmin[i] = rec[2 * size_t(i)];
    d_tmax[i] = rec[2 * size_t(i) + 1];
  }
  if (blockIdx.x == 0 && threadIdx.x < kSegs) {
    const uint64_t first = uint64_t(threadIdx.x) * per;
    seg_cnt[threadIdx.x * kLine] = first >= n ? 0u : uint32_t(n - first < per ? n - first : per);
  }
}

}  // namespace

extern "C" {

trt_status trt_trace_closest_noalloc(trt_context* ctx, const float* rays, uint32_t n, float* hits, uint32_t* inst) {
  if (!ctx || (n && (!rays || !hits))) return TRT_INVALID_ARG;
  if (!ctx->built) return fail(ctx, TRT_NOT_BUILT, "scene not built");
  if (n > kNoAllocMaxRays)
    return fail(ctx, TRT_INVALID_ARG, "trt_trace_closest_noalloc takes at most " + std::to_string(kNoAllocMaxRays) +
                                          " rays per call");
  if (n == 0) return TRT_OK;
  for (uint32_t i = 0; i < n; ++i) {
    const float* r = &rays[size_t(i) * 8];
    if (!finite_all(r, 3) || !finite_all(r + 4, 3) || r[3] != r[3] || r[7] != r[7])
      return fail(ctx, TRT_INVALID_ARG, "ray " + std::to_string(i) + " has a non-finite origin or direction or a NaN bound");
    if (!(r[3] >= 0.f)) return fail(ctx, TRT_INVALID_ARG, "ray " + std::to_string(i) + " has a negative tmin");
  }
  HIPC(hipSetDevice(ctx->device), "hipSetDevice");
  uint8_t* base = nullptr;
  HIPC(scratch_ptr(ctx->device, &base), "scratch address");
  float4* rec = reinterpret_cast<float4*>(base);
  float4* o_tmin = rec + 2 * size_t(kNoAllocMaxRays);
  float4* d_tmax = o_tmin + kNoAllocMaxRays;
  float4* thit = d_tmax + kNoAllocMaxRays;
  uint32_t* tinst = reinterpret_cast<uint32_t*>(thit + kNoAllocMaxRays);
  hipStream_t stm = ctx->stream;
  uint32_t* ctr = ctx->counters.as<uint32_t>();
  HIPC(hipMemcpyAsync(rec, rays, size_t(n) * 32, hipMemcpyHostToDevice, stm), "upload rays");
  launch_zero_counters(ctr, stm);
  const uint32_t per = (n + kSegs - 1) / kSegs;
  hipLaunchKernelGGL(k_noalloc_prepare, dim3((n + 255) / 256), dim3(256), 0, stm, rec, n, per, o_tmin, d_tmax,
                     ctr + seg_ctr(kSegExt, 0));
  const SegIn in{ctr + seg_ctr(kSegExt, 0), ctr + seg_ctr(kSegWorkExt, 0), per};
  launch_extend(accel_of(ctx), RayQueue{o_tmin, d_tmax, nullptr, nullptr}, HitBuf{thit, tinst}, in, ctr,
                ctx->grid_trace, false, stm);
  HIPC(hipGetLastError(), "trace launch");
  uint32_t overflow = 0;
  HIPC(hipMemcpyAsync(hits, thit, size_t(n) * 16, hipMemcpyDeviceToHost, stm), "read hits");
  if (inst && ctx->two_level) HIPC(hipMemcpyAsync(inst, tinst, size_t(n) * 4, hipMemcpyDeviceToHost, stm), "read hits");
  HIPC(hipMemcpyAsync(&overflow, ctr + kCtrOverflow, 4, hipMemcpyDeviceToHost, stm), "read overflow");
  HIPC(hipStreamSynchronize(stm), "trace");
  if (overflow) return fail(ctx, TRT_DEVICE, "traversal stack overflow");
  if (inst && !ctx->two_level)
    for (uint32_t i = 0; i < n; ++i) inst[i] = kNoInstance;
  return TRT_OK;
}

trt_status trt_trace_any(trt_context* ctx, const float* rays, uint32_t n, uint8_t* occluded) {
  if (!occluded) return TRT_INVALID_ARG;
  return trace_common(ctx, rays, n, nullptr, nullptr, occluded, true);
}

trt_status trt_pick(trt_context* ctx, uint32_t width, uint32_t height, uint32_t x, uint32_t y, trt_pick_result* out) {
  if (!ctx || !out) return TRT_INVALID_ARG;
  if (!ctx->camera_set) return fail(ctx, TRT_INVALID_ARG, "camera not set");
  if (x >= width || y >= height) return fail(ctx, TRT_INVALID_ARG, "pixel outside the film");
  V3 o, d;
  camera_ray_from(ctx->cam, x, y, width, height, 0.5f, 0.5f, 0.5f, 0.5f, 0.f, o, d);
  const float ray[8] = {o.x, o.y, o.z, 0.f, d.x, d.y, d.z, 1e30f};
  float hit[4];
  uint32_t inst = kNoInstance;
  trt_status s = trt_trace_closest(ctx, ray, 1, hit, &inst);
  if (s != TRT_OK) return s;
  out->origin[0] = o.x; out->origin[1] = o.y; out->origin[2] = o.z;
  out->direction[0] = d.x; out->direction[1] = d.y; out->direction[2] = d.z;
  std::memcpy(&out->primitive, &hit[3], 4);
  out->t = out->primitive == kNoHit ? 1e30f : hit[0];
  out->instance = inst;
  return TRT_OK;
}


trt_status trt_set_material_params(trt_context* ctx, const float* params, uint32_t n, uint32_t which) {
  if (!ctx || (n && !params)) return TRT_INVALID_ARG;
  if (!ctx->built) return fail(ctx, TRT_NOT_BUILT, "scene not built");
  if (n != ctx->materials.size()) return fail(ctx, TRT_INVALID_ARG, "n must equal the number of materials");
  if (which < 1 || which > 3) return fail(ctx, TRT_INVALID_ARG, "which must be 1 (evaluation), 2 (sampling) or 3");
  if (!finite_all(params, size_t(n) * 5)) return fail(ctx, TRT_INVALID_ARG, "non-finite parameter");
  HIPC(hipSetDevice(ctx->device), "hipSetDevice");
  if (ctx->sampling_materials.empty()) ctx->sampling_materials = ctx->materials;
  auto apply = [&](std::vector<Material>& mats) {
    for (uint32_t i = 0; i < n; ++i) {
      for (int k = 0; k < 3; ++k) mats[i].base[k] = params[size_t(i) * 5 + k];
      mats[i].roughness = params[size_t(i) * 5 + 3];
      mats[i].metallic = params[size_t(i) * 5 + 4];
    }
  };
  if (which & 1) apply(ctx->materials);
  if (which & 2) apply(ctx->sampling_materials);
  HIPC(upload(ctx->d_materials, ctx->materials.data(), ctx->materials.size() * sizeof(Material), ctx->stream),
       "upload materials");
  HIPC(upload(ctx->d_mat_sampling, ctx->sampling_materials.data(), ctx->sampling_materials.size() * sizeof(Material),
              ctx->stream),
       "upload sampling materials");
  HIPC(hipStreamSynchronize(ctx->stream), "upload materials");
  reset_accumulation(ctx);
  invalidate_frame(ctx);
  return TRT_OK;
}

trt_status trt_backward(trt_context* ctx, const trt_render_opts* o, const float* dldi, float* grad) {
  trt_status s = check_render(ctx, o);
  if (s != TRT_OK) return s;
  if (!dldi || !grad) return TRT_INVALID_ARG;
  const trt_render_opts& L = ctx->last_opts;
  if (!ctx->has_frame)
    return fail(ctx, TRT_INVALID_ARG,
                "trt_backward replays the last trt_render, and the scene, camera, sky or materials changed since "
                "(or nothing was rendered): render the frame again first");
  if (L.width != o->width || L.height != o->height || L.spp != o->spp || L.max_depth != o->max_depth ||
      L.rr_depth != o->rr_depth || L.seed != o->seed || L.ray_epsilon != o->ray_epsilon || L.clamp != o->clamp ||
      L.mode != o->mode)
    return fail(ctx, TRT_INVALID_ARG, "trt_backward replays the last trt_render: call it with the same options");
  if (o->clamp != 0.f) return fail(ctx, TRT_UNSUPPORTED, "the adjoint needs clamp = 0");
  if (o->mode != kModePathTracing) return fail(ctx, TRT_UNSUPPORTED, "the adjoint needs the path-tracing mode");
  HIPC(hipSetDevice(ctx->device), "hipSetDevice");
  const uint32_t W = o->width, H = o->height, D = std::min<uint32_t>(o->max_depth, kMaxDepth);
  const uint64_t n = uint64_t(W) * H * o->spp;
  const uint64_t R = n * D;
  const uint32_t nm = uint32_t(ctx->materials.size());
  HIPC(ctx->a_lrem.ensure(n * 16), "alloc adjoint");
  HIPC(ctx->a_pend.ensure(n * 64), "alloc adjoint");
  HIPC(ctx->a_key.ensure(R * 4), "alloc adjoint");
  HIPC(ctx->a_g.ensure(R * 4 * kNumParams), "alloc adjoint");
  HIPC(ctx->a_sga.ensure(ctx->queue_slots * 16), "alloc adjoint");  // per shadow-queue slot
  HIPC(ctx->a_sgb.ensure(ctx->queue_slots * 4), "alloc adjoint");
  HIPC(ctx->a_dldi.ensure(uint64_t(W) * H * 12), "alloc adjoint");
  HIPC(ctx->a_partial.ensure(size_t(kGradBlocks) * kGradMatsPerPass * kNumParams * 4), "alloc adjoint");
  HIPC(ctx->a_out.ensure(size_t(nm) * kNumParams * 4 + 16), "alloc adjoint");
  HIPC(hipMemcpyAsync(ctx->a_dldi.p, dldi, uint64_t(W) * H * 12, hipMemcpyHostToDevice, ctx->stream), "upload dL/dI");

  RenderParams rp = params_of_opts(ctx, *o);
  rp.sample_offset = ctx->last_sample_offset;
  const DevAccel acc = accel_of(ctx);
  const SceneView sv = view_of(ctx);
  uint32_t* ctr = ctx->counters.as<uint32_t>();
  AdjBuffers ab;
  ab.dldi = ctx->a_dldi.as<float>();
  ab.lrem = ctx->a_lrem.as<float4>();
  ab.pend = ctx->a_pend.as<float4>();
  ab.rec_key = ctx->a_key.as<uint32_t>();
  ab.rec_g = ctx->a_g.as<float>();
  ab.s_ga = ctx->a_sga.as<float4>();
  ab.s_gb = ctx->a_sgb.as<float>();
  ab.n_records = R;
  ab.depth = D;
  RayQueue q[2];
  for (int i = 0; i < 2; ++i) {
    q[i].o_tmin = ctx->q_o[i].as<float4>();
    q[i].d_tmax = ctx->q_d[i].as<float4>();
    q[i].t_pdf = ctx->q_t[i].as<float4>();
    q[i].meta = ctx->q_m[i].as<uint32_t>();
  }
  HitBuf hb{ctx->h_t.as<float4>(), ctx->h_i.as<uint32_t>()};
  ShadowQueue sq{ctx->s_o[0].as<float4>(), ctx->s_d[0].as<float4>(), ctx->s_c[0].as<float4>(), nullptr};
  hipStream_t stm = ctx->stream;
  const uint32_t cap = ctx->seg_cap;
  launch_zero_counters(ctr, stm);
  launch_raygen_adj(rp, uint32_t(n), cap, q[0], ctx->radiance.as<float4>(), ab, ctr, stm);
  for (uint32_t d = 0; d < D; ++d) {
    const int cur = int(d & 1u), nxt = cur ^ 1;
    const int di = int(d);
    launch_extend(acc, q[cur], hb, SegIn{ctr + seg_ctr(kSegExt, di), ctr + seg_ctr(kSegWorkExt, di), cap}, ctr,
                  ctx->grid_trace, false, stm);
    launch_shade_adj(sv, sampling_mats(ctx), rp, q[cur], hb, q[nxt], sq, ab,
                     SegIn{ctr + seg_ctr(kSegExt, di), ctr + seg_ctr(kSegWorkShade, di), cap},
                     ctr + seg_ctr(kSegExt, di + 1), ctr + seg_ctr(kSegShadow, di), ctr, d, ctx->grid_shade, stm);
    launch_shadow_adj(acc, sq, ab, SegIn{ctr + seg_ctr(kSegShadow, di), ctr + seg_ctr(kSegWorkShadow, di), cap}, ctr,
                      d, ctx->grid_trace, stm);
  }
  for (uint32_t lo = 0; lo < nm; lo += kGradMatsPerPass) {
    const uint32_t cnt = std::min(kGradMatsPerPass, nm - lo);
    launch_grad_reduce(ab, lo, cnt, ctx->a_partial.as<float>(), ctx->a_out.as<float>() + size_t(lo) * kNumParams, stm);
  }
  HIPC(hipGetLastError(), "adjoint launch");
  HIPC(hipStreamSynchronize(stm), "adjoint");
  uint32_t overflow = 0;
  HIPC(hipMemcpy(&overflow, ctr + kCtrOverflow, 4, hipMemcpyDeviceToHost), "read overflow");
  if (overflow) return fail(ctx, TRT_DEVICE, "traversal stack overflow");
  HIPC(hipMemcpy(grad, ctx->a_out.p, size_t(nm) * kNumParams * 4, hipMemcpyDeviceToHost), "read gradient");
  return TRT_OK;
}


// ---------------------------------------------------------------------------------
// Denoiser
// ---------------------------------------------------------------------------------
static uint16_t host_f2bf(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return uint16_t(u >> 16);
}

static trt_status need_unet(trt_context* ctx) {
  if (!ctx) return TRT_INVALID_ARG;
  if (!ctx->unet) return fail(ctx, TRT_NOT_BUILT, "denoiser not initialised (trt_denoiser_init)");
  return TRT_OK;
}

trt_status trt_denoiser_init(trt_context* ctx, uint32_t seed) {
  if (!ctx) return TRT_INVALID_ARG;
  HIPC(hipSetDevice(ctx->device), "hipSetDevice");
  if (!ctx->unet) ctx->unet.reset(new Unet());
  HIPC(ctx->unet->init(seed, ctx->stream), "denoiser init");
  return TRT_OK;
}

uint64_t trt_denoiser_num_params(void) {
  static const uint64_t n = Unet().num_params();
  return n;
}

trt_status trt_denoiser_get_params(trt_context* ctx, float* out, uint64_t n) {
  trt_status s = need_unet(ctx);
  if (s != TRT_OK) return s;
  if (!out || n != ctx->unet->num_params()) return fail(ctx, TRT_INVALID_ARG, "n must equal trt_denoiser_num_params()");
  HIPC(ctx->unet->get_params(out, ctx->stream), "denoiser params");
  return TRT_OK;
}

trt_status trt_denoiser_set_params(trt_context* ctx, const float* in, uint64_t n) {
  if (!ctx || !in) return TRT_INVALID_ARG;
  if (!ctx->unet) {
    trt_status s = trt_denoiser_init(ctx, 0);
    if (s != TRT_OK) return s;
  }
  if (n != ctx->unet->num_params()) return fail(ctx, TRT_INVALID_ARG, "n must equal trt_denoiser_num_params()");
  if (!finite_all(in, n)) return fail(ctx, TRT_INVALID_ARG, "non-finite denoiser parameter");
  HIPC(ctx->unet->set_params(in, ctx->stream), "denoiser params");
  return TRT_OK;
}

// Own versioned format: magic "TRTDN001", uint64 count, count float32 values.
trt_status trt_denoiser_save(trt_context* ctx, const char* path) {
  trt_status s = need_unet(ctx);
  if (s != TRT_OK) return s;
  if (!path) return TRT_INVALID_ARG;
  std::vector<float> w(ctx->unet->num_params());
  HIPC(ctx->unet->get_params(w.data(), ctx->stream), "denoiser params");
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(ctx, TRT_INVALID_ARG, std::string("cannot open ") + path);
  const uint64_t n = w.size();
  bool ok = std::fwrite("TRTDN001", 1, 8, f) == 8 && std::fwrite(&n, 8, 1, f) == 1 &&
            std::fwrite(w.data(), 4, w.size(), f) == w.size();
  ok = (std::fclose(f) == 0) && ok;
  return ok ? TRT_OK : fail(ctx, TRT_DEVICE, "write failed");
}

trt_status trt_denoiser_load(trt_context* ctx, const char* path) {
  if (!ctx || !path) return TRT_INVALID_ARG;
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(ctx, TRT_INVALID_ARG, std::string("cannot open ") + path);
  char magic[8];
  uint64_t n = 0;
  bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, "TRTDN001", 8) == 0 && std::fread(&n, 8, 1, f) == 1 &&
            n == trt_denoiser_num_params();
  std::vector<float> w(ok ? n : 0);
  ok = ok && std::fread(w.data(), 4, n, f) == n;
  std::fclose(f);
  if (!ok) return fail(ctx, TRT_INVALID_ARG, "not a denoiser weight file of this network");
  return trt_denoiser_set_params(ctx, w.data(), n);
}

trt_status trt_denoiser_dataset(trt_context* ctx, uint32_t n, uint32_t width, uint32_t height) {
  trt_status s = need_unet(ctx);
  if (s != TRT_OK) return s;
  if (n == 0 || width == 0 || height == 0) return fail(ctx, TRT_INVALID_ARG, "empty dataset");
  const uint64_t P = uint64_t(n) * width * height;
  HIPC(ctx->dn_ds_feat.ensure(P * kFeatures * 2), "alloc dataset");
  HIPC(ctx->dn_ds_tgt.ensure(P * 3 * 4), "alloc dataset");
  ctx->dn_ds_n = n;
  ctx->dn_ds_w = width;
  ctx->dn_ds_h = height;
  ctx->dn_ds_have.assign(n, 0);
  return TRT_OK;
}

trt_status trt_denoiser_capture(trt_context* ctx, uint32_t slot, uint32_t kind) {
  trt_status s = need_unet(ctx);
  if (s != TRT_OK) return s;
  if (slot >= ctx->dn_ds_n || kind > 1) return fail(ctx, TRT_INVALID_ARG, "slot or kind out of range");
  if (ctx->width != ctx->dn_ds_w || ctx->height != ctx->dn_ds_h || ctx->sample_count == 0)
    return fail(ctx, TRT_INVALID_ARG, "render a frame of the dataset's size first");
  const uint64_t P = uint64_t(ctx->width) * ctx->height;
  if (kind == 0)
    dn_features(ctx->color.as<float4>(), ctx->albedo.as<float4>(), ctx->normal.as<float4>(),
                ctx->dn_ds_feat.as<uint16_t>() + size_t(slot) * P * kFeatures, P, ctx->stream);
  else
    dn_target(ctx->color.as<float4>(), ctx->dn_ds_tgt.as<float>() + size_t(slot) * P * 3, P, ctx->stream);
  HIPC(hipGetLastError(), "capture");
  HIPC(hipStreamSynchronize(ctx->stream), "capture");
  ctx->dn_ds_have[slot] |= uint8_t(1u << kind);
  return TRT_OK;
}

trt_status trt_denoiser_train(trt_context* ctx, uint32_t steps, uint32_t batch, uint32_t crop, float lr, uint32_t seed,
                              float* losses) {
  trt_status s = need_unet(ctx);
  if (s != TRT_OK) return s;
  if (ctx->dn_ds_n == 0) return fail(ctx, TRT_INVALID_ARG, "no dataset");
  for (uint8_t h : ctx->dn_ds_have)
    if (h != 3) return fail(ctx, TRT_INVALID_ARG, "every dataset slot needs an input and a target");
  if (batch == 0 || crop % 128 || crop > ctx->dn_ds_w || crop > ctx->dn_ds_h || !(lr > 0.f))
    return fail(ctx, TRT_INVALID_ARG, "crop must be a multiple of 128 within the dataset images; batch, lr > 0");
  HIPC(hipSetDevice(ctx->device), "hipSetDevice");
  const uint64_t P = uint64_t(batch) * crop * crop;
  HIPC(ctx->dn_batch_feat.ensure(P * kFeatures * 2), "alloc batch");
  HIPC(ctx->dn_batch_tgt.ensure(P * 3 * 4), "alloc batch");
  HIPC(ctx->dn_crops.ensure(size_t(batch) * 16), "alloc batch");
  std::mt19937 rng(seed);
  std::vector<int> crops(size_t(batch) * 4);
  for (uint32_t it = 0; it < steps; ++it) {
    for (uint32_t b = 0; b < batch; ++b) {
      crops[b * 4 + 0] = int(rng() % ctx->dn_ds_n);
      crops[b * 4 + 1] = int(rng() % (ctx->dn_ds_w - crop + 1));
      crops[b * 4 + 2] = int(rng() % (ctx->dn_ds_h - crop + 1));
      crops[b * 4 + 3] = 0;
    }
    HIPC(hipMemcpyAsync(ctx->dn_crops.p, crops.data(), crops.size() * 4, hipMemcpyHostToDevice, ctx->stream),
         "upload crops");
    dn_gather_crop(ctx->dn_ds_feat.as<uint16_t>(), ctx->dn_ds_tgt.as<float>(), int(ctx->dn_ds_h), int(ctx->dn_ds_w),
                   ctx->dn_crops.as<int4>(), int(batch), int(crop), ctx->dn_batch_feat.as<uint16_t>(),
                   ctx->dn_batch_tgt.as<float>(), ctx->stream);
    float loss = 0.f;
    HIPC(ctx->unet->train_step(ctx->dn_batch_feat.as<uint16_t>(), ctx->dn_batch_tgt.as<float>(), int(batch), int(crop),
                               int(crop), lr, ctx->stream, &loss),
         "train step");
    if (losses) losses[it] = loss;
  }
  return TRT_OK;
}

static trt_status denoise_impl(trt_context* ctx, trt_denoise_stats* st, float blend);

trt_status trt_denoise(trt_context* ctx, trt_denoise_stats* st) { return denoise_impl(ctx, st, 0.f); }

trt_status trt_denoiser_set_policy(trt_context* ctx, uint32_t enabled, uint32_t sample_threshold) {
  if (!ctx) return TRT_INVALID_ARG;
  ctx->dn_enabled = enabled ? 1u : 0u;
  ctx->dn_threshold = sample_threshold;
  return TRT_OK;
}

trt_status trt_denoise_frame(trt_context* ctx, trt_denoise_stats* st, uint32_t* ran) {
  if (!ctx) return TRT_INVALID_ARG;
  if (ran) *ran = 0;
  if (st) *st = trt_denoise_stats{};
  // Renderer::ShouldDenoise: enabled, path tracing or ambient occlusion, enough samples
  const uint32_t mode = ctx->last_opts.mode;
  const bool mode_ok = mode == kModePathTracing || mode == kModeAmbientOcclusion || mode == kModeAmbientOcclusionShading;
  if (!ctx->dn_enabled || !mode_ok || ctx->sample_count == 0 || ctx->sample_count < ctx->dn_threshold) return TRT_OK;
  // Renderer::DenoiseFrame: again once the sample count reaches 1.5 x the last denoised one
  const uint32_t again = ctx->dn_last_spp + (ctx->dn_last_spp >> 1);
  if (ctx->dn_last_spp != 0 && ctx->sample_count < again) return TRT_OK;
  // Denoiser::DenoiseFrame: blend factor 1 / (spp + 1) towards the noisy input
  const trt_status s = denoise_impl(ctx, st, 1.f / float(ctx->sample_count + 1));
  if (s != TRT_OK) return s;
  ctx->dn_last_spp = ctx->sample_count;
  if (ran) *ran = 1;
  return TRT_OK;
}

static trt_status denoise_impl(trt_context* ctx, trt_denoise_stats* st, float blend) {
  trt_status s = need_unet(ctx);
  if (s != TRT_OK) return s;
  if (ctx->width == 0 || ctx->sample_count == 0) return fail(ctx, TRT_INVALID_ARG, "nothing rendered yet");
  if (!ctx->unet->shape_ok(int(ctx->height), int(ctx->width)))
    return fail(ctx, TRT_UNSUPPORTED, "denoising needs width % 128 == 0 and height % 8 == 0");
  HIPC(hipSetDevice(ctx->device), "hipSetDevice");
  const uint64_t P = uint64_t(ctx->width) * ctx->height;
  HIPC(ctx->dn_feat.ensure(P * kFeatures * 2), "alloc features");
  HIPC(ctx->dn_denoised.ensure(P * 16), "alloc denoised");
  hipEvent_t* ev = ctx->ev;
  HIPC(hipEventRecord(ev[0], ctx->stream), "event");
  dn_features(ctx->color.as<float4>(), ctx->albedo.as<float4>(), ctx->normal.as<float4>(), ctx->dn_feat.as<uint16_t>(),
              P, ctx->stream);
  HIPC(hipEventRecord(ev[1], ctx->stream), "event");
  HIPC(ctx->unet->forward(ctx->dn_feat.as<uint16_t>(), 1, int(ctx->height), int(ctx->width), ctx->stream,
                          ctx->dn_denoised.as<float4>()),
       "forward");
  if (blend > 0.f) dn_blend(ctx->dn_denoised.as<float4>(), ctx->color.as<float4>(), blend, P, ctx->stream);
  HIPC(hipEventRecord(ev[2], ctx->stream), "event");
  HIPC(hipStreamSynchronize(ctx->stream), "denoise");
  if (st) {
    (void)hipEventElapsedTime(&st->features_ms, ev[0], ev[1]);
    (void)hipEventElapsedTime(&st->network_ms, ev[1], ev[2]);
    (void)hipEventElapsedTime(&st->total_ms, ev[0], ev[2]);
  }
  return TRT_OK;
}

trt_status trt_tonemap(trt_context* ctx, const trt_tonemap_opts* opts, uint8_t* rgba8) {
  if (!ctx || !opts || !rgba8) return TRT_INVALID_ARG;
  if (!(opts->exposure >= 0.f) || !(opts->exposure < 1e30f) || !(opts->gamma > 0.f) || !(opts->gamma < 1e6f))
    return fail(ctx, TRT_INVALID_ARG, "exposure must be >= 0 and gamma > 0");
  const uint64_t npix = uint64_t(ctx->width) * ctx->height;
  if (npix == 0) return fail(ctx, TRT_NOT_BUILT, "no frame");
  const DevBuf* src = &ctx->color;
  if (opts->source == 1) {
    if (!ctx->dn_denoised.p || ctx->dn_denoised.bytes < npix * 16) return fail(ctx, TRT_NOT_BUILT, "no denoised frame");
    src = &ctx->dn_denoised;
  } else if (opts->source != 0) {
    return fail(ctx, TRT_INVALID_ARG, "source must be 0 (colour) or 1 (denoised)");
  }
  HIPC(hipSetDevice(ctx->device), "hipSetDevice");
  HIPC(ctx->tonemapped.ensure(npix * 4), "alloc tonemap");
  TonemapParams tp{opts->method, opts->exposure, opts->gamma};
  launch_tonemap(src->as<float4>(), uint32_t(npix), tp, ctx->tonemapped.as<uchar4>(), ctx->stream);
  HIPC(hipGetLastError(), "tonemap launch");
  HIPC(hipMemcpyAsync(rgba8, ctx->tonemapped.p, npix * 4, hipMemcpyDeviceToHost, ctx->stream), "read tonemap");
  HIPC(hipStreamSynchronize(ctx->stream), "tonemap");
  return TRT_OK;
}

trt_status trt_read_denoised(trt_context* ctx, float* rgba) {
  if (!ctx || !rgba) return TRT_INVALID_ARG;
  if (!ctx->dn_denoised.p || ctx->dn_denoised.bytes < size_t(ctx->width) * ctx->height * 16)
    return fail(ctx, TRT_NOT_BUILT, "no denoised frame");
  return read_pixels(ctx, ctx->dn_denoised, rgba);
}

// Test hooks on host arrays: features fp32 [N][H][W][16] (rounded to bf16 here),
// output fp32 [N][H][W][32] (log space; channels >= 3 are padding).
trt_status trt_denoiser_forward(trt_context* ctx, const float* features, uint32_t N, uint32_t H, uint32_t W,
                                float* out) {
  trt_status s = need_unet(ctx);
  if (s != TRT_OK) return s;
  if (!features || !out || N == 0 || !ctx->unet->shape_ok(int(H), int(W)))
    return fail(ctx, TRT_INVALID_ARG, "need width % 128 == 0, height % 8 == 0");
  HIPC(hipSetDevice(ctx->device), "hipSetDevice");
  const uint64_t P = uint64_t(N) * H * W;
  std::vector<uint16_t> hb(P * kFeatures);
  for (uint64_t i = 0; i < hb.size(); ++i) hb[i] = host_f2bf(features[i]);
  HIPC(ctx->dn_io.ensure(P * kFeatures * 2), "alloc");
  HIPC(hipMemcpyAsync(ctx->dn_io.p, hb.data(), hb.size() * 2, hipMemcpyHostToDevice, ctx->stream), "upload");
  HIPC(ctx->unet->forward(ctx->dn_io.as<uint16_t>(), int(N), int(H), int(W), ctx->stream), "forward");
  HIPC(hipMemcpyAsync(out, ctx->unet->output(), P * 32 * 4, hipMemcpyDeviceToHost, ctx->stream), "read");
  HIPC(hipStreamSynchronize(ctx->stream), "forward");
  return TRT_OK;
}

// Inference path on host features (as trt_denoise runs it on the rendered buffers):
// radiance [N][H][W][4] = (expm1(max(feature c + output c, 0)) for c < 3, 1).
trt_status trt_denoiser_apply(trt_context* ctx, const float* features, uint32_t N, uint32_t H, uint32_t W,
                              float* radiance) {
  trt_status s = need_unet(ctx);
  if (s != TRT_OK) return s;
  if (!features || !radiance || N == 0 || !ctx->unet->shape_ok(int(H), int(W)))
    return fail(ctx, TRT_INVALID_ARG, "need width % 128 == 0, height % 8 == 0");
  HIPC(hipSetDevice(ctx->device), "hipSetDevice");
  const uint64_t P = uint64_t(N) * H * W;
  std::vector<uint16_t> hb(P * kFeatures);
  for (uint64_t i = 0; i < hb.size(); ++i) hb[i] = host_f2bf(features[i]);
  DevBuf out;
  HIPC(out.ensure(P * 16), "alloc");
  HIPC(ctx->dn_io.ensure(P * kFeatures * 2), "alloc");
  HIPC(hipMemcpyAsync(ctx->dn_io.p, hb.data(), hb.size() * 2, hipMemcpyHostToDevice, ctx->stream), "upload");
  HIPC(ctx->unet->forward(ctx->dn_io.as<uint16_t>(), int(N), int(H), int(W), ctx->stream, out.as<float4>()),
       "forward");
  HIPC(hipMemcpyAsync(radiance, out.p, P * 16, hipMemcpyDeviceToHost, ctx->stream), "read");
  HIPC(hipStreamSynchronize(ctx->stream), "apply");
  return TRT_OK;
}

// Post-activation output of convolution `index` (0..13) of the last forward, as floats.
trt_status trt_denoiser_activation(trt_context* ctx, uint32_t index, float* out, uint64_t n) {
  trt_status s = need_unet(ctx);
  if (s != TRT_OK) return s;
  const size_t m = ctx->unet->activation_elems(int(index));
  if (!out || m == 0 || n != m) return fail(ctx, TRT_INVALID_ARG, "index out of range or wrong element count");
  std::vector<uint16_t> h(m);
  HIPC(hipMemcpy(h.data(), ctx->unet->activation(int(index)), m * 2, hipMemcpyDeviceToHost), "read activation");
  for (size_t i = 0; i < m; ++i) {
    const uint32_t u = uint32_t(h[i]) << 16;
    std::memcpy(&out[i], &u, 4);
  }
  return TRT_OK;
}

// Gradient of sum(dout * output) w.r.t. the parameters for the last trt_denoiser_forward.
trt_status trt_denoiser_backward(trt_context* ctx, const float* dout, uint32_t N, uint32_t H, uint32_t W, float* grads) {
  trt_status s = need_unet(ctx);
  if (s != TRT_OK) return s;
  if (!dout || !grads) return TRT_INVALID_ARG;
  HIPC(hipSetDevice(ctx->device), "hipSetDevice");
  const uint64_t P = uint64_t(N) * H * W;
  DevBuf d;
  HIPC(d.ensure(P * 32 * 4), "alloc");
  HIPC(hipMemcpyAsync(d.p, dout, P * 32 * 4, hipMemcpyHostToDevice, ctx->stream), "upload");
  HIPC(ctx->unet->backward(d.as<float>(), int(N), int(H), int(W), ctx->stream), "backward");
  HIPC(ctx->unet->get_grads(grads, ctx->stream), "read grads");
  return TRT_OK;
}

}  // extern "C"
