// trt — C++ façade over the C ABI (include/trt.h), with the verbs of the reference's
// Tracer::Renderer (Tracer/Renderer/Renderer.h:34-171): UpdateScene, RenderFrame, Reset,
// SetCamera, SetRenderMode / SetMaterialPropertyId, Settings / SetSettings, Statistics,
// SampleCount, PickRay, and the denoiser (Renderer::GetDenoiser, Tracer/Optix/Denoiser.h:14-38).
// RenderFrame accumulates into the context's buffers instead of a GL texture; the frame
// is read back with Color() / Denoised(). Failures throw trt::Error (the reference shows
// a message box and exits, Tracer/Utility/Errors.cpp:15-37). Header only; link libtrt.so.
#ifndef TRT_HPP
#define TRT_HPP

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "trt.h"

namespace trt {

class Error : public std::runtime_error {
 public:
  Error(trt_status status, const std::string& msg) : std::runtime_error(msg), status_(status) {}
  trt_status status() const { return status_; }

 private:
  trt_status status_;
};

// One mesh of a scene (the reference's Model, Tracer/Resources/Model.h): triangles as 9
// floats each, optional per-vertex normals (9 floats per triangle), a material id per
// triangle (indexing Scene::materials, offset by an instance's material offset).
struct Mesh {
  std::vector<float> positions;
  std::vector<float> normals;  // empty: face normals
  std::vector<uint32_t> material_ids;
};

// The reference's Scene (Tracer/Renderer/Scene.h): meshes, materials, optional instances
// (none: every mesh is placed as given), optional sky.
struct Scene {
  std::vector<Mesh> meshes;
  std::vector<trt_material> materials;
  std::vector<trt_instance> instances;
  bool has_sky = false;
  trt_sky sky{};
};

// The reference's KernelSettings (Tracer/Common/CommonStructs.h:145-153) with the film size.
inline trt_render_opts DefaultSettings() {
  trt_render_opts o{};
  o.width = 1920;
  o.height = 1080;
  o.spp = 1;
  o.max_depth = 8;
  o.rr_depth = 3;
  o.seed = 0;
  o.ray_epsilon = 1e-3f;
  o.clamp = 0.f;
  o.mode = TRT_MODE_PATH_TRACING;
  o.mode_flags = 0;
  o.ao_dist = 1.f;
  o.zdepth_max = 1.f;
  return o;
}

class Renderer {
 public:
  explicit Renderer(int device = 0) {
    const trt_status s = trt_create(device, &ctx_);
    if (s != TRT_OK) throw Error(s, "trt_create failed (no usable HIP device?)");
  }
  ~Renderer() { trt_destroy(ctx_); }
  Renderer(const Renderer&) = delete;
  Renderer& operator=(const Renderer&) = delete;

  trt_context* Context() { return ctx_; }

  // Replaces the scene and builds its BVHs (Renderer::UpdateScene).
  void UpdateScene(const Scene& scene, const trt_build_opts* opts = nullptr) {
    check(trt_clear_scene(ctx_));
    check(trt_set_materials(ctx_, scene.materials.data(), uint32_t(scene.materials.size())));
    for (const Mesh& m : scene.meshes) {
      const uint32_t n = uint32_t(m.material_ids.size());
      if (m.positions.size() != size_t(n) * 9 || (!m.normals.empty() && m.normals.size() != size_t(n) * 9))
        throw Error(TRT_INVALID_ARG, "mesh arrays must hold 9 floats per triangle");
      uint32_t id = 0;
      check(trt_add_mesh(ctx_, m.positions.data(), m.normals.empty() ? nullptr : m.normals.data(),
                         m.material_ids.data(), n, &id));
    }
    if (!scene.instances.empty())
      check(trt_set_instances(ctx_, scene.instances.data(), uint32_t(scene.instances.size())));
    check(trt_build(ctx_, opts));
    if (scene.has_sky) check(trt_set_sky(ctx_, &scene.sky));
  }

  // Loads a scene file of the reference's JSON format (SceneFile::Load); its camera and
  // render settings (multisample, maxdepth, ray-epsilon) are applied.
  trt_scene_file_info LoadScene(const std::string& path, const trt_build_opts* opts = nullptr) {
    trt_scene_file_info info{};
    check(trt_load_scene_file(ctx_, path.c_str(), opts, &info));
    settings_.spp = info.spp;
    settings_.max_depth = info.max_depth;
    settings_.ray_epsilon = info.ray_epsilon;
    settings_.ao_dist = info.ao_dist;
    settings_.zdepth_max = info.zdepth_max;
    Reset();
    return info;
  }

  void SetCamera(const trt_camera& cam) { check(trt_set_camera(ctx_, &cam)); }

  void SetRenderMode(uint32_t mode) {
    settings_.mode = mode;
    Reset();
  }
  void SetMaterialPropertyId(uint32_t id) {
    settings_.mode_flags = id;
    Reset();
  }

  const trt_render_opts& Settings() const { return settings_; }
  void SetSettings(const trt_render_opts& s) {
    settings_ = s;
    Reset();
  }

  // Traces settings.spp samples per pixel and adds them to the accumulator.
  const trt_stats& RenderFrame() {
    check(trt_render(ctx_, &settings_, &stats_));
    return stats_;
  }
  void Reset() { check(trt_reset(ctx_)); }
  // Per-phase times in Statistics() (on by default; off, only counts and frame_ms).
  void SetPhaseTiming(bool on) { check(trt_set_phase_timing(ctx_, on ? 1 : 0)); }
  const trt_stats& Statistics() const { return stats_; }
  uint32_t SampleCount() const { return trt_sample_count(ctx_); }

  // Camera ray through the centre of pixel (x, y) and its closest hit (Renderer::PickRay).
  trt_pick_result PickRay(uint32_t x, uint32_t y) {
    trt_pick_result r{};
    check(trt_pick(ctx_, settings_.width, settings_.height, x, y, &r));
    return r;
  }

  // Frame buffers, width x height x 4 floats, row 0 at the bottom.
  std::vector<float> Color() { return pixels(trt_read_color); }
  std::vector<float> Accumulator() { return pixels(trt_read_accum); }
  std::vector<float> Albedo() { return pixels(trt_read_albedo); }
  std::vector<float> Normal() { return pixels(trt_read_normal); }
  std::vector<float> Denoised() { return pixels(trt_read_denoised); }

  // Denoiser (Renderer::GetDenoiser): weights, the reference's enable flag and sample
  // threshold, and the policy step Renderer::DenoiseFrame runs after every frame.
  void LoadDenoiser(const std::string& weights) { check(trt_denoiser_load(ctx_, weights.c_str())); }
  void SetDenoiser(bool enabled, uint32_t sample_threshold) {
    check(trt_denoiser_set_policy(ctx_, enabled ? 1u : 0u, sample_threshold));
  }
  // Returns whether the denoiser ran for the current sample count.
  bool DenoiseFrame(trt_denoise_stats* stats = nullptr) {
    uint32_t ran = 0;
    check(trt_denoise_frame(ctx_, stats, &ran));
    return ran != 0;
  }

 private:
  void check(trt_status s) const {
    if (s != TRT_OK) throw Error(s, trt_last_error(ctx_));
  }
  std::vector<float> pixels(trt_status (*fn)(trt_context*, float*)) {
    std::vector<float> out(size_t(settings_.width) * settings_.height * 4);
    check(fn(ctx_, out.data()));
    return out;
  }

  trt_context* ctx_ = nullptr;
  trt_render_opts settings_ = DefaultSettings();
  trt_stats stats_{};
};

}  // namespace trt

#endif  // TRT_HPP
