// C ABI of the scene-file loader (trt_scene_file_read, trt_load_scene_file): the host
// parse of scene_file.cpp, handed to the context through the public scene entry points.
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/trt.h"
#include "api_internal.h"
#include "scene_file.h"

namespace {

constexpr float kPi = 3.14159265358979323846f;

trt_material to_trt(const trt::FileMaterial& m) {
  trt_material o;
  std::memset(&o, 0, sizeof(o));
  for (int k = 0; k < 3; ++k) {
    o.base_color[k] = m.diffuse[k];
    o.emission[k] = m.emissive[k] * m.diffuse[k];
  }
  o.roughness = m.roughness;
  o.metallic = m.metallic;
  o.specular = m.specular;
  o.specular_tint = m.specular_tint;
  o.sheen = m.sheen;
  o.sheen_tint = m.sheen_tint;
  o.clearcoat = m.clearcoat;
  o.clearcoat_gloss = m.clearcoat_gloss;
  o.subsurface = m.subsurface;
  o.anisotropic = m.anisotropic;
  o.model = 0;
  return o;
}

void fill_info(const trt::SceneFile& sf, trt_scene_file_info* info) {
  std::memset(info, 0, sizeof(*info));
  info->n_models = uint32_t(sf.models.size());
  info->n_instances = uint32_t(sf.instances.size());
  for (const trt::FileModel& m : sf.models) {
    info->n_materials += uint32_t(m.materials.size());
    info->n_triangles += uint32_t(m.mat.size());
  }
  info->has_camera = sf.has_camera ? 1u : 0u;
  for (int k = 0; k < 3; ++k) {
    info->camera.position[k] = sf.cam_pos[k];
    info->camera.target[k] = sf.cam_target[k];
    info->camera.up[k] = sf.cam_up[k];
    info->sky_dir[k] = sf.sky_dir[k];
  }
  info->camera.fov = sf.fov_deg * (kPi / 180.f);
  info->camera.aperture = sf.aperture;
  info->camera.focal_dist = sf.focal_dist;
  info->camera.distortion = sf.distortion;
  info->camera.bokeh_sides = sf.bokeh_sides;
  info->camera.bokeh_rotation = sf.bokeh_rotation;
  info->spp = sf.spp;
  info->max_depth = sf.max_depth;
  info->ray_epsilon = sf.ray_epsilon;
  info->ao_dist = sf.ao_dist;
  info->zdepth_max = sf.zdepth_max;
  info->sky_enabled = sf.sky_enabled ? 1u : 0u;
  info->sky_draw_sun = sf.sky_draw_sun ? 1u : 0u;
  info->sky_angular_diameter = sf.sky_angular_diameter;
  info->sky_intensity = sf.sky_intensity;
  info->sky_turbidity = sf.sky_turbidity;
  info->sky_bias = sf.sky_bias;
  info->n_warnings = uint32_t(sf.warnings.size());
}

void put_msg(char* msg, size_t len, const std::string& s) {
  if (!msg || len == 0) return;
  const size_t n = s.size() < len - 1 ? s.size() : len - 1;
  std::memcpy(msg, s.data(), n);
  msg[n] = 0;
}

std::string joined(const std::vector<std::string>& w) {
  std::string s;
  for (const std::string& x : w) s += x + "\n";
  return s;
}

}  // namespace

extern "C" {

trt_status trt_scene_file_read(const char* path, trt_scene_file_info* info, float* positions, float* normals,
                               uint32_t* mat_ids, uint32_t* model_tri_offset, uint32_t* model_tri_count,
                               uint32_t* model_mat_offset, trt_material* materials, trt_instance* instances,
                               char* msg, size_t msg_len) {
  if (!path || !info) return TRT_INVALID_ARG;
  trt::SceneFile sf;
  std::string err;
  if (!trt::load_scene_file(path, sf, err)) {
    put_msg(msg, msg_len, err);
    return TRT_INVALID_ARG;
  }
  fill_info(sf, info);
  uint32_t tri = 0, mat = 0;
  std::vector<uint32_t> mat_base(sf.models.size());
  for (size_t mi = 0; mi < sf.models.size(); ++mi) {
    const trt::FileModel& m = sf.models[mi];
    const uint32_t nt = uint32_t(m.mat.size());
    if (positions) std::memcpy(positions + size_t(tri) * 9, m.pos.data(), size_t(nt) * 36);
    if (normals) std::memcpy(normals + size_t(tri) * 9, m.nrm.data(), size_t(nt) * 36);
    if (mat_ids) std::memcpy(mat_ids + tri, m.mat.data(), size_t(nt) * 4);
    if (model_tri_offset) model_tri_offset[mi] = tri;
    if (model_tri_count) model_tri_count[mi] = nt;
    if (model_mat_offset) model_mat_offset[mi] = mat;
    mat_base[mi] = mat;
    if (materials)
      for (size_t k = 0; k < m.materials.size(); ++k) materials[mat + k] = to_trt(m.materials[k]);
    tri += nt;
    mat += uint32_t(m.materials.size());
  }
  if (instances)
    for (size_t i = 0; i < sf.instances.size(); ++i) {
      std::memcpy(instances[i].transform, sf.instances[i].rows, 48);
      instances[i].mesh = sf.instances[i].model;
      instances[i].material_offset = mat_base[sf.instances[i].model];
    }
  put_msg(msg, msg_len, joined(sf.warnings));
  return TRT_OK;
}

trt_status trt_load_scene_file(trt_context* ctx, const char* path, const trt_build_opts* opts,
                               trt_scene_file_info* info) {
  if (!ctx || !path) return TRT_INVALID_ARG;
  trt::SceneFile sf;
  std::string err;
  if (!trt::load_scene_file(path, sf, err)) return trt_internal_fail(ctx, TRT_INVALID_ARG, err);
  trt_scene_file_info local;
  fill_info(sf, &local);
  std::vector<trt_material> mats;
  std::vector<uint32_t> mat_base;
  for (const trt::FileModel& m : sf.models) {
    mat_base.push_back(uint32_t(mats.size()));
    for (const trt::FileMaterial& fm : m.materials) mats.push_back(to_trt(fm));
  }
  trt_status s = trt_clear_scene(ctx);
  if (s != TRT_OK) return s;
  if ((s = trt_set_materials(ctx, mats.data(), uint32_t(mats.size()))) != TRT_OK) return s;
  for (const trt::FileModel& m : sf.models) {
    uint32_t id = 0;
    s = trt_add_mesh(ctx, m.pos.data(), m.nrm.data(), m.mat.data(), uint32_t(m.mat.size()), &id);
    if (s != TRT_OK) return s;
  }
  std::vector<trt_instance> inst(sf.instances.size());
  for (size_t i = 0; i < inst.size(); ++i) {
    std::memcpy(inst[i].transform, sf.instances[i].rows, 48);
    inst[i].mesh = sf.instances[i].model;
    inst[i].material_offset = mat_base[sf.instances[i].model];
  }
  if ((s = trt_set_instances(ctx, inst.data(), uint32_t(inst.size()))) != TRT_OK) return s;
  if ((s = trt_build(ctx, opts)) != TRT_OK) return s;
  if (sf.has_camera && (s = trt_set_camera(ctx, &local.camera)) != TRT_OK) return s;
  trt_sky sky;
  sky.enabled = local.sky_enabled;
  sky.draw_sun = local.sky_draw_sun;
  for (int k = 0; k < 3; ++k) sky.sun_dir[k] = local.sky_dir[k];
  sky.angular_diameter = local.sky_angular_diameter;
  sky.intensity = local.sky_intensity;
  sky.turbidity = local.sky_turbidity;
  sky.sun_nee = 0;  // the reference's setting (its scene files carry no such switch)
  if ((s = trt_set_sky(ctx, &sky)) != TRT_OK) return s;
  if (info) *info = local;
  return TRT_OK;
}

}  // extern "C"
