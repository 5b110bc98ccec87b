// Host-side construction of the emissive-triangle table and its Vose alias table.
//
// Replaces Model::BuildLights (Resources/Model.cpp:171-236) + Scene::GatherLights
// (Renderer/Scene.cpp:275-302): every triangle whose material emits becomes a
// world-space light record with area and power = sum(radiance) x area. Selection is
// proportional to power through an alias table (Vose 1991, O(1) per sample) instead of
// the reference's binary search over an exclusive prefix sum (CUDA/GPU/CudaLights.h:38-59),
// and every light is selectable (the reference's prefix excludes the last one).
// Area and normal use exactly the operations of surface_at(), so the pick
// probability of a light equals the one recomputed at a BSDF-sampled emitter hit.
#pragma once
#include <vector>

#include "scene.h"
#include "sky.h"

namespace trt {

// Vose's alias method. Builds prob/alias for weights w (all >= 0, sum > 0).
inline void build_alias_table(const std::vector<double>& w, std::vector<float>& prob, std::vector<uint32_t>& alias) {
  const size_t n = w.size();
  prob.assign(n, 1.f);
  alias.resize(n);
  for (size_t i = 0; i < n; ++i) alias[i] = uint32_t(i);
  if (n == 0) return;
  double sum = 0.0;
  for (double x : w) sum += x;
  if (!(sum > 0.0)) return;
  std::vector<double> q(n);
  std::vector<uint32_t> small, large;
  for (size_t i = 0; i < n; ++i) {
    q[i] = w[i] * double(n) / sum;
    (q[i] < 1.0 ? small : large).push_back(uint32_t(i));
  }
  while (!small.empty() && !large.empty()) {
    const uint32_t s = small.back();
    small.pop_back();
    const uint32_t l = large.back();
    prob[s] = float(q[s]);
    alias[s] = l;
    q[l] = (q[l] + q[s]) - 1.0;
    if (q[l] < 1.0) {
      large.pop_back();
      small.push_back(l);
    }
  }
  for (uint32_t l : large) prob[l] = 1.f;
  for (uint32_t s : small) prob[s] = 1.f;  // numerical leftovers
}

// World-space light record of triangle v0,v1,v2 with radiance le.
inline LightTri make_light(V3 v0, V3 v1, V3 v2, V3 le) {
  LightTri L{};
  const V3 c = cross(v1 - v0, v2 - v0);
  const float cl = length(c);
  const V3 n = cl > 0.f ? c * (1.f / cl) : v3(0.f, 0.f, 1.f);
  const V3 e1 = v1 - v0, e2 = v2 - v0;
  L.p0[0] = v0.x; L.p0[1] = v0.y; L.p0[2] = v0.z;
  L.e1[0] = e1.x; L.e1[1] = e1.y; L.e1[2] = e1.z;
  L.e2[0] = e2.x; L.e2[1] = e2.y; L.e2[2] = e2.z;
  L.n[0] = n.x; L.n[1] = n.y; L.n[2] = n.z;
  L.le[0] = le.x; L.le[1] = le.y; L.le[2] = le.z;
  L.area = 0.5f * cl;
  L.power = light_power(le, L.area);
  return L;
}

struct LightTable {
  std::vector<LightTri> lights;
  std::vector<float> prob;
  std::vector<uint32_t> alias;
  float total_power = 0.f;
};

// Collects emissive triangles. Single-level scenes pass instances == nullptr;
// two-level scenes list every (instance, triangle of its mesh) pair.
inline LightTable gather_lights(const float* tri_pos, const uint32_t* tri_mat, uint32_t n_tris,
                                const Material* mats, uint32_t n_mats, const Instance* instances, uint32_t n_inst,
                                const uint32_t* mesh_tri_count) {
  LightTable t;
  auto emit = [&](uint32_t g, uint32_t mat_id, const Instance* inst) {
    if (mat_id >= n_mats) return;
    const Material& m = mats[mat_id];
    const V3 le = v3(m.emission[0], m.emission[1], m.emission[2]);
    if (is_black(le)) return;
    const float* p = tri_pos + size_t(g) * 9;
    V3 v0 = v3(p[0], p[1], p[2]), v1 = v3(p[3], p[4], p[5]), v2 = v3(p[6], p[7], p[8]);
    if (inst) {
      v0 = xform_point(inst->o2w, v0);
      v1 = xform_point(inst->o2w, v1);
      v2 = xform_point(inst->o2w, v2);
    }
    LightTri L = make_light(v0, v1, v2, le);
    if (L.power > 0.f) t.lights.push_back(L);
  };
  if (!instances) {
    for (uint32_t g = 0; g < n_tris; ++g) emit(g, tri_mat[g], nullptr);
  } else {
    for (uint32_t i = 0; i < n_inst; ++i) {
      const Instance& in = instances[i];
      const uint32_t cnt = mesh_tri_count[in.mesh];
      for (uint32_t k = 0; k < cnt; ++k) emit(in.tri_offset + k, tri_mat[in.tri_offset + k] + in.mat_offset, &in);
    }
  }
  std::vector<double> w(t.lights.size());
  double tot = 0.0;
  for (size_t i = 0; i < t.lights.size(); ++i) {
    w[i] = t.lights[i].power;
    tot += w[i];
  }
  t.total_power = float(tot);
  build_alias_table(w, t.prob, t.alias);
  return t;
}

// Half-diagonal of the world-space bounds of the scene: all vertices of a single-level
// scene; for instanced scenes the transformed corners of each instance's mesh bounds.
// Sets the sun's selection weight (sky.h), so the GPU context and the oracle compute
// it with this same function.
inline float scene_radius(const float* tri_pos, uint32_t n_tris, const Instance* instances, uint32_t n_inst,
                          const uint32_t* mesh_tri_offset, const uint32_t* mesh_tri_count, uint32_t n_meshes) {
  float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
  auto grow = [&](V3 p) {
    const float c[3] = {p.x, p.y, p.z};
    for (int k = 0; k < 3; ++k) {
      lo[k] = fmin(lo[k], c[k]);
      hi[k] = fmax(hi[k], c[k]);
    }
  };
  if (!instances) {
    for (size_t i = 0; i < size_t(n_tris) * 3; ++i) grow(v3(tri_pos[i * 3], tri_pos[i * 3 + 1], tri_pos[i * 3 + 2]));
  } else {
    std::vector<float> mb(size_t(n_meshes) * 6);
    for (uint32_t m = 0; m < n_meshes; ++m) {
      float ml[3] = {1e30f, 1e30f, 1e30f}, mh[3] = {-1e30f, -1e30f, -1e30f};
      for (size_t i = size_t(mesh_tri_offset[m]) * 3; i < size_t(mesh_tri_offset[m] + mesh_tri_count[m]) * 3; ++i)
        for (int k = 0; k < 3; ++k) {
          ml[k] = fmin(ml[k], tri_pos[i * 3 + k]);
          mh[k] = fmax(mh[k], tri_pos[i * 3 + k]);
        }
      for (int k = 0; k < 3; ++k) {
        mb[m * 6 + k] = ml[k];
        mb[m * 6 + 3 + k] = mh[k];
      }
    }
    for (uint32_t i = 0; i < n_inst; ++i) {
      const float* b = &mb[size_t(instances[i].mesh) * 6];
      if (b[0] > b[3]) continue;  // empty mesh
      for (int c = 0; c < 8; ++c)
        grow(xform_point(instances[i].o2w, v3(b[(c & 1) ? 3 : 0], b[(c & 2) ? 4 : 1], b[(c & 4) ? 5 : 2])));
    }
  }
  if (lo[0] > hi[0]) return 0.f;
  return 0.5f * length(v3(hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]));
}

}  // namespace trt
