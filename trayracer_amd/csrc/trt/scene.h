// Scene records shared by the device kernels and the host oracle, and the
// hit-attribute fetch.
//
// Semantics follow the reference's GetIntersectionAttributes / FixNormals
// (CUDA/GPU/CudaMaterial.h:103-212): shading normal = normalised barycentric blend of
// vertex normals (face normal when the mesh has none), geometric normal from the
// triangle edges, both flipped to face the incoming ray (two-sided surfaces), the
// instance's inverse-transpose applied to normals. Deliberate deviations: triangle
// area and light power are measured in world space (the reference uses object-space
// area, Resources/Model.cpp:218-222, CudaMaterial.h:142), and emission is the
// material's radiance (the reference multiplies it by the diffuse colour,
// CudaMaterial.h:158).
#pragma once
#include "bsdf.h"
#include "common.h"
#include "intersect.h"
#include "sampling.h"
#include "sky.h"

namespace trt {

constexpr uint32_t kNoHit = 0xffffffffu;
constexpr uint32_t kNoInstance = 0xffffffffu;

// Instance of a mesh (two-level scenes). Row-major 3x4 affine transforms:
// p_world = o2w * (p_obj, 1). Replaces Instance::InstanceData + the inverse transform
// rows uploaded at Renderer/Renderer.cpp:476-493.
struct alignas(16) Instance {
  float o2w[12];
  float w2o[12];
  uint32_t mesh;        // mesh index
  uint32_t tri_offset;  // first triangle of the mesh in the global triangle arrays
  uint32_t mat_offset;  // added to the mesh's per-triangle material ids
  uint32_t blas_root;   // byte offset of the root of the mesh's bottom-level BVH in the node array
  uint32_t id;          // caller's instance index (hit reports and tie order use it)
  uint32_t pad[3];
};

// world->object inverse of a row-major 3x4 affine transform (computed in double)
inline void invert_affine(const float* m, float* inv) {
  const double a = m[0], b = m[1], c = m[2], d = m[4], e = m[5], f = m[6], g = m[8], h = m[9], i = m[10];
  const double A = e * i - f * h, B = -(d * i - f * g), C = d * h - e * g;
  const double det = a * A + b * B + c * C;
  const double id = det != 0.0 ? 1.0 / det : 0.0;
  double r[9] = {A * id, -(b * i - c * h) * id, (b * f - c * e) * id,
                 B * id, (a * i - c * g) * id, -(a * f - c * d) * id,
                 C * id, -(a * h - b * g) * id, (a * e - b * d) * id};
  const double tx = m[3], ty = m[7], tz = m[11];
  for (int row = 0; row < 3; ++row) {
    inv[row * 4 + 0] = float(r[row * 3 + 0]);
    inv[row * 4 + 1] = float(r[row * 3 + 1]);
    inv[row * 4 + 2] = float(r[row * 3 + 2]);
    inv[row * 4 + 3] = float(-(r[row * 3 + 0] * tx + r[row * 3 + 1] * ty + r[row * 3 + 2] * tz));
  }
}

TRT_FI V3 xform_point(const float* m, V3 p) {
  return v3(fma(m[2], p.z, fma(m[1], p.y, m[0] * p.x)) + m[3], fma(m[6], p.z, fma(m[5], p.y, m[4] * p.x)) + m[7],
            fma(m[10], p.z, fma(m[9], p.y, m[8] * p.x)) + m[11]);
}
TRT_FI V3 xform_vector(const float* m, V3 v) {
  return v3(fma(m[2], v.z, fma(m[1], v.y, m[0] * v.x)), fma(m[6], v.z, fma(m[5], v.y, m[4] * v.x)),
            fma(m[10], v.z, fma(m[9], v.y, m[8] * v.x)));
}
// normal transform: transpose of the world->object linear part
TRT_FI V3 xform_normal(const float* w2o, V3 n) {
  return v3(fma(w2o[8], n.z, fma(w2o[4], n.y, w2o[0] * n.x)), fma(w2o[9], n.z, fma(w2o[5], n.y, w2o[1] * n.x)),
            fma(w2o[10], n.z, fma(w2o[6], n.y, w2o[2] * n.x)));
}

// Emissive triangle in world space (replaces LightTriangle, Common/CommonStructs.h:257-273).
struct alignas(16) LightTri {
  float p0[3];
  float area;
  float e1[3];
  float power;  // sum of radiance channels x area (the reference's light "energy")
  float e2[3];
  float pad0;
  float n[3];   // unit geometric normal; emission leaves on this side
  float pad1;
  float le[3];  // radiance
  float pad2;
};

TRT_FI float light_power(V3 le, float area) { return sum3(le) * area; }

// Everything the shading step reads, as raw pointers (device or host memory).
struct SceneView {
  const float* tri_pos;         // 9 floats per triangle (v0, v1, v2), object space
  const float* tri_nrm;         // 9 floats per triangle or nullptr (face normals)
  const uint32_t* tri_mat;      // material id per triangle
  const Material* materials;    // per material id
  const Instance* instances;    // nullptr for single-level scenes
  const LightTri* lights;       // emissive triangles
  const float* alias_prob;      // Vose table over light power
  const uint32_t* alias_idx;
  uint32_t n_lights;
  float total_power;
  uint32_t n_tris;
  uint32_t n_materials;
  const SkyParams* sky;          // radiance of rays that leave the scene (disabled: black); never null.
                                // A pointer, not a copy: read only on misses and sun samples, its
                                // 18 words would otherwise sit in a kernel's scalar registers
};

struct Hit {
  float t, b1, b2;
  uint32_t prim;  // kNoHit on a miss
  uint32_t inst;  // kNoInstance for single-level scenes
};

struct SurfacePoint {
  V3 p;       // world position
  V3 ng;      // world geometric normal, facing the incoming ray
  V3 ns;      // world shading normal, same side as ng
  V3 ng_raw;  // world geometric normal by winding (emission side)
  V3 ns_raw;  // world shading normal before it is turned to the geometric normal's side
  float area; // world-space area
  uint32_t mat;
};

TRT_FI void load_tri(const SceneView& s, uint32_t g, V3& v0, V3& v1, V3& v2) {
  const float* p = s.tri_pos + static_cast<size_t>(g) * 9u;
  v0 = v3(p[0], p[1], p[2]);
  v1 = v3(p[3], p[4], p[5]);
  v2 = v3(p[6], p[7], p[8]);
}

// Hit attributes in world space. d = ray direction (world).
TRT_HD inline SurfacePoint surface_at(const SceneView& s, const Hit& h, V3 ray_o, V3 d) {
  SurfacePoint sp;
  uint32_t g = h.prim;
  uint32_t mat_off = 0;
  const Instance* inst = nullptr;
  if (h.inst != kNoInstance) {
    inst = &s.instances[h.inst];
    g = inst->tri_offset + h.prim;
    mat_off = inst->mat_offset;
  }
  V3 v0, v1, v2;
  load_tri(s, g, v0, v1, v2);
  V3 n_sh;
  if (s.tri_nrm) {
    const float* q = s.tri_nrm + static_cast<size_t>(g) * 9u;
    const float b0 = (1.f - h.b1) - h.b2;
    n_sh = v3(fma(h.b2, q[6], fma(h.b1, q[3], b0 * q[0])), fma(h.b2, q[7], fma(h.b1, q[4], b0 * q[1])),
              fma(h.b2, q[8], fma(h.b1, q[5], b0 * q[2])));
  }
  if (inst) {
    v0 = xform_point(inst->o2w, v0);
    v1 = xform_point(inst->o2w, v1);
    v2 = xform_point(inst->o2w, v2);
    if (s.tri_nrm) n_sh = xform_normal(inst->w2o, n_sh);
  }
  const V3 c = cross(v1 - v0, v2 - v0);
  const float cl = length(c);
  sp.area = 0.5f * cl;
  sp.ng_raw = cl > 0.f ? c * (1.f / cl) : v3(0.f, 0.f, 1.f);
  sp.p = bary_point(v0, v1, v2, h.b1, h.b2);
  (void)ray_o;
  V3 ng = sp.ng_raw;
  if (dot(ng, d) > 0.f) ng = -ng;
  V3 ns = s.tri_nrm ? normalize(n_sh) : sp.ng_raw;
  sp.ns_raw = ns;
  if (dot(ns, ng) < 0.f) ns = -ns;
  sp.ng = ng;
  sp.ns = ns;
  sp.mat = s.tri_mat[g] + mat_off;
  return sp;
}

}  // namespace trt
