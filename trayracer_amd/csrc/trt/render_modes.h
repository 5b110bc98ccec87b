// Debug / AOV render modes (replaces the reference's non-path-tracing shade kernels,
// Tracer/CUDA/GPU/Shade/*.h, dispatched by Shade(RenderModes, ...),
// Tracer/CUDA/GPU/cuda.cu:22-84). Each mode writes a colour per camera sample that the
// frame averages like path-traced radiance; the two ambient-occlusion modes trace one
// cosine-distributed bounce ray first. Mode numbers are this library's own (path
// tracing is 0); the reference's RenderModes enum orders them alphabetically.
//
// Deviations, for lack of texture coordinates in the scene data: the tangent and
// bitangent modes show the shading frame built from the shading normal (the frame
// every BSDF here uses), and the texture-coordinate mode shows the hit's barycentric
// coordinates (b1, b2, 0). The wireframe mode marks hits within 3% (in barycentric
// distance) of a triangle edge (the reference compares the primitive ids of one ray
// traced twice, which marks every pixel).
#pragma once
#include "integrator.h"

namespace trt {

enum RenderMode : uint32_t {
  kModePathTracing = 0,
  kModeAmbientOcclusion = 1,
  kModeAmbientOcclusionShading = 2,
  kModeBitangent = 3,
  kModeGeometricNormal = 4,
  kModeMaterialId = 5,
  kModeMaterialProperty = 6,
  kModeObjectId = 7,
  kModeShadingNormal = 8,
  kModeTangent = 9,
  kModeTextureCoordinate = 10,
  kModeWireframe = 11,
  kModeZDepth = 12,
  kNumRenderModes = 13
};

// Material property ids of the material-property mode, in the reference's order
// (Tracer/Common/CommonStructs.h:78-96).
enum MaterialPropertyId : uint32_t {
  kPropAnisotropic = 0,
  kPropClearcoat,
  kPropClearcoatGloss,
  kPropDiffuse,
  kPropEmissive,
  kPropMetallic,
  kPropNormal,
  kPropRoughness,
  kPropSheen,
  kPropSheenTint,
  kPropSpecular,
  kPropSpecularTint,
  kPropSubsurface,
  kNumMaterialProperties
};

// Distinct colour per id: bit 3k + c of the id becomes bit 7 - k of channel c, so
// consecutive ids differ in the high bits of the first channels.
TRT_FI V3 id_color(uint32_t id) {
  float ch[3];
  for (uint32_t c = 0; c < 3; ++c) {
    uint32_t v = 0;
    for (uint32_t k = 0; k < 8; ++k) v |= ((id >> (3u * k + c)) & 1u) << (7u - k);
    ch[c] = static_cast<float>(v) * (1.f / 255.f);
  }
  return v3(ch[0], ch[1], ch[2]);
}

TRT_FI V3 to_unit_color(V3 n) { return (n + v3(1.f)) * 0.5f; }

TRT_FI V3 material_property(const Material& m, uint32_t id) {
  switch (id) {
    case kPropAnisotropic: return v3(m.anisotropic);
    case kPropClearcoat: return v3(m.clearcoat);
    case kPropClearcoatGloss: return v3(m.clearcoat_gloss);
    case kPropDiffuse: return v3(m.base[0], m.base[1], m.base[2]);
    case kPropEmissive: return v3(m.emission[0], m.emission[1], m.emission[2]);
    case kPropMetallic: return v3(m.metallic);
    case kPropRoughness: return v3(m.roughness);
    case kPropSheen: return v3(m.sheen);
    case kPropSheenTint: return v3(m.sheen_tint);
    case kPropSpecular: return v3(m.specular);
    case kPropSpecularTint: return v3(m.specular_tint);
    case kPropSubsurface: return v3(m.subsurface);
    default: return v3(0.f);  // normal maps are not supported
  }
}

// One vertex of a debug mode; same contract as shade(). Misses stay black.
TRT_HD inline void shade_debug(const SceneView& s, const RenderParams& rp, const PathState& ps, const Hit& h,
                               uint32_t depth, ShadeOut& out) {
  out.L = v3(0.f);
  out.has_shadow = 0;
  out.has_ext = 0;
  out.albedo = v3(0.f);
  out.normal = v3(0.f);
  out.vdepth = 0.f;
  const uint32_t mode = rp.mode;
  const bool ao = mode == kModeAmbientOcclusion || mode == kModeAmbientOcclusionShading;
  if (depth > 0) {  // the occlusion ray: distance to the first hit, as a fraction of ao_dist
    if (ao) {
      const float t = h.prim == kNoHit ? 1e30f : h.t;
      const float z = t > rp.ao_dist ? 1.f : t / rp.ao_dist;
      out.L = ps.T * z;
    }
    return;
  }
  if (h.prim == kNoHit) return;
  const SurfacePoint sp = surface_at(s, h, ps.o, ps.d);
  const Material& m = s.materials[sp.mat];
  const V3 base = v3(m.base[0], m.base[1], m.base[2]);
  out.albedo = mode == kModeAmbientOcclusion ? v3(1.f) : base;
  out.normal = sp.ns;
  out.vdepth = h.t * dot(ps.d, v3(rp.cam.fwd[0], rp.cam.fwd[1], rp.cam.fwd[2]));
  switch (mode) {
    case kModeAmbientOcclusion:
    case kModeAmbientOcclusionShading: {
      uint32_t pixel, sample;
      path_pixel_sample(rp, ps.path, pixel, sample);
      Pcg32 rng = rng_for(rp.seed, pixel, sample, 0u);
      const float u0 = rng_float(rng), u1 = rng_float(rng);
      const Frame fr = make_frame(sp.ns);
      out.has_ext = 1;
      out.next.o = sp.p;
      out.next.d = normalize(to_world(fr, sample_cosine_hemisphere(u0, u1)));
      out.next.T = mode == kModeAmbientOcclusion ? v3(1.f) : base;
      out.next.pdf = 1.f;
      out.next.flags = 0u;
      out.next.path = ps.path;
      break;
    }
    case kModeBitangent: out.L = to_unit_color(make_frame(sp.ns_raw).b); break;
    case kModeGeometricNormal: out.L = to_unit_color(sp.ng_raw); break;
    case kModeMaterialId: out.L = id_color(sp.mat + 1u); break;
    case kModeMaterialProperty: out.L = material_property(m, rp.mode_flags); break;
    case kModeObjectId: out.L = id_color((h.inst == kNoInstance ? 0u : h.inst) + 1u); break;
    case kModeShadingNormal: out.L = to_unit_color(sp.ns_raw); break;
    case kModeTangent: out.L = to_unit_color(make_frame(sp.ns_raw).t); break;
    case kModeTextureCoordinate: out.L = v3(h.b1, h.b2, 0.f); break;
    case kModeWireframe: {
      const float b0 = (1.f - h.b1) - h.b2;
      out.L = fmin(b0, fmin(h.b1, h.b2)) < 0.03f ? v3(1.f) : v3(0.f);
      break;
    }
    case kModeZDepth: out.L = v3(clamp(out.vdepth / rp.zdepth_max, 0.f, 1.f)); break;
    default: break;
  }
}

// Dispatch on the render mode (host oracle; the device launches separate kernels).
TRT_HD inline void shade_mode(const SceneView& s, const Material* sampling_mats, const RenderParams& rp,
                              const PathState& ps, const Hit& h, uint32_t depth, ShadeOut& out) {
  if (rp.mode == kModePathTracing) shade(s, sampling_mats, rp, ps, h, depth, out);
  else shade_debug(s, rp, ps, h, depth, out);
}

}  // namespace trt
