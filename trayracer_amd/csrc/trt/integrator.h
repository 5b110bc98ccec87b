// One path vertex of the wavefront integrator, shared by the GPU Shade kernel and the
// host oracle.
//
// Flow follows the reference's PathTracingKernel (CUDA/GPU/Shade/PathTracing.h:8-175):
// miss -> sky (none here), first-hit AOVs, emitter hit (full emission for camera rays,
// BSDF-sampled hits weighted by MIS against light sampling) and termination at
// emitters, next-event estimation towards a power-sampled emissive triangle
// (CudaLights.h:13-124), BSDF sampling of the extension ray, Russian roulette
// survival probability min(1, max throughput). The reference's quirks are not
// reproduced (no double-counted primary emission, albedo included in NEE, light
// points uniformly distributed on the triangle, the last light is selectable,
// balance-heuristic MIS weights on both strategies, RR from a configurable depth,
// counter-based RNG keys per (pixel, sample, depth)).
//
// Random-number dimensions of vertex stream (pixel, sample, depth):
//   0,1 light choice (alias column, coin)   2,3 point on the light
//   4 lobe choice   5,6 BSDF direction   7 Russian roulette
#pragma once
#include "bsdf.h"
#include "camera.h"
#include "common.h"
#include "intersect.h"
#include "rng.h"
#include "sampling.h"
#include "scene.h"

namespace trt {

constexpr uint32_t kFlagSpecular = 1u;  // previous bounce was a delta lobe (full emission)

struct alignas(16) RenderParams {
  CameraParams cam;
  uint32_t width, height, spp, max_depth;
  uint32_t rr_depth, seed, sample_offset, flags;
  float eps;        // ray epsilon (reference KernelSettings::rayEpsilon)
  float clamp;      // firefly clamp on the max component of a contribution; 0 = off
  float ao_dist;    // ambient-occlusion distance (KernelSettings::aoDist)
  float zdepth_max; // depth of white in the z-depth mode (KernelSettings::zDepthMax)
  uint32_t mode;    // render mode (render_modes.h); 0 = path tracing
  uint32_t mode_flags;  // material property shown by the material-property mode
  uint32_t separate_sampling;  // 1: the sampling materials are a set of their own
                               // (differentiable renders), see the Russian roulette
  uint32_t strategy;  // emitter estimators: 0 = both, MIS-weighted (the renderer always);
                      // 1 = light sampling only, 2 = BSDF sampling only (each alone is
                      // unbiased too: the oracle's tests compare the three means)
};

struct PathState {
  V3 o, d;      // ray of the next segment
  V3 T;         // throughput
  float pdf;    // solid-angle pdf of the BSDF sample that produced d
  uint32_t flags;
  uint32_t path;  // path index = sample_in_frame * W * H + pixel
};

TRT_FI void path_pixel_sample(const RenderParams& rp, uint32_t path, uint32_t& pixel, uint32_t& sample) {
  const uint32_t np = rp.width * rp.height;
  pixel = path % np;
  sample = rp.sample_offset + path / np;
}

TRT_FI V3 clamp_contrib(V3 c, float mx) {
  if (mx <= 0.f) return fix_nan(c);
  const float m = max3(c);
  return fix_nan(m > mx ? c * (mx / m) : c);
}

// Camera ray of a path; returns false when the path carries no weight.
TRT_FI bool raygen(const RenderParams& rp, uint32_t path, PathState& ps) {
  uint32_t pixel, sample;
  path_pixel_sample(rp, path, pixel, sample);
  Pcg32 rng = rng_for(rp.seed, pixel, sample, kCameraDepth);
  const uint32_t x = pixel % rp.width, y = pixel / rp.width;
  V3 o, d;
  const float w = generate_camera_ray(rp.cam, x, y, rp.width, rp.height, rng, o, d);
  ps.o = o;
  ps.d = d;
  ps.T = v3(w);
  ps.pdf = 1.f;
  ps.flags = kFlagSpecular;
  ps.path = path;
  return w > 0.f;
}

struct ShadeOut {
  V3 L;  // emission at this vertex: added to the path radiance before s_contrib
  uint32_t has_shadow;
  uint32_t has_ext;
  V3 s_d;
  float s_tmax;
  V3 s_contrib;  // added to the path radiance when the shadow ray is unoccluded
  V3 s_o;
  PathState next;
  // first-hit AOVs (valid at depth 0)
  V3 albedo, normal;
  float vdepth;
};

// Solid-angle pdf of light sampling towards an emitter hit at distance t (total: the
// selection weight of all lights, the sun's included).
TRT_FI float light_pdf_hit(V3 le, float area, float cos_l, float t, float total) {
  const float pick = light_power(le, area) / total;
  return pick * (t * t) / (area * cos_l);
}

// Selection weight of the sun disc in next-event estimation (0 when it is not a light).
TRT_FI float sun_weight(const SceneView& s) { return s.sky->sun_nee ? s.sky->sun_power : 0.f; }

// Adjoint extras of one vertex (path-replay backpropagation, Vicini et al. 2021):
// derivatives of the vertex's terms with respect to its material's differentiable
// parameters theta = (baseColor r, g, b, roughness, metallic). Sampling decisions and
// MIS weights use the sampling materials and are treated as constants ("detached"),
// so the derivative of a BSDF-sampled weight f cos / pdf is (df/dtheta) cos / pdf.
constexpr int kNumParams = 5;
struct AdjOut {
  uint32_t mat;                 // material of the vertex (kNoHit when none)
  float nee_g[kNumParams];      // sum_c delta_c d(NEE contribution_c)/d theta_k
  float r[kNumParams][3];       // (d f_c / d theta_k) / f_c at the sampled extension
};

// SEP: whether the sampling materials are a set of their own, -1 = as rp.separate_sampling
// says (the device kernels fix it at compile time, so that with one set the material is
// read once and the separate-set code is left out).
template <bool ADJ, int SEP = -1>
TRT_HD inline void shade_impl(const SceneView& s, const Material* sampling_mats, const RenderParams& rp,
                              const PathState& ps, const Hit& h, uint32_t depth, ShadeOut& out, V3 delta,
                              AdjOut* adj) {
  const bool sep = SEP < 0 ? rp.separate_sampling != 0u : SEP == 1;
  out.L = v3(0.f);
  out.has_shadow = 0;
  out.has_ext = 0;
  out.albedo = v3(0.f);
  out.normal = v3(0.f);
  out.vdepth = 0.f;
  if (ADJ) adj->mat = kNoHit;
  if (h.prim == kNoHit) {  // the ray left the scene: sky radiance (PathTracing.h:40-55)
    if (s.sky->enabled) {
      const bool sun = s.sky->draw_sun != 0u && (depth == 0 || (ps.flags & kFlagSpecular));
      V3 sky = sky_radiance(*s.sky, ps.d, sun);
      // a BSDF-sampled ray that reaches the sun disc when the disc is also a light:
      // its radiance with the MIS weight against the sun's next-event estimate
      if (!sun && s.sky->sun_nee && rp.strategy != 1u) {
        const V3 ls = sun_radiance(*s.sky, ps.d);
        if (!is_black(ls)) {
          const float sw = sun_weight(s);
          const float pl = (sw / (s.total_power + sw)) * s.sky->sun_pdf;
          const float w = rp.strategy == 2u ? 1.f : ps.pdf / (ps.pdf + pl);
          sky = sky + ls * w;
        }
      }
      out.L = clamp_contrib(ps.T * sky, rp.clamp);
      if (depth == 0) out.albedo = sky;
    }
    return;
  }

  // The light pick depends only on the vertex's random numbers: draw them and start the
  // alias-table and light-record loads before the surface's, so that on the GPU the two
  // chains of dependent loads overlap (values are unchanged by the order).
  uint32_t pixel, sample;
  path_pixel_sample(rp, ps.path, pixel, sample);
  Pcg32 rng = rng_for(rp.seed, pixel, sample, depth);
  float u[8];
  for (int i = 0; i < 8; ++i) u[i] = rng_float(rng);
  // light choice: the sun disc with probability sun / total when it is a light (u[0]
  // rescaled for the triangle choice otherwise), else an emissive triangle by power
  const float sun_w = sun_weight(s);
  const float total = s.total_power + sun_w;
  const bool nee = (s.n_lights > 0 && s.total_power > 0.f) || sun_w > 0.f;
  bool pick_sun = false;
  LightTri L{};
  if (nee) {
    float u0 = u[0];
    if (sun_w > 0.f) {
      const float ps_sun = sun_w / total;
      pick_sun = u0 < ps_sun;
      u0 = fmin((u0 - ps_sun) / (1.f - ps_sun), 0x1.fffffep-1f);
    }
    if (!pick_sun) L = s.lights[alias_sample(s.alias_prob, s.alias_idx, s.n_lights, u0, u[1])];
  }

  const SurfacePoint sp = surface_at(s, h, ps.o, ps.d);
  const Material& m = s.materials[sp.mat];
  const Material& ms = sep ? sampling_mats[sp.mat] : m;  // the same values when not separate
  const V3 le = v3(m.emission[0], m.emission[1], m.emission[2]);
  if (depth == 0) {
    out.albedo = is_black(le) ? v3(m.base[0], m.base[1], m.base[2]) : le;
    out.normal = sp.ns;
    out.vdepth = h.t * dot(ps.d, v3(rp.cam.fwd[0], rp.cam.fwd[1], rp.cam.fwd[2]));
  }

  if (!is_black(le)) {
    const float cos_l = -dot(ps.d, sp.ng_raw);
    if (cos_l > 0.f) {
      if (depth == 0 || (ps.flags & kFlagSpecular)) {
        out.L = clamp_contrib(ps.T * le, rp.clamp);
      } else if (rp.strategy != 1u) {
        const float pl = light_pdf_hit(le, sp.area, cos_l, h.t, total);
        const float w = rp.strategy == 2u ? 1.f : ps.pdf / (ps.pdf + pl);
        out.L = clamp_contrib(ps.T * le * w, rp.clamp);
      }
    }
    return;  // paths end at emitters
  }

  const Frame fr = make_frame(sp.ns);
  const V3 wo = to_local(fr, -ps.d);
  if (!(wo.z > 0.f)) return;
  const BsdfParams<float> bp = params_of(m);
  const BsdfParams<float> bs = params_of(ms);
  // per-vertex BSDF terms of the one-pass evaluations (one parameter set)
  BsdfPre pre{};
  if (!sep) pre = bsdf_pre(bp, wo);
  BsdfParams<Dual<kNumParams>> bd;
  if (ADJ) {
    adj->mat = sp.mat;
    for (int k = 0; k < kNumParams; ++k) adj->nee_g[k] = 0.f;
    for (int k = 0; k < kNumParams; ++k) adj->r[k][0] = adj->r[k][1] = adj->r[k][2] = 0.f;
    bd.base = t3(dvar<kNumParams>(m.base[0], 0), dvar<kNumParams>(m.base[1], 1), dvar<kNumParams>(m.base[2], 2));
    bd.roughness = dvar<kNumParams>(m.roughness, 3);
    bd.metallic = dvar<kNumParams>(m.metallic, 4);
    bd.specular = bp.specular;
    bd.specular_tint = bp.specular_tint;
    bd.sheen = bp.sheen;
    bd.sheen_tint = bp.sheen_tint;
    bd.clearcoat = bp.clearcoat;
    bd.clearcoat_gloss = bp.clearcoat_gloss;
    bd.subsurface = bp.subsurface;
    bd.anisotropic = bp.anisotropic;
    bd.model = bp.model;
  }

  // next-event estimation: a point on the chosen triangle, or a direction in the sun's cone
  if (nee && rp.strategy != 2u) {
    V3 wl = v3(0.f), lrad = v3(0.f);
    float pl = 0.f, s_tmax = 0.f;
    bool ok = false;
    if (pick_sun) {
      wl = sample_sun(*s.sky, u[2], u[3]);
      if (dot(wl, sp.ng) > 0.f) {
        lrad = sun_radiance(*s.sky, wl);
        pl = (sun_w / total) * s.sky->sun_pdf;
        s_tmax = 1e30f;
        ok = !is_black(lrad);
      }
    } else {
      float b1, b2;
      sample_triangle(u[2], u[3], b1, b2);
      const V3 p0 = v3(L.p0[0], L.p0[1], L.p0[2]);
      const V3 e1 = v3(L.e1[0], L.e1[1], L.e1[2]);
      const V3 e2 = v3(L.e2[0], L.e2[1], L.e2[2]);
      const V3 q = v3(fma(b2, e2.x, fma(b1, e1.x, p0.x)), fma(b2, e2.y, fma(b1, e1.y, p0.y)),
                      fma(b2, e2.z, fma(b1, e1.z, p0.z)));
      const V3 dir = q - sp.p;
      const float dist2 = dot(dir, dir);
      const float dist = sqrt(dist2);
      if (dist > 0.f) {
        wl = dir * (1.f / dist);
        const float cos_l = -dot(wl, v3(L.n[0], L.n[1], L.n[2]));
        if (cos_l > 0.f && dot(wl, sp.ng) > 0.f) {
          lrad = v3(L.le[0], L.le[1], L.le[2]);
          pl = (L.power / total) * dist2 / (L.area * cos_l);
          s_tmax = dist - 2.f * rp.eps;
          ok = true;
        }
      }
    }
    if (ok) {
      const V3 wi = to_local(fr, wl);
      if (wi.z > 0.f) {
        float pb;
        const T3<float> f = sep ? bsdf_f(bp, wo, wi) : bsdf_f_pdf(bp, pre, wo, wi, pb);
        if (sep) pb = bsdf_pdf(bs, wo, wi);
        // balance heuristic against BSDF sampling, whose sample would be traced by the
        // extension ray; the last vertex traces none, so there the light sample stands
        // alone (a partial weight would drop the BSDF share of the last bounce's light)
        const float w = (rp.strategy == 1u || depth + 1 >= rp.max_depth) ? 1.f : pl / (pl + pb);
        const float k = wi.z * (w / pl);
        const V3 c = v3(ps.T.x * f.x * lrad.x * k, ps.T.y * f.y * lrad.y * k, ps.T.z * f.z * lrad.z * k);
        if (!is_black(c)) {
          out.has_shadow = 1;
          out.s_o = sp.p;
          out.s_d = wl;
          out.s_tmax = s_tmax;
          out.s_contrib = clamp_contrib(c, rp.clamp);
          if (ADJ) {
            const T3<Dual<kNumParams>> fd = bsdf_f(bd, wo, wi);
            const float kx = delta.x * ps.T.x * lrad.x * k, ky = delta.y * ps.T.y * lrad.y * k,
                        kz = delta.z * ps.T.z * lrad.z * k;
            for (int j = 0; j < kNumParams; ++j)
              adj->nee_g[j] = (kx * fd.x.d[j] + ky * fd.y.d[j]) + kz * fd.z.d[j];
          }
        }
      }
    }
  }

  // extension
  if (depth + 1 < rp.max_depth) {
    // with one parameter set, the sample's pdf and f come from one fused evaluation
    BsdfSample smp;
    T3<float> f1 = t3(0.f, 0.f, 0.f);
    if (sep) {
      smp = bsdf_sample(bs, wo, u[4], u[5], u[6]);
    } else {
      smp.pdf = 0.f;
      smp.wi = v3(0.f, 0.f, 1.f);
      V3 wi;
      if (bsdf_sample_dir(bs, pre, wo, u[4], u[5], u[6], wi)) {
        smp.wi = wi;
        f1 = bsdf_f_pdf(bp, pre, wo, wi, smp.pdf);
      }
    }
    if (smp.pdf > 0.f) {
      const V3 wiw = normalize(to_world(fr, smp.wi));
      if (dot(wiw, sp.ng) > 0.f) {
        const T3<float> f = sep ? bsdf_f(bp, wo, smp.wi) : f1;
        const float k = smp.wi.z / smp.pdf;
        V3 Tn = v3(ps.T.x * f.x * k, ps.T.y * f.y * k, ps.T.z * f.z * k);
        bool alive = !is_black(Tn);
        if (alive && depth >= rp.rr_depth) {
          // Russian roulette (PathTracing.h:148-154): survival probability
          // min(1, max throughput). With a set of sampling materials of their own
          // (differentiable renders) it is a sampling decision like the others: this
          // vertex's BSDF weight under the sampling materials alone (the throughput
          // carries earlier vertices' evaluation parameters), which the adjoint holds
          // fixed, so the continuation's derivative stays (df/dtheta) / f
          float q;
          if (sep) {
            const T3<float> fs = bsdf_f(bs, wo, smp.wi);
            q = fmin(1.f, fmax(fmax(fs.x, fs.y), fs.z) * k);
          } else {
            q = fmin(1.f, max3(Tn));
          }
          if (u[7] >= q) alive = false;
          else Tn = Tn * (1.f / q);
        }
        if (alive) {
          out.has_ext = 1;
          out.next.o = sp.p;
          out.next.d = wiw;
          out.next.T = fix_nan(Tn);
          out.next.pdf = smp.pdf;
          out.next.flags = 0u;
          out.next.path = ps.path;
          if (ADJ) {
            const T3<Dual<kNumParams>> fd = bsdf_f(bd, wo, smp.wi);
            for (int j = 0; j < kNumParams; ++j) {
              adj->r[j][0] = f.x != 0.f ? fd.x.d[j] / f.x : 0.f;
              adj->r[j][1] = f.y != 0.f ? fd.y.d[j] / f.y : 0.f;
              adj->r[j][2] = f.z != 0.f ? fd.z.d[j] / f.z : 0.f;
            }
          }
        }
      }
    }
  }
}

// One vertex: ps = incoming path state, h = closest hit of ray (ps.o, ps.d).
// sampling_mats may differ from s.materials only in differentiable renders
// (sampling decisions use it, BSDF values use s.materials).
template <int SEP = -1>
TRT_HD inline void shade(const SceneView& s, const Material* sampling_mats, const RenderParams& rp,
                         const PathState& ps, const Hit& h, uint32_t depth, ShadeOut& out) {
  shade_impl<false, SEP>(s, sampling_mats, rp, ps, h, depth, out, v3(0.f), nullptr);
}

// Same vertex replayed for the adjoint: delta = dLoss/d(path radiance).
TRT_HD inline void shade_adjoint(const SceneView& s, const Material* sampling_mats, const RenderParams& rp,
                                 const PathState& ps, const Hit& h, uint32_t depth, V3 delta, ShadeOut& out,
                                 AdjOut& adj) {
  shade_impl<true>(s, sampling_mats, rp, ps, h, depth, out, delta, &adj);
}

// Gradient of the BSDF-sampled continuation of a vertex once the radiance that
// arrived through it (l_after) is known: sum_c delta_c r_kc l_after_c.
TRT_FI void continuation_grad(const AdjOut& a, V3 delta, V3 l_after, float* g) {
  const float x = delta.x * l_after.x, y = delta.y * l_after.y, z = delta.z * l_after.z;
  for (int k = 0; k < kNumParams; ++k) g[k] = (a.r[k][0] * x + a.r[k][1] * y) + a.r[k][2] * z;
}

}  // namespace trt
