// BSDFs: Disney "principled" BRDF and Lambert.
//
// The reference's lobe set and parameterisation (CUDA/GPU/Shade/BSDF/Disney.h:28-579,
// Microfacet.h:28-326; material properties Common/CommonStructs.h:76-91) is
// re-derived here from the published models, not from those files:
//   * Burley 2012, "Physically-Based Shading at Disney": retro-reflective diffuse,
//     Hanrahan-Krueger-like subsurface blend, sheen, anisotropic GGX (GTR2) specular
//     with Schlick Fresnel from a tinted F0, GTR1 clearcoat (alpha = lerp(.1,.001,gloss),
//     fixed Smith alpha 0.25, F0 = 0.04, weight 0.25);
//   * Heitz 2014 for the height-correlated Smith masking-shadowing G2 = 1/(1+L(o)+L(i));
//   * Heitz 2018, "Sampling the GGX distribution of visible normals" for VNDF sampling.
// Deliberate deviations from the reference (which is not enabled there,
// PathTracing.h:127-129): no reuse of one random number for lobe choice and direction,
// sheen is not flagged specular, GTR1 pdf uses |m.z|, the VNDF frame test uses v.z.
//
// Local frame: z = shading normal. f() excludes the cosine term.
// The evaluation is templated on the scalar type so the adjoint can run it on
// dual numbers (baseColor.rgb, roughness, metallic are the differentiable inputs).
#pragma once
#include "ad.h"
#include "common.h"
#include "sampling.h"

namespace trt {

enum MaterialModel : uint32_t { kModelDisney = 0, kModelLambert = 1 };

// Device/host material record, 80 bytes (five 16-byte rows).
struct alignas(16) Material {
  float base[3];
  float roughness;
  float emission[3];  // emitted radiance (front face)
  float metallic;
  float specular;
  float specular_tint;
  float sheen;
  float sheen_tint;
  float clearcoat;
  float clearcoat_gloss;
  float subsurface;
  float anisotropic;
  uint32_t model;
  uint32_t pad[3];
};

template <class S>
struct BsdfParams {
  T3<S> base;
  S roughness;
  S metallic;
  float specular, specular_tint, sheen, sheen_tint, clearcoat, clearcoat_gloss, subsurface, anisotropic;
  uint32_t model;
};

TRT_FI BsdfParams<float> params_of(const Material& m) {
  BsdfParams<float> p;
  p.base = t3(m.base[0], m.base[1], m.base[2]);
  p.roughness = m.roughness;
  p.metallic = m.metallic;
  p.specular = m.specular;
  p.specular_tint = m.specular_tint;
  p.sheen = m.sheen;
  p.sheen_tint = m.sheen_tint;
  p.clearcoat = m.clearcoat;
  p.clearcoat_gloss = m.clearcoat_gloss;
  p.subsurface = m.subsurface;
  p.anisotropic = m.anisotropic;
  p.model = m.model;
  return p;
}

TRT_FI float schlick_weight(float c) {
  const float m = clamp(1.f - c, 0.f, 1.f);
  const float m2 = m * m;
  return m2 * m2 * m;
}

// alpha of the specular lobe (Burley 2012, addendum: alpha = roughness^2, anisotropic
// aspect = sqrt(1 - 0.9 anisotropic)), clamped to 1e-3 like the reference
template <class S>
TRT_FI void ggx_alpha(const S& roughness, float anisotropic, S& ax, S& ay) {
  const float aspect = sqrt(1.f - 0.9f * anisotropic);
  const S r2 = roughness * roughness;
  ax = fmax(r2 / aspect, 1e-3f);
  ay = fmax(r2 * aspect, 1e-3f);
}

// anisotropic GTR2/GGX normal distribution D(m)
template <class S>
TRT_FI S ggx_d(V3 m, const S& ax, const S& ay) {
  const S t = sqT(m.x / ax) + sqT(m.y / ay) + m.z * m.z;
  return 1.f / (kPi * ax * ay * t * t);
}

// Smith Lambda for the anisotropic GGX distribution
template <class S>
TRT_FI S ggx_lambda(V3 w, const S& ax, const S& ay) {
  const float z2 = w.z * w.z;
  const S a2t = (sqT(ax * w.x) + sqT(ay * w.y)) / z2;
  return (sqrt(1.f + a2t) - 1.f) * 0.5f;
}

// GTR1 (clearcoat) distribution, a < 1, at the unit half vector m (local frame). The
// denominator 1 + (a^2 - 1) cos^2 is formed as sin^2 + a^2 cos^2 from m's components:
// near the peak the first form cancels to ~a^2 and loses most of its float precision
// at glossy settings (a = 0.001).
TRT_FI float gtr1_d(V3 m, float a) {
  const float a2 = a * a;
  const float t = fma(a2, m.z * m.z, fma(m.x, m.x, m.y * m.y));
  return (a2 - 1.f) / (kPi * log(a2) * t);
}

TRT_FI float smith_g1_ggx_iso(float cos_w, float a) {
  const float a2 = a * a;
  const float c2 = cos_w * cos_w;
  return 2.f * cos_w / (cos_w + sqrt(fma(a2, 1.f - c2, c2)));
}

template <class S>
TRT_FI S lum3(const T3<S>& c) {
  return 0.2126729f * c.x + 0.7151522f * c.y + 0.0721750f * c.z;
}

// BSDF value f(wo, wi) (no cosine), RGB
template <class S>
TRT_HD T3<S> bsdf_f(const BsdfParams<S>& p, V3 wo, V3 wi) {
  const S zero = sc<S>(0.f);
  T3<S> out = t3(zero, zero, zero);
  if (wo.z <= 0.f || wi.z <= 0.f) return out;
  if (p.model == kModelLambert) {
    return p.base * kInvPi;
  }
  const V3 h = normalize(wo + wi);
  const float cl = wi.z, cv = wo.z, cd = dot(wi, h);
  const float FL = schlick_weight(cl), FV = schlick_weight(cv), FH = schlick_weight(cd);
  const S one_m_metal = 1.f - p.metallic;

  // tint: base colour normalised by luminance
  const S lum = lum3(p.base);
  T3<S> tint = t3(sc<S>(1.f), sc<S>(1.f), sc<S>(1.f));
  if (val(lum) > 0.f) tint = t3(p.base.x / lum, p.base.y / lum, p.base.z / lum);

  // diffuse (Burley) with subsurface blend
  const S rcd2 = p.roughness * (cd * cd);
  const S fd90 = 0.5f + 2.f * rcd2;
  const S fd = (1.f + (fd90 - 1.f) * FL) * (1.f + (fd90 - 1.f) * FV);
  S dw = fd;
  if (p.subsurface > 0.f) {
    const S fss = (1.f + (rcd2 - 1.f) * FL) * (1.f + (rcd2 - 1.f) * FV);
    const S ss = 1.25f * (fss * (1.f / (cl + cv) - 0.5f) + 0.5f);
    dw = (1.f - p.subsurface) * fd + p.subsurface * ss;
  }
  const S kd = dw * kInvPi * one_m_metal;
  out = p.base * kd;

  // sheen
  if (p.sheen > 0.f) {
    const S ks = one_m_metal * (FH * p.sheen);
    const float st = p.sheen_tint;
    out = out + t3(((1.f - st) + st * tint.x) * ks, ((1.f - st) + st * tint.y) * ks, ((1.f - st) + st * tint.z) * ks);
  }

  // specular (GGX + Schlick with tinted dielectric F0 mixed towards base by metallic)
  S ax, ay;
  ggx_alpha(p.roughness, p.anisotropic, ax, ay);
  const S D = ggx_d(h, ax, ay);
  const S G = 1.f / (1.f + ggx_lambda(wo, ax, ay) + ggx_lambda(wi, ax, ay));
  const S spec_common = D * G / (4.f * cl * cv);
  const float sp = p.specular * 0.08f, stt = p.specular_tint;
  const T3<S> c0 = t3(mixT(0.f, p.base.x, p.metallic) + (1.f - p.metallic) * (sp * ((1.f - stt) + stt * tint.x)),
                      mixT(0.f, p.base.y, p.metallic) + (1.f - p.metallic) * (sp * ((1.f - stt) + stt * tint.y)),
                      mixT(0.f, p.base.z, p.metallic) + (1.f - p.metallic) * (sp * ((1.f - stt) + stt * tint.z)));
  const T3<S> F = t3(c0.x + (1.f - c0.x) * FH, c0.y + (1.f - c0.y) * FH, c0.z + (1.f - c0.z) * FH);
  out = out + F * spec_common;

  // clearcoat
  if (p.clearcoat > 0.f) {
    const float a = lerp(0.1f, 0.001f, p.clearcoat_gloss);
    const float Dr = gtr1_d(h, a);
    const float Fr = lerp(0.04f, 1.f, FH);
    const float Gr = smith_g1_ggx_iso(cl, 0.25f) * smith_g1_ggx_iso(cv, 0.25f);
    const float kc = 0.25f * p.clearcoat * Dr * Fr * Gr / (4.f * cl * cv);
    out = out + t3(zero + kc, zero + kc, zero + kc);
  }
  return out;
}

// Lobe selection probabilities (diffuse, specular, clearcoat) for sampling.
TRT_FI void lobe_probs(const BsdfParams<float>& p, float cos_o, float& pd, float& ps, float& pc) {
  const float FV = schlick_weight(cos_o);
  const float lum = lum3(p.base);
  float wd = (1.f - p.metallic) * fmax(lum, 0.02f);
  if (p.metallic >= 1.f) wd = 0.f;
  const float sp = p.specular * 0.08f;
  const float c0 = lerp(sp, lum, p.metallic);
  const float ws = fmax(c0 + (1.f - c0) * FV, 0.02f);
  const float wc = p.clearcoat > 0.f ? 0.25f * p.clearcoat * lerp(0.04f, 1.f, FV) : 0.f;
  const float inv = 1.f / ((wd + ws) + wc);
  pd = wd * inv;
  ps = ws * inv;
  pc = wc * inv;
}

// solid-angle pdf of bsdf_sample producing wi
TRT_HD inline float bsdf_pdf(const BsdfParams<float>& p, V3 wo, V3 wi) {
  if (wo.z <= 0.f || wi.z <= 0.f) return 0.f;
  if (p.model == kModelLambert) return wi.z * kInvPi;
  float pd, ps, pc;
  lobe_probs(p, wo.z, pd, ps, pc);
  const V3 h = normalize(wo + wi);
  const float oh = dot(wo, h);
  float pdf = pd * (wi.z * kInvPi);
  if (oh > 0.f) {
    float ax, ay;
    ggx_alpha(p.roughness, p.anisotropic, ax, ay);
    const float g1 = 1.f / (1.f + ggx_lambda(wo, ax, ay));
    pdf += ps * (g1 * ggx_d(h, ax, ay) / (4.f * wo.z));
    if (pc > 0.f) {
      const float a = lerp(0.1f, 0.001f, p.clearcoat_gloss);
      pdf += pc * (gtr1_d(h, a) * h.z / (4.f * oh));
    }
  }
  return pdf;
}

// Per-vertex constants of one parameter set and outgoing direction wo (wo.z > 0) that
// bsdf_f, bsdf_pdf and bsdf_sample would each form again: the tint, the GGX alphas,
// 1 + Lambda(wo) and the lobe selection probabilities (Disney model; unused for Lambert).
struct BsdfPre {
  T3<float> tint;
  float ax, ay, l1;
  float pd, ps, pc;
};

TRT_HD inline BsdfPre bsdf_pre(const BsdfParams<float>& p, V3 wo) {
  BsdfPre q;
  q.tint = t3(1.f, 1.f, 1.f);
  q.ax = q.ay = q.l1 = q.pd = q.ps = q.pc = 0.f;
  if (p.model == kModelLambert || !(wo.z > 0.f)) return q;
  const float lum = lum3(p.base);
  if (lum > 0.f) q.tint = t3(p.base.x / lum, p.base.y / lum, p.base.z / lum);
  ggx_alpha(p.roughness, p.anisotropic, q.ax, q.ay);
  q.l1 = 1.f + ggx_lambda(wo, q.ax, q.ay);
  lobe_probs(p, wo.z, q.pd, q.ps, q.pc);
  return q;
}

// bsdf_f(p, wo, wi) and bsdf_pdf(p, wo, wi) of one parameter set in one pass, for the
// integrator where the evaluation and sampling parameters coincide: the half vector, D
// and the clearcoat D are formed once instead of twice, and the per-vertex terms come
// from q = bsdf_pre(p, wo). Every value is formed by the same operations as in the two
// separate functions, so both results equal theirs bit for bit.
TRT_HD inline T3<float> bsdf_f_pdf(const BsdfParams<float>& p, const BsdfPre& q, V3 wo, V3 wi, float& pdf) {
  pdf = 0.f;
  const float zero = 0.f;
  T3<float> out = t3(zero, zero, zero);
  if (wo.z <= 0.f || wi.z <= 0.f) return out;
  if (p.model == kModelLambert) {
    pdf = wi.z * kInvPi;
    return p.base * kInvPi;
  }
  const V3 h = normalize(wo + wi);
  const float cl = wi.z, cv = wo.z, cd = dot(wi, h);
  const float FL = schlick_weight(cl), FV = schlick_weight(cv), FH = schlick_weight(cd);
  const float one_m_metal = 1.f - p.metallic;
  const T3<float> tint = q.tint;

  const float rcd2 = p.roughness * (cd * cd);
  const float fd90 = 0.5f + 2.f * rcd2;
  const float fd = (1.f + (fd90 - 1.f) * FL) * (1.f + (fd90 - 1.f) * FV);
  float dw = fd;
  if (p.subsurface > 0.f) {
    const float fss = (1.f + (rcd2 - 1.f) * FL) * (1.f + (rcd2 - 1.f) * FV);
    const float ss = 1.25f * (fss * (1.f / (cl + cv) - 0.5f) + 0.5f);
    dw = (1.f - p.subsurface) * fd + p.subsurface * ss;
  }
  const float kd = dw * kInvPi * one_m_metal;
  out = p.base * kd;

  if (p.sheen > 0.f) {
    const float ks = one_m_metal * (FH * p.sheen);
    const float st = p.sheen_tint;
    out = out + t3(((1.f - st) + st * tint.x) * ks, ((1.f - st) + st * tint.y) * ks, ((1.f - st) + st * tint.z) * ks);
  }

  const float ax = q.ax, ay = q.ay;
  const float D = ggx_d(h, ax, ay);
  const float G = 1.f / (q.l1 + ggx_lambda(wi, ax, ay));
  const float spec_common = D * G / (4.f * cl * cv);
  const float sp = p.specular * 0.08f, stt = p.specular_tint;
  const T3<float> c0 = t3(mixT(0.f, p.base.x, p.metallic) + (1.f - p.metallic) * (sp * ((1.f - stt) + stt * tint.x)),
                          mixT(0.f, p.base.y, p.metallic) + (1.f - p.metallic) * (sp * ((1.f - stt) + stt * tint.y)),
                          mixT(0.f, p.base.z, p.metallic) + (1.f - p.metallic) * (sp * ((1.f - stt) + stt * tint.z)));
  const T3<float> F = t3(c0.x + (1.f - c0.x) * FH, c0.y + (1.f - c0.y) * FH, c0.z + (1.f - c0.z) * FH);
  out = out + F * spec_common;

  const float a = lerp(0.1f, 0.001f, p.clearcoat_gloss);
  float Dr = 0.f;
  if (p.clearcoat > 0.f) {
    Dr = gtr1_d(h, a);
    const float Fr = lerp(0.04f, 1.f, FH);
    const float Gr = smith_g1_ggx_iso(cl, 0.25f) * smith_g1_ggx_iso(cv, 0.25f);
    const float kc = 0.25f * p.clearcoat * Dr * Fr * Gr / (4.f * cl * cv);
    out = out + t3(zero + kc, zero + kc, zero + kc);
  }

  // bsdf_pdf's terms
  const float oh = dot(wo, h);
  float pr = q.pd * (wi.z * kInvPi);
  if (oh > 0.f) {
    const float g1 = 1.f / q.l1;
    pr += q.ps * (g1 * D / (4.f * wo.z));
    if (q.pc > 0.f) pr += q.pc * ((p.clearcoat > 0.f ? Dr : gtr1_d(h, a)) * h.z / (4.f * oh));
  }
  pdf = pr;
  return out;
}

// Heitz 2018 VNDF sample of the anisotropic GGX distribution (wo.z > 0)
TRT_FI V3 sample_ggx_vndf(V3 wo, float ax, float ay, float u0, float u1) {
  const V3 vh = normalize(v3(ax * wo.x, ay * wo.y, wo.z));
  const float lensq = fma(vh.x, vh.x, vh.y * vh.y);
  V3 T1 = lensq > 0.f ? v3(-vh.y, vh.x, 0.f) * (1.f / sqrt(lensq)) : v3(1.f, 0.f, 0.f);
  const V3 T2 = cross(vh, T1);
  const float r = sqrt(u0);
  float s, c;
  sincos_turns(u1, s, c);
  const float t1 = r * c;
  float t2 = r * s;
  const float sv = 0.5f * (1.f + vh.z);
  t2 = fma(1.f - sv, safe_sqrt(1.f - t1 * t1), sv * t2);
  const float tz = safe_sqrt((1.f - t1 * t1) - t2 * t2);
  const V3 nh = v3(fma(tz, vh.x, fma(t2, T2.x, t1 * T1.x)), fma(tz, vh.y, fma(t2, T2.y, t1 * T1.y)),
                   fma(tz, vh.z, fma(t2, T2.z, t1 * T1.z)));
  return normalize(v3(ax * nh.x, ay * nh.y, fmax(nh.z, 0.f)));
}

// GTR1 half-vector sample: pdf(m) = D(m) cos(theta_m)
TRT_FI V3 sample_gtr1(float a, float u0, float u1) {
  const float a2 = a * a;
  const float cos2 = (1.f - pow(a2, 1.f - u0)) / (1.f - a2);
  const float ct = safe_sqrt(cos2);
  const float st = safe_sqrt(1.f - cos2);
  float s, c;
  sincos_turns(u1, s, c);
  return v3(st * c, st * s, ct);
}

struct BsdfSample {
  V3 wi;
  float pdf;  // 0 = no sample
};

// Direction of bsdf_sample (false: no sample), with q = bsdf_pre(p, wo).
TRT_HD inline bool bsdf_sample_dir(const BsdfParams<float>& p, const BsdfPre& q, V3 wo, float ul, float u0, float u1,
                                   V3& wi) {
  if (wo.z <= 0.f) return false;
  if (p.model == kModelLambert) {
    wi = sample_cosine_hemisphere(u0, u1);
    return wi.z > 0.f;
  }
  if (ul < q.pd) {
    wi = sample_cosine_hemisphere(u0, u1);
  } else {
    V3 m;
    if (ul < q.pd + q.ps || q.pc <= 0.f) {
      m = sample_ggx_vndf(wo, q.ax, q.ay, u0, u1);
    } else {
      m = sample_gtr1(lerp(0.1f, 0.001f, p.clearcoat_gloss), u0, u1);
    }
    wi = reflect(wo, m);
  }
  return wi.z > 0.f;
}

// Draws wi given three uniforms: ul selects the lobe, (u0,u1) the direction.
TRT_HD inline BsdfSample bsdf_sample(const BsdfParams<float>& p, V3 wo, float ul, float u0, float u1) {
  BsdfSample s;
  s.pdf = 0.f;
  s.wi = v3(0.f, 0.f, 1.f);
  if (wo.z <= 0.f) return s;
  if (p.model == kModelLambert) {
    s.wi = sample_cosine_hemisphere(u0, u1);
    s.pdf = s.wi.z * kInvPi;
    if (!(s.wi.z > 0.f)) s.pdf = 0.f;
    return s;
  }
  V3 wi;
  if (!bsdf_sample_dir(p, bsdf_pre(p, wo), wo, ul, u0, u1, wi)) return s;
  s.wi = wi;
  s.pdf = bsdf_pdf(p, wo, wi);
  return s;
}

}  // namespace trt
