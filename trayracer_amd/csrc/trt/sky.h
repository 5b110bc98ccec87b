// Analytic daylight sky seen by rays that leave the scene (replaces SampleSky,
// Tracer/CUDA/GPU/CudaSky.h:44-156, with Sky::Build, Tracer/Renderer/Sky.cpp:81-109).
//
// The model is Preetham, Shirley and Smits (1999, "A practical analytic model for
// daylight") in the real-time form that the reference adapts (the public "Sky-Shader"
// of three.js lineage): Rayleigh and Mie total scattering for three primary
// wavelengths, Rayleigh and Henyey-Greenstein phase functions, an optical-length
// approximation per zenith angle, the sun's horizon falloff, a sun disc, and that
// shader's tonemap, saturation and sunfade exponent, so values match the reference's
// radiance. Everything is evaluated with the project's own exp/log/atan (common.h) on
// both host and device, so the oracle and the GPU agree bit for bit.
//
// By default, as in the reference (CudaLights.h:11, ENABLE_SKY_NEE false), the sky is
// not sampled by next-event estimation: it is reached by BSDF-sampled rays that miss,
// and the sun disc is drawn only for camera rays and specular bounces
// (PathTracing.h:40-55). With `sun_nee` the sun disc becomes a light of next-event
// estimation, chosen against the emissive triangles by energy as the reference's
// disabled branch does (CudaLights.h:24, 66-88): its direction is sampled uniformly in
// the disc's cone, and BSDF-sampled rays that reach the disc add its radiance with the
// balance-heuristic MIS weight. Its selection weight is the disc's radiance x solid
// angle x the squared radius of the scene's bounds (the power a sphere around the scene
// intercepts, in the triangles' radiance x area units).
#pragma once
#include "common.h"
#include "sampling.h"

namespace trt {

struct SkyParams {
  uint32_t enabled;
  uint32_t draw_sun;
  float sun_dir[3];    // unit vector towards the sun
  float cos_disc;      // cosine of the sun's angular diameter (the disc test's threshold)
  float sun_e;         // sun intensity after the horizon falloff
  float beta_r[3];     // total Rayleigh scattering x Rayleigh coefficient, per primary
  float beta_m[3];     // total Mie scattering x Mie coefficient, per primary
  float mix_w;         // saturate((1 - cos(sun, up))^5): weight of the in-scattering correction
  float gamma;         // final exponent 1 / (1 + sunfade)
  uint32_t sun_nee;    // 1: the sun disc is a light of next-event estimation
  float sun_pdf;       // 1 / solid angle of the disc's cone (uniform direction sampling)
  float sun_power;     // selection weight against the emissive triangles (set per scene)
};

constexpr float kSkyPi = 3.14159265358979f;

// acos for x in [-1, 1] via atan (sqrt((1 - x)(1 + x)) / x); acos(0) = pi / 2.
TRT_FI float acos_sky(float x) {
  const float a = fabs(x);
  const float s = sqrt(fmax((1.f - a) * (1.f + a), 0.f));
  const float r = a > 0.f ? atan(s / a) : 1.57079632679f;
  return x < 0.f ? kSkyPi - r : r;
}

TRT_FI float saturate(float x) { return clamp(x, 0.f, 1.f); }

TRT_FI float smoothstep_sky(float e0, float e1, float x) {
  const float t = saturate((x - e0) / (e1 - e0));
  return t * t * (3.f - 2.f * t);
}

// Precomputes the direction-independent terms (host side, once per scene).
// angular_diameter in arc minutes (the reference's unit), intensity of the sun,
// turbidity of the atmosphere.
TRT_HD inline SkyParams make_sky(bool enabled, bool draw_sun, const float dir[3], float angular_diameter,
                                 float intensity, float turbidity, bool sun_nee = false) {
  SkyParams s{};
  s.enabled = enabled ? 1u : 0u;
  s.draw_sun = draw_sun ? 1u : 0u;
  s.sun_nee = (enabled && draw_sun && sun_nee) ? 1u : 0u;
  V3 d = v3(dir[0], dir[1], dir[2]);
  const float l = length(d);
  d = l > 0.f ? d * (1.f / l) : v3(0.f, 1.f, 0.f);
  s.sun_dir[0] = d.x;
  s.sun_dir[1] = d.y;
  s.sun_dir[2] = d.z;
  const float diam = angular_diameter * (kSkyPi / (60.f * 180.f));
  float sn, cs;
  sincos_rad(diam, sn, cs);
  s.cos_disc = cs;
  // the disc's cone: 1 - cos = 2 sin^2(diam / 2), free of cancellation for small discs
  float sh, ch;
  sincos_rad(0.5f * diam, sh, ch);
  const float one_m_cos = 2.f * sh * sh;
  s.sun_pdf = one_m_cos > 0.f ? 1.f / (2.f * kSkyPi * one_m_cos) : 0.f;
  if (!(one_m_cos > 0.f)) s.sun_nee = 0u;
  const float cos_up = d.y;
  // sun intensity falling off below the horizon ("earth shadow"): cutoff pi/1.95, steepness 1.5
  s.sun_e = intensity * fmax(0.f, 1.f - exp(-((kSkyPi / 1.95f - acos_sky(cos_up)) / 1.5f)));
  float sunfade = 1.f - saturate(1.f - exp(-d.z / 500.f));
  // Rayleigh: 8 pi^3 (n^2 - 1)^2 (6 + 3 p) / (3 N lambda^4 (6 - 7 p)), n = 1.0003,
  // depolarisation p = 0.031, N = 2.545e25 molecules per m^3
  const float n2m1 = 3e-4f * 2.0003f;  // n^2 - 1 = (n - 1)(n + 1), exact to float rounding
  const float num = 8.f * (kSkyPi * kSkyPi * kSkyPi) * (n2m1 * n2m1) * (6.f + 3.f * 0.031f);
  const float rayleigh = 2.5f - (1.f - sunfade);
  // Mie: 0.434 c pi (2 pi / lambda)^(v - 2) K, c = 0.2 T 1e-17, v = 4, x 0.005
  const float c = 0.2f * turbidity * 10e-18f;
  const float lambda[3] = {620e-9f, 492e-9f, 390e-9f};
  const float K[3] = {0.686f, 0.678f, 0.666f};
  for (int i = 0; i < 3; ++i) {
    const float l2 = lambda[i] * lambda[i];
    s.beta_r[i] = num / (3.f * 2.545e25f * (l2 * l2) * (6.f - 7.f * 0.031f)) * rayleigh;
    const float k = 2.f * kSkyPi / lambda[i];
    s.beta_m[i] = 0.434f * c * kSkyPi * (k * k) * K[i] * 0.005f;
  }
  const float omc = 1.f - cos_up;
  s.mix_w = saturate((omc * omc) * (omc * omc) * omc);
  sunfade = sunfade * lerp(1.f, 0.9f, fmax(d.y, 0.f));
  s.gamma = 1.f / (1.f + sunfade);
  return s;
}

// Radiance of the sky along unit direction `dir` (RGB); `sun` draws the sun disc.
TRT_FI V3 sky_radiance(const SkyParams& s, V3 dir, bool sun) {
  if (!s.enabled) return v3(0.f);
  const V3 sd = v3(s.sun_dir[0], s.sun_dir[1], s.sun_dir[2]);
  // optical length along dir (cut off at the horizon to avoid the singularity)
  const float up = fmax(dir.y, 0.f);
  const float zen_deg = acos_sky(up) * (180.f / kSkyPi);
  const float denom = up + 0.15f * pow(93.885f - zen_deg, -1.253f);
  const float sr = 8.4e3f / denom, sm = 1.25e3f / denom;
  // phases at the angle between dir and the sun
  const float ct = dot(dir, sd);
  const float rc = fma(ct, 0.5f, 0.5f);
  const float rphase = (3.f / (16.f * kSkyPi)) * (1.f + rc * rc);
  const float g = 0.8f;
  const float hg = (1.f / (4.f * kSkyPi)) * ((1.f - g * g) / pow(1.f - 2.f * g * ct + g * g, 1.5f));
  const float disc = sun ? smoothstep_sky(s.cos_disc, s.cos_disc + 2e-5f, ct) : 0.f;
  float lin[3], l0[3];
  for (int i = 0; i < 3; ++i) {
    const float br = s.beta_r[i], bm = s.beta_m[i];
    const float fex = exp(-(br * sr + bm * sm));
    const float ratio = (br * rphase + bm * hg) / (br + bm);
    float li = pow(s.sun_e * ratio * (1.f - fex), 1.5f);
    li = li * lerp(1.f, pow(s.sun_e * ratio * fex, 0.5f), s.mix_w);
    lin[i] = li;
    l0[i] = 0.1f * fex + s.sun_e * 19000.f * fex * disc;
  }
  // saturation: halfway between the luma grey and the colour
  const float luma = 0.299f * l0[0] + 0.587f * l0[1] + 0.114f * l0[2];
  float out[3];
  for (int i = 0; i < 3; ++i) {
    const float col = lin[i] + lerp(luma, l0[i], 0.5f);
    const float tm = fmax(0.f, (col * (1.f + col / 1000.f)) / (col + 128.f));
    out[i] = pow(tm, s.gamma);
  }
  return v3(out[0], out[1], out[2]);
}

// Radiance the sun disc adds along `dir` (zero outside the disc).
TRT_FI V3 sun_radiance(const SkyParams& s, V3 dir) {
  const V3 sd = v3(s.sun_dir[0], s.sun_dir[1], s.sun_dir[2]);
  if (!(dot(dir, sd) > s.cos_disc)) return v3(0.f);
  const V3 a = sky_radiance(s, dir, true), b = sky_radiance(s, dir, false);
  return v3(fmax(a.x - b.x, 0.f), fmax(a.y - b.y, 0.f), fmax(a.z - b.z, 0.f));
}

// Direction uniformly distributed in the disc's cone (pdf = sun_pdf).
TRT_FI V3 sample_sun(const SkyParams& s, float u0, float u1) {
  const float omc = 1.f / (2.f * kSkyPi * s.sun_pdf);  // 1 - cos of the cone
  const float z = 1.f - u0 * omc;
  const float r = sqrt(fmax(u0 * omc * (2.f - u0 * omc), 0.f));  // sin = sqrt(1 - z^2)
  float sn, cs;
  sincos_turns(u1, sn, cs);
  const Frame f = make_frame(v3(s.sun_dir[0], s.sun_dir[1], s.sun_dir[2]));
  return normalize(to_world(f, v3(r * cs, r * sn, z)));
}

// Host: the sun's selection weight for a scene whose bounds have half-diagonal `radius`.
inline float sun_selection_power(const SkyParams& s, float radius) {
  if (!s.sun_nee) return 0.f;
  const V3 l = sun_radiance(s, v3(s.sun_dir[0], s.sun_dir[1], s.sun_dir[2]));
  return (l.x + l.y + l.z) / s.sun_pdf * (radius * radius);
}

}  // namespace trt
