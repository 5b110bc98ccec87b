// Display transform of the accumulated image (replaces the reference's tone-mapping
// post pass, Tracer/glsl/Tonemap.frag): colour x exposure, one of seven published
// tone curves, gamma correction, 8-bit quantisation. Curves, with the method numbers
// of the reference's `tonemapMethod` uniform (Tracer/glsl/Tonemap.json):
//   0 ACES fit (Narkowicz 2015), 1 filmic (Hejl and Burgess-Dawson, returned in linear
//   space as the reference does), 2 Lottes 2016, 3 Reinhard, 4 Reinhard with white
//   point 4, 5 Uchimura 2017 ("GT"), 6 Uncharted 2 (Hable 2010, exposure bias 2,
//   white 11.2); any other number passes the colour through.
// Shared by the device kernel and the host oracle; uses the project's own pow/exp.
#pragma once
#include "common.h"

namespace trt {

struct TonemapParams {
  uint32_t method;
  float exposure;
  float gamma;
};

TRT_FI float tm_aces(float x) {
  const float a = 2.51f, b = 0.03f, c = 2.43f, d = 0.59f, e = 0.14f;
  return clamp((x * (a * x + b)) / (x * (c * x + d) + e), 0.f, 1.f);
}

TRT_FI float tm_filmic(float x) {
  const float y = fmax(0.f, x - 0.004f);
  return pow((y * (6.2f * y + 0.5f)) / (y * (6.2f * y + 1.7f) + 0.06f), 2.2f);
}

// Lottes: x^a / (x^(a d) b + c) with b, c chosen so that mid grey 0.18 -> 0.267 and
// hdr_max 8 -> 1 (a = 1.6 contrast, d = 0.977 shoulder).
TRT_FI float tm_lottes(float x) {
  const float a = 1.6f, d = 0.977f, hdr_max = 8.f, mid_in = 0.18f, mid_out = 0.267f;
  const float ma = pow(mid_in, a), mad = pow(mid_in, a * d), ha = pow(hdr_max, a), had = pow(hdr_max, a * d);
  const float b = (-ma + ha * mid_out) / ((had - mad) * mid_out);
  const float c = (had * ma - ha * mad * mid_out) / ((had - mad) * mid_out);
  return pow(x, a) / (pow(x, a * d) * b + c);
}

TRT_FI float tm_reinhard(float x) { return x / (1.f + x); }

TRT_FI float tm_reinhard2(float x) {
  const float w = 4.f;
  return (x * (1.f + x / (w * w))) / (1.f + x);
}

// Uchimura's "GT" curve: toe m (x/m)^c, linear section of length l from m, and an
// exponential shoulder to the display maximum P, blended by smoothstep/step weights.
TRT_FI float tm_uchimura(float x) {
  const float P = 1.f, a = 1.f, m = 0.22f, l = 0.4f, c = 1.33f, b = 0.f;
  const float l0 = ((P - m) * l) / a;
  const float S0 = m + l0, S1 = m + a * l0;
  const float C2 = (a * P) / (P - S1), CP = -C2 / P;
  const float ts = clamp(x / m, 0.f, 1.f);
  const float w0 = 1.f - ts * ts * (3.f - 2.f * ts);
  const float w2 = x >= m + l0 ? 1.f : 0.f;
  const float w1 = 1.f - w0 - w2;
  const float T = m * pow(x / m, c) + b;
  const float S = P - (P - S1) * exp(CP * (x - S0));
  const float L = m + a * (x - m);
  return T * w0 + L * w1 + S * w2;
}

TRT_FI float tm_hable_partial(float x) {
  const float A = 0.15f, B = 0.50f, C = 0.10f, D = 0.20f, E = 0.02f, F = 0.30f;
  return ((x * (A * x + C * B) + D * E) / (x * (A * x + B) + D * F)) - E / F;
}

TRT_FI float tm_uncharted2(float x) { return tm_hable_partial(2.f * x) * (1.f / tm_hable_partial(11.2f)); }

TRT_FI float tonemap_channel(uint32_t method, float x) {
  switch (method) {
    case 0: return tm_aces(x);
    case 1: return tm_filmic(x);
    case 2: return tm_lottes(x);
    case 3: return tm_reinhard(x);
    case 4: return tm_reinhard2(x);
    case 5: return tm_uchimura(x);
    case 6: return tm_uncharted2(x);
    default: return x;
  }
}

// Linear colour -> 8-bit display value of one channel.
TRT_FI uint32_t tonemap_u8(const TonemapParams& p, float linear) {
  float v = tonemap_channel(p.method, linear * p.exposure);
  v = pow(fmax(v, 0.f), 1.f / p.gamma);  // gamma correction (pow(0, .) = 0)
  v = clamp(v, 0.f, 1.f);
  return static_cast<uint32_t>(fma(v, 255.f, 0.5f));
}

}  // namespace trt
