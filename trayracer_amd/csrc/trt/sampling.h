// Sampling constructions and shading frames.
//
// Follows the reference's helpers in meaning (CUDA/GPU/CudaUtility.h:181-292:
// OrthonormalBase, WorldToTangent, TangentToWorld, SampleCosineHemisphere; the lens
// samplers at Optix/GPU/optix.cu:51-100), written from the textbook constructions:
// Malley's method for the cosine hemisphere, Shirley-Chiu concentric disk mapping,
// the square-root parametrisation of uniform triangle sampling (Osada et al. 2002),
// Duff et al. 2017 orthonormal basis, and Vose's alias method (1991) for discrete
// power-proportional light selection.
#pragma once
#include "common.h"

namespace trt {

struct Frame {
  V3 t, b, n;
};

// Duff, Burgess, Christensen, Hery, Kensler, Liani, Villemin 2017 (branchless ONB)
TRT_FI Frame make_frame(V3 n) {
  const float sign = n.z >= 0.f ? 1.f : -1.f;
  const float a = -1.f / (sign + n.z);
  const float b = n.x * n.y * a;
  Frame f;
  f.t = v3(fma(sign * n.x * n.x, a, 1.f), sign * b, -sign * n.x);
  f.b = v3(b, fma(n.y * n.y, a, sign), -n.y);
  f.n = n;
  return f;
}
TRT_FI V3 to_local(const Frame& f, V3 v) { return v3(dot(v, f.t), dot(v, f.b), dot(v, f.n)); }
TRT_FI V3 to_world(const Frame& f, V3 v) {
  return v3(fma(v.z, f.n.x, fma(v.y, f.b.x, v.x * f.t.x)), fma(v.z, f.n.y, fma(v.y, f.b.y, v.x * f.t.y)),
            fma(v.z, f.n.z, fma(v.y, f.b.z, v.x * f.t.z)));
}

// Malley: uniform disk point lifted to the hemisphere; pdf = cos(theta)/pi
TRT_FI V3 sample_cosine_hemisphere(float u0, float u1) {
  float s, c;
  sincos_turns(u0, s, c);
  const float r = sqrt(u1);
  return v3(r * c, r * s, safe_sqrt(1.f - u1));
}

// Shirley-Chiu concentric mapping of [0,1)^2 onto the unit disk
TRT_FI void sample_concentric_disk(float u0, float u1, float& dx, float& dy) {
  const float a = fma(2.f, u0, -1.f);
  const float b = fma(2.f, u1, -1.f);
  if (a == 0.f && b == 0.f) {
    dx = 0.f;
    dy = 0.f;
    return;
  }
  float r, turns;
  if (fabs(a) > fabs(b)) {
    r = a;
    turns = 0.125f * (b / a);
  } else {
    r = b;
    turns = 0.25f - 0.125f * (a / b);
  }
  float s, c;
  sincos_turns(turns, s, c);
  dx = r * c;
  dy = r * s;
}

// uniform point on a triangle: barycentrics (b1, b2) of v1, v2 (Osada et al. 2002)
TRT_FI void sample_triangle(float u0, float u1, float& b1, float& b2) {
  const float s = sqrt(u0);
  b1 = s * (1.f - u1);
  b2 = s * u1;
}

// Vose alias table lookup: column from u0, coin from u1.
TRT_FI uint32_t alias_sample(const float* prob, const uint32_t* alias, uint32_t n, float u0, float u1) {
  uint32_t i = static_cast<uint32_t>(u0 * static_cast<float>(n));
  if (i >= n) i = n - 1;
  return u1 < prob[i] ? i : alias[i];
}

TRT_FI V3 reflect(V3 wo, V3 n) {  // mirror of wo about n (both pointing away)
  const float d = 2.f * dot(wo, n);
  return v3(fma(d, n.x, -wo.x), fma(d, n.y, -wo.y), fma(d, n.z, -wo.z));
}

}  // namespace trt
