// Shared host/device core: vector type, scalar helpers and the few transcendental
// functions the path needs, written so that the host build (g++) and the gfx950
// build (hipcc) produce bit-identical results.
//
// Rules every file under trt/ follows (both compilers run with -ffp-contract=off):
//   * only +, -, *, /, sqrt (all correctly rounded IEEE-754 on both targets) and
//     explicit fused multiply-add trt::fma (correctly rounded on both);
//   * min/max are written as selects, so NaN/signed-zero handling is identical;
//   * no libm transcendental calls: sin/cos/tan/log/exp below are polynomial
//     evaluations in fixed order (Taylor series on reduced ranges, Cody-Waite
//     style reductions), so both targets evaluate the same operation sequence.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define TRT_HD __host__ __device__
#define TRT_DEVICE_CODE 1
#else
#define TRT_HD
#endif

#define TRT_FI TRT_HD inline __attribute__((always_inline))

namespace trt {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 0.318309886183790671538f;
constexpr float kTwoPi = 6.28318530717958647692f;
constexpr float kInf = __builtin_huge_valf();

TRT_FI float fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
TRT_FI float sqrt(float x) { return __builtin_sqrtf(x); }
TRT_FI double sqrtd(double x) { return __builtin_sqrt(x); }
TRT_FI float fmin(float a, float b) { return a < b ? a : b; }
TRT_FI float fmax(float a, float b) { return a > b ? a : b; }
TRT_FI float fabs(float a) { return a < 0.f ? -a : a; }
TRT_FI float clamp(float x, float lo, float hi) { return fmin(fmax(x, lo), hi); }
TRT_FI float sq(float x) { return x * x; }
TRT_FI float lerp(float a, float b, float t) { return (1.f - t) * a + t * b; }
TRT_FI float safe_sqrt(float x) { return sqrt(fmax(x, 0.f)); }

TRT_FI uint32_t f2u(float f) { return __builtin_bit_cast(uint32_t, f); }
TRT_FI float u2f(uint32_t u) { return __builtin_bit_cast(float, u); }

// ---------------------------------------------------------------------------------
// Vector
// ---------------------------------------------------------------------------------
struct V3 {
  float x, y, z;
};
TRT_FI V3 v3(float x, float y, float z) { return V3{x, y, z}; }
TRT_FI V3 v3(float s) { return V3{s, s, s}; }
TRT_FI V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
TRT_FI V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
TRT_FI V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
TRT_FI V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
TRT_FI V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
TRT_FI V3 operator*(float s, V3 a) { return {a.x * s, a.y * s, a.z * s}; }
TRT_FI V3 operator/(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
TRT_FI V3& operator+=(V3& a, V3 b) { a = a + b; return a; }
TRT_FI V3& operator*=(V3& a, V3 b) { a = a * b; return a; }
TRT_FI V3& operator*=(V3& a, float s) { a = a * s; return a; }
// dot/cross evaluate in a fixed order with explicit fma (identical on both targets)
TRT_FI float dot(V3 a, V3 b) { return fma(a.z, b.z, fma(a.y, b.y, a.x * b.x)); }
TRT_FI V3 cross(V3 a, V3 b) {
  return {fma(a.y, b.z, -(a.z * b.y)), fma(a.z, b.x, -(a.x * b.z)), fma(a.x, b.y, -(a.y * b.x))};
}
TRT_FI float length(V3 a) { return sqrt(dot(a, a)); }
TRT_FI V3 normalize(V3 a) {
  const float l = length(a);
  return l > 0.f ? a * (1.f / l) : a;
}
TRT_FI float max3(V3 a) { return fmax(a.x, fmax(a.y, a.z)); }
TRT_FI float min3(V3 a) { return fmin(a.x, fmin(a.y, a.z)); }
TRT_FI float sum3(V3 a) { return (a.x + a.y) + a.z; }
TRT_FI float luminance(V3 c) { return fma(0.0721750f, c.z, fma(0.7151522f, c.y, 0.2126729f * c.x)); }
TRT_FI bool is_black(V3 c) { return c.x == 0.f && c.y == 0.f && c.z == 0.f; }
TRT_FI float comp(V3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }
// point on triangle from barycentrics (b1, b2) relative to v0
TRT_FI V3 bary_point(V3 v0, V3 v1, V3 v2, float b1, float b2) {
  const float b0 = (1.f - b1) - b2;
  return v3(fma(b2, v2.x, fma(b1, v1.x, b0 * v0.x)), fma(b2, v2.y, fma(b1, v1.y, b0 * v0.y)),
            fma(b2, v2.z, fma(b1, v1.z, b0 * v0.z)));
}
// replaces NaN/Inf components with 0 (the reference's fixnan, CudaLinearMath.h:138-142)
TRT_FI V3 fix_nan(V3 c) {
  auto f = [](float v) { return (v == v && fabs(v) < kInf) ? v : 0.f; };
  return v3(f(c.x), f(c.y), f(c.z));
}

// ---------------------------------------------------------------------------------
// Transcendentals (own evaluation; identical op order on host and device)
// ---------------------------------------------------------------------------------
// sin(2*pi*r), cos(2*pi*r) for |r| <= 1/8 by Taylor series in r (truncation error
// below 3e-9 relative, i.e. well under one float ulp).
TRT_FI void sincos_turn_small(float r, float& s, float& c) {
  const float r2 = r * r;
  float p = 42.058693944897634f;                // r^9
  p = fma(p, r2, -76.70585975306136f);          // r^7
  p = fma(p, r2, 81.60524927607504f);           // r^5
  p = fma(p, r2, -41.341702240399755f);         // r^3
  p = fma(p, r2, 6.283185307179586f);           // r^1
  s = p * r;
  float q = -26.426256783374388f;               // r^10
  q = fma(q, r2, 60.24464137187664f);           // r^8
  q = fma(q, r2, -85.45681720669371f);          // r^6
  q = fma(q, r2, 64.93939402266828f);           // r^4
  q = fma(q, r2, -19.739208802178716f);         // r^2
  c = fma(q, r2, 1.0f);
}

// sin/cos of 2*pi*u for any finite u ("turns"); exact range reduction.
TRT_FI void sincos_turns(float u, float& s, float& c) {
  // reduce to [0,1): u - floor(u)
  float f = u - __builtin_floorf(u);
  if (f >= 1.f) f = 0.f;
  // quadrant: f = q/4 + r, r in [-1/8, 1/8)
  float q4 = __builtin_floorf(f * 4.f + 0.5f);   // 0..4
  float r = f - q4 * 0.25f;                        // exact (Sterbenz / power-of-two grid)
  int q = static_cast<int>(q4) & 3;
  float ss, cc;
  sincos_turn_small(r, ss, cc);
  if (q == 0) { s = ss; c = cc; }
  else if (q == 1) { s = cc; c = -ss; }
  else if (q == 2) { s = -ss; c = -cc; }
  else { s = -cc; c = ss; }
}

constexpr float kInvTwoPi = 0.159154943091895335768f;
TRT_FI void sincos_rad(float x, float& s, float& c) { sincos_turns(x * kInvTwoPi, s, c); }
TRT_FI float tan_rad(float x) {
  float s, c;
  sincos_rad(x, s, c);
  return s / c;
}

// natural log for x > 0 (finite): x = m * 2^e, m in [sqrt(1/2), sqrt(2));
// log(m) = 2 atanh(t), t = (m-1)/(m+1), |t| <= 0.1716; series to t^11.
TRT_FI float log(float x) {
  if (!(x > 0.f)) return x == 0.f ? -kInf : __builtin_nanf("");
  if (x == kInf) return kInf;
  uint32_t b = f2u(x);
  int e;
  if ((b >> 23) == 0) {  // subnormal: scale up by 2^23
    x = x * 8388608.f;
    b = f2u(x);
    e = static_cast<int>(b >> 23) - 127 - 23;
  } else {
    e = static_cast<int>(b >> 23) - 127;
  }
  float m = u2f((b & 0x007fffffu) | 0x3f800000u);  // [1,2)
  if (m > 1.41421356f) {
    m = m * 0.5f;
    e += 1;
  }
  const float t = (m - 1.f) / (m + 1.f);
  const float t2 = t * t;
  float p = 1.f / 11.f;
  p = fma(p, t2, 1.f / 9.f);
  p = fma(p, t2, 1.f / 7.f);
  p = fma(p, t2, 1.f / 5.f);
  p = fma(p, t2, 1.f / 3.f);
  p = fma(p, t2, 1.f);
  const float lm = 2.f * t * p;
  // e * ln2 split in two parts (Cody-Waite) for accuracy
  const float fe = static_cast<float>(e);
  return fma(fe, 0.693145752f, fma(fe, 1.42860677e-06f, lm));
}

// exp(x): x = k ln2 + r, |r| <= ln2/2, Taylor to r^8, scaled by 2^k.
TRT_FI float exp(float x) {
  if (x > 88.7f) return kInf;
  if (x < -103.9f) return 0.f;
  const float k = __builtin_floorf(fma(x, 1.44269504088896341f, 0.5f));
  const float r = fma(-k, 1.42860677e-06f, fma(-k, 0.693145752f, x));
  float p = 1.f / 40320.f;
  p = fma(p, r, 1.f / 5040.f);
  p = fma(p, r, 1.f / 720.f);
  p = fma(p, r, 1.f / 120.f);
  p = fma(p, r, 1.f / 24.f);
  p = fma(p, r, 1.f / 6.f);
  p = fma(p, r, 0.5f);
  p = fma(p, r, 1.f);
  p = fma(p, r, 1.f);
  int ki = static_cast<int>(k);
  // scale by 2^k in two steps so that subnormal results stay accurate
  int k1 = ki / 2;
  int k2 = ki - k1;
  const float s1 = u2f(static_cast<uint32_t>(k1 + 127) << 23);
  const float s2 = u2f(static_cast<uint32_t>(k2 + 127) << 23);
  return (p * s1) * s2;
}

TRT_FI float pow(float x, float y) {  // x > 0
  if (x <= 0.f) return 0.f;
  return exp(y * log(x));
}

// atan for the camera distortion model: atan(x) = 2 atan(x / (1 + sqrt(1 + x^2)))
// twice (argument <= tan(pi/16)), then Taylor to x^15.
TRT_FI float atan(float x) {
  const bool neg = x < 0.f;
  float a = neg ? -x : x;
  if (a == kInf) return neg ? -1.57079632679f : 1.57079632679f;
  a = a / (1.f + sqrt(fma(a, a, 1.f)));
  a = a / (1.f + sqrt(fma(a, a, 1.f)));
  const float a2 = a * a;
  float p = -1.f / 15.f;
  p = fma(p, a2, 1.f / 13.f);
  p = fma(p, a2, -1.f / 11.f);
  p = fma(p, a2, 1.f / 9.f);
  p = fma(p, a2, -1.f / 7.f);
  p = fma(p, a2, 1.f / 5.f);
  p = fma(p, a2, -1.f / 3.f);
  p = fma(p, a2, 1.f);
  const float r = 4.f * (a * p);
  return neg ? -r : r;
}

}  // namespace trt
