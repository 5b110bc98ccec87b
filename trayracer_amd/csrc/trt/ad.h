// Forward-mode dual numbers with N tangent directions, used to differentiate the BSDF
// evaluation with respect to per-triangle material parameters in the adjoint pass.
// The value part of every operation performs exactly the float operation that the
// plain (float) instantiation performs, so the primal of the adjoint replay is
// bit-identical to the forward render.
#pragma once
#include "common.h"

namespace trt {

template <int N>
struct Dual {
  float v;
  float d[N];
};

template <int N>
TRT_FI Dual<N> dconst(float v) {
  Dual<N> r;
  r.v = v;
  for (int i = 0; i < N; ++i) r.d[i] = 0.f;
  return r;
}
template <int N>
TRT_FI Dual<N> dvar(float v, int k) {
  Dual<N> r = dconst<N>(v);
  r.d[k] = 1.f;
  return r;
}

// scalar traits: value(), make constant
TRT_FI float val(float x) { return x; }
template <int N>
TRT_FI float val(const Dual<N>& x) { return x.v; }

template <int N>
TRT_FI Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r;
  r.v = a.v + b.v;
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i];
  return r;
}
template <int N>
TRT_FI Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r;
  r.v = a.v - b.v;
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i];
  return r;
}
template <int N>
TRT_FI Dual<N> operator-(const Dual<N>& a) {
  Dual<N> r;
  r.v = -a.v;
  for (int i = 0; i < N; ++i) r.d[i] = -a.d[i];
  return r;
}
template <int N>
TRT_FI Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r;
  r.v = a.v * b.v;
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
  return r;
}
template <int N>
TRT_FI Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r;
  r.v = a.v / b.v;
  const float ib = 1.f / b.v;
  for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) * ib;
  return r;
}
// mixed with float
template <int N>
TRT_FI Dual<N> operator+(const Dual<N>& a, float b) {
  Dual<N> r = a;
  r.v = a.v + b;
  return r;
}
template <int N>
TRT_FI Dual<N> operator+(float a, const Dual<N>& b) {
  Dual<N> r = b;
  r.v = a + b.v;
  return r;
}
template <int N>
TRT_FI Dual<N> operator-(const Dual<N>& a, float b) {
  Dual<N> r = a;
  r.v = a.v - b;
  return r;
}
template <int N>
TRT_FI Dual<N> operator-(float a, const Dual<N>& b) {
  Dual<N> r;
  r.v = a - b.v;
  for (int i = 0; i < N; ++i) r.d[i] = -b.d[i];
  return r;
}
template <int N>
TRT_FI Dual<N> operator*(const Dual<N>& a, float b) {
  Dual<N> r;
  r.v = a.v * b;
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b;
  return r;
}
template <int N>
TRT_FI Dual<N> operator*(float a, const Dual<N>& b) {
  Dual<N> r;
  r.v = a * b.v;
  for (int i = 0; i < N; ++i) r.d[i] = a * b.d[i];
  return r;
}
template <int N>
TRT_FI Dual<N> operator/(const Dual<N>& a, float b) {
  Dual<N> r;
  r.v = a.v / b;
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] / b;
  return r;
}
template <int N>
TRT_FI Dual<N> operator/(float a, const Dual<N>& b) {
  Dual<N> r;
  r.v = a / b.v;
  const float ib = 1.f / b.v;
  for (int i = 0; i < N; ++i) r.d[i] = -r.v * b.d[i] * ib;
  return r;
}
template <int N>
TRT_FI Dual<N> sqrt(const Dual<N>& a) {
  Dual<N> r;
  r.v = sqrt(a.v);
  const float k = r.v > 0.f ? 0.5f / r.v : 0.f;
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * k;
  return r;
}
// fma on duals: value uses the fused op, derivative by the product rule
template <int N>
TRT_FI Dual<N> fma(const Dual<N>& a, const Dual<N>& b, const Dual<N>& c) {
  Dual<N> r;
  r.v = fma(a.v, b.v, c.v);
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i] + c.d[i];
  return r;
}
template <int N>
TRT_FI Dual<N> fmax(const Dual<N>& a, float b) {
  return a.v > b ? a : dconst<N>(b);
}
template <int N>
TRT_FI Dual<N> fmin(const Dual<N>& a, float b) {
  return a.v < b ? a : dconst<N>(b);
}
template <int N>
TRT_FI Dual<N> fmax(const Dual<N>& a, const Dual<N>& b) {
  return a.v > b.v ? a : b;
}
template <int N>
TRT_FI Dual<N> clamp(const Dual<N>& x, float lo, float hi) {
  return fmin(fmax(x, lo), hi);
}

// constant of scalar type S (float or Dual<N>)
template <class S>
struct ScalarOps {
  TRT_FI static S c(float v) { return v; }
};
template <int N>
struct ScalarOps<Dual<N>> {
  TRT_FI static Dual<N> c(float v) { return dconst<N>(v); }
};
template <class S>
TRT_FI S sc(float v) {
  return ScalarOps<S>::c(v);
}
// (1-t) a + t b with constant a
template <class S>
TRT_FI S mixT(float a, const S& b, const S& t) {
  return (1.f - t) * a + t * b;
}

// generic helpers usable with float and Dual
template <class S>
TRT_FI S sqT(const S& x) {
  return x * x;
}
template <class S>
TRT_FI S lerpT(const S& a, const S& b, const S& t) {
  return (1.f - t) * a + t * b;
}

// 3-vector over a scalar type S (float or Dual<N>)
template <class S>
struct T3 {
  S x, y, z;
};
template <class S>
TRT_FI T3<S> t3(const S& x, const S& y, const S& z) {
  return T3<S>{x, y, z};
}
template <class S>
TRT_FI T3<S> operator+(const T3<S>& a, const T3<S>& b) {
  return {a.x + b.x, a.y + b.y, a.z + b.z};
}
template <class S>
TRT_FI T3<S> operator*(const T3<S>& a, const T3<S>& b) {
  return {a.x * b.x, a.y * b.y, a.z * b.z};
}
template <class S>
TRT_FI T3<S> operator*(const T3<S>& a, const S& s) {
  return {a.x * s, a.y * s, a.z * s};
}
template <int N>
TRT_FI T3<Dual<N>> operator*(const T3<Dual<N>>& a, float s) {
  return {a.x * s, a.y * s, a.z * s};
}

}  // namespace trt
