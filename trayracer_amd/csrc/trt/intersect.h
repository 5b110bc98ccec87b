// Ray/box and watertight ray/triangle intersection.
//
// Replaces the closest-hit / any-hit semantics of the reference's OptiX programs
// (Optix/GPU/optix.cu:195-218 for the hit programs, :264-265 closest-hit traces with
// t in [rayEpsilon, 1e30), :333-335 shadow traces with t in [eps, dist - 2 eps)).
// OptiX's own triangle test is vendor code; this build uses the watertight test of
// Woop, Benthin and Wald (JCGT 2013) with a double-precision re-evaluation of the
// edge functions when any of them is exactly zero, as the paper prescribes.
// Triangles are two-sided (the reference builds with face culling disabled,
// Resources/Model.cpp:270).
//
// Tie rule: a candidate replaces the current closest hit when t < t_best, or when
// t == t_best and its (instance, primitive) key is smaller. The result is therefore
// independent of traversal order and of the acceleration structure.
#pragma once
#include "common.h"

namespace trt {

struct Ray {
  V3 o;
  float tmin;
  V3 d;
  float tmax;
};

// Per-ray constants of the watertight test (Woop et al. 2013, section 3).
struct RayTri {
  V3 o;
  int kx, ky, kz;
  float sx, sy, sz;
};

TRT_FI RayTri ray_tri_setup(V3 o, V3 d) {
  RayTri r;
  r.o = o;
  const float ax = fabs(d.x), ay = fabs(d.y), az = fabs(d.z);
  int kz = 2;
  if (ax > ay && ax > az) kz = 0;
  else if (ay > az) kz = 1;
  int kx = kz + 1 == 3 ? 0 : kz + 1;
  int ky = kx + 1 == 3 ? 0 : kx + 1;
  const float dz = comp(d, kz);
  if (dz < 0.f) {  // keep the winding direction
    const int t = kx;
    kx = ky;
    ky = t;
  }
  r.kx = kx;
  r.ky = ky;
  r.kz = kz;
  r.sx = comp(d, kx) / dz;
  r.sy = comp(d, ky) / dz;
  r.sz = 1.f / dz;
  return r;
}

struct TriHit {
  float t, b1, b2;  // barycentric weights of v1 and v2
};

// Returns true and fills h when the ray hits the (two-sided) triangle with
// tmin <= t <= tmax.
TRT_FI bool intersect_triangle(const RayTri& r, V3 v0, V3 v1, V3 v2, float tmin, float tmax, TriHit& h) {
  const V3 A = v0 - r.o, B = v1 - r.o, C = v2 - r.o;
  const float Akz = comp(A, r.kz), Bkz = comp(B, r.kz), Ckz = comp(C, r.kz);
  const float Ax = comp(A, r.kx) - r.sx * Akz;
  const float Ay = comp(A, r.ky) - r.sy * Akz;
  const float Bx = comp(B, r.kx) - r.sx * Bkz;
  const float By = comp(B, r.ky) - r.sy * Bkz;
  const float Cx = comp(C, r.kx) - r.sx * Ckz;
  const float Cy = comp(C, r.ky) - r.sy * Ckz;
  float U = Cx * By - Cy * Bx;
  float V = Ax * Cy - Ay * Cx;
  float W = Bx * Ay - By * Ax;
  if (U == 0.f || V == 0.f || W == 0.f) {
    const double CxBy = static_cast<double>(Cx) * static_cast<double>(By);
    const double CyBx = static_cast<double>(Cy) * static_cast<double>(Bx);
    U = static_cast<float>(CxBy - CyBx);
    const double AxCy = static_cast<double>(Ax) * static_cast<double>(Cy);
    const double AyCx = static_cast<double>(Ay) * static_cast<double>(Cx);
    V = static_cast<float>(AxCy - AyCx);
    const double BxAy = static_cast<double>(Bx) * static_cast<double>(Ay);
    const double ByAx = static_cast<double>(By) * static_cast<double>(Ax);
    W = static_cast<float>(BxAy - ByAx);
  }
  if ((U < 0.f || V < 0.f || W < 0.f) && (U > 0.f || V > 0.f || W > 0.f)) return false;
  const float det = U + V + W;
  if (det == 0.f) return false;
  const float Az = r.sz * Akz, Bz = r.sz * Bkz, Cz = r.sz * Ckz;
  const float T = U * Az + V * Bz + W * Cz;
  const float t = T / det;
  if (!(t >= tmin && t <= tmax)) return false;
  const float inv = 1.f / det;
  h.t = t;
  h.b1 = V * inv;
  h.b2 = W * inv;
  return true;
}

// (t, key) lexicographic order used by every closest-hit search
TRT_FI bool closer(float t, uint32_t key, float tbest, uint32_t keybest) {
  return t < tbest || (t == tbest && key < keybest);
}

// Slack of the traversal's conservative slab test (Ize 2013, "Robust BVH ray traversal";
// the test itself is in the device traversal): each slab distance is (bound - o) * inv
// with inv = 1/d, so its relative error is a few ulp whatever the magnitude of o. The
// far distance is widened by kBoxSlack (~128 ulp), which covers both the slab error and
// the error of a triangle's computed t, so a triangle lying on a box face (axis-aligned
// geometry, equal-t ties) is never culled by rounding. The traversal result therefore
// equals a brute-force loop over the same triangles.
constexpr float kBoxSlack = 1.0000153f;  // 1 + 2^-16

}  // namespace trt
