// Camera model and primary-ray generation.
//
// Same model as the reference's GenerateCameraRay (Optix/GPU/optix.cu:142-174) with
// its lens samplers (:51-112) and barrel distortion (:116-138): pixel + jitter mapped
// to [-1,1] with y divided by the aspect ratio (so `fov` is the horizontal field of
// view), thin lens with a disk or polygonal (bokeh) aperture, focal distance, barrel
// distortion and cos^4 vignetting applied to the path throughput. The basis follows
// CameraNode::Transform (Resources/CameraNode.cpp:49-62): forward = normalize(target -
// pos), side = normalize(cross(forward, up)), up' = normalize(cross(side, forward)).
// Pixel (x, y) has index y*W + x; y = 0 is the bottom row.
#pragma once
#include "common.h"
#include "rng.h"
#include "sampling.h"

namespace trt {

struct CameraParams {
  float pos[3];
  float tan_half_fov;
  float fwd[3];
  float aperture;
  float side[3];
  float focal_dist;
  float up[3];
  float distortion;
  float distortion_k;  // atan(d)/d with d = tan_half_fov * distortion / 2 (host precomputed)
  int bokeh_sides;
  float bokeh_rotation;
  float pad;
};

TRT_FI void camera_basis(V3 pos, V3 target, V3 up, V3& fwd, V3& side, V3& upv) {
  fwd = normalize(target - pos);
  side = normalize(cross(fwd, up));
  upv = normalize(cross(side, fwd));
  (void)pos;
}

// Builds CameraParams (host and oracle use the same function). fov in radians.
TRT_FI CameraParams make_camera(V3 pos, V3 target, V3 up, float fov, float aperture, float focal_dist,
                                float distortion, int bokeh_sides, float bokeh_rotation) {
  CameraParams c;
  V3 f, s, u;
  camera_basis(pos, target, up, f, s, u);
  c.pos[0] = pos.x; c.pos[1] = pos.y; c.pos[2] = pos.z;
  c.fwd[0] = f.x; c.fwd[1] = f.y; c.fwd[2] = f.z;
  c.side[0] = s.x; c.side[1] = s.y; c.side[2] = s.z;
  c.up[0] = u.x; c.up[1] = u.y; c.up[2] = u.z;
  c.tan_half_fov = tan_rad(0.5f * fov);
  c.aperture = aperture;
  c.focal_dist = focal_dist > 0.f ? focal_dist : 1.f;
  c.distortion = distortion;
  const float d = c.tan_half_fov * distortion * 0.5f;
  c.distortion_k = d > 0.f ? atan(d) / d : 1.f;
  c.bokeh_sides = bokeh_sides;
  c.bokeh_rotation = bokeh_rotation;
  c.pad = 0.f;
  return c;
}

// Lens sample in [-1,1]^2 aperture units: concentric disk, or a uniformly chosen
// triangle of the regular polygon (centre + two adjacent corners).
TRT_FI void sample_aperture(const CameraParams& c, float u0, float u1, float u2, float& lx, float& ly) {
  if (c.bokeh_sides < 3) {
    sample_concentric_disk(u0, u1, lx, ly);
    return;
  }
  const float n = static_cast<float>(c.bokeh_sides);
  int idx = static_cast<int>(u2 * n);
  if (idx >= c.bokeh_sides) idx = c.bokeh_sides - 1;
  const float step = kTwoPi / n;
  float s0, c0, s1, c1;
  sincos_rad(fma(static_cast<float>(idx), step, c.bokeh_rotation), s0, c0);
  sincos_rad(fma(static_cast<float>(idx + 1), step, c.bokeh_rotation), s1, c1);
  float b1, b2;
  sample_triangle(u0, u1, b1, b2);
  lx = fma(b2, c1, b1 * c0);
  ly = fma(b2, s1, b1 * s0);
}

// Primary ray for pixel (x, y) of a W x H film. Returns the vignetting weight
// (0 when the distorted ray misses the sensor).
// Same with explicit sample values: jitter (jx, jy) in [0,1)^2, lens (l0, l1, l2).
TRT_FI float camera_ray_from(const CameraParams& c, uint32_t x, uint32_t y, uint32_t W, uint32_t H, float jx,
                             float jy, float l0, float l1, float l2, V3& o, V3& d) {
  const float fw = static_cast<float>(W), fh = static_cast<float>(H);
  const float aspect = fw / fh;
  const float sx = fma((static_cast<float>(x) + jx) / fw, 2.f, -1.f);
  const float sy = fma((static_cast<float>(y) + jy) / fh, 2.f, -1.f) / aspect;
  float cx = 0.f, cy = 0.f;
  if (c.aperture > 0.f) {
    sample_aperture(c, l0, l1, l2, cx, cy);
    cx *= c.aperture;
    cy *= c.aperture;
  }
  float lx = c.tan_half_fov * sx - cx / c.focal_dist;
  float ly = c.tan_half_fov * sy - cy / c.focal_dist;
  if (c.distortion > 0.f && !(lx == 0.f && ly == 0.f)) {
    const float scale = sqrt(fma(lx, lx, ly * ly));
    const float distorted = scale * c.distortion * c.distortion_k;
    if (distorted >= 1.57079632679f) return 0.f;
    const float k = tan_rad(distorted) / (scale * c.distortion);
    lx *= k;
    ly *= k;
  }
  const V3 F = v3(c.fwd[0], c.fwd[1], c.fwd[2]);
  const V3 S = v3(c.side[0], c.side[1], c.side[2]);
  const V3 U = v3(c.up[0], c.up[1], c.up[2]);
  d = normalize(v3(fma(ly, U.x, fma(lx, S.x, F.x)), fma(ly, U.y, fma(lx, S.y, F.y)), fma(ly, U.z, fma(lx, S.z, F.z))));
  o = v3(fma(cy, U.x, fma(cx, S.x, c.pos[0])), fma(cy, U.y, fma(cx, S.y, c.pos[1])),
         fma(cy, U.z, fma(cx, S.z, c.pos[2])));
  const float v = fmax(0.f, dot(d, F));
  const float v2 = v * v;
  return v2 * v2;
}

TRT_FI float generate_camera_ray(const CameraParams& c, uint32_t x, uint32_t y, uint32_t W, uint32_t H, Pcg32& rng,
                                 V3& o, V3& d) {
  const float jx = rng_float(rng), jy = rng_float(rng);
  const float l0 = rng_float(rng), l1 = rng_float(rng), l2 = rng_float(rng);
  return camera_ray_from(c, x, y, W, H, jx, jy, l0, l1, l2, o, d);
}

}  // namespace trt
