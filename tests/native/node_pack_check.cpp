// Property check of trt::pack_nodes, the 48-byte device node layout (centre and half
// extent per axis; host code only). Prints "ok <nodes>" or the first violation and exits
// non-zero.
//   1. the child references decoded from every packed node equal the 64-byte node's,
//      internal children as byte offsets (index x 48), leaf codes unchanged;
//   2. every stored box [c - h, c + h] contains the exact one, with h at least
//      (1 + 2^-10) x the exact half extent about the stored centre; a centre that carries
//      a byte lies within 255 ulp of the midpoint, and h exceeds the needed value by at
//      most 256 ulp (+ its rounding);
//   3. rays that touch an exact child box in FP64 (aimed at its faces, edges and
//      interior, from inside it just short of a face, some with a zero direction
//      component) pass the device's float slab test (x = c - o, s = sign(inv), entry
//      (x - s h) inv, exit (x + s h) inv, NaN-ignoring combine, far distance times
//      kBoxSlack), and rays with a zero component whose origin lies outside the box's
//      slab on that axis fail it;
//   4. bounds of either sign, at zero and at large magnitudes pack; an infinite bound
//      is refused.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "bvh_build.h"
#include "trt/intersect.h"

static int fail(const char* what, size_t i) {
  std::printf("FAIL %s at %zu\n", what, i);
  return 1;
}

// distance in representable floats (ordered integer line)
static long long ulp_dist(float a, float b) {
  auto ord = [](float f) {
    int32_t i;
    std::memcpy(&i, &f, 4);
    return i < 0 ? -static_cast<long long>(i & 0x7fffffff) : static_cast<long long>(i);
  };
  return std::llabs(ord(a) - ord(b));
}

int main(int argc, char** argv) {
  const uint32_t n = argc > 1 ? uint32_t(std::atoi(argv[1])) : 20000;
  std::mt19937 rng(11);
  std::uniform_real_distribution<float> u(-1.f, 1.f);
  std::vector<float> box(size_t(n) * 6);
  for (uint32_t i = 0; i < n; ++i) {
    const float scale = (i % 4 == 0) ? 1e4f : ((i % 4 == 1) ? 1e-3f : 1.f);
    for (int k = 0; k < 3; ++k) {
      const float c = u(rng) * scale, e = std::fabs(u(rng)) * scale * 0.01f;
      box[size_t(i) * 6 + k] = (i % 7 == 0) ? 0.f : c - e;  // some bounds exactly at zero
      box[size_t(i) * 6 + 3 + k] = (i % 7 == 0) ? 0.f : c + e;
    }
  }
  const trt::BuildResult b = trt::build_binned_sah(box.data(), n, 32, 4, 2, trt::kTravCostTriangles);
  const std::vector<trt::DevNode2> dn = trt::to_device_layout(b);
  std::vector<trt::DevNode48> pk;
  if (!trt::pack_nodes(dn, pk)) return fail("pack refused finite boxes", 0);
  if (pk.size() != dn.size()) return fail("size", pk.size());
  std::mt19937 rr(5);
  std::uniform_real_distribution<double> ur(-1.0, 1.0);
  uint64_t rays = 0, touching = 0, culled = 0;
  for (size_t i = 0; i < dn.size(); ++i) {
    const trt::DevNode2& a = dn[i];
    const trt::DevNode48& p = pk[i];
    if (trt::node_child_ref(p.r0[0], p.r0[1], p.r0[2], p.r0[3]) != trt::device_child_ref(a.c[0])) return fail("child0 ref", i);
    if (trt::node_child_ref(p.r1[0], p.r1[1], p.r1[2], p.r1[3]) != trt::device_child_ref(a.c[1])) return fail("child1 ref", i);
    const float lo[2][3] = {{a.r0[0], a.r0[2], a.r2[0]}, {a.r1[0], a.r1[2], a.r2[2]}};
    const float hi[2][3] = {{a.r0[1], a.r0[3], a.r2[1]}, {a.r1[1], a.r1[3], a.r2[3]}};
    const float cs[2][3] = {{p.r0[0], p.r0[2], p.r2[0]}, {p.r1[0], p.r1[2], p.r2[2]}};
    const float hs[2][3] = {{p.r0[1], p.r0[3], p.r2[1]}, {p.r1[1], p.r1[3], p.r2[3]}};
    for (int c = 0; c < 2; ++c) {
      for (int k = 0; k < 3; ++k) {
        const double need = std::max(double(hi[c][k]) - cs[c][k], double(cs[c][k]) - lo[c][k]);
        if (!(double(cs[c][k]) - hs[c][k] <= lo[c][k]) || !(double(cs[c][k]) + hs[c][k] >= hi[c][k]))
          return fail("stored box does not contain the exact one", i);
        if (!(double(hs[c][k]) >= need * (1.0 + trt::kHalfExtentMargin))) return fail("half-extent margin", i);
        const float mid = float(0.5 * (double(lo[c][k]) + double(hi[c][k])));
        if (ulp_dist(cs[c][k], mid) > (k < 2 ? 255 : 0)) return fail("centre moved too far", i);
        const float hneed = float(need * (1.0 + trt::kHalfExtentMargin));
        if (ulp_dist(hs[c][k], hneed) > (k < 2 ? 258 : 2)) return fail("half extent larger than needed", i);
      }
      for (int r = 0; r < 12; ++r) {
        double o[3], d[3];
        for (int k = 0; k < 3; ++k) {
          const double ctr = 0.5 * (double(lo[c][k]) + double(hi[c][k]));
          const double ext = std::max(double(hi[c][k]) - double(lo[c][k]), 1e-6 * (1.0 + std::fabs(ctr)));
          o[k] = ctr + ur(rr) * ext * 4.0;
          const double tgt = (r & 1) ? (ur(rr) > 0 ? double(hi[c][k]) : double(lo[c][k])) : ctr + 0.5 * ur(rr) * ext;
          d[k] = tgt - o[k];
        }
        if (r == 7) d[r % 3] = 0.0;  // a zero direction component
        if (r >= 8) {  // from inside the box, just short of a face, towards it: a tiny exit distance
          const int k = r % 3;
          for (int j = 0; j < 3; ++j) {
            o[j] = double(lo[c][j]) + (0.25 + 0.5 * std::fabs(ur(rr))) * (double(hi[c][j]) - double(lo[c][j]));
            d[j] = 1e-3 * ur(rr);
          }
          const double w = double(hi[c][k]) - double(lo[c][k]);
          o[k] = double(hi[c][k]) - w * 1e-6;
          d[k] = 1.0;
        }
        const float fo[3] = {float(o[0]), float(o[1]), float(o[2])};
        const float fd[3] = {float(d[0]), float(d[1]), float(d[2])};
        double tn = 0.0, tf = 1e30;
        bool hit = true;
        for (int k = 0; k < 3; ++k) {
          if (fd[k] == 0.f) {
            if (double(fo[k]) < lo[c][k] || double(fo[k]) > hi[c][k]) hit = false;
            continue;
          }
          double t0 = (double(lo[c][k]) - fo[k]) / fd[k], t1 = (double(hi[c][k]) - fo[k]) / fd[k];
          if (t0 > t1) std::swap(t0, t1);
          tn = std::max(tn, t0);
          tf = std::min(tf, t1);
        }
        ++rays;
        if (r == 7 && !hit) {  // parallel to an axis, outside the stored slab on it: culled
          const int k = r % 3;
          if (double(fo[k]) < double(cs[c][k]) - hs[c][k] || double(fo[k]) > double(cs[c][k]) + hs[c][k]) {
            const float inv = 1.f / fd[k], sg = std::copysign(1.f, inv);
            const float x = cs[c][k] - fo[k], sh = hs[c][k] * sg;
            const float ne = (x - sh) * inv, fa = (x + sh) * inv;
            if (!(ne > fa) && !(ne == fa && std::isinf(ne))) return fail("parallel ray outside its slab not culled", i);
            ++culled;
          }
        }
        if (!(hit && tn <= tf)) continue;
        ++touching;
        float nmax = 0.f, fmin = 1e30f;  // tmin = 0, tmax = 1e30
        for (int k = 0; k < 3; ++k) {
          const float inv = 1.f / fd[k], sg = std::copysign(1.f, inv);
          const float x = cs[c][k] - fo[k], sh = hs[c][k] * sg;
          const float ne = (x - sh) * inv, fa = (x + sh) * inv;
          if (!std::isnan(ne)) nmax = std::max(nmax, ne);
          if (!std::isnan(fa)) fmin = std::min(fmin, fa);
        }
        if (!(nmax <= fmin * trt::kBoxSlack)) return fail("slab test culled a box the ray touches", i);
      }
    }
  }
  if (touching < rays / 8) return fail("too few touching rays", touching);
  if (culled < 100) return fail("too few parallel rays outside their slab", culled);
  // an infinite bound cannot carry a byte other than its own low byte (0)
  std::vector<trt::DevNode2> bad(1);
  std::memset(&bad[0], 0, sizeof(bad[0]));
  bad[0].r0[1] = HUGE_VALF;
  bad[0].c[0] = 0x00000500;
  std::vector<trt::DevNode48> out;
  if (trt::pack_nodes(bad, out)) return fail("infinite bound accepted", 0);
  std::printf("ok %zu\n", pk.size());
  return 0;
}
