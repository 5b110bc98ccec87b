// Property check of trt::presplit_triangles and the BVH built over its references
// (host code only). Prints "ok" or the first violation and exits non-zero.
//   1. every piece box lies inside its triangle's box;
//   2. the pieces of a triangle cover it: its vertices, edge points and interior
//      points (barycentric, evaluated in double) each lie in at least one piece box;
//   3. the reference count is within the budget n * (1 + beta) (+ rounding);
//   4. every triangle keeps at least one reference;
//   5. the binned-SAH tree over the references holds each reference exactly once, and
//      each leaf box contains its references' boxes.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "bvh_build.h"

static int fail(const char* what, size_t i) {
  std::printf("FAIL %s at %zu\n", what, i);
  return 1;
}

int main(int argc, char** argv) {
  const float beta = argc > 1 ? float(std::atof(argv[1])) : 1.0f;
  const uint32_t n = argc > 2 ? uint32_t(std::atoi(argv[2])) : 20000;
  std::mt19937 rng(7);
  std::uniform_real_distribution<float> u(0.f, 1.f);
  std::vector<float> pos(size_t(n) * 9);
  for (uint32_t i = 0; i < n; ++i) {
    float* p = &pos[size_t(i) * 9];
    const int kind = i % 5;
    for (int k = 0; k < 9; ++k) p[k] = u(rng);
    if (kind == 1) {  // axis-aligned (z = const) large triangle
      p[2] = p[5] = p[8] = 0.25f;
    } else if (kind == 2) {  // thin sliver
      for (int k = 0; k < 3; ++k) p[6 + k] = p[k] + 1e-3f * (p[3 + k] - p[k]) + 1e-6f;
    } else if (kind == 3) {  // degenerate: repeated vertex
      for (int k = 0; k < 3; ++k) p[6 + k] = p[k];
    } else if (kind == 4) {  // small triangle
      for (int k = 3; k < 9; ++k) p[k] = p[k % 3] + 0.01f * (u(rng) - 0.5f);
    }
  }
  std::vector<float> box;
  std::vector<uint32_t> refp;
  trt::presplit_triangles(pos.data(), n, beta, box, refp);
  if (box.size() != refp.size() * 6) return fail("box/ref size", 0);
  if (double(refp.size()) > double(n) * (1.0 + beta) + 1.0) return fail("budget", refp.size());
  std::vector<std::vector<uint32_t>> pieces(n);
  for (size_t r = 0; r < refp.size(); ++r) {
    if (refp[r] >= n) return fail("ref prim range", r);
    pieces[refp[r]].push_back(uint32_t(r));
  }
  for (uint32_t i = 0; i < n; ++i) {
    if (pieces[i].empty()) return fail("triangle without reference", i);
    const float* p = &pos[size_t(i) * 9];
    for (uint32_t r : pieces[i]) {
      const float* b = &box[size_t(r) * 6];
      for (int k = 0; k < 3; ++k) {
        const float lo = std::fmin(p[k], std::fmin(p[3 + k], p[6 + k]));
        const float hi = std::fmax(p[k], std::fmax(p[3 + k], p[6 + k]));
        if (b[k] < lo || b[3 + k] > hi || b[k] > b[3 + k]) return fail("piece outside triangle box", i);
      }
    }
    for (int s = 0; s < 64; ++s) {
      double w0, w1;
      if (s < 3) {
        w0 = s == 0, w1 = s == 1;
      } else if (s < 12) {  // edges
        const double t = u(rng);
        const int e = s % 3;
        w0 = e == 0 ? t : (e == 1 ? 0.0 : 1.0 - t);
        w1 = e == 0 ? 1.0 - t : (e == 1 ? t : 0.0);
      } else {
        w0 = u(rng), w1 = u(rng);
        if (w0 + w1 > 1.0) w0 = 1.0 - w0, w1 = 1.0 - w1;
      }
      const double w2 = 1.0 - w0 - w1;
      double q[3];
      for (int k = 0; k < 3; ++k) q[k] = w0 * p[k] + w1 * p[3 + k] + w2 * p[6 + k];
      bool in = false;
      for (uint32_t r : pieces[i]) {
        const float* b = &box[size_t(r) * 6];
        bool all = true;
        // tolerance: the double evaluation of q itself carries ~1e-16 relative error
        for (int k = 0; k < 3; ++k) all = all && q[k] >= double(b[k]) - 1e-12 && q[k] <= double(b[3 + k]) + 1e-12;
        in = in || all;
      }
      if (!in) return fail("point not covered", i);
    }
  }
  trt::BuildResult br = trt::build_binned_sah(box.data(), uint32_t(refp.size()), 32, 4, 4, trt::kTravCostTriangles);
  std::vector<int> seen(refp.size(), 0);
  for (const trt::BuildNode& nd : br.nodes) {
    if (nd.count == 0) continue;
    for (uint32_t j = nd.first; j < nd.first + nd.count; ++j) {
      const uint32_t r = br.prim_order[j];
      seen[r]++;
      for (int k = 0; k < 3; ++k)
        if (box[size_t(r) * 6 + k] < nd.lo[k] || box[size_t(r) * 6 + 3 + k] > nd.hi[k]) return fail("leaf box", r);
    }
  }
  for (size_t r = 0; r < seen.size(); ++r)
    if (seen[r] != 1) return fail("reference not exactly once in leaves", r);
  std::printf("ok refs %zu (x%.3f) sah %.1f depth %d\n", refp.size(), double(refp.size()) / n, br.sah_cost, br.depth);
  return 0;
}
