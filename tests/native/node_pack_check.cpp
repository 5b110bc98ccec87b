// Property check of trt::pack_nodes, the 48-byte device node layout (host code only).
// Prints "ok <nodes>" or the first violation and exits non-zero.
//   1. the child references decoded from every packed node equal the 64-byte node's;
//   2. every packed bound lies outward of the exact one (lower bounds not above,
//      upper bounds not below) and within 256 ulp of it; z bounds are unchanged;
//   3. bounds of either sign, at zero and at large magnitudes pack; an infinite bound
//      is refused.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "bvh_build.h"

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
  for (size_t i = 0; i < dn.size(); ++i) {
    const trt::DevNode2& a = dn[i];
    const trt::DevNode48& p = pk[i];
    if (trt::node_child_ref(p.r0[0], p.r0[1], p.r0[2], p.r0[3]) != a.c[0]) return fail("child0 ref", i);
    if (trt::node_child_ref(p.r1[0], p.r1[1], p.r1[2], p.r1[3]) != a.c[1]) return fail("child1 ref", i);
    for (int k = 0; k < 4; ++k) {
      const bool upper = (k & 1) != 0;
      const float e[2] = {a.r0[k], a.r1[k]}, s[2] = {p.r0[k], p.r1[k]};
      for (int r = 0; r < 2; ++r) {
        if (upper ? !(s[r] >= e[r]) : !(s[r] <= e[r])) return fail("bound moved inward", i);
        if (ulp_dist(s[r], e[r]) > 256) return fail("bound moved more than 256 ulp", i);
        // nearest: stepping outward one ulp at a time, the first float with the carried
        // low byte is the packed bound
        float x = e[r];
        const uint32_t want = trt::f2u(s[r]) & 0xffu;
        while ((trt::f2u(x) & 0xffu) != want) x = std::nextafter(x, upper ? HUGE_VALF : -HUGE_VALF);
        if (trt::f2u(x) != trt::f2u(s[r])) return fail("packed bound is not the nearest carrying its byte", i);
      }
      if (std::memcmp(&p.r2[k], &a.r2[k], 4) != 0) return fail("z bound changed", i);
    }
  }
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
