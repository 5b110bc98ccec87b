// Binned-SAH BVH builder (host side). See bvh_build.h.
#include "bvh_build.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <future>
#include <limits>

namespace trt {
namespace {

struct Box {
  float lo[3], hi[3];
  void reset() {
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::numeric_limits<float>::infinity();
      hi[k] = -std::numeric_limits<float>::infinity();
    }
  }
  void grow(const Box& b) {
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], b.lo[k]);
      hi[k] = std::max(hi[k], b.hi[k]);
    }
  }
  void grow_pt(const float* p) {
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], p[k]);
      hi[k] = std::max(hi[k], p[k]);
    }
  }
  double area() const {
    const double dx = double(hi[0]) - lo[0], dy = double(hi[1]) - lo[1], dz = double(hi[2]) - lo[2];
    if (dx < 0 || dy < 0 || dz < 0) return 0.0;
    return 2.0 * (dx * dy + dy * dz + dz * dx);
  }
};

struct Ctx {
  const float* box;  // 6 per prim
  std::vector<float> cen;  // 3 per prim
  std::vector<uint32_t>* order;
  int bins;
  int max_leaf;
  double trav_cost;
  std::atomic<int> threads_left;
};

struct SubTree {
  std::vector<BuildNode> nodes;
  int depth = 0;
};

Box prim_box(const Ctx& c, uint32_t p) {
  Box b;
  std::memcpy(b.lo, c.box + size_t(p) * 6, 12);
  std::memcpy(b.hi, c.box + size_t(p) * 6 + 3, 12);
  return b;
}

// Builds the subtree over order[begin, end) into st.nodes (local indices); returns its
// root's local index.
int build_rec(Ctx& c, uint32_t begin, uint32_t end, SubTree& st, int level);

int make_leaf(Ctx& c, uint32_t begin, uint32_t end, const Box& bb, SubTree& st) {
  BuildNode nd;
  std::memcpy(nd.lo, bb.lo, 12);
  std::memcpy(nd.hi, bb.hi, 12);
  nd.left = nd.right = -1;
  nd.first = begin;
  nd.count = end - begin;
  st.nodes.push_back(nd);
  (void)c;
  return int(st.nodes.size()) - 1;
}

int build_rec(Ctx& c, uint32_t begin, uint32_t end, SubTree& st, int level) {
  std::vector<uint32_t>& ord = *c.order;
  const uint32_t n = end - begin;
  Box bb, cb;
  bb.reset();
  cb.reset();
  for (uint32_t i = begin; i < end; ++i) {
    bb.grow(prim_box(c, ord[i]));
    cb.grow_pt(&c.cen[size_t(ord[i]) * 3]);
  }
  if (n <= 1) return make_leaf(c, begin, end, bb, st);

  // binned SAH over all three axes
  const int B = c.bins;
  double best_cost = std::numeric_limits<double>::infinity();
  int best_axis = -1, best_split = -1;
  std::vector<Box> bin_box(B);
  std::vector<uint32_t> bin_cnt(B);
  std::vector<double> right_area(B);
  std::vector<uint32_t> right_cnt(B);
  for (int axis = 0; axis < 3; ++axis) {
    const float lo = cb.lo[axis], hi = cb.hi[axis];
    if (!(hi > lo)) continue;
    const double scale = double(B) / (double(hi) - double(lo));
    for (int b = 0; b < B; ++b) {
      bin_box[b].reset();
      bin_cnt[b] = 0;
    }
    for (uint32_t i = begin; i < end; ++i) {
      const uint32_t p = ord[i];
      int b = int((double(c.cen[size_t(p) * 3 + axis]) - lo) * scale);
      b = std::min(std::max(b, 0), B - 1);
      bin_box[b].grow(prim_box(c, p));
      bin_cnt[b]++;
    }
    Box acc;
    acc.reset();
    uint32_t cnt = 0;
    for (int b = B - 1; b > 0; --b) {
      acc.grow(bin_box[b]);
      cnt += bin_cnt[b];
      right_area[b] = acc.area();
      right_cnt[b] = cnt;
    }
    acc.reset();
    cnt = 0;
    for (int s = 1; s < B; ++s) {  // split between bin s-1 and s
      acc.grow(bin_box[s - 1]);
      cnt += bin_cnt[s - 1];
      if (cnt == 0 || right_cnt[s] == 0) continue;
      const double cost = acc.area() * cnt + right_area[s] * right_cnt[s];
      if (cost < best_cost) {
        best_cost = cost;
        best_axis = axis;
        best_split = s;
      }
    }
  }
  const double parent_area = bb.area();
  const double leaf_cost = double(n);
  const double split_cost =
      parent_area > 0 ? c.trav_cost + best_cost / parent_area : std::numeric_limits<double>::infinity();
  if (n <= uint32_t(c.max_leaf) && (best_axis < 0 || leaf_cost <= split_cost)) return make_leaf(c, begin, end, bb, st);

  uint32_t mid;
  if (best_axis < 0) {
    mid = begin + n / 2;  // all centroids coincide: split the range in half
  } else {
    const float lo = cb.lo[best_axis], hi = cb.hi[best_axis];
    const double scale = double(B) / (double(hi) - double(lo));
    auto it = std::partition(ord.begin() + begin, ord.begin() + end, [&](uint32_t p) {
      int b = int((double(c.cen[size_t(p) * 3 + best_axis]) - lo) * scale);
      b = std::min(std::max(b, 0), B - 1);
      return b < best_split;
    });
    mid = uint32_t(it - ord.begin());
    if (mid == begin || mid == end) mid = begin + n / 2;
  }

  BuildNode nd;
  std::memcpy(nd.lo, bb.lo, 12);
  std::memcpy(nd.hi, bb.hi, 12);
  nd.first = 0;
  nd.count = 0;
  st.nodes.push_back(nd);
  const int me = int(st.nodes.size()) - 1;

  const bool par = n > 65536 && c.threads_left.fetch_sub(1) > 0;
  if (par) {
    SubTree right_tree;
    auto fut = std::async(std::launch::async, [&]() { return build_rec(c, mid, end, right_tree, level + 1); });
    const int l = build_rec(c, begin, mid, st, level + 1);
    const int r_local = fut.get();
    const int off = int(st.nodes.size());
    for (BuildNode x : right_tree.nodes) {
      if (x.count == 0) {
        x.left += off;
        x.right += off;
      }
      st.nodes.push_back(x);
    }
    st.nodes[me].left = l;
    st.nodes[me].right = r_local + off;
  } else {
    const int l = build_rec(c, begin, mid, st, level + 1);
    const int r = build_rec(c, mid, end, st, level + 1);
    st.nodes[me].left = l;
    st.nodes[me].right = r;
  }
  return me;
}

void finish_stats(BuildResult& r) {
  // depth in internal nodes, SAH cost, leaf count (iterative DFS)
  struct It { int node, d; };
  std::vector<It> stack{{0, 0}};
  const double root_area = [&] {
    Box b;
    std::memcpy(b.lo, r.nodes[0].lo, 12);
    std::memcpy(b.hi, r.nodes[0].hi, 12);
    return b.area();
  }();
  r.depth = 0;
  r.sah_cost = 0;
  r.n_leaves = 0;
  while (!stack.empty()) {
    const It it = stack.back();
    stack.pop_back();
    const BuildNode& nd = r.nodes[size_t(it.node)];
    Box b;
    std::memcpy(b.lo, nd.lo, 12);
    std::memcpy(b.hi, nd.hi, 12);
    const double a = root_area > 0 ? b.area() / root_area : 1.0;
    if (nd.count > 0 || nd.left < 0) {
      r.n_leaves++;
      r.sah_cost += a * nd.count;
      r.depth = std::max(r.depth, it.d);
    } else {
      r.sah_cost += a;
      stack.push_back({nd.left, it.d + 1});
      stack.push_back({nd.right, it.d + 1});
    }
  }
}

struct Poly {
  int n = 0;
  double v[12][3];
};

void clip_plane(const Poly& in, int axis, double c, bool keep_below, Poly& out) {
  out.n = 0;
  for (int i = 0; i < in.n; ++i) {
    const double* a = in.v[i];
    const double* b = in.v[(i + 1) % in.n];
    const double da = keep_below ? c - a[axis] : a[axis] - c;
    const double db = keep_below ? c - b[axis] : b[axis] - c;
    if (da >= 0 && out.n < 12) {
      std::memcpy(out.v[out.n++], a, sizeof(double) * 3);
    }
    if ((da < 0) != (db < 0) && out.n < 12) {
      const double t = da / (da - db);
      double* o = out.v[out.n++];
      for (int k = 0; k < 3; ++k) o[k] = a[k] + t * (b[k] - a[k]);
      o[axis] = c;
    }
  }
}

void poly_box(const Poly& p, double lo[3], double hi[3]) {
  for (int k = 0; k < 3; ++k) {
    lo[k] = 1e300;
    hi[k] = -1e300;
  }
  for (int i = 0; i < p.n; ++i)
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], p.v[i][k]);
      hi[k] = std::max(hi[k], p.v[i][k]);
    }
}

double box_area(const double lo[3], const double hi[3]) {
  const double dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
  return 2.0 * (dx * dy + dy * dz + dz * dx);
}

void emit_box(const double lo[3], const double hi[3], const float* tlo, const float* thi, uint32_t prim,
              std::vector<float>& box, std::vector<uint32_t>& refs) {
  float b[6];
  for (int k = 0; k < 3; ++k) {
    // outward rounding, clamped to the triangle's own box
    float l = float(lo[k]), h = float(hi[k]);
    if (double(l) > lo[k]) l = std::nextafter(l, -std::numeric_limits<float>::infinity());
    if (double(h) < hi[k]) h = std::nextafter(h, std::numeric_limits<float>::infinity());
    b[k] = std::max(l, tlo[k]);
    b[3 + k] = std::min(h, thi[k]);
  }
  box.insert(box.end(), b, b + 6);
  refs.push_back(prim);
}

void split_rec(const Poly& p, int splits, const float* tlo, const float* thi, uint32_t prim, std::vector<float>& box,
               std::vector<uint32_t>& refs) {
  double lo[3], hi[3];
  poly_box(p, lo, hi);
  if (splits <= 0 || p.n < 3) {
    emit_box(lo, hi, tlo, thi, prim, box, refs);
    return;
  }
  int axis = 0;
  for (int k = 1; k < 3; ++k)
    if (hi[k] - lo[k] > hi[axis] - lo[axis]) axis = k;
  const double c = 0.5 * (lo[axis] + hi[axis]);
  Poly a, b;
  clip_plane(p, axis, c, true, a);
  clip_plane(p, axis, c, false, b);
  if (a.n < 3 || b.n < 3) {
    emit_box(lo, hi, tlo, thi, prim, box, refs);
    return;
  }
  double alo[3], ahi[3], blo[3], bhi[3];
  poly_box(a, alo, ahi);
  poly_box(b, blo, bhi);
  const double aa = box_area(alo, ahi), ab = box_area(blo, bhi);
  const int rest = splits - 1;
  const int sa = aa + ab > 0 ? int(std::lround(rest * aa / (aa + ab))) : rest / 2;
  split_rec(a, sa, tlo, thi, prim, box, refs);
  split_rec(b, rest - sa, tlo, thi, prim, box, refs);
}

}  // namespace

void presplit_triangles(const float* pos, uint32_t n, float beta, std::vector<float>& ref_box,
                        std::vector<uint32_t>& ref_prim) {
  ref_box.clear();
  ref_prim.clear();
  std::vector<double> excess(n);
  double total = 0.0;
  for (uint32_t i = 0; i < n; ++i) {
    const float* p = pos + size_t(i) * 9;
    double lo[3], hi[3];
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min<double>(p[k], std::min<double>(p[3 + k], p[6 + k]));
      hi[k] = std::max<double>(p[k], std::max<double>(p[3 + k], p[6 + k]));
    }
    const double e1[3] = {double(p[3]) - p[0], double(p[4]) - p[1], double(p[5]) - p[2]};
    const double e2[3] = {double(p[6]) - p[0], double(p[7]) - p[1], double(p[8]) - p[2]};
    const double cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2],
                 cz = e1[0] * e2[1] - e1[1] * e2[0];
    const double tri_area = 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz);
    excess[i] = std::max(0.0, box_area(lo, hi) - 2.0 * tri_area);
    total += excess[i];
  }
  const double budget = double(beta) * n;
  ref_box.reserve(size_t((n + budget) * 6 * 1.05));
  ref_prim.reserve(size_t((n + budget) * 1.05));
  for (uint32_t i = 0; i < n; ++i) {
    const float* p = pos + size_t(i) * 9;
    float tlo[3], thi[3];
    for (int k = 0; k < 3; ++k) {
      tlo[k] = std::min(p[k], std::min(p[3 + k], p[6 + k]));
      thi[k] = std::max(p[k], std::max(p[3 + k], p[6 + k]));
    }
    const int splits = total > 0 ? int(budget * excess[i] / total) : 0;
    if (splits == 0) {
      ref_box.insert(ref_box.end(), tlo, tlo + 3);
      ref_box.insert(ref_box.end(), thi, thi + 3);
      ref_prim.push_back(i);
      continue;
    }
    Poly poly;
    poly.n = 3;
    for (int v = 0; v < 3; ++v)
      for (int k = 0; k < 3; ++k) poly.v[v][k] = p[v * 3 + k];
    split_rec(poly, splits, tlo, thi, i, ref_box, ref_prim);
  }
}

BuildResult build_binned_sah(const float* prim_box_data, uint32_t n, int bins, int max_leaf, int n_threads,
                             double trav_cost) {
  BuildResult res;
  res.prim_order.resize(n);
  for (uint32_t i = 0; i < n; ++i) res.prim_order[i] = i;
  if (n == 0) {
    BuildNode nd{};
    nd.left = nd.right = -1;
    nd.first = 0;
    nd.count = 0;
    res.nodes.push_back(nd);
    res.n_leaves = 1;
    return res;
  }
  Ctx c;
  c.box = prim_box_data;
  c.cen.resize(size_t(n) * 3);
  for (uint32_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k)
      c.cen[size_t(i) * 3 + k] = 0.5f * (prim_box_data[size_t(i) * 6 + k] + prim_box_data[size_t(i) * 6 + 3 + k]);
  c.order = &res.prim_order;
  c.bins = std::max(2, bins);
  c.max_leaf = std::max(1, std::min(7, max_leaf));
  c.trav_cost = trav_cost;
  c.threads_left = std::max(0, n_threads - 1);
  SubTree st;
  st.nodes.reserve(size_t(n) * 2 / std::max(1, c.max_leaf / 2) + 16);
  build_rec(c, 0, n, st, 0);
  res.nodes = std::move(st.nodes);
  finish_stats(res);
  return res;
}

namespace {

// The float nearest to v in the outward direction (down for a lower bound, up for an
// upper one) whose low byte is `byte`; false if none is finite. Stepping outward moves
// the bit pattern by one per step (up for a positive value going up or a negative one
// going down, where the magnitude grows; down otherwise), so the answer is found
// directly unless the steps would cross zero, which the step-by-step walk handles.
bool carry_byte(float v, uint32_t byte, bool upper, float& out) {
  if (std::isfinite(v)) {
    const uint32_t u = f2u(v), sign = u & 0x80000000u, mag = u & 0x7fffffffu;
    const bool grows = (sign == 0u) == upper;  // magnitude grows with each step
    if (grows) {
      const uint32_t m = mag + ((byte - mag) & 0xffu);
      if (m < 0x7f800000u) {
        out = u2f(sign | m);
        return true;
      }
      return false;  // the walk reaches infinity first
    }
    const uint32_t k = (mag - byte) & 0xffu;
    if (k <= mag) {  // k = mag lands on a zero of v's sign, as the walk does
      out = u2f(sign | (mag - k));
      return true;
    }
  }
  float x = v;
  for (int k = 0; k <= 512; ++k) {
    if ((f2u(x) & 0xffu) == byte && std::isfinite(x)) {
      out = x;
      return true;
    }
    x = std::nextafter(x, upper ? HUGE_VALF : -HUGE_VALF);
    if (!std::isfinite(x)) return false;
  }
  return false;
}

}  // namespace

namespace {

// Smallest float not below v.
float float_up(double v) {
  float f = float(v);
  if (double(f) < v) f = std::nextafter(f, HUGE_VALF);
  return f;
}

// Centre and half extent of [lo, hi] as stored (see DevNode48); cb / hb: the low bytes the
// centre and the half extent carry, or -1 for none.
bool centre_half(float lo, float hi, int cb, int hb, float& c, float& h) {
  if (!std::isfinite(lo) || !std::isfinite(hi)) return false;
  c = float(0.5 * (double(lo) + double(hi)));
  if (cb >= 0 && !carry_byte(c, uint32_t(cb), false, c)) return false;
  const double half = std::max(double(hi) - double(c), double(c) - double(lo));
  h = float_up(float_up(half) * (1.0 + kHalfExtentMargin));
  if (!std::isfinite(h)) return false;
  if (hb >= 0 && !carry_byte(h, uint32_t(hb), true, h)) return false;
  return true;
}

}  // namespace

bool pack_nodes(const std::vector<DevNode2>& in, std::vector<DevNode48>& out) {
  static_assert(sizeof(DevNode48) == kNodeBytes, "node size");
  if (in.size() >= kMaxDeviceNodes) return false;
  out.resize(in.size());
  for (size_t i = 0; i < in.size(); ++i) {
    const DevNode2& n = in[i];
    DevNode48& d = out[i];
    const float* xy[2] = {n.r0, n.r1};  // lo.x, hi.x, lo.y, hi.y per child
    float* row[2] = {d.r0, d.r1};
    for (int c = 0; c < 2; ++c) {
      const uint32_t ref = uint32_t(device_child_ref(n.c[c]));
      for (int k = 0; k < 2; ++k) {
        const int cb = int((ref >> (16 * k)) & 0xffu), hb = int((ref >> (16 * k + 8)) & 0xffu);
        if (!centre_half(xy[c][2 * k], xy[c][2 * k + 1], cb, hb, row[c][2 * k], row[c][2 * k + 1])) return false;
      }
      if (!centre_half(n.r2[2 * c], n.r2[2 * c + 1], -1, -1, d.r2[2 * c], d.r2[2 * c + 1])) return false;
    }
  }
  return true;
}

std::vector<DevNode2> to_device_layout(const BuildResult& b) {
  std::vector<DevNode2> out;
  const BuildNode& root = b.nodes[0];
  auto is_leaf = [&](const BuildNode& nd) { return nd.count > 0 || nd.left < 0; };
  auto fill_child = [&](DevNode2& d, int slot, const BuildNode& ch, int32_t ref) {
    if (slot == 0) {
      d.r0[0] = ch.lo[0]; d.r0[1] = ch.hi[0]; d.r0[2] = ch.lo[1]; d.r0[3] = ch.hi[1];
      d.r2[0] = ch.lo[2]; d.r2[1] = ch.hi[2];
    } else {
      d.r1[0] = ch.lo[0]; d.r1[1] = ch.hi[0]; d.r1[2] = ch.lo[1]; d.r1[3] = ch.hi[1];
      d.r2[2] = ch.lo[2]; d.r2[3] = ch.hi[2];
    }
    d.c[slot] = ref;
  };
  if (is_leaf(root)) {
    DevNode2 d{};
    fill_child(d, 0, root, leaf_code(root.first, root.count));
    BuildNode empty{};
    empty.count = 0;
    fill_child(d, 1, empty, leaf_code(0, 0));
    out.push_back(d);
    return out;
  }
  // depth-first numbering of internal nodes, left child first
  std::vector<int> dev_index(b.nodes.size(), -1);
  std::vector<int> order;
  order.reserve(b.nodes.size() / 2 + 1);
  std::vector<int> stack{0};
  while (!stack.empty()) {
    const int i = stack.back();
    stack.pop_back();
    dev_index[size_t(i)] = int(order.size());
    order.push_back(i);
    const BuildNode& nd = b.nodes[size_t(i)];
    if (!is_leaf(b.nodes[size_t(nd.right)])) stack.push_back(nd.right);
    if (!is_leaf(b.nodes[size_t(nd.left)])) stack.push_back(nd.left);
  }
  out.resize(order.size());
  for (size_t k = 0; k < order.size(); ++k) {
    const BuildNode& nd = b.nodes[size_t(order[k])];
    DevNode2 d{};
    const int ch[2] = {nd.left, nd.right};
    for (int s = 0; s < 2; ++s) {
      const BuildNode& cn = b.nodes[size_t(ch[s])];
      const int32_t ref = is_leaf(cn) ? leaf_code(cn.first, cn.count) : dev_index[size_t(ch[s])];
      fill_child(d, s, cn, ref);
    }
    out[k] = d;
  }
  return out;
}

}  // namespace trt
