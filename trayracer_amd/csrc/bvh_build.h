// Binned-SAH BVH builder (host) and the device node layout.
//
// Replaces the OptiX acceleration-structure builds of the reference (BLAS per model,
// Resources/Model.cpp:240-306; TLAS over instances, Optix/OptixRenderer.cpp:90-123),
// which are vendor code. The builder is the textbook binned SAH (Wald 2007, "On fast
// construction of SAH-based bounding volume hierarchies"): 32 centroid bins per axis,
// all three axes evaluated, leaves of at most `max_leaf` primitives, subtrees built
// in parallel on host threads.
#pragma once
#include <stdint.h>

#include <vector>

#include "trt/common.h"

namespace trt {

struct BuildNode {
  float lo[3], hi[3];
  int32_t left, right;  // children (internal nodes)
  uint32_t first, count;  // leaf range in prim_order (count > 0 => leaf)
};

struct BuildResult {
  std::vector<BuildNode> nodes;  // nodes[0] = root
  std::vector<uint32_t> prim_order;
  int depth = 0;          // max number of internal nodes on a root-to-leaf path
  double sah_cost = 0.0;  // SAH cost of the tree (traversal 1, intersection 1)
  uint32_t n_leaves = 0;
};

// prim_box: 6 floats per primitive (lo xyz, hi xyz). A node of at most max_leaf
// primitives becomes a leaf when its cost (one per primitive) is no more than the
// split's, trav_cost + the children's area-weighted primitive counts.
//   kTravCostTriangles: triangle trees, whose leaf triangles the traversal tests
//   spread over a whole wave (1.5 measured 2% faster than 1 on the 1M soup: leaves
//   fill up, 43 instead of 46 node visits per ray for 9 instead of 6 triangle tests);
//   kTravCostInstances: the top level over instances, where a leaf primitive costs a
//   whole bottom-level traversal (1.5 there cost 8%).
constexpr double kTravCostTriangles = 1.5;
constexpr double kTravCostInstances = 1.0;
BuildResult build_binned_sah(const float* prim_box, uint32_t n, int bins, int max_leaf, int n_threads,
                             double trav_cost);

// Spatial pre-splitting of triangles (after Karras & Aila 2013, "Fast parallel
// construction of high-quality bounding volume hierarchies", section 5): a budget of
// beta * n extra references is distributed over the triangles in proportion to how much
// their bounding box exceeds their own area; each selected triangle is cut recursively
// at the spatial median of its piece's longest axis (Sutherland-Hodgman clipping in
// double precision), giving several tighter references to the same triangle. Boxes are
// rounded outward to float so their union always covers the triangle.
// Outputs 6 floats per reference and the triangle index of each reference.
void presplit_triangles(const float* tri_pos, uint32_t n, float beta, std::vector<float>& ref_box,
                        std::vector<uint32_t>& ref_prim);

// Device layout: one node stores both children's boxes (64 bytes, four float4 rows):
//   row0 = (c0.lo.x, c0.hi.x, c0.lo.y, c0.hi.y)
//   row1 = (c1.lo.x, c1.hi.x, c1.lo.y, c1.hi.y)
//   row2 = (c0.lo.z, c0.hi.z, c1.lo.z, c1.hi.z)
//   row3 = (child0, child1, 0, 0) as int32
// child >= 0: index of an internal node; child < 0: leaf ~((first << 3) | count),
// first indexing the reordered triangle array, count in [0, 7].
struct alignas(16) DevNode2 {
  float r0[4], r1[4], r2[4];
  int32_t c[4];
};

inline int32_t leaf_code(uint32_t first, uint32_t count) { return ~static_cast<int32_t>((first << 3) | count); }

// Traversal layout on the device: 48 bytes, each child box as a centre and a half
// extent per axis,
//   row0 = (c0.cx, c0.hx, c0.cy, c0.hy), row1 = (c1.cx, c1.hx, c1.cy, c1.hy),
//   row2 = (c0.cz, c0.hz, c1.cz, c1.hz),
// so that a slab's near and far distances are (x - s h) inv and (x + s h) inv with
// x = c - o and s = sign(inv): floating-point adds and multiplies only, with no per-axis
// min / max (which issue at half the rate of FP32 arithmetic on gfx950). The two child
// references ride in the low bytes of the x and y values (child0 in row0's four floats,
// child1 in row1's, least significant byte first), so a node is three 16-byte loads. An
// internal child is referenced by its byte offset in the node array (index x 48, so the
// traversal forms the address with one add; offsets stay below 2^31), a leaf by its
// leaf code. A centre that carries a byte is moved down to the nearest float with that
// low byte (at most 255 ulp); the half extent then covers the exact box from the stored
// centre on both sides, is enlarged by 2^-10 of itself (which covers the rounding of the
// slab distances beyond the relative error the traversal's far-distance slack covers)
// and, if it carries a byte, moved up to the nearest float with that low byte. Every
// stored box therefore contains the exact one and the slab test stays conservative.
struct alignas(16) DevNode48 {
  float r0[4], r1[4], r2[4];
};

// Child reference carried by the low bytes of one row (cx, hx, cy, hy).
TRT_FI int32_t node_child_ref(float a, float b, float c, float d) {
  return static_cast<int32_t>((f2u(a) & 0xffu) | ((f2u(b) & 0xffu) << 8) | ((f2u(c) & 0xffu) << 16) |
                              (f2u(d) << 24));
}

// Device reference of a DevNode2 child: the byte offset of an internal node, a leaf code
// unchanged.
constexpr uint32_t kNodeBytes = 48;
constexpr uint64_t kMaxDeviceNodes = (uint64_t(1) << 31) / kNodeBytes;  // offsets below 2^31
inline int32_t device_child_ref(int32_t c) { return c >= 0 ? c * static_cast<int32_t>(kNodeBytes) : c; }

// Relative enlargement of every stored half extent.
constexpr double kHalfExtentMargin = 1.0 / 1024.0;

// Converts nodes to the 48-byte layout. False for a non-finite bound, one that cannot
// carry its byte, or kMaxDeviceNodes nodes or more.
bool pack_nodes(const std::vector<DevNode2>& in, std::vector<DevNode48>& out);

// Converts a build tree into the device layout (depth-first order, node 0 = root).
// Returns the device nodes; a root leaf is wrapped in a node with an empty second child.
std::vector<DevNode2> to_device_layout(const BuildResult& b);

}  // namespace trt
