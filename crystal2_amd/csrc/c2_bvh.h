// BVH layouts shared by the host builder (c2_bvh.cpp) and the gfx950 kernels.
//
// Build: binned SAH over entity bounding boxes into a binary tree with at most
// kLeafSize entities per leaf, then collapsed into a 4-wide tree whose nodes are
// exactly one 128-byte cache line (or, for C2_ACCEL_BVH8, an 8-wide tree, Node8 below):
//
//   Node4 (128 B, 128-B aligned)
//     lo_x[4] hi_x[4] lo_y[4] hi_y[4] lo_z[4] hi_z[4]   fp32 child boxes (96 B)
//     child[4]   >= 0 : index of an inner Node4
//                <  0 : leaf, ~child = first primitive slot | (count - 1) << kSlotBits
//     count      4 x 8-bit primitive counts of leaf children (0 for inner / empty); the
//                kernels read the count from the child word and never load this one
//   Empty child slots carry an inverted box (+inf / -inf) and never pass the slab test.
//
// Child boxes are the exact fp32 unions (min/max) of the entity boxes below them,
// so a node test computed with the reference's slab arithmetic
// ((b - o) * invD, then max/min, src/Core/Geometry.h:136-157) passes whenever
// the test of any entity box below it would pass: node culling never removes a
// primitive the reference's per-primitive pre-clip would keep.
//
// Primitive slots (leaf order), 48 B = 3 x float4 each:
//   triangle: (v0.x v0.y v0.z v1.x) (v1.y v1.z v2.x v2.y) (v2.z entity kind=0 -)
//   sphere:   (c.x  c.y  c.z  r   ) (sphere_index - - -)  (-    entity kind=1 -)
// Vertices are stored unmodified so e1 = v1 - v0, e2 = v2 - v0 and the entity box
// (component-wise min/max of v0, v1, v2) are recomputed bit-exactly in the kernel.
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

namespace c2 {

constexpr int kLeafSize = 4;        // largest leaf (2-bit count field); the ray-cast accelerator's leaves
constexpr int32_t kSlotBits = 28;   // leaf child word: slot in bits 0-27, count - 1 in bits 28-29
constexpr int kLeafSizeRender = 2;  // leaves of the renderer's trees (fewer primitive tests per ray)
constexpr int kStackSize = 64;      // per-lane traversal stack (entries)
constexpr int kMaxBvh4Depth = 21;   // 3 * depth + 1 <= kStackSize
constexpr int kMaxBvh2Depth = 42;

struct alignas(128) Node4 {
    float lo_x[4], hi_x[4], lo_y[4], hi_y[4], lo_z[4], hi_z[4];
    int32_t child[4];
    uint32_t count;  // 8 bits per child
    uint32_t pad[3];
};
static_assert(sizeof(Node4) == 128, "Node4 must be one 128-byte line");
// The kernels read a node as 8 float4: planes lo_x hi_x lo_y hi_y lo_z hi_z, then
// child[4], then count (the near plane of axis a is float4 2a + (direction < 0)).
static_assert(offsetof(Node4, hi_x) == 16 && offsetof(Node4, lo_y) == 32 && offsetof(Node4, hi_z) == 80 &&
                  offsetof(Node4, child) == 96 && offsetof(Node4, count) == 112,
              "Node4 plane layout");

// 8-wide variant (the accelerator C2_ACCEL_BVH8): the same binary tree collapsed into
// 256-byte nodes, two 128-byte lines; the kernels read a node as 14 float4: plane p of the
// children 0..3 at byte 32 p and of the children 4..7 at 32 p + 16 (planes lo_x hi_x lo_y hi_y
// lo_z hi_z, so the near plane of axis a is at 64 a + 32 (direction < 0)), then child[8].
// The stack holds at most 7 pushes per level, so an 8-wide tree is at most kMaxBvh8Depth deep.
constexpr int kStackSize8 = 96;
constexpr int kMaxBvh8Depth = 13;  // 7 * depth + 1 <= kStackSize8
struct alignas(128) Node8 {
    float lo_x[8], hi_x[8], lo_y[8], hi_y[8], lo_z[8], hi_z[8];
    int32_t child[8];
    uint32_t pad[8];
};
static_assert(sizeof(Node8) == 256, "Node8 must be two 128-byte lines");
static_assert(offsetof(Node8, hi_x) == 32 && offsetof(Node8, lo_y) == 64 && offsetof(Node8, hi_z) == 160 &&
                  offsetof(Node8, child) == 192,
              "Node8 plane layout");

struct Box3 {
    float lo[3], hi[3];
};

struct Bvh2Node {
    Box3 box;
    int32_t left, right;  // inner: children; leaf: left = first slot, right = -count
    bool leaf() const { return right < 0; }
};

struct BuiltBvh {
    std::vector<Bvh2Node> bvh2;        // root at 0
    std::vector<int32_t> order;        // slot -> entity
    std::vector<Node4> nodes;          // root at 0
    std::vector<Node8> nodes8;         // root at 0 (built by collapse_bvh8)
    int32_t depth2 = 0, depth4 = 0, depth8 = 0;
};

// Builds both trees from per-entity boxes. Throws c2::Error on depth overflow.
void build_bvh(const std::vector<Box3>& boxes, BuiltBvh& out, int leaf_size = kLeafSize);
// Collapses out.bvh2 into out.nodes8 (the same leaves and primitive slots as out.nodes).
// Throws c2::Error (C2_ERR_UNSUPPORTED) when the 8-wide tree exceeds kMaxBvh8Depth.
void collapse_bvh8(BuiltBvh& out);

}  // namespace c2
