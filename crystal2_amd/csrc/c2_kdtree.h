// KD-tree accelerator ("Acceleration": "KDTree"): the reference's tree over entity boxes,
// flattened for the kernels.
//
// KdNode (16 B, int4): axis (-1 = leaf), split position (float bits), children.
//   inner: (axis, splitPos, child below, child above)   child 0 = none
//   leaf:  (-1, -, first entry in objs, entry count)
// Node 0 is unused, so that 0 means "no child" as in the reference's node pool.
// objs holds the entities of each leaf in entity order; an entity whose box straddles a
// split plane is listed under both children.
#pragma once
#include <cstdint>
#include <vector>

#include "c2_bvh.h"

namespace c2 {

struct KdNode {
    int32_t axis;
    float split;
    int32_t a, b;
};
static_assert(sizeof(KdNode) == 16, "KdNode is one int4");

struct KdTree {
    std::vector<KdNode> nodes;   // [0] unused, root at `root`
    std::vector<int32_t> objs;   // leaf entity lists
    int32_t root = 0;            // 0: no entities
    Box3 master{};               // union of every entity box
    int32_t depth = 0;
};

// KDTree::Build / _build (src/Accelerators/KDTree.cpp:59-193): SAH over box edges (cost
// 0.05 + (1 - eb)(pb nBelow + pa nAbove), empty-side bonus eb = 0.2), leaves of <= 2
// entities, depth <= 20, an entity goes below when min <= split and above when max > split.
void build_kdtree(const std::vector<Box3>& boxes, KdTree& out);

}  // namespace c2
