// Host BVH builder: binned SAH binary BVH (leaf <= 4) collapsed into 4-wide
// 128-byte nodes or 8-wide 256-byte nodes (layouts documented in c2_bvh.h).
//
// The reference builds a median-split binary BVH with leaf <= 2
// (src/Accelerators/BVH.cpp:156-188, its SAH path at :281-325 is compiled out).
// The tree shape decides only which of two exactly equal-t hits is found first;
// the kernels break such ties by lowest entity index, so any valid tree gives the
// same closest hit. This builder is written for traversal cost on gfx950 instead.
#include "c2_bvh.h"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <stdexcept>
#include <thread>

#include "c2_host.h"

namespace c2 {

namespace {

inline float area(const Box3& b) {
    float dx = b.hi[0] - b.lo[0], dy = b.hi[1] - b.lo[1], dz = b.hi[2] - b.lo[2];
    if (!(dx >= 0) || !(dy >= 0) || !(dz >= 0)) return 0.f;
    return 2.f * (dx * dy + dy * dz + dz * dx);
}
inline void grow(Box3& a, const Box3& b) {
    for (int k = 0; k < 3; k++) {
        a.lo[k] = std::min(a.lo[k], b.lo[k]);
        a.hi[k] = std::max(a.hi[k], b.hi[k]);
    }
}
inline Box3 empty_box() {
    Box3 b;
    for (int k = 0; k < 3; k++) {
        b.lo[k] = INFINITY;
        b.hi[k] = -INFINITY;
    }
    return b;
}

constexpr int kBins = 32;

struct Builder {
    const std::vector<Box3>& boxes;
    std::vector<float> cen;  // [n][3]
    std::vector<int32_t>& order;
    std::vector<Bvh2Node>& nodes;
    int leafSize;
    int maxDepth = 0;

    Builder(const std::vector<Box3>& b, std::vector<int32_t>& o, std::vector<Bvh2Node>& n, int leaf)
        : boxes(b), order(o), nodes(n), leafSize(leaf) {}

    int ceil_log2(int64_t x) {
        int r = 0;
        while ((int64_t(1) << r) < x) r++;
        return r;
    }

    // Builds node `idx` over order[begin, end).
    void build(int idx, int begin, int end, int depth) {
        maxDepth = std::max(maxDepth, depth);
        Box3 bb = empty_box(), cb = empty_box();
        for (int i = begin; i < end; i++) {
            int e = order[i];
            grow(bb, boxes[e]);
            Box3 c;
            for (int k = 0; k < 3; k++) c.lo[k] = c.hi[k] = cen[(size_t)e * 3 + k];
            grow(cb, c);
        }
        nodes[idx].box = bb;
        int count = end - begin;
        if (count <= leafSize) {
            nodes[idx].left = begin;
            nodes[idx].right = -count;
            return;
        }
        int mid = -1, axis = 0;
        bool forceMedian = depth + 1 + ceil_log2((count + leafSize - 1) / leafSize) >= kMaxBvh2Depth;
        float ext[3];
        for (int k = 0; k < 3; k++) ext[k] = cb.hi[k] - cb.lo[k];
        axis = (ext[0] >= ext[1] && ext[0] >= ext[2]) ? 0 : (ext[1] >= ext[2] ? 1 : 2);
        if (!forceMedian && ext[axis] > 0) {
            float bestCost = INFINITY;
            int bestAxis = -1, bestBin = -1;
            for (int k = 0; k < 3; k++) {
                if (!(ext[k] > 0)) continue;
                Box3 binBox[kBins];
                int binCnt[kBins] = {0};
                for (int b = 0; b < kBins; b++) binBox[b] = empty_box();
                float scale = kBins / ext[k];
                for (int i = begin; i < end; i++) {
                    int e = order[i];
                    int b = (int)((cen[(size_t)e * 3 + k] - cb.lo[k]) * scale);
                    b = std::min(std::max(b, 0), kBins - 1);
                    binCnt[b]++;
                    grow(binBox[b], boxes[e]);
                }
                float rightArea[kBins];
                int rightCnt[kBins];
                Box3 acc = empty_box();
                int c = 0;
                for (int b = kBins - 1; b > 0; b--) {
                    grow(acc, binBox[b]);
                    c += binCnt[b];
                    rightArea[b] = area(acc);
                    rightCnt[b] = c;
                }
                acc = empty_box();
                c = 0;
                for (int b = 0; b < kBins - 1; b++) {
                    grow(acc, binBox[b]);
                    c += binCnt[b];
                    if (c == 0 || rightCnt[b + 1] == 0) continue;
                    float cost = area(acc) * c + rightArea[b + 1] * rightCnt[b + 1];
                    if (cost < bestCost) {
                        bestCost = cost;
                        bestAxis = k;
                        bestBin = b;
                    }
                }
            }
            if (bestAxis >= 0) {
                float scale = kBins / ext[bestAxis];
                float lo = cb.lo[bestAxis];
                auto it = std::partition(order.begin() + begin, order.begin() + end, [&](int e) {
                    int b = (int)((cen[(size_t)e * 3 + bestAxis] - lo) * scale);
                    b = std::min(std::max(b, 0), kBins - 1);
                    return b <= bestBin;
                });
                mid = (int)(it - order.begin());
                if (mid == begin || mid == end) mid = -1;
                axis = bestAxis;
            }
        }
        if (mid < 0) {  // object median on the widest centroid axis (deterministic tie order)
            mid = begin + count / 2;
            std::nth_element(order.begin() + begin, order.begin() + mid, order.begin() + end, [&](int a, int b) {
                float ca = cen[(size_t)a * 3 + axis], cbv = cen[(size_t)b * 3 + axis];
                return ca < cbv || (ca == cbv && a < b);
            });
        }
        int l = (int)nodes.size();
        nodes.push_back(Bvh2Node{});
        nodes.push_back(Bvh2Node{});
        nodes[idx].left = l;
        nodes[idx].right = l + 1;
        build(l, begin, mid, depth + 1);
        build(l + 1, mid, end, depth + 1);
    }
};

// Collapse of the binary tree into W-wide nodes (Node4 or Node8). Open the shallowest
// inner child first (ties: largest surface area), so every W-wide level spans log2(W)
// binary levels wherever the tree allows.
template <int W, class Node>
struct Collapser {
    const std::vector<Bvh2Node>& b2;
    std::vector<Node>& out;
    int maxDepth = 0;

    int collapse(int i, int depth) {
        maxDepth = std::max(maxDepth, depth);
        std::vector<int> kids, kdepth;
        if (b2[i].leaf()) {
            kids.push_back(i);
            kdepth.push_back(0);
        } else {
            kids.push_back(b2[i].left);
            kids.push_back(b2[i].right);
            kdepth.push_back(1);
            kdepth.push_back(1);
        }
        while ((int)kids.size() < W) {
            int best = -1;
            float bestA = -1;
            for (size_t k = 0; k < kids.size(); k++) {
                if (b2[kids[k]].leaf()) continue;
                float a = area(b2[kids[k]].box);
                if (best < 0 || kdepth[k] < kdepth[best] || (kdepth[k] == kdepth[best] && a > bestA)) {
                    bestA = a;
                    best = (int)k;
                }
            }
            if (best < 0) break;
            int c = kids[best], d = kdepth[best];
            kids[best] = b2[c].left;
            kdepth[best] = d + 1;
            kids.insert(kids.begin() + best + 1, b2[c].right);
            kdepth.insert(kdepth.begin() + best + 1, d + 1);
        }
        int me = (int)out.size();
        out.push_back(Node{});
        Node n{};
        for (int s = 0; s < W; s++) {
            n.lo_x[s] = n.lo_y[s] = n.lo_z[s] = INFINITY;
            n.hi_x[s] = n.hi_y[s] = n.hi_z[s] = -INFINITY;
            n.child[s] = -1;
        }
        uint32_t count = 0;
        for (size_t s = 0; s < kids.size(); s++) {
            const Bvh2Node& c = b2[kids[s]];
            n.lo_x[s] = c.box.lo[0]; n.lo_y[s] = c.box.lo[1]; n.lo_z[s] = c.box.lo[2];
            n.hi_x[s] = c.box.hi[0]; n.hi_y[s] = c.box.hi[1]; n.hi_z[s] = c.box.hi[2];
            if (c.leaf()) {
                n.child[s] = ~(int32_t)((uint32_t)c.left | ((uint32_t)(-c.right - 1) << kSlotBits));
                if (s < 4) count |= (uint32_t)(-c.right) << (8 * s);
            } else {
                n.child[s] = -1;  // patched below
            }
        }
        if constexpr (W == 4) n.count = count;
        for (size_t s = 0; s < kids.size(); s++) {
            const Bvh2Node& c = b2[kids[s]];
            if (!c.leaf()) n.child[s] = collapse(kids[s], depth + 1);
        }
        out[me] = n;
        return me;
    }
};

}  // namespace

void build_bvh(const std::vector<Box3>& boxes, BuiltBvh& out, int leaf_size) {
    if (leaf_size < 1 || leaf_size > kLeafSize) throw std::invalid_argument("BVH leaf size out of range");
    size_t n = boxes.size();
    out.order.resize(n);
    std::iota(out.order.begin(), out.order.end(), 0);
    out.bvh2.clear();
    out.nodes.clear();
    if (n == 0) return;
    Builder b(boxes, out.order, out.bvh2, leaf_size);
    b.cen.resize(n * 3);
    for (size_t e = 0; e < n; e++)
        for (int k = 0; k < 3; k++) b.cen[e * 3 + k] = 0.5f * boxes[e].lo[k] + 0.5f * boxes[e].hi[k];
    out.bvh2.reserve(2 * n / leaf_size + 4);
    out.bvh2.push_back(Bvh2Node{});
    b.build(0, 0, (int)n, 0);
    out.depth2 = b.maxDepth;
    Collapser<4, Node4> c{out.bvh2, out.nodes};
    out.nodes.reserve(out.bvh2.size() / 2 + 2);
    c.collapse(0, 0);
    out.depth4 = c.maxDepth;
    if (out.depth4 > kMaxBvh4Depth)
        fail(C2_ERR_UNSUPPORTED, "BVH4 depth " + std::to_string(out.depth4) + " exceeds the traversal stack bound");
}

void collapse_bvh8(BuiltBvh& out) {
    out.nodes8.clear();
    out.depth8 = 0;
    if (out.bvh2.empty()) return;
    Collapser<8, Node8> c{out.bvh2, out.nodes8};
    out.nodes8.reserve(out.bvh2.size() / 6 + 2);
    c.collapse(0, 0);
    out.depth8 = c.maxDepth;
    if (out.depth8 > kMaxBvh8Depth)
        fail(C2_ERR_UNSUPPORTED, "BVH8 depth " + std::to_string(out.depth8) + " exceeds the traversal stack bound");
}

}  // namespace c2
