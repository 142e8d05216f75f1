// Host build of the KD-tree accelerator (layout in c2_kdtree.h).
#include "c2_kdtree.h"

#include <algorithm>
#include <cmath>
#include <limits>

#include "c2_host.h"

namespace c2 {
namespace {

constexpr int kMaxDepth = 20;  // KDTree.cpp:44

struct Edge {
    float pos;
    bool end;
    // equal positions: ends before starts (KDTree.cpp:31-35)
    bool operator<(const Edge& e) const { return pos == e.pos ? end > e.end : pos < e.pos; }
};

// BoundingBox::SurfaceArea (src/Core/Geometry.h:101-109), the same sum order
float surface_area(const Box3& b) {
    float area = 0;
    for (int i = 0; i < 3; i++) area += 2 * (b.hi[i] - b.lo[i]) * (b.hi[(i + 1) % 3] - b.lo[(i + 1) % 3]);
    return area;
}

struct Builder {
    const std::vector<Box3>& boxes;
    KdTree& t;

    int32_t leaf(const std::vector<int32_t>& objs) {
        KdNode n{-1, 0.f, (int32_t)t.objs.size(), (int32_t)objs.size()};
        t.objs.insert(t.objs.end(), objs.begin(), objs.end());
        t.nodes.push_back(n);
        return (int32_t)t.nodes.size() - 1;
    }

    int32_t build(const Box3& outer, const std::vector<int32_t>& objs, int depth) {
        t.depth = std::max(t.depth, depth);
        if (objs.size() <= 2 || depth == kMaxDepth) return leaf(objs);
        int split = 0;
        float splitPos = 0;
        float best = std::numeric_limits<float>::infinity();
        const float totalSA = surface_area(outer);
        float diff[3];
        for (int c = 0; c < 3; c++) diff[c] = outer.hi[c] - outer.lo[c];
        std::vector<Edge> edges;
        for (int d = 0; d < 3; d++) {
            edges.clear();
            for (int32_t e : objs) {
                edges.push_back(Edge{boxes[e].lo[d], false});
                edges.push_back(Edge{boxes[e].hi[d], true});
            }
            std::sort(edges.begin(), edges.end());
            int nBelow = 0, nAbove = (int)objs.size();
            const float L = outer.lo[d], R = outer.hi[d];
            for (const Edge& e : edges) {
                if (e.end) --nAbove;
                if (e.pos > L && e.pos < R) {
                    float LA = diff[(d + 1) % 3], LB = diff[(d + 2) % 3];
                    float belowSA = 2 * (LA * LB + (e.pos - L) * (LA + LB));
                    float aboveSA = 2 * (LA * LB + (R - e.pos) * (LA + LB));
                    float eb = (nAbove == 0 || nBelow == 0) ? 0.2 : 0;
                    float pb = belowSA / totalSA, pa = aboveSA / totalSA;
                    // 0.05 is a double: the sum is formed in double and rounded once
                    float cost = 0.05 + (1 - eb) * (pb * nBelow + pa * nAbove);
                    if (cost < best) {
                        best = cost;
                        split = d;
                        splitPos = e.pos;
                    }
                }
                if (!e.end) ++nBelow;
            }
        }
        if (best >= (float)objs.size()) return leaf(objs);
        Box3 below = outer, above = outer;
        below.hi[split] = splitPos;
        above.lo[split] = splitPos;
        std::vector<int32_t> lo, hi;
        for (int32_t e : objs) {
            if (boxes[e].lo[split] <= splitPos) lo.push_back(e);
            if (boxes[e].hi[split] > splitPos) hi.push_back(e);
        }
        t.nodes.push_back(KdNode{split, splitPos, 0, 0});
        int32_t self = (int32_t)t.nodes.size() - 1;
        int32_t a = build(below, lo, depth + 1);
        int32_t b = build(above, hi, depth + 1);
        t.nodes[self].a = a;
        t.nodes[self].b = b;
        return self;
    }
};

}  // namespace

void build_kdtree(const std::vector<Box3>& boxes, KdTree& out) {
    out = KdTree{};
    out.nodes.push_back(KdNode{-1, 0.f, 0, 0});  // node 0: "none"
    for (int c = 0; c < 3; c++) {
        out.master.lo[c] = std::numeric_limits<float>::infinity();
        out.master.hi[c] = -std::numeric_limits<float>::infinity();
    }
    if (boxes.empty()) return;
    // Bound3f::Union with glm::min / glm::max, KDTree.cpp:62-68
    for (const Box3& b : boxes)
        for (int c = 0; c < 3; c++) {
            out.master.lo[c] = (b.lo[c] < out.master.lo[c]) ? b.lo[c] : out.master.lo[c];
            out.master.hi[c] = (out.master.hi[c] < b.hi[c]) ? b.hi[c] : out.master.hi[c];
        }
    std::vector<int32_t> all(boxes.size());
    for (size_t i = 0; i < boxes.size(); i++) all[i] = (int32_t)i;
    Builder bld{boxes, out};
    out.root = bld.build(out.master, all, 0);
}

}  // namespace c2
