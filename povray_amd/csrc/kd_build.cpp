// Host kd-tree builder with the cost model of the reference's BSP tree
// (BSPTree::BuildRecursive, core/bounding/bsptree.cpp:452-705, after Eric
// Haines, Ray Tracing News 17.1): a node with c objects costs TK1 + c TP
// unsplit; a split plane at relative position r costs
//   TK1 + (1 + cph) TK3 + TP (pab + cph/2 ((1+miss) pa + (1+miss) pb)
//                              + r ch pa + (1 - r) ch pb)
// with TK1 = 1, TK3 = 5, TP = 150, miss = 0.2 (bsptree.cpp:81-85), object
// extents widened by BSP_TOLERANCE = 1e-5, and every object-boundary plane
// inside the cell tried on each axis.  Objects straddling the chosen plane
// go to both children.  Own implementation: event lists per node, a stable
// sort (plane, then event order), deterministic output.
#include "kd_build.h"

#include <algorithm>
#include <cmath>

namespace pvx {

namespace {

constexpr float kTK1 = 1.0f, kTK3 = 5.0f, kTP = 150.0f, kMiss1 = 1.2f, kTol = 0.00001f;

struct Event {
    float plane;
    uint32_t obj;
    uint32_t max; // 0: object starts (min side), 1: object ends (max side)
};

struct Builder {
    const float *lo, *hi;
    KdTree &t;
    std::vector<Event> ev[3];

    void leaf(uint32_t inode, const std::vector<uint32_t> &idx) {
        KdNode &nd = t.nodes[inode];
        nd.info = 0x80000000u | (uint32_t)idx.size();
        nd.u.first = (uint32_t)t.list.size();
        t.list.insert(t.list.end(), idx.begin(), idx.end());
    }

    void rec(uint32_t inode, std::vector<uint32_t> idx, float cmin[3], float cmax[3], uint32_t level_left,
             uint32_t depth) {
        t.depth = std::max(t.depth, depth);
        const uint32_t cnt = (uint32_t)idx.size();
        if (cnt <= 1 || level_left == 0) {
            leaf(inode, idx);
            return;
        }
        float best = kTK1 + cnt * kTP;
        int best_axis = -1;
        uint32_t best_split = 0, best_cnt = 0;
        float size[5];
        size[0] = size[3] = cmax[0] - cmin[0];
        size[1] = size[4] = cmax[1] - cmin[1];
        size[2] = cmax[2] - cmin[2];
        const float enh = size[0] * size[1] + size[0] * size[2] + size[1] * size[2];
        const float enhinv = 1.0f / enh;
        for (int axis = 0; axis < 3; axis++) {
            uint32_t pa = 0, pb = cnt, pab = 0;
            const float bmin = cmin[axis], bmax = cmax[axis];
            float cph = size[axis + 1] * size[axis + 2] * enhinv;
            const float relmul = 1.0f / size[axis];
            const float chmul = size[axis] * (size[axis + 1] + size[axis + 2]) * enhinv;
            const float constcost = kTK1 + (1.0f + cph) * kTK3;
            cph *= 0.5f;
            std::vector<Event> &e = ev[axis];
            e.clear();
            for (uint32_t i : idx) {
                const float smin = lo[3 * (size_t)i + axis] - kTol, smax = hi[3 * (size_t)i + axis] + kTol;
                if (smin >= bmax || smax <= bmin) continue;
                if (smin < bmin) {
                    pab++;
                    pb--;
                }
                e.push_back(Event{smin, i, 0});
                e.push_back(Event{smax, i, 1});
            }
            std::stable_sort(e.begin(), e.end(), [](const Event &a, const Event &b) { return a.plane < b.plane; });
            for (uint32_t k = 0; k < (uint32_t)e.size(); k++) {
                const float plane = e[k].plane;
                if (e[k].max) {
                    pa++;
                    pab--;
                }
                if (plane > bmin && plane < bmax) {
                    const float r = (plane - bmin) * relmul;
                    const float cah = r * chmul, cbh = (1.0f - r) * chmul;
                    const float cost =
                        constcost + kTP * (pab + cph * (kMiss1 * pa + kMiss1 * pb) + cah * pa + cbh * pb);
                    if (cost < best) {
                        best = cost;
                        best_axis = axis;
                        best_split = k;
                        best_cnt = (uint32_t)e.size();
                    }
                }
                if (!e[k].max) {
                    pab++;
                    pb--;
                }
            }
        }
        if (best_axis < 0 || level_left == 0) {
            leaf(inode, idx);
            return;
        }
        std::vector<Event> &e = ev[best_axis]; // still sorted from the sweep
        (void)best_cnt;
        const float plane = e[best_split].plane;
        uint32_t split = best_split + (e[best_split].max ? 1u : 0u); // a closing plane goes left
        std::vector<uint32_t> left, right;
        std::vector<uint8_t> seen_l, seen_r;
        for (uint32_t k = 0; k < split; k++) left.push_back(e[k].obj);
        for (uint32_t k = split; k < (uint32_t)e.size(); k++) right.push_back(e[k].obj);
        auto uniq = [](std::vector<uint32_t> &v) {
            std::sort(v.begin(), v.end());
            v.erase(std::unique(v.begin(), v.end()), v.end());
        };
        uniq(left);
        uniq(right);
        const uint32_t child = (uint32_t)t.nodes.size();
        t.nodes.push_back(KdNode{});
        t.nodes.push_back(KdNode{});
        KdNode &nd = t.nodes[inode];
        nd.info = ((uint32_t)best_axis << 29) | child;
        nd.u.plane = plane;
        std::vector<Event>().swap(e); // free before recursing
        float keep = cmax[best_axis];
        cmax[best_axis] = plane;
        rec(child, std::move(left), cmin, cmax, level_left - 1, depth + 1);
        cmax[best_axis] = keep;
        keep = cmin[best_axis];
        cmin[best_axis] = plane;
        rec(child + 1, std::move(right), cmin, cmax, level_left - 1, depth + 1);
        cmin[best_axis] = keep;
    }
};

} // namespace

KdTree build_kd(const float *lo, const float *hi, uint32_t n, uint32_t max_depth) {
    KdTree t;
    t.nodes.push_back(KdNode{});
    float cmin[3] = {INFINITY, INFINITY, INFINITY}, cmax[3] = {-INFINITY, -INFINITY, -INFINITY};
    std::vector<uint32_t> idx(n);
    for (uint32_t i = 0; i < n; i++) {
        idx[i] = i;
        for (int a = 0; a < 3; a++) {
            cmin[a] = std::min(cmin[a], lo[3 * (size_t)i + a]);
            cmax[a] = std::max(cmax[a], hi[3 * (size_t)i + a]);
        }
    }
    if (n == 0) {
        t.nodes[0].info = 0x80000000u;
        t.nodes[0].u.first = 0;
        return t;
    }
    for (int a = 0; a < 3; a++) {
        t.bmin[a] = cmin[a];
        t.bmax[a] = cmax[a];
    }
    Builder b{lo, hi, t, {}};
    b.rec(0, std::move(idx), cmin, cmax, std::min<uint32_t>(max_depth, kKdMaxDepth), 0);
    return t;
}

} // namespace pvx
