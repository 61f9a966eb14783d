// Host BVH2 builder: binned SAH over conservative FP32 primitive boxes.
//
// Replaces the reference's bounding-slab build (Build_BBox_Tree /
// sort_and_split, core/bounding/boundingbox.cpp:262-430, and the per-mesh
// tree of Build_Mesh_BBox_Tree, core/shape/mesh.cpp:1376-1409) with one flat
// binary tree over every bounded primitive (mesh triangles expanded), laid
// out as 64 B nodes for coalesced fetches.  Top levels are split serially;
// subtrees below kParallelGrain primitives are built on worker threads and
// concatenated in a fixed order, so the layout is deterministic.
#include "bvh_build.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <future>
#include <thread>

namespace pvx {

namespace {

constexpr uint32_t kParallelGrain = 1u << 15;
constexpr int kSahDepthLimit = 40; // below this depth: median splits (bounds tree depth)
constexpr int kMaxBins = 64;

struct Box3 {
    float lo[3], hi[3];
    void empty() {
        for (int k = 0; k < 3; k++) {
            lo[k] = INFINITY;
            hi[k] = -INFINITY;
        }
    }
    void grow(const Box3 &b) {
        for (int k = 0; k < 3; k++) {
            lo[k] = std::min(lo[k], b.lo[k]);
            hi[k] = std::max(hi[k], b.hi[k]);
        }
    }
    float half_area() const {
        float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        if (dx < 0 || dy < 0 || dz < 0) return 0.0f;
        return dx * dy + dy * dz + dz * dx;
    }
};

class Builder {
  public:
    Builder(const float *lo, const float *hi, uint32_t n, int max_leaf, int bins)
        : lo_(lo), hi_(hi), n_(n), max_leaf_(std::max(1, std::min(8, max_leaf))),
          bins_(std::max(4, std::min(kMaxBins, bins))) {
        idx_.resize(n);
        cen_.resize(3 * (size_t)n);
        for (uint32_t i = 0; i < n; i++) {
            idx_[i] = i;
            for (int k = 0; k < 3; k++)
                cen_[3 * (size_t)i + k] = 0.5f * (lo[3 * (size_t)i + k] + hi[3 * (size_t)i + k]);
        }
    }

    BvhResult run() {
        BvhResult res;
        if (n_ == 0) {
            res.root = INT32_MIN;
            return res;
        }
        // Phase 1: serial top.
        std::vector<Node> top;
        std::vector<PendingTask> tasks;
        int maxd = 0;
        int32_t root = build_top(0, n_, 0, top, tasks, maxd);
        // Phase 2: subtrees in parallel.
        std::vector<std::vector<Node>> sub(tasks.size());
        std::vector<int32_t> sub_root(tasks.size());
        std::vector<int> sub_depth(tasks.size(), 0);
        unsigned hw = std::max(1u, std::thread::hardware_concurrency());
        std::atomic_size_t next(0);
        auto worker = [&]() {
            for (;;) {
                size_t t = next.fetch_add(1);
                if (t >= tasks.size()) break;
                int md = 0;
                sub_root[t] = build_rec(tasks[t].b, tasks[t].e, tasks[t].depth, sub[t], md);
                sub_depth[t] = md;
            }
        };
        std::vector<std::thread> pool;
        for (unsigned i = 0; i < std::min<size_t>(hw, tasks.size()); i++) pool.emplace_back(worker);
        for (auto &th : pool) th.join();
        // Phase 3: concatenate in task order and patch links.
        res.nodes = std::move(top);
        auto fix = [](int32_t link, int32_t off) { return link >= 0 ? link + off : link; };
        for (size_t t = 0; t < tasks.size(); t++) {
            int32_t off = (int32_t)res.nodes.size();
            for (Node nd : sub[t]) {
                nd.c0 = fix(nd.c0, off);
                nd.c1 = fix(nd.c1, off);
                res.nodes.push_back(nd);
            }
            int32_t link = fix(sub_root[t], off);
            if (tasks[t].node == SIZE_MAX) {
                root = link;
            } else {
                Node &p = res.nodes[tasks[t].node];
                if (tasks[t].which == 0)
                    p.c0 = link;
                else
                    p.c1 = link;
            }
            maxd = std::max(maxd, sub_depth[t]);
        }
        res.root = root;
        res.depth = (uint32_t)maxd;
        res.order = std::move(idx_);
        return res;
    }

  private:
    struct PendingTask {
        uint32_t b, e;
        int depth;
        size_t node; // parent node index in top array, SIZE_MAX = root
        int which;
    };

    const float *lo_, *hi_;
    uint32_t n_;
    int max_leaf_, bins_;
    std::vector<uint32_t> idx_;
    std::vector<float> cen_;

    Box3 prim_box(uint32_t p) const {
        Box3 b;
        for (int k = 0; k < 3; k++) {
            b.lo[k] = lo_[3 * (size_t)p + k];
            b.hi[k] = hi_[3 * (size_t)p + k];
        }
        return b;
    }

    static int32_t leaf_link(uint32_t start, uint32_t count) {
        return ~(int32_t)((start << 3) | (count - 1));
    }

    // Choose a split for [b, e); returns the split position m (b < m < e).
    uint32_t split(uint32_t b, uint32_t e, int depth, Box3 &lb, Box3 &rb) {
        Box3 cb;
        cb.empty();
        for (uint32_t i = b; i < e; i++) {
            const float *c = &cen_[3 * (size_t)idx_[i]];
            for (int k = 0; k < 3; k++) {
                cb.lo[k] = std::min(cb.lo[k], c[k]);
                cb.hi[k] = std::max(cb.hi[k], c[k]);
            }
        }
        int axis = 0;
        float ext = cb.hi[0] - cb.lo[0];
        for (int k = 1; k < 3; k++)
            if (cb.hi[k] - cb.lo[k] > ext) {
                ext = cb.hi[k] - cb.lo[k];
                axis = k;
            }
        uint32_t m = b + (e - b) / 2;
        bool done = false;
        if (ext > 0.0f && depth < kSahDepthLimit) {
            // binned SAH over all three axes
            float best_cost = INFINITY;
            int best_axis = -1, best_bin = -1;
            for (int a = 0; a < 3; a++) {
                float extent = cb.hi[a] - cb.lo[a];
                if (!(extent > 0.0f)) continue;
                Box3 bb[kMaxBins];
                uint32_t cnt[kMaxBins] = {0};
                for (int i = 0; i < bins_; i++) bb[i].empty();
                float scale = bins_ / extent;
                for (uint32_t i = b; i < e; i++) {
                    uint32_t p = idx_[i];
                    int bi = (int)((cen_[3 * (size_t)p + a] - cb.lo[a]) * scale);
                    bi = std::min(bins_ - 1, std::max(0, bi));
                    cnt[bi]++;
                    bb[bi].grow(prim_box(p));
                }
                float right_area[kMaxBins];
                uint32_t right_cnt[kMaxBins];
                Box3 acc;
                acc.empty();
                uint32_t c = 0;
                for (int i = bins_ - 1; i > 0; i--) {
                    acc.grow(bb[i]);
                    c += cnt[i];
                    right_area[i] = acc.half_area();
                    right_cnt[i] = c;
                }
                acc.empty();
                c = 0;
                for (int i = 0; i < bins_ - 1; i++) {
                    acc.grow(bb[i]);
                    c += cnt[i];
                    if (c == 0 || right_cnt[i + 1] == 0) continue;
                    float cost = acc.half_area() * c + right_area[i + 1] * right_cnt[i + 1];
                    if (cost < best_cost) {
                        best_cost = cost;
                        best_axis = a;
                        best_bin = i;
                    }
                }
            }
            if (best_axis >= 0) {
                float scale = bins_ / (cb.hi[best_axis] - cb.lo[best_axis]);
                float lo_a = cb.lo[best_axis];
                auto mid = std::partition(idx_.begin() + b, idx_.begin() + e, [&](uint32_t p) {
                    int bi = (int)((cen_[3 * (size_t)p + best_axis] - lo_a) * scale);
                    bi = std::min(bins_ - 1, std::max(0, bi));
                    return bi <= best_bin;
                });
                m = (uint32_t)(mid - idx_.begin());
                done = (m > b && m < e);
            }
        }
        if (!done) {
            // object median along the widest centroid axis
            m = b + (e - b) / 2;
            std::nth_element(idx_.begin() + b, idx_.begin() + m, idx_.begin() + e,
                             [&](uint32_t p, uint32_t q) {
                                 float cp = cen_[3 * (size_t)p + axis], cq = cen_[3 * (size_t)q + axis];
                                 return cp < cq || (cp == cq && p < q);
                             });
        }
        lb.empty();
        rb.empty();
        for (uint32_t i = b; i < m; i++) lb.grow(prim_box(idx_[i]));
        for (uint32_t i = m; i < e; i++) rb.grow(prim_box(idx_[i]));
        return m;
    }

    static void set_child(Node &nd, int which, const Box3 &bx, int32_t link) {
        if (which == 0) {
            nd.lo0x = bx.lo[0]; nd.hi0x = bx.hi[0];
            nd.lo0y = bx.lo[1]; nd.hi0y = bx.hi[1];
            nd.lo0z = bx.lo[2]; nd.hi0z = bx.hi[2];
            nd.c0 = link;
        } else {
            nd.lo1x = bx.lo[0]; nd.hi1x = bx.hi[0];
            nd.lo1y = bx.lo[1]; nd.hi1y = bx.hi[1];
            nd.lo1z = bx.lo[2]; nd.hi1z = bx.hi[2];
            nd.c1 = link;
        }
    }

    int32_t build_rec(uint32_t b, uint32_t e, int depth, std::vector<Node> &out, int &maxd) {
        maxd = std::max(maxd, depth);
        uint32_t n = e - b;
        if (n <= (uint32_t)max_leaf_) return leaf_link(b, n);
        Box3 lb, rb;
        uint32_t m = split(b, e, depth, lb, rb);
        size_t me = out.size();
        out.push_back(Node{});
        out[me].pad0 = out[me].pad1 = 0;
        int32_t l = build_rec(b, m, depth + 1, out, maxd);
        int32_t r = build_rec(m, e, depth + 1, out, maxd);
        set_child(out[me], 0, lb, l);
        set_child(out[me], 1, rb, r);
        return (int32_t)me;
    }

    int32_t build_top(uint32_t b, uint32_t e, int depth, std::vector<Node> &out,
                      std::vector<PendingTask> &tasks, int &maxd) {
        // the root itself may be small: build it directly
        if (e - b <= kParallelGrain) {
            tasks.push_back(PendingTask{b, e, depth, SIZE_MAX, 0});
            return INT32_MIN;
        }
        return top_rec(b, e, depth, out, tasks, maxd);
    }

    int32_t top_rec(uint32_t b, uint32_t e, int depth, std::vector<Node> &out,
                    std::vector<PendingTask> &tasks, int &maxd) {
        maxd = std::max(maxd, depth);
        Box3 lb, rb;
        uint32_t m = split(b, e, depth, lb, rb);
        size_t me = out.size();
        out.push_back(Node{});
        out[me].pad0 = out[me].pad1 = 0;
        uint32_t ranges[2][2] = {{b, m}, {m, e}};
        for (int w = 0; w < 2; w++) {
            uint32_t cb = ranges[w][0], ce = ranges[w][1];
            int32_t link;
            if (ce - cb <= kParallelGrain) {
                tasks.push_back(PendingTask{cb, ce, depth + 1, me, w});
                link = 0; // patched after the subtree is built
            } else {
                link = top_rec(cb, ce, depth + 1, out, tasks, maxd);
            }
            set_child(out[me], w, w == 0 ? lb : rb, link);
        }
        return (int32_t)me;
    }
};

} // namespace

BvhResult build_bvh(const float *lo, const float *hi, uint32_t n, int max_leaf, int bins) {
    Builder b(lo, hi, n, max_leaf, bins);
    return b.run();
}

namespace {

struct Cand {
    int32_t link;
    float lo[3], hi[3];
};

float cand_area(const Cand &c) {
    float d[3];
    for (int k = 0; k < 3; k++) d[k] = std::max(0.0f, c.hi[k] - c.lo[k]);
    return d[0] * d[1] + d[1] * d[2] + d[2] * d[0];
}

void node_children(const Node &m, Cand &a, Cand &b) {
    a = Cand{m.c0, {m.lo0x, m.lo0y, m.lo0z}, {m.hi0x, m.hi0y, m.hi0z}};
    b = Cand{m.c1, {m.lo1x, m.lo1y, m.lo1z}, {m.hi1x, m.hi1y, m.hi1z}};
}

int32_t collapse_rec(const std::vector<Node> &nodes, int32_t n, uint32_t d, std::vector<Node4> &out,
                     uint32_t &depth) {
    depth = std::max(depth, d);
    Cand c[4];
    int cnt = 2;
    node_children(nodes[n], c[0], c[1]);
    while (cnt < 4) {
        int best = -1;
        float ba = -1.0f;
        for (int k = 0; k < cnt; k++)
            if (c[k].link >= 0 && cand_area(c[k]) > ba) {
                ba = cand_area(c[k]);
                best = k;
            }
        if (best < 0) break;
        Cand a, b;
        node_children(nodes[c[best].link], a, b);
        c[best] = a;
        c[cnt++] = b;
    }
    const int32_t idx = (int32_t)out.size();
    out.emplace_back();
    int32_t links[4];
    for (int k = 0; k < 4; k++) {
        if (k >= cnt)
            links[k] = INT32_MIN;
        else
            links[k] = c[k].link >= 0 ? collapse_rec(nodes, c[k].link, d + 1, out, depth) : c[k].link;
    }
    Node4 &w = out[(size_t)idx];
    for (int k = 0; k < 4; k++) {
        const bool v = k < cnt;
        w.lox[k] = v ? c[k].lo[0] : 0.0f;
        w.loy[k] = v ? c[k].lo[1] : 0.0f;
        w.loz[k] = v ? c[k].lo[2] : 0.0f;
        w.hix[k] = v ? c[k].hi[0] : 0.0f;
        w.hiy[k] = v ? c[k].hi[1] : 0.0f;
        w.hiz[k] = v ? c[k].hi[2] : 0.0f;
        w.c[k] = links[k];
        w.pad[k] = 0;
    }
    return idx;
}

} // namespace

void collapse_bvh4(const std::vector<Node> &nodes, int32_t root, std::vector<Node4> &nodes4, int32_t &root4,
                   uint32_t &depth4) {
    nodes4.clear();
    depth4 = 0;
    if (root < 0) { // empty, or a single leaf
        root4 = root;
        return;
    }
    nodes4.reserve(nodes.size() / 2 + 1);
    root4 = collapse_rec(nodes, root, 0, nodes4, depth4);
}

} // namespace pvx
