// FP64 CPU ORACLE — TEST INFRASTRUCTURE ONLY (see oracle.hpp).
// Restatement of the "bounding slabs" BBox tree: build, priority queue and
// slab tests (core/bounding/boundingbox.cpp, boundingbox.h,
// core/scene/object.cpp:1074-1108).
#include <algorithm>
#include <functional>

#include "oracle.hpp"

namespace orc {

static const int kBunchingFactor = 4; // boundingbox.cpp:68

V3 mtrans_point(const double (*m)[4], V3 p) {
    V3 r;
    for (int i = 0; i < 3; i++) r[i] = p.x * m[0][i] + p.y * m[1][i] + p.z * m[2][i] + m[3][i];
    return r;
}
V3 mtrans_direction(const double (*m)[4], V3 p) {
    V3 r;
    for (int i = 0; i < 3; i++) r[i] = p.x * m[0][i] + p.y * m[1][i] + p.z * m[2][i];
    return r;
}
V3 minv_trans_normal(const double (*m)[4], V3 p) {
    V3 r;
    for (int i = 0; i < 3; i++) r[i] = p.x * m[i][0] + p.y * m[i][1] + p.z * m[i][2];
    return r;
}

// boundingbox.h:114-118: lowerLeft = float(mins), size = float(maxs - mins).
FBox make_bbox_from_min_max(V3 mins, V3 maxs) {
    FBox b;
    for (int k = 0; k < 3; k++) {
        b.ll[k] = (float)mins[k];
        b.size[k] = (float)(maxs[k] - mins[k]);
    }
    return b;
}

// --- build ----------------------------------------------------------------
namespace {

struct Builder {
    std::vector<BNode> &nodes;
    std::vector<int32_t> finite; // node ids, grows with composite nodes
    std::vector<float> area_left, area_right;
    int32_t root = -1;

    explicit Builder(std::vector<BNode> &n) : nodes(n) {}

    // find_axis (boundingbox.cpp:693-749): FP32 extents, first max wins.
    int find_axis(ptrdiff_t first, ptrdiff_t last) {
        float mins[3], maxs[3];
        for (int k = 0; k < 3; k++) {
            mins[k] = (float)kBoundHuge;
            maxs[k] = (float)-kBoundHuge;
        }
        for (ptrdiff_t i = first; i < last; i++) {
            const FBox &b = nodes[finite[i]].bbox;
            for (int k = 0; k < 3; k++) {
                if (b.ll[k] < mins[k]) mins[k] = b.ll[k];
                float hi = b.ll[k] + b.size[k];
                if (hi > maxs[k]) maxs[k] = hi;
            }
        }
        int which = 0;
        float d = (float)-kBoundHuge, e;
        e = maxs[0] - mins[0];
        if (e > d) { d = e; which = 0; }
        e = maxs[1] - mins[1];
        if (e > d) { d = e; which = 1; }
        e = maxs[2] - mins[2];
        if (e > d) which = 2;
        return which;
    }

    // calc_bbox (boundingbox.cpp:751-785): DBL accumulation, then FP32 box.
    FBox calc_bbox(const std::vector<int32_t> &ids, ptrdiff_t first, ptrdiff_t last) {
        V3 bmin(kBoundHuge, kBoundHuge, kBoundHuge), bmax(-kBoundHuge, -kBoundHuge, -kBoundHuge);
        for (ptrdiff_t i = first; i < last; i++) {
            const FBox &b = nodes[ids[i]].bbox;
            for (int k = 0; k < 3; k++) {
                double tmin = b.ll[k];
                double tmax = tmin + b.size[k];
                if (tmin < bmin[k]) bmin[k] = tmin;
                if (tmax > bmax[k]) bmax[k] = tmax;
            }
        }
        return make_bbox_from_min_max(bmin, bmax);
    }

    // build_area_table (boundingbox.cpp:787-832): running FP32 box, half area.
    void build_area_table(ptrdiff_t a, ptrdiff_t b, float *areas) {
        ptrdiff_t imin, dir;
        if (a < b) { imin = a; dir = 1; } else { imin = b; dir = -1; }
        float bmin[3], bmax[3];
        for (int k = 0; k < 3; k++) {
            bmin[k] = (float)kBoundHuge;
            bmax[k] = (float)-kBoundHuge;
        }
        for (ptrdiff_t i = a; i != b + dir; i += dir) {
            const FBox &bb = nodes[finite[i]].bbox;
            for (int k = 0; k < 3; k++) {
                float tmin = bb.ll[k];
                float tmax = tmin + bb.size[k];
                if (tmin < bmin[k]) bmin[k] = tmin;
                if (tmax > bmax[k]) bmax[k] = tmax;
            }
            float lx = bmax[0] - bmin[0], ly = bmax[1] - bmin[1], lz = bmax[2] - bmin[2];
            areas[i - imin] = lx * (ly + lz) + ly * lz;
        }
    }

    // sort_and_split (boundingbox.cpp:834-927).  The reference sorts with
    // qsort on the box-centre key; glibc's qsort is a merge sort (stable), so
    // a stable sort reproduces the element order.
    bool sort_and_split(ptrdiff_t first, ptrdiff_t last) {
        ptrdiff_t size = last - first;
        ptrdiff_t best_loc = -1;
        if (size <= 0) return false;
        if (size > kBunchingFactor) {
            int axis = find_axis(first, last);
            auto key = [&](int32_t id) {
                const FBox &b = nodes[id].bbox;
                return (float)(2.0 * (double)b.ll[axis] + (double)b.size[axis]);
            };
            std::stable_sort(finite.begin() + first, finite.begin() + last,
                             [&](int32_t a, int32_t b) { return key(a) < key(b); });
            if ((ptrdiff_t)area_left.size() < size) {
                area_left.resize(size);
                area_right.resize(size);
            }
            build_area_table(first, last - 1, area_left.data());
            build_area_table(last - 1, first, area_right.data());
            float best_index = area_right[0] * float(size - 3);
            for (ptrdiff_t i = 1; i < size; i++) {
                float new_index = float(i) * area_left[i - 1] + float(size - i) * area_right[i];
                if (new_index < best_index) {
                    best_index = new_index;
                    best_loc = i + first;
                }
            }
        }
        if (best_loc < 0) {
            BNode cd;
            cd.kids.assign(finite.begin() + first, finite.begin() + last);
            cd.bbox = calc_bbox(finite, first, last);
            int32_t id = (int32_t)nodes.size();
            nodes.push_back(std::move(cd));
            root = id;
            finite.push_back(id);
            return false;
        }
        sort_and_split(first, best_loc);
        sort_and_split(best_loc, last);
        return true;
    }
};

} // namespace

void BBoxTree::build(const std::vector<FBox> &finite_boxes, const std::vector<int32_t> &finite_ids,
                     const std::vector<FBox> &infinite_boxes,
                     const std::vector<int32_t> &infinite_ids) {
    nodes.clear();
    root = -1;
    Builder b(nodes);
    for (size_t i = 0; i < finite_boxes.size(); i++) {
        BNode n;
        n.bbox = finite_boxes[i];
        n.element = finite_ids[i];
        b.finite.push_back((int32_t)nodes.size());
        nodes.push_back(std::move(n));
    }
    std::vector<int32_t> inf;
    for (size_t i = 0; i < infinite_boxes.size(); i++) {
        BNode n;
        n.bbox = infinite_boxes[i];
        n.element = infinite_ids[i];
        n.infinite = true;
        inf.push_back((int32_t)nodes.size());
        nodes.push_back(std::move(n));
    }
    // Build_BBox_Tree (boundingbox.cpp:262-323)
    if (!finite_boxes.empty()) {
        ptrdiff_t low = 0, high = (ptrdiff_t)finite_boxes.size();
        while (b.sort_and_split(low, high)) {
            low = high;
            high = (ptrdiff_t)b.finite.size();
        }
        root = b.root;
        if (!inf.empty()) {
            BNode cd;
            cd.kids = inf;
            cd.bbox = b.calc_bbox(inf, 0, (ptrdiff_t)inf.size());
            cd.infinite = true;
            int32_t cid = (int32_t)nodes.size();
            nodes.push_back(std::move(cd));
            BNode &r = nodes[root];
            r.kids.insert(r.kids.begin(), cid);
            r.bbox = b.calc_bbox(r.kids, 0, (ptrdiff_t)r.kids.size());
            r.infinite = true;
        }
    } else if (!inf.empty()) {
        BNode cd;
        cd.kids = inf;
        cd.bbox = b.calc_bbox(inf, 0, (ptrdiff_t)inf.size());
        cd.infinite = true;
        root = (int32_t)nodes.size();
        nodes.push_back(std::move(cd));
    }
}

int BBoxTree::depth() const {
    if (root < 0) return 0;
    std::function<int(int32_t)> rec = [&](int32_t id) -> int {
        int m = 0;
        for (int32_t k : nodes[id].kids) m = std::max(m, rec(k));
        return m + 1;
    };
    return rec(root);
}

// --- traversal helpers ------------------------------------------------------
RayInfo::RayInfo(const Ray &r) {
    for (int k = 0; k < 3; k++) {
        origin[k] = (float)r.o[k];
        double t = r.d[k];
        nonzero[k] = (t != 0.0);
        inv_dir[k] = 0.0f;
        positive[k] = false;
        if (nonzero[k]) {
            inv_dir[k] = (float)(1.0 / t);
            positive[k] = (t > 0.0);
        }
    }
}

void PQueue::insert(double depth, int32_t node) {
    size_t size = q.size();
    q.resize(size + 1);
    size_t i = size;
    while (i > 1 && depth < q[i / 2].depth) {
        q[i] = q[i / 2];
        i /= 2;
    }
    q[i].depth = depth;
    q[i].node = node;
}

bool PQueue::remove_min(double &depth, int32_t &node) {
    size_t size = q.size() - 1;
    if (size == 0) return false;
    depth = q[1].depth;
    node = q[1].node;
    size_t i = 1, j;
    while (i <= size / 2) {
        if (2 * i == size || q[2 * i].depth < q[2 * i + 1].depth)
            j = 2 * i;
        else
            j = 2 * i + 1;
        if (q[size].depth <= q[j].depth) break;
        q[i] = q[j];
        i = j;
    }
    if (i != size) q[i] = q[size];
    q.pop_back();
    return true;
}

void check_and_enqueue(PQueue &queue, const BNode &node, int32_t node_id, const RayInfo &ri) {
    double dmin, dmax;
    if (!node.infinite) {
        const FBox &b = node.bbox;
        dmin = -kBoundHuge;
        dmax = kBoundHuge;
        for (int dim = 0; dim < 3; ++dim) {
            if (ri.nonzero[dim]) {
                double tmin, tmax;
                if (ri.positive[dim]) {
                    tmax = (float)((b.ll[dim] + b.size[dim] - ri.origin[dim]) * ri.inv_dir[dim]);
                    if (tmax < kEpsilon) return;
                    tmin = (float)((b.ll[dim] - ri.origin[dim]) * ri.inv_dir[dim]);
                } else {
                    tmax = (float)((b.ll[dim] - ri.origin[dim]) * ri.inv_dir[dim]);
                    if (tmax < kEpsilon) return;
                    tmin = (float)((b.ll[dim] + b.size[dim] - ri.origin[dim]) * ri.inv_dir[dim]);
                }
                if (tmax < dmax) {
                    if (tmin > dmin) {
                        if (tmin > tmax) return;
                        dmin = tmin;
                    } else {
                        if (dmin > tmax) return;
                    }
                    dmax = tmax;
                } else {
                    if (tmin > dmin) {
                        if (tmin > dmax) return;
                        dmin = tmin;
                    }
                }
            } else {
                float lo = b.ll[dim], hi = b.ll[dim] + b.size[dim];
                if (!(ri.origin[dim] >= lo && ri.origin[dim] <= hi)) return;
            }
        }
    } else {
        dmin = -kMaxDistance;
    }
    queue.insert(dmin, node_id);
}

// Intersect_BBox_Dir<BX,BY,BZ> (object.cpp:1074-1108) with the caller's
// variant/origin/invdir set-up (object.cpp:125-131): FP32 throughout.
bool intersect_bbox_dir(const FBox &bb, const Ray &r, float mind, float maxd) {
    float origin[3], invdir[3];
    for (int k = 0; k < 3; k++) {
        origin[k] = (float)r.o[k];
        invdir[k] = (float)(1.0 / r.d[k]);
    }
    int bx = invdir[0] < 0.0f, by = invdir[1] < 0.0f, bz = invdir[2] < 0.0f;
    float bounds[2][3];
    for (int k = 0; k < 3; k++) {
        bounds[0][k] = bb.ll[k];
        bounds[1][k] = bb.ll[k] + bb.size[k];
    }
    float tmin = (bounds[bx][0] - origin[0]) * invdir[0];
    float tmax = (bounds[1 - bx][0] - origin[0]) * invdir[0];
    float tymin = (bounds[by][1] - origin[1]) * invdir[1];
    float tymax = (bounds[1 - by][1] - origin[1]) * invdir[1];
    if ((tmin > tymax) || (tymin > tmax)) return false;
    if (tymin > tmin) tmin = tymin;
    if (tymax < tmax) tmax = tymax;
    float tzmin = (bounds[bz][2] - origin[2]) * invdir[2];
    float tzmax = (bounds[1 - bz][2] - origin[2]) * invdir[2];
    if ((tmin > tzmax) || (tzmin > tmax)) return false;
    if (tzmin > tmin) tmin = tzmin;
    if (tzmax < tmax) tmax = tzmax;
    return (tmin < maxd) && (tmax > mind);
}

} // namespace orc
