// Host BVH2 builder (binned SAH) producing the 64 B node layout of pvx_types.h.
#ifndef PVX_BVH_BUILD_H
#define PVX_BVH_BUILD_H

#include <vector>

#include "pvx_types.h"

namespace pvx {

struct BvhResult {
    std::vector<Node> nodes;
    std::vector<uint32_t> order; // leaf order: order[k] = input primitive id
    int32_t root = INT32_MIN;    // node index, or a leaf link when the tree is one leaf
    uint32_t depth = 0;          // max node depth (root = 0)
};

// lo/hi: 3*n conservative FP32 primitive bounds.
BvhResult build_bvh(const float *lo, const float *hi, uint32_t n, int max_leaf, int bins);

// Collapses a binary tree into a 4-wide one over the same leaves: each wide
// node takes its binary node's two children and keeps opening the inner
// child of largest surface area until it holds four.  depth4 = max wide
// node depth (root = 0).
void collapse_bvh4(const std::vector<Node> &nodes, int32_t root, std::vector<Node4> &nodes4, int32_t &root4,
                   uint32_t &depth4);

} // namespace pvx

#endif
