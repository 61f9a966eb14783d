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

} // namespace pvx

#endif
