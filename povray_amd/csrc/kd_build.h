// Host kd-tree (binary space partition) builder: the alternative
// acceleration structure of the reference's +BM2 mode.
#ifndef PVX_KD_BUILD_H
#define PVX_KD_BUILD_H

#include <vector>

#include "pvx_types.h"

namespace pvx {

struct KdTree {
    std::vector<KdNode> nodes; // node 0 = root
    std::vector<uint32_t> list; // primitive indices of the leaves
    float bmin[3] = {0, 0, 0}, bmax[3] = {0, 0, 0};
    uint32_t depth = 0;
};

// lo/hi: 3*n conservative FP32 primitive bounds; max_depth <= kKdMaxDepth.
KdTree build_kd(const float *lo, const float *hi, uint32_t n, uint32_t max_depth);

} // namespace pvx

#endif
