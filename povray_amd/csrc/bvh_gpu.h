// GPU BVH builder: the bounding-slab tree (the reference's Build_BBox_Tree,
// core/bounding/boundingbox.cpp:262-430) rebuilt as a linear BVH on the
// device, emitted in the same 64 B binary and 128 B 4-wide node layouts as
// the host builder.
#ifndef PVX_BVH_GPU_H
#define PVX_BVH_GPU_H

#include <hip/hip_runtime.h>

#include <vector>

#include "pvx_types.h"

namespace pvx {

struct GpuBvh {
    Node *nodes = nullptr; // binary tree (device)
    uint32_t n_nodes = 0;
    int32_t root = INT32_MIN;
    uint32_t depth = 0;
    Node4 *nodes4 = nullptr; // 4-wide tree over the same leaves (when asked for)
    uint32_t n_nodes4 = 0;
    int32_t root4 = INT32_MIN;
    uint32_t depth4 = 0;
    uint64_t device_bytes = 0; // bytes of the node arrays
};

// Builds the tree over n primitives.  d_lo / d_hi: 3n conservative FP32
// bounds (device); d_prims_in: the records in input order; d_prims_out
// (device, n records) receives them in leaf order.  Leaves hold up to
// max_leaf (1..8) primitives.  The node arrays are allocated with hipMalloc
// and appended to `allocs`; temporaries are freed before returning.
// Runs on `st` and synchronizes it.  Returns a PVX_* status.
int gpu_build_bvh(const float *d_lo, const float *d_hi, const PrimRec *d_prims_in, PrimRec *d_prims_out,
                  uint32_t n, int max_leaf, bool wide, hipStream_t st, GpuBvh &out, std::vector<void *> &allocs);

} // namespace pvx

#endif
