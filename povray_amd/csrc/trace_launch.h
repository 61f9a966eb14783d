// Closest-hit kernel launches shared by the pvx translation units.
//
// The traversal kernels are templates over the walk, the object-flag
// precondition, the counting mode and single/multi-scene batches.  They are
// instantiated in trace_kernels.hip, which the Makefile compiles once per
// part (PVX_TRACE_PART = 0..16), each part holding a few instantiations, so
// that they build in parallel; trace_launch() picks the part.
#ifndef PVX_TRACE_LAUNCH_H
#define PVX_TRACE_LAUNCH_H

#include <hip/hip_runtime.h>

#include "../../include/pvx.h"
#include "pvx_types.h"

namespace pvx {

constexpr int kBlock = 256;      // threads per block of every traversal kernel
constexpr int kStackLds = 24;    // traversal-stack entries per lane in LDS
constexpr int kStackSpill = 40;  // further entries per lane in the spill area
constexpr int kStackMax = kStackLds + kStackSpill;
constexpr int32_t kEmpty = INT32_MIN;

// Spill-area bytes per resident thread of a traversal grid.
constexpr size_t kSpillPerThread = (size_t)kStackSpill * sizeof(int32_t);

// Traversal walks: the 4-wide tree one node or leaf per pass (triangle and
// sphere scenes, light rays), the 4-wide tree while-while (camera and bounce
// rays), the binary tree (+ boxes and stand-alone triangles), the binary tree
// with every primitive kind (+ tori, CSG, cones, ...), the kd-tree, the
// binary tree with lathes and sors as well.
enum TraceWalk : uint32_t {
    kWalkStep4 = 0,
    kWalkWW4 = 1,
    kWalkBinary = 2,
    kWalkComplex = 3,
    kWalkKd = 4,
    kWalkRevolution = 5,
    kWalkLightWW = 6, // light rays made in the kernel, 4-wide tree while-while (occlusion)
    kWalkLightDyn = 7, // the same, each lane taking a new ray as soon as its ray ends
    kWalkBinary0 = 8,  // level-0 scene (triangles, spheres) through the binary tree
    kWalkLightBin = 9,  // light rays made in the kernel, binary tree (occlusion)
    kWalkBinary0WW = 10, // kWalkBinary0 in while-while passes
    kWalkLightDynBin = 11 // kWalkLightDyn through the binary tree
};

struct TraceKernel {
    uint32_t walk;
    bool flags; // object-flag precondition per ray kind
    bool count; // per-ray node-visit / primitive-test counters
    bool multi; // rays of many scenes (chunk table), no kd walk
};

// Arguments of every traversal kernel (passed by value as the kernel
// argument).  Single-scene launches use `sv` (and `kd`); multi-scene ones
// `views`, `chunks` and `clocks`.
struct TraceArgs {
    SceneView sv;
    KdView kd;
    const SceneView *views;
    const uint4 *chunks;
    unsigned long long *clocks;
    const pvx_ray *rays;
    const double *max_depth;
    uint32_t kind;
    const uint8_t *kinds;
    uint32_t n;
    pvx_hit *hits;
    int32_t *spill; // kSpillPerThread bytes for each thread of the grid
    uint32_t *node_visits;
    uint32_t *prim_tests;
    // kWalkLightWW: ray k goes from gen_groups[k / gen_S].ip to grid point
    // k mod gen_S of light gen_lights[gen_groups[k / gen_S].light] (rays past
    // that light's grid are skipped) with that distance as its max depth; its
    // visibility (1: nothing blocks) is written to gen_vis[k]
    const LightGroup *gen_groups;
    const LightRec *gen_lights;
    uint32_t gen_S;
    uint8_t *gen_vis;
    uint32_t *gen_next; // kWalkLightDyn: next ray to hand out (zero at launch)
    const uint32_t *gen_count; // when set: n = gen_count[0] * gen_S, read on the device
    uint32_t gen_refill_below; // kWalkLightDyn: refill once no more than this many lanes hold rays
};

// Launches kernel `k` as `nblocks` blocks of kBlock threads on `st`; with
// `occupancy` non-null it stores the kernel's resident blocks per CU there
// instead of launching.
hipError_t trace_launch(const TraceKernel &k, const TraceArgs &a, uint32_t nblocks, hipStream_t st,
                        int *occupancy = nullptr);

// Per-object queries (ObjectBase::All_Intersections, Inside_Object).
hipError_t launch_all_intersections(const SceneView &sv, const pvx_ray *rays, const uint32_t *objs, uint32_t n,
                                    uint32_t max_hits, double *depths, int32_t *prims, uint32_t *counts,
                                    const uint32_t *obj_type, const uint32_t *obj_index, hipStream_t st);
hipError_t launch_inside(const SceneView &sv, const double *pts, const uint32_t *objs, uint32_t n, uint8_t *out,
                         const uint32_t *obj_type, const uint32_t *obj_index, const uint32_t *obj_flags,
                         hipStream_t st);

// The parts (one per build of trace_kernels.hip).
#define PVX_TRACE_PART_DECL(n)                                                                              \
    hipError_t trace_part##n(const TraceKernel &k, const TraceArgs &a, uint32_t nblocks, hipStream_t st, \
                             int *occupancy)
PVX_TRACE_PART_DECL(0);
PVX_TRACE_PART_DECL(1);
PVX_TRACE_PART_DECL(2);
PVX_TRACE_PART_DECL(3);
PVX_TRACE_PART_DECL(4);
PVX_TRACE_PART_DECL(5);
PVX_TRACE_PART_DECL(6);
PVX_TRACE_PART_DECL(7);
PVX_TRACE_PART_DECL(8);
PVX_TRACE_PART_DECL(9);
PVX_TRACE_PART_DECL(10);
PVX_TRACE_PART_DECL(11);
PVX_TRACE_PART_DECL(13);
PVX_TRACE_PART_DECL(14);
PVX_TRACE_PART_DECL(15);
PVX_TRACE_PART_DECL(16);

} // namespace pvx

#endif
