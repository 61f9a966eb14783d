// pvx: MI355X (gfx950) closest-hit / per-shape / camera kernels and the C-ABI.
//
// Closest hit (replaces Trace::FindIntersection, core/render/trace.cpp:285-344):
//   * one lane per ray, persistent grid sized to the device (wave64);
//   * FP32 traversal of a BVH2 (64 B nodes): slab test in FMA form against
//     conservative boxes, near child first, far child pushed on a per-lane
//     LDS stack (32 entries, [depth][lane] so lanes never share a bank) that
//     overflows into a per-lane global spill area;
//   * leaves hold 64 B primitive records: mesh triangles and spheres inline,
//     boxes/tori/CSG by reference.  Each candidate is decided with the
//     reference's own FP64 arithmetic, so hit ids and depths follow it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pvx.h"
#include "bvh_gpu.h"
#include "internal.h"
#include "scene_prep.h"
#include "shapes_device.h"

using namespace pvx;

namespace pvx {
thread_local std::string g_err;
int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}
} // namespace pvx

extern "C" const char *pvx_last_error(void) { return pvx::g_err.c_str(); }

// ----------------------------------------------------------------------------
// kernels
// ----------------------------------------------------------------------------
namespace {

constexpr int kBlock = 256;
constexpr int kStackLds = 24;
constexpr int kStackSpill = 40;
constexpr int kStackMax = kStackLds + kStackSpill;
constexpr int32_t kEmpty = INT32_MIN;

struct RayF {
    float ix, iy, iz;    // 1 / direction (FP32, zero components clamped)
    float oix, oiy, oiz; // origin * inverse direction
};

__device__ inline float safe_inv(double d) {
    float f = (float)d;
    if (fabsf(f) < 1.0e-20f) f = f < 0.0f ? -1.0e-20f : 1.0e-20f;
    return 1.0f / f;
}

__device__ inline bool better(double dep, int32_t obj, int32_t prim, double bt, int32_t bo,
                              int32_t bp) {
    // strictly closer wins; an exact tie between two hits goes to the lower
    // (object, primitive) key so results do not depend on traversal order.
    if (dep < bt) return true;
    return dep == bt && bo >= 0 && (obj < bo || (obj == bo && prim < bp));
}

// Find_Intersection's depth filter (object.cpp:149-161) with the shadow-ray
// postcondition dist > SMALL_TOLERANCE (trace.cpp:1946-1949).
__device__ inline bool depth_ok(double dep, bool shadow) {
    return dep >= kMinIsectDepth && (!shadow || dep > kSmallTolerance);
}

// NoSomethingFlagRayObjectCondition (trace.cpp:84-95) /
// NoShadowFlagRayObjectCondition (trace.cpp:1941-1944).
__device__ inline bool precondition(uint32_t flags, uint32_t kind) {
    if (kind == PVX_RAY_SHADOW) return !(flags & PVX_FLAG_NO_SHADOW);
    if ((kind & PVX_RAY_IMAGE) && (flags & PVX_FLAG_NO_IMAGE)) return false;
    if ((kind & PVX_RAY_REFLECTION) && (flags & PVX_FLAG_NO_REFLECTION)) return false;
    if ((kind & PVX_RAY_RADIOSITY) && (flags & PVX_FLAG_NO_RADIOSITY)) return false;
    return true;
}

// Closest valid hit of one CSG object (its leaves' hits that survive every
// CSG filter up to the root); returns false when none beats `limit`.
__device__ inline bool csg_closest(const SceneView &sv, uint32_t root, D3 o, D3 d, bool shadow,
                                   double limit, double &dep_out, int32_t &slot_out) {
    uint32_t b = sv.csg_leaf_range[2 * root], e = sv.csg_leaf_range[2 * root + 1];
    bool found = false;
    double best = limit;
    for (uint32_t k = b; k < e; k++) {
        uint32_t slot = sv.csg_leaves[k];
        const CsgChild &c = sv.csg_children[slot];
        double hs[kMaxLeafHits];
        int nh = leaf_hits(sv, c.type, c.index, o, d, hs);
        for (int j = 0; j < nh; j++) {
            double dep = hs[j];
            if (!depth_ok(dep, shadow)) continue;
            if (!(dep < best || (found && dep == best && (int32_t)slot < slot_out))) continue;
            D3 ip = add3(o, mul3(d, dep));
            if (csg_survives(sv, slot, ip)) {
                best = dep;
                slot_out = (int32_t)slot;
                found = true;
            }
        }
    }
    dep_out = best;
    return found;
}

// Closest valid hit of a quadric (Find_Intersection over its IStack).
__device__ inline bool quadric_closest(const QuadricRec &q, D3 o, D3 d, bool shadow, double &dep_out) {
    double hs[2];
    const int nh = quadric_hits(q, o, d, hs);
    bool found = false;
    double best = 0.0;
    for (int j = 0; j < nh; j++)
        if (depth_ok(hs[j], shadow) && (!found || hs[j] < best)) {
            best = hs[j];
            found = true;
        }
    dep_out = best;
    return found;
}

// Closest valid hit of a box / torus / stand-alone triangle object
// (Find_Intersection over its IStack).
__device__ inline bool simple_closest(const SceneView &sv, uint32_t kind, uint32_t index, D3 o,
                                      D3 d, bool shadow, double &dep_out) {
    if (kind == kPrimTriangle64 || kind == kPrimDisc || kind == kPrimPolygon) {
        double dep;
        const bool hit = kind == kPrimDisc      ? disc_hit(sv.discs[index], sv.xforms, o, d, dep)
                         : kind == kPrimPolygon ? polygon_hit(sv.polygons[index], sv.poly_pts, sv.xforms, o, d, dep)
                                                : tri_rec_hit(sv.triangles[index], o, d, dep);
        if (!hit || !depth_ok(dep, shadow)) return false;
        dep_out = dep;
        return true;
    }
    double hs[kMaxLeafHits];
    int side[4];
    int nh = kind == kPrimBox    ? box_hits(sv.boxes[index], sv.xforms, o, d, hs, side)
             : kind == kPrimCone ? cone_hits(sv.cones[index], sv.xforms, o, d, hs, side)
                                 : torus_hits(sv.tori[index], sv.xforms, o, d, hs);
    bool found = false;
    double best = 0.0;
    for (int j = 0; j < nh; j++) {
        if (depth_ok(hs[j], shadow) && (!found || hs[j] < best)) {
            best = hs[j];
            found = true;
        }
    }
    dep_out = best;
    return found;
}

// kLevel 0: triangles and spheres; 1: also boxes; 2: also tori and CSG.
// kWW: while-while traversal (triangle scenes, camera and bounce rays).
// kMulti: rays of many scenes in one launch.  `n` is then 64 x the number of
// chunks; chunk c = {scene, first ray, ray count <= 64} covers the wavefront
// whose loop index is 64c..64c+63, so the scene (and its SceneView, read from
// `views`) is uniform across the wavefront.  With `clocks`, lane 0 of each
// wavefront records the wall clock at its start (minimum into clocks[0]) and
// at the end of each chunk (maximum into clocks[1 + scene]).
template <int kLevel, bool kFlags, bool kCount, bool kWW, bool kMulti = false>
__device__ __forceinline__ void trace_closest_body(const SceneView &sv0, const SceneView *__restrict__ views,
                                                   const uint4 *__restrict__ chunks,
                                                   unsigned long long *__restrict__ clocks,
                                                   const pvx_ray *__restrict__ rays,
                                                   const double *__restrict__ max_depth, uint32_t kind,
                                                   const uint8_t *__restrict__ kinds, uint32_t n,
                                                   pvx_hit *__restrict__ hits, int32_t *__restrict__ spill,
                                                   uint32_t *__restrict__ node_visits,
                                                   uint32_t *__restrict__ prim_tests, int32_t *s_stack) {
    const uint32_t lane = threadIdx.x;
    const uint32_t gtid = blockIdx.x * kBlock + lane;
    int32_t *my_spill = spill + (size_t)gtid * kStackSpill;
    // kRayOcclusion on a light-ray launch: stop at the first hit that the
    // caller counts as blocking (closer than max depth - SHADOW_TOLERANCE)
    const bool occl = (kind & kRayOcclusion) != 0;
    kind &= 0xFFu;
    const bool shadow = (kind == PVX_RAY_SHADOW);
    if constexpr (kMulti) {
        if (clocks && (lane & 63u) == 0) atomicMin(&clocks[0], (unsigned long long)wall_clock64());
    }

    for (uint32_t i = gtid; i < n; i += gridDim.x * kBlock) {
        uint32_t ri = i, scene = 0;
        if constexpr (kMulti) {
            const uint4 ch = chunks[__builtin_amdgcn_readfirstlane(i >> 6)];
            scene = ch.x;
            if ((i & 63u) >= ch.z) continue; // lane 0 always has a ray
            ri = ch.y + (i & 63u);
        }
        const SceneView sv = kMulti ? views[scene] : sv0;
        const pvx_ray r = rays[ri];
        // per-ray kind (precondition only; shadow rays come in launches of their own)
        const uint32_t rk = (kFlags && kinds) ? (uint32_t)kinds[ri] : kind;
        D3 o = d3(r.o[0], r.o[1], r.o[2]);
        D3 d = d3(r.d[0], r.d[1], r.d[2]);
        double bt = kBoundHuge;
        if (max_depth) {
            double md = max_depth[ri];
            if (md >= kEpsilon) bt = md; // trace.cpp:157-158
        }
        const double occl_lim = bt - kShadowTolerance;
        int32_t bo = -1, bp = 0;

        // objects outside the BVH (planes, unbounded CSG)
        for (uint32_t k = 0; k < sv.n_inf; k++) {
            const InfRec inf = sv.inf[k];
            if (kFlags && !precondition(sv.obj_flags[inf.obj], rk)) continue;
            double dep;
            int32_t prim = 0;
            bool ok;
            if (inf.kind == kPrimPlane) {
                ok = plane_hit(sv.planes[inf.index], o, d, dep) && depth_ok(dep, shadow);
            } else if (kLevel >= 2) {
                ok = inf.kind == kPrimQuadric ? quadric_closest(sv.quadrics[inf.index], o, d, shadow, dep)
                                              : csg_closest(sv, inf.index, o, d, shadow, bt, dep, prim);
            } else {
                ok = false;
            }
            if (ok && better(dep, inf.obj, prim, bt, bo, bp)) {
                bt = dep;
                bo = inf.obj;
                bp = prim;
            }
        }

        RayF rf;
        rf.ix = safe_inv(d.x);
        rf.iy = safe_inv(d.y);
        rf.iz = safe_inv(d.z);
        rf.oix = (float)o.x * rf.ix;
        rf.oiy = (float)o.y * rf.iy;
        rf.oiz = (float)o.z * rf.iz;
        float tb = (float)bt * 1.000001f;

        int32_t node = (occl && bo >= 0 && bt < occl_lim) ? kEmpty : kLevel == 0 ? sv.root4 : sv.root;
        int sp = 0;
        uint32_t cnt_nodes = 0, cnt_prims = 0;
        // Triangle / sphere scenes traverse the 4-wide tree; camera and
        // bounce rays with while-while passes (the lanes of a wavefront visit
        // inner nodes together, then leaves together), light rays one node or
        // leaf per pass.  Scenes with boxes, tori and CSG, whose leaf tests
        // are long, take the binary tree one step per pass.
#define PVX_POP()                                                                       \
    do {                                                                                \
        if (sp == 0) {                                                                  \
            node = kEmpty;                                                              \
        } else {                                                                        \
            sp--;                                                                       \
            node = sp < kStackLds ? s_stack[sp * kBlock + lane] : my_spill[sp - kStackLds]; \
        }                                                                               \
    } while (0)
        // one 4-wide inner-node step (triangle / sphere scenes): the hit
        // children sorted by entry distance, the nearest taken, the others
        // pushed farthest first; pops when no child is hit
        auto step_node4 = [&]() {
            if (kCount) cnt_nodes++;
            const float4 *np = reinterpret_cast<const float4 *>(sv.nodes4 + node);
            const float4 xl = np[0], xh = np[1], yl = np[2], yh = np[3], zl = np[4], zh = np[5];
            const int4 cl = reinterpret_cast<const int4 *>(np)[6];
            auto slab = [&](float lx, float hx, float ly, float hy, float lz, float hz, int32_t link) {
                const float x0 = fmaf(lx, rf.ix, -rf.oix), x1 = fmaf(hx, rf.ix, -rf.oix);
                const float y0 = fmaf(ly, rf.iy, -rf.oiy), y1 = fmaf(hy, rf.iy, -rf.oiy);
                const float z0 = fmaf(lz, rf.iz, -rf.oiz), z1 = fmaf(hz, rf.iz, -rf.oiz);
                const float tn = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), 0.0f));
                const float tf = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fminf(fmaxf(z0, z1), tb));
                return (link != kEmpty && tn <= tf * 1.000001f) ? tn : INFINITY;
            };
            float d0 = slab(xl.x, xh.x, yl.x, yh.x, zl.x, zh.x, cl.x);
            float d1 = slab(xl.y, xh.y, yl.y, yh.y, zl.y, zh.y, cl.y);
            float d2 = slab(xl.z, xh.z, yl.z, yh.z, zl.z, zh.z, cl.z);
            float d3 = slab(xl.w, xh.w, yl.w, yh.w, zl.w, zh.w, cl.w);
            int32_t l0 = cl.x, l1 = cl.y, l2 = cl.z, l3 = cl.w;
            auto cx = [](float &da, int32_t &la, float &db, int32_t &lb) {
                if (db < da) {
                    const float td = da;
                    da = db;
                    db = td;
                    const int32_t tl = la;
                    la = lb;
                    lb = tl;
                }
            };
            cx(d0, l0, d1, l1);
            cx(d2, l2, d3, l3);
            cx(d0, l0, d2, l2);
            cx(d1, l1, d3, l3);
            cx(d1, l1, d2, l2);
            if (d0 == INFINITY) {
                PVX_POP();
                return;
            }
            auto push = [&](int32_t x) {
                if (sp < kStackLds)
                    s_stack[sp * kBlock + lane] = x;
                else
                    my_spill[sp - kStackLds] = x;
                sp++;
            };
            if (d3 != INFINITY) push(l3);
            if (d2 != INFINITY) push(l2);
            if (d1 != INFINITY) push(l1);
            node = l0;
        };
        // one leaf step of a triangle / sphere scene (tests its primitives, then pops)
        auto step_leaf = [&]() {
            // leaf: ~node = first << 3 | (count - 1)
            const uint32_t leaf = (uint32_t)(~node);
            const uint32_t first = leaf >> 3, count = (leaf & 7u) + 1u;
            if (kCount) cnt_prims += count;
            for (uint32_t k = 0; k < count; k++) {
                const PrimRec &p = sv.prims[first + k];
                const uint32_t ka = p.kind_axis;
                const uint32_t pk = ka & 15u;
                const int32_t obj = p.obj;
                if (kFlags && !precondition(sv.obj_flags[obj], rk)) continue;
                if (pk == kPrimTriangle) {
                    double dep;
                    if (tri_hit64(p.tri, ka >> 4, o, d, dep) && depth_ok(dep, shadow) &&
                        better(dep, obj, p.tri.prim, bt, bo, bp)) {
                        bt = dep;
                        bo = obj;
                        bp = p.tri.prim;
                    }
                } else if (pk == kPrimSphere) {
                    double hs[2];
                    int nh = sphere_hits(d3(p.sph.c[0], p.sph.c[1], p.sph.c[2]), p.sph.r, o, d, hs);
                    for (int j = 0; j < nh; j++)
                        if (depth_ok(hs[j], shadow) && better(hs[j], obj, 0, bt, bo, bp)) {
                            bt = hs[j];
                            bo = obj;
                            bp = 0;
                        }
                }
            }
            tb = (float)bt * 1.000001f;
            if (occl && bo >= 0 && bt < occl_lim) {
                node = kEmpty;
                return;
            }
            PVX_POP();
        };
        if constexpr (kLevel == 0) {
            if constexpr (kWW) {
                while (node != kEmpty) {
                    while (node >= 0) step_node4();
                    while (node < 0 && node != kEmpty) step_leaf();
                }
            } else {
                while (node != kEmpty) {
                    if (node >= 0)
                        step_node4();
                    else
                        step_leaf();
                }
            }
        } else {
            while (node != kEmpty) {
                if (node >= 0) {
                    if (kCount) cnt_nodes++;
                    const float4 *np = reinterpret_cast<const float4 *>(sv.nodes + node);
                    const float4 a = np[0], b = np[1], c = np[2];
                    const int4 l = reinterpret_cast<const int4 *>(np)[3];
                    float x0 = fmaf(a.x, rf.ix, -rf.oix), x1 = fmaf(a.y, rf.ix, -rf.oix);
                    float y0 = fmaf(a.z, rf.iy, -rf.oiy), y1 = fmaf(a.w, rf.iy, -rf.oiy);
                    float z0 = fmaf(c.x, rf.iz, -rf.oiz), z1 = fmaf(c.y, rf.iz, -rf.oiz);
                    float tn0 = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), 0.0f));
                    float tf0 = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fminf(fmaxf(z0, z1), tb));
                    x0 = fmaf(b.x, rf.ix, -rf.oix);
                    x1 = fmaf(b.y, rf.ix, -rf.oix);
                    y0 = fmaf(b.z, rf.iy, -rf.oiy);
                    y1 = fmaf(b.w, rf.iy, -rf.oiy);
                    z0 = fmaf(c.z, rf.iz, -rf.oiz);
                    z1 = fmaf(c.w, rf.iz, -rf.oiz);
                    float tn1 = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), 0.0f));
                    float tf1 = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fminf(fmaxf(z0, z1), tb));
                    bool h0 = tn0 <= tf0 * 1.000001f;
                    bool h1 = tn1 <= tf1 * 1.000001f;
                    if (h0 && h1) {
                        int32_t nearc = l.x, farc = l.y;
                        if (tn1 < tn0) {
                            nearc = l.y;
                            farc = l.x;
                        }
                        if (sp < kStackLds)
                            s_stack[sp * kBlock + lane] = farc;
                        else
                            my_spill[sp - kStackLds] = farc;
                        sp++;
                        node = nearc;
                        continue;
                    }
                    if (h0) {
                        node = l.x;
                        continue;
                    }
                    if (h1) {
                        node = l.y;
                        continue;
                    }
                } else {
                    // leaf: ~node = first << 3 | (count - 1)
                    const uint32_t leaf = (uint32_t)(~node);
                    const uint32_t first = leaf >> 3, count = (leaf & 7u) + 1u;
                    if (kCount) cnt_prims += count;
                    for (uint32_t k = 0; k < count; k++) {
                        const PrimRec &p = sv.prims[first + k];
                        const uint32_t ka = p.kind_axis;
                        const uint32_t pk = ka & 15u;
                        const int32_t obj = p.obj;
                        if (kFlags && !precondition(sv.obj_flags[obj], rk)) continue;
                        if (pk == kPrimTriangle) {
                            double dep;
                            if (tri_hit64(p.tri, ka >> 4, o, d, dep) && depth_ok(dep, shadow) &&
                                better(dep, obj, p.tri.prim, bt, bo, bp)) {
                                bt = dep;
                                bo = obj;
                                bp = p.tri.prim;
                            }
                        } else if (pk == kPrimSphere) {
                            double hs[2];
                            int nh = sphere_hits(d3(p.sph.c[0], p.sph.c[1], p.sph.c[2]), p.sph.r, o, d, hs);
                            for (int j = 0; j < nh; j++)
                                if (depth_ok(hs[j], shadow) && better(hs[j], obj, 0, bt, bo, bp)) {
                                    bt = hs[j];
                                    bo = obj;
                                    bp = 0;
                                }
                        } else if (kLevel == 1) { // boxes and stand-alone triangles
                            if (pk == kPrimTriangle64) {
                                double dep;
                                if (tri_rec_hit(sv.triangles[p.ref.index], o, d, dep) && depth_ok(dep, shadow) &&
                                    better(dep, obj, 0, bt, bo, bp)) {
                                    bt = dep;
                                    bo = obj;
                                    bp = 0;
                                }
                            } else {
                                double hs[2];
                                int side[2];
                                int nh = box_hits(sv.boxes[p.ref.index], sv.xforms, o, d, hs, side);
                                for (int j = 0; j < nh; j++)
                                    if (depth_ok(hs[j], shadow) && better(hs[j], obj, 0, bt, bo, bp)) {
                                        bt = hs[j];
                                        bo = obj;
                                        bp = 0;
                                    }
                            }
                        } else if (kLevel >= 2) {
                            double dep;
                            int32_t prim = 0;
                            bool ok;
                            if (pk == kPrimCsg)
                                ok = csg_closest(sv, p.ref.index, o, d, shadow, bt, dep, prim);
                            else
                                ok = simple_closest(sv, pk, p.ref.index, o, d, shadow, dep);
                            if (ok && better(dep, obj, prim, bt, bo, bp)) {
                                bt = dep;
                                bo = obj;
                                bp = prim;
                            }
                        }
                    }
                    tb = (float)bt * 1.000001f;
                    if (occl && bo >= 0 && bt < occl_lim) {
                        node = kEmpty;
                        continue;
                    }
                }
                // pop
                if (sp == 0) {
                    node = kEmpty;
                } else {
                    sp--;
                    node = sp < kStackLds ? s_stack[sp * kBlock + lane] : my_spill[sp - kStackLds];
                }
            }
        }
#undef PVX_POP
        hits[ri] = pvx_hit{bt, bo, bp};
        if (kCount) {
            node_visits[ri] = cnt_nodes;
            prim_tests[ri] = cnt_prims;
        }
        if constexpr (kMulti) {
            if (clocks && (lane & 63u) == 0) atomicMax(&clocks[1 + scene], (unsigned long long)wall_clock64());
        }
    }
}

#define PVX_TRACE_ARGS                                                                                      \
    SceneView sv, const pvx_ray *__restrict__ rays, const double *__restrict__ max_depth, uint32_t kind,    \
        const uint8_t *__restrict__ kinds, uint32_t n, pvx_hit *__restrict__ hits,                          \
        int32_t *__restrict__ spill, uint32_t *__restrict__ node_visits, uint32_t *__restrict__ prim_tests

template <int kLevel, bool kFlags, bool kCount>
__global__ __launch_bounds__(kBlock) void k_trace_closest(PVX_TRACE_ARGS) {
    __shared__ int32_t s_stack[kStackLds * kBlock];
    trace_closest_body<kLevel, kFlags, kCount, false>(sv, nullptr, nullptr, nullptr, rays, max_depth, kind, kinds, n,
                                                      hits, spill, node_visits, prim_tests, s_stack);
}

// Triangle scenes, camera and bounce rays: while-while traversal (78 VGPRs
// and a 24 KB LDS stack per block: 6 wavefronts per SIMD; forcing 7 or 8
// costs more in register pressure than it gains).
template <bool kFlags, bool kCount>
__global__ __launch_bounds__(kBlock) void k_trace_closest_ww(
    PVX_TRACE_ARGS) {
    __shared__ int32_t s_stack[kStackLds * kBlock];
    trace_closest_body<0, kFlags, kCount, true>(sv, nullptr, nullptr, nullptr, rays, max_depth, kind, kinds, n, hits,
                                                spill, node_visits, prim_tests, s_stack);
}
#undef PVX_TRACE_ARGS

// The same kernels over a batch of scenes (chunk table and SceneView table).
#define PVX_MULTI_ARGS                                                                                      \
    const SceneView *__restrict__ views, const uint4 *__restrict__ chunks,                                  \
        unsigned long long *__restrict__ clocks, const pvx_ray *__restrict__ rays,                          \
        const double *__restrict__ max_depth, uint32_t kind, const uint8_t *__restrict__ kinds, uint32_t n, \
        pvx_hit *__restrict__ hits, int32_t *__restrict__ spill, uint32_t *__restrict__ node_visits,        \
        uint32_t *__restrict__ prim_tests

template <int kLevel, bool kFlags, bool kCount>
__global__ __launch_bounds__(kBlock) void k_trace_closest_multi(PVX_MULTI_ARGS) {
    __shared__ int32_t s_stack[kStackLds * kBlock];
    trace_closest_body<kLevel, kFlags, kCount, false, true>(views[0], views, chunks, clocks, rays, max_depth, kind,
                                                            kinds, n, hits, spill, node_visits, prim_tests, s_stack);
}

template <bool kFlags, bool kCount>
__global__ __launch_bounds__(kBlock) void k_trace_closest_ww_multi(PVX_MULTI_ARGS) {
    __shared__ int32_t s_stack[kStackLds * kBlock];
    trace_closest_body<0, kFlags, kCount, true, true>(views[0], views, chunks, clocks, rays, max_depth, kind, kinds,
                                                      n, hits, spill, node_visits, prim_tests, s_stack);
}
#undef PVX_MULTI_ARGS

__global__ void k_clock_init(unsigned long long *clocks, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) clocks[i] = i == 0 ? ~0ull : 0ull;
}

// Closest valid hit among one primitive record of any kind, folded into
// (bt, bo, bp) as the leaf loops of k_trace_closest do.
template <bool kFlags>
__device__ __forceinline__ void prim_test_any(const SceneView &sv, const PrimRec &p, D3 o, D3 d, bool shadow,
                                              uint32_t rk, double &bt, int32_t &bo, int32_t &bp) {
    const uint32_t ka = p.kind_axis;
    const uint32_t pk = ka & 15u;
    const int32_t obj = p.obj;
    if (kFlags && !precondition(sv.obj_flags[obj], rk)) return;
    if (pk == kPrimTriangle) {
        double dep;
        if (tri_hit64(p.tri, ka >> 4, o, d, dep) && depth_ok(dep, shadow) && better(dep, obj, p.tri.prim, bt, bo, bp)) {
            bt = dep;
            bo = obj;
            bp = p.tri.prim;
        }
    } else if (pk == kPrimSphere) {
        double hs[2];
        int nh = sphere_hits(d3(p.sph.c[0], p.sph.c[1], p.sph.c[2]), p.sph.r, o, d, hs);
        for (int j = 0; j < nh; j++)
            if (depth_ok(hs[j], shadow) && better(hs[j], obj, 0, bt, bo, bp)) {
                bt = hs[j];
                bo = obj;
                bp = 0;
            }
    } else {
        double dep;
        int32_t prim = 0;
        bool ok;
        if (pk == kPrimCsg)
            ok = csg_closest(sv, p.ref.index, o, d, shadow, bt, dep, prim);
        else
            ok = simple_closest(sv, pk, p.ref.index, o, d, shadow, dep);
        if (ok && better(dep, obj, prim, bt, bo, bp)) {
            bt = dep;
            bo = obj;
            bp = prim;
        }
    }
}

// Closest hit through the kd-tree (BSPTree::operator()(ray),
// core/bounding/bsptree.cpp:128-226): the ray's range inside the tree's box,
// the near child first with the far one remembered with its [entry, exit)
// range, every primitive of a reached leaf tested, and a stop once the next
// range starts beyond the closest hit.  Each primitive is decided as in
// k_trace_closest; one listed in several leaves is tested again (the
// reference's mailbox only saves that work).
template <bool kFlags, bool kCount>
__global__ __launch_bounds__(kBlock) void k_trace_closest_kd(SceneView sv, KdView kd, const pvx_ray *__restrict__ rays,
                                                           const double *__restrict__ max_depth, uint32_t kind,
                                                           const uint8_t *__restrict__ kinds, uint32_t n,
                                                           pvx_hit *__restrict__ hits, uint32_t *__restrict__ node_visits,
                                                           uint32_t *__restrict__ prim_tests) {
    struct Entry {
        uint32_t node;
        float rentry, rexit;
    };
    kind &= 0xFFu; // closest hits only (the occlusion bit is a hint)
    const bool shadow = (kind == PVX_RAY_SHADOW);
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const pvx_ray r = rays[i];
        const uint32_t rk = (kFlags && kinds) ? (uint32_t)kinds[i] : kind;
        const D3 o = d3(r.o[0], r.o[1], r.o[2]), d = d3(r.d[0], r.d[1], r.d[2]);
        double bt = kBoundHuge;
        if (max_depth) {
            const double md = max_depth[i];
            if (md >= kEpsilon) bt = md;
        }
        int32_t bo = -1, bp = 0;
        uint32_t cn = 0, cp = 0;
        for (uint32_t k = 0; k < sv.n_inf; k++) {
            const InfRec inf = sv.inf[k];
            if (kFlags && !precondition(sv.obj_flags[inf.obj], rk)) continue;
            double dep;
            int32_t prim = 0;
            bool ok = inf.kind == kPrimPlane     ? (plane_hit(sv.planes[inf.index], o, d, dep) && depth_ok(dep, shadow))
                      : inf.kind == kPrimQuadric ? quadric_closest(sv.quadrics[inf.index], o, d, shadow, dep)
                                                 : csg_closest(sv, inf.index, o, d, shadow, bt, dep, prim);
            if (ok && better(dep, inf.obj, prim, bt, bo, bp)) {
                bt = dep;
                bo = inf.obj;
                bp = prim;
            }
        }
        // the ray's range inside the tree's box (FP64 slabs)
        double t0 = -INFINITY, t1 = INFINITY;
        bool inbox = true;
        for (int a = 0; a < 3; a++) {
            const double oa = comp(o, a), da = comp(d, a), lo = kd.bmin[a], hi = kd.bmax[a];
            if (da == 0.0) {
                if (oa < lo || oa > hi) inbox = false;
                continue;
            }
            double ta = (lo - oa) / da, tb2 = (hi - oa) / da;
            if (ta > tb2) {
                const double x = ta;
                ta = tb2;
                tb2 = x;
            }
            t0 = fmax(t0, ta);
            t1 = fmin(t1, tb2);
        }
        if (inbox && t0 <= t1 && t1 >= 0.0) {
            double rentry = fmax(t0, 0.0), rexit = t1;
            Entry st[kKdMaxDepth + 1];
            int sp = 0;
            uint32_t inode = 0;
            while (rentry < bt) {
                const KdNode nd = kd.nodes[inode];
                if (!(nd.info >> 31)) { // split
                    if (kCount) cn++;
                    const uint32_t axis = (nd.info >> 29) & 3u;
                    uint32_t left = nd.info & 0x1FFFFFFFu, right = left + 1;
                    const float plane = nd.u.plane;
                    const double oa = comp(o, axis), da = comp(d, axis);
                    if (da == 0.0) { // parallel to the plane: one side, or both when on it
                        if (oa == (double)plane) {
                            st[sp++] = Entry{right, (float)rentry, (float)rexit};
                            inode = left;
                        } else {
                            inode = oa > (double)plane ? right : left;
                        }
                        continue;
                    }
                    const float rdist = (float)(((double)plane - oa) / da);
                    if ((oa > (double)plane) || ((rdist == 0.0f) && (da < 0))) {
                        const uint32_t x = left;
                        left = right;
                        right = x;
                    }
                    if ((rdist < 0.0f) || (rdist > rexit)) {
                        inode = left;
                    } else if (rdist < rentry) {
                        inode = right;
                    } else {
                        st[sp++] = Entry{right, rdist, (float)rexit};
                        inode = left;
                        rexit = rdist;
                    }
                } else { // leaf
                    const uint32_t count = nd.info & 0x7FFFFFFFu, first = nd.u.first;
                    if (kCount) cp += count;
                    for (uint32_t k = 0; k < count; k++)
                        prim_test_any<kFlags>(sv, sv.prims[kd.list[first + k]], o, d, shadow, rk, bt, bo, bp);
                    if (sp == 0) break;
                    sp--;
                    inode = st[sp].node;
                    rentry = st[sp].rentry;
                    rexit = st[sp].rexit;
                }
            }
        }
        hits[i] = pvx_hit{bt, bo, bp};
        if (kCount) {
            node_visits[i] = cn;
            prim_tests[i] = cp;
        }
    }
}

// Perspective camera rays (tracepixel.cpp:357-372, 918-926).
__global__ void k_camera_rays(pvx_camera cam, pvx_ray *rays) {
    uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t y = blockIdx.y;
    if (x >= cam.width || y >= cam.height) return;
    double x0 = ((double)x + cam.jitter_x) / (double)cam.width - 0.5;
    double y0 = 0.5 - ((double)y + cam.jitter_y) / (double)cam.height;
    double dx = cam.direction[0] + x0 * cam.right[0] + y0 * cam.up[0];
    double dy = cam.direction[1] + x0 * cam.right[1] + y0 * cam.up[1];
    double dz = cam.direction[2] + x0 * cam.right[2] + y0 * cam.up[2];
    double len = sqrt(dx * dx + dy * dy + dz * dz);
    pvx_ray r;
    r.o[0] = cam.location[0];
    r.o[1] = cam.location[1];
    r.o[2] = cam.location[2];
    r.d[0] = dx / len;
    r.d[1] = dy / len;
    r.d[2] = dz / len;
    rays[(size_t)y * cam.width + x] = r;
}

// All intersections of one object per ray, sorted ascending
// (ObjectBase::All_Intersections, object.h:238).
__global__ void k_all_intersections(SceneView sv, const pvx_ray *rays, const uint32_t *objs,
                                    uint32_t n, uint32_t max_hits, double *depths, int32_t *prims,
                                    uint32_t *counts, const uint32_t *obj_type,
                                    const uint32_t *obj_index) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const pvx_ray r = rays[i];
    D3 o = d3(r.o[0], r.o[1], r.o[2]);
    D3 d = d3(r.d[0], r.d[1], r.d[2]);
    uint32_t ob = objs[i];
    uint32_t type = obj_type[ob], index = obj_index[ob];
    double buf[16];
    int32_t pbuf[16];
    uint32_t cnt = 0;
    auto push = [&](double dep, int32_t prim) {
        if (cnt < 16) {
            // insertion sort keeps the list ascending
            uint32_t j = cnt;
            while (j > 0 && buf[j - 1] > dep) {
                buf[j] = buf[j - 1];
                pbuf[j] = pbuf[j - 1];
                j--;
            }
            buf[j] = dep;
            pbuf[j] = prim;
        }
        cnt++;
    };
    if (type == PVX_PLANE) {
        double dep;
        if (plane_hit(sv.planes[index], o, d, dep)) push(dep, 0);
    } else if (type == PVX_CSG) {
        uint32_t b = sv.csg_leaf_range[2 * index], e = sv.csg_leaf_range[2 * index + 1];
        for (uint32_t k = b; k < e; k++) {
            uint32_t slot = sv.csg_leaves[k];
            const CsgChild &c = sv.csg_children[slot];
            double hs[kMaxLeafHits];
            int nh = leaf_hits(sv, c.type, c.index, o, d, hs);
            for (int j = 0; j < nh; j++)
                if (csg_survives(sv, slot, add3(o, mul3(d, hs[j])))) push(hs[j], (int32_t)slot);
        }
    } else if (type == PVX_MESH) {
        // every triangle of the mesh: walk the whole BVH without culling by depth
        int32_t stack[kStackMax];
        int sp = 0;
        int32_t node = sv.root;
        float ix = safe_inv(d.x), iy = safe_inv(d.y), iz = safe_inv(d.z);
        float oix = (float)o.x * ix, oiy = (float)o.y * iy, oiz = (float)o.z * iz;
        while (node != kEmpty) {
            if (node >= 0) {
                const Node &nd = sv.nodes[node];
                const float lo0[3] = {nd.lo0x, nd.lo0y, nd.lo0z}, hi0[3] = {nd.hi0x, nd.hi0y, nd.hi0z};
                const float lo1[3] = {nd.lo1x, nd.lo1y, nd.lo1z}, hi1[3] = {nd.hi1x, nd.hi1y, nd.hi1z};
                const float iv[3] = {ix, iy, iz}, oi[3] = {oix, oiy, oiz};
                float tn0 = 0.0f, tf0 = 3.0e38f, tn1 = 0.0f, tf1 = 3.0e38f;
                for (int k = 0; k < 3; k++) {
                    float a0 = fmaf(lo0[k], iv[k], -oi[k]), b0 = fmaf(hi0[k], iv[k], -oi[k]);
                    float a1 = fmaf(lo1[k], iv[k], -oi[k]), b1 = fmaf(hi1[k], iv[k], -oi[k]);
                    tn0 = fmaxf(tn0, fminf(a0, b0));
                    tf0 = fminf(tf0, fmaxf(a0, b0));
                    tn1 = fmaxf(tn1, fminf(a1, b1));
                    tf1 = fminf(tf1, fmaxf(a1, b1));
                }
                bool h0 = tn0 <= tf0 * 1.000001f, h1 = tn1 <= tf1 * 1.000001f;
                if (h0 && h1 && sp < kStackMax) {
                    stack[sp++] = nd.c1;
                    node = nd.c0;
                    continue;
                }
                if (h0) { node = nd.c0; continue; }
                if (h1) { node = nd.c1; continue; }
            } else {
                uint32_t leaf = (uint32_t)(~node), first = leaf >> 3, count = (leaf & 7u) + 1u;
                for (uint32_t k = 0; k < count; k++) {
                    const PrimRec &p = sv.prims[first + k];
                    if ((p.kind_axis & 15u) != kPrimTriangle || p.obj != (int32_t)ob) continue;
                    double dep;
                    if (tri_hit64(p.tri, p.kind_axis >> 4, o, d, dep)) push(dep, p.tri.prim);
                }
            }
            node = sp == 0 ? kEmpty : stack[--sp];
        }
    } else {
        double hs[kMaxLeafHits];
        int nh = leaf_hits(sv, type, index, o, d, hs);
        for (int j = 0; j < nh; j++) push(hs[j], 0);
    }
    counts[i] = cnt;
    for (uint32_t k = 0; k < max_hits; k++) {
        bool have = k < cnt && k < 16;
        depths[(size_t)i * max_hits + k] = have ? buf[k] : HUGE_VAL;
        prims[(size_t)i * max_hits + k] = have ? pbuf[k] : -1;
    }
}

// Inside_Object per (point, object) (object.cpp:346-355).
__global__ void k_inside(SceneView sv, const double *pts, const uint32_t *objs, uint32_t n,
                         uint8_t *out, const uint32_t *obj_type, const uint32_t *obj_index,
                         const uint32_t *obj_flags) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    D3 p = d3(pts[3 * (size_t)i], pts[3 * (size_t)i + 1], pts[3 * (size_t)i + 2]);
    uint32_t ob = objs[i];
    out[i] = shape_inside<kCsgMaxDepth>(sv, obj_type[ob], obj_index[ob], obj_flags[ob], p) ? 1 : 0;
}

__global__ void k_copy16(uint4 *__restrict__ dst, const uint4 *__restrict__ src, size_t n16) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16;
         i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

} // namespace

template <class T>
static int upload(pvx_scene *s, const std::vector<T> &v, const T **dst) {
    *dst = nullptr;
    if (v.empty()) return PVX_OK;
    void *p = nullptr;
    HIP_TRY(hipMalloc(&p, v.size() * sizeof(T)));
    s->allocs.push_back(p);
    HIP_TRY(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    s->device_bytes += v.size() * sizeof(T);
    *dst = static_cast<const T *>(p);
    return PVX_OK;
}

static void scene_free(pvx_scene *s);

// PVX_BUILDER_GPU: upload the records (input order) and their bounds, build
// the tree on the device, keep the records in leaf order.
static int build_on_gpu(pvx_scene *s, const pvx_build_opts *opts) {
    HostScene &h = s->host;
    const auto t0 = std::chrono::steady_clock::now();
    const uint32_t n = (uint32_t)h.prims.size();
    s->n_prims = n;
    void *d_in = nullptr, *d_lo = nullptr, *d_hi = nullptr, *d_out = nullptr;
    int r = PVX_OK;
    auto cleanup = [&]() {
        if (d_in) (void)hipFree(d_in);
        if (d_lo) (void)hipFree(d_lo);
        if (d_hi) (void)hipFree(d_hi);
    };
    if (n > 0) {
        hipError_t e = hipMalloc(&d_out, (size_t)n * sizeof(PrimRec)); // kept: freed with the scene
        if (e == hipSuccess) {
            s->allocs.push_back(d_out);
            s->device_bytes += (size_t)n * sizeof(PrimRec);
            e = hipMalloc(&d_in, (size_t)n * sizeof(PrimRec));
        }
        if (e == hipSuccess) e = hipMalloc(&d_lo, (size_t)n * 12);
        if (e == hipSuccess) e = hipMalloc(&d_hi, (size_t)n * 12);
        if (e == hipSuccess) e = hipMemcpy(d_in, h.prims.data(), (size_t)n * sizeof(PrimRec), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(d_lo, h.prim_lo.data(), (size_t)n * 12, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(d_hi, h.prim_hi.data(), (size_t)n * 12, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            cleanup();
            return fail(PVX_EDEVICE, std::string("GPU BVH build upload: ") + hipGetErrorString(e));
        }
    }
    const auto tb = std::chrono::steady_clock::now();
    GpuBvh g;
    const int leaf = opts->max_leaf_size ? (int)opts->max_leaf_size : 4;
    r = gpu_build_bvh(static_cast<const float *>(d_lo), static_cast<const float *>(d_hi),
                      static_cast<const PrimRec *>(d_in), static_cast<PrimRec *>(d_out), n, leaf, h.has_complex == 0,
                      nullptr, g, s->allocs);
    const auto te = std::chrono::steady_clock::now();
    cleanup();
    if (r) return r;
    s->device_bytes += g.device_bytes;
    s->view.nodes = g.nodes;
    s->view.prims = static_cast<const PrimRec *>(d_out);
    s->view.nodes4 = g.nodes4;
    h.root = g.root;
    h.depth = g.depth;
    h.root4 = g.root4;
    h.depth4 = g.depth4;
    s->n_nodes2 = g.n_nodes;
    s->n_nodes4 = g.n_nodes4;
    h.bvh_ms = std::chrono::duration<double, std::milli>(te - tb).count();
    h.build_ms += std::chrono::duration<double, std::milli>(te - t0).count();
    // the host copies are no longer needed
    std::vector<PrimRec>().swap(h.prims);
    std::vector<float>().swap(h.prim_lo);
    std::vector<float>().swap(h.prim_hi);
    return PVX_OK;
}

static void scene_free(pvx_scene *s) {
    if (s->rad_engine_free && s->rad_engine) s->rad_engine_free(s->rad_engine);
    s->rad_engine = nullptr;
    for (void *p : s->allocs) (void)hipFree(p);
    s->allocs.clear();
    if (s->shading.d_mats) (void)hipFree(s->shading.d_mats);
    if (s->shading.d_lights) (void)hipFree(s->shading.d_lights);
    s->shading.d_mats = nullptr;
    s->shading.d_lights = nullptr;
}

// Kernel family of a scene: 0 (triangles / spheres through the 4-wide
// tree), 1 (+ boxes) or 2 (+ tori, CSG).  A triangle scene whose 4-wide tree
// is too deep for the traversal stack takes the binary tree (family 1).
static uint32_t kernel_level(const SceneView &v) { return v.has_complex == 0 && !v.wide ? 1u : v.has_complex; }

extern "C" int pvx_init(int device) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) return fail(PVX_EDEVICE, "no HIP device available");
    if (device < 0 || device >= n) return fail(PVX_EINVAL, "device index out of range");
    HIP_TRY(hipSetDevice(device));
    return PVX_OK;
}

extern "C" int pvx_scene_create(const pvx_scene_desc *desc, const pvx_build_opts *opts,
                                pvx_scene **out) {
    if (!desc || !out) return fail(PVX_EINVAL, "null argument");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(PVX_EDEVICE, "no HIP device available");
    pvx_scene *s = new pvx_scene();
    HIP_TRY(hipGetDevice(&s->device));
    std::string err;
    int rc = prepare_scene(*desc, opts, s->host, err);
    if (rc != PVX_OK) {
        delete s;
        return fail(rc, err);
    }
    HostScene &h = s->host;
    SceneView &v = s->view;
    int r = PVX_OK;
    if (h.gpu_build) {
        if ((r = build_on_gpu(s, opts)) != PVX_OK) {
            scene_free(s);
            delete s;
            return r;
        }
    } else {
        s->n_nodes2 = h.nodes.size();
        s->n_nodes4 = h.nodes4.size();
        s->n_prims = h.prims.size();
    }
    if (h.depth + 1 >= (uint32_t)kStackMax) {
        scene_free(s);
        delete s;
        return fail(PVX_EUNSUPPORTED, "BVH deeper than the traversal stack");
    }
#define UP(vec, field)                                          \
    if ((r = upload(s, vec, &v.field)) != PVX_OK) {             \
        scene_free(s);                                          \
        delete s;                                               \
        return r;                                               \
    }
    if (!h.gpu_build) {
        UP(h.nodes, nodes);
        UP(h.prims, prims);
        UP(h.nodes4, nodes4);
    }
    UP(h.inf, inf);
    UP(h.obj_flags, obj_flags);
    UP(h.planes, planes);
    UP(h.boxes, boxes);
    UP(h.tori, tori);
    UP(h.cones, cones);
    UP(h.discs, discs);
    UP(h.quadrics, quadrics);
    UP(h.polygons, polygons);
    UP(h.poly_pts, poly_pts);
    UP(h.xforms, xforms);
    UP(h.sph, sph);
    UP(h.csg, csg);
    UP(h.csg_children, csg_children);
    UP(h.csg_leaves, csg_leaves);
    UP(h.csg_leaf_range, csg_leaf_range);
    UP(h.shade_tris, shade_tris);
    UP(h.obj_tri_base, obj_tri_base);
    UP(h.mesh_vtx, mesh_vtx);
    UP(h.mesh_nrm, mesh_nrm);
    UP(h.triangles, triangles);
#undef UP
    if (h.use_kd) {
        if ((r = upload(s, h.kd_nodes, &s->kd.nodes)) != PVX_OK || (r = upload(s, h.kd_list, &s->kd.list)) != PVX_OK) {
            scene_free(s);
            delete s;
            return r;
        }
        for (int a = 0; a < 3; a++) {
            s->kd.bmin[a] = h.kd_bmin[a];
            s->kd.bmax[a] = h.kd_bmax[a];
        }
        s->use_kd = true;
    }
    const uint32_t *tmp = nullptr;
    if ((r = upload(s, h.obj_type, &tmp)) != PVX_OK) { scene_free(s); delete s; return r; }
    s->d_obj_type = const_cast<uint32_t *>(tmp);
    std::vector<uint32_t> obj_index(desc->n_objects);
    for (uint32_t i = 0; i < desc->n_objects; i++) obj_index[i] = desc->objects[i].index;
    if ((r = upload(s, obj_index, &tmp)) != PVX_OK) { scene_free(s); delete s; return r; }
    s->d_obj_index = const_cast<uint32_t *>(tmp);
    s->n_objects = desc->n_objects;
    v.root = h.root;
    v.root4 = h.root4;
    // the 4-wide walk pushes up to three links per level
    v.wide = (h.has_complex == 0 && 3 * h.depth4 + 1 < (uint32_t)kStackMax) ? 1u : 0u;
    v.n_inf = (uint32_t)h.inf.size();
    v.any_flags = h.any_flags;
    v.has_complex = h.has_complex;

    // persistent grid: enough resident blocks to fill every CU
    int cus = 0;
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, s->device));
    int per_cu = 0;
    auto pick = [](uint32_t level, bool flags) -> const void * {
        if (level >= 2) return flags ? (const void *)k_trace_closest<2, true, false> : (const void *)k_trace_closest<2, false, false>;
        if (level == 1) return flags ? (const void *)k_trace_closest<1, true, false> : (const void *)k_trace_closest<1, false, false>;
        return flags ? (const void *)k_trace_closest_ww<true, false> : (const void *)k_trace_closest_ww<false, false>;
    };
    auto kfn = pick(kernel_level(v), h.any_flags != 0);
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, kBlock, 0));
    if (per_cu < 1) per_cu = 1;
    s->grid = (uint32_t)(cus * per_cu);
    s->spill_bytes = (size_t)2 * s->grid * kBlock * kStackSpill * sizeof(int32_t); // grids up to 2x (autotuner)
    *out = s;
    return PVX_OK;
}

// The scene's spill area for launches on `st` (allocated on the stream's
// first launch; freed with the scene).
static int scene_spill(pvx_scene *s, hipStream_t st, int32_t **out) {
    std::lock_guard<std::mutex> lock(s->spill_mu);
    for (const auto &e : s->spills)
        if (e.first == st) {
            *out = e.second;
            return PVX_OK;
        }
    void *p = nullptr;
    HIP_TRY(hipMalloc(&p, s->spill_bytes));
    s->allocs.push_back(p);
    s->device_bytes += s->spill_bytes;
    s->spills.emplace_back(st, static_cast<int32_t *>(p));
    *out = static_cast<int32_t *>(p);
    return PVX_OK;
}

extern "C" int pvx_scene_destroy(pvx_scene *s) {
    if (!s) return PVX_OK;
    (void)hipDeviceSynchronize();
    scene_free(s);
    delete s;
    return PVX_OK;
}

extern "C" int pvx_scene_get_stats(const pvx_scene *s, pvx_scene_stats *out) {
    if (!s || !out) return fail(PVX_EINVAL, "null argument");
    const bool wide = s->view.wide != 0; // the tree the closest-hit kernels walk
    out->n_nodes = s->use_kd ? s->host.kd_nodes.size() : wide ? s->n_nodes4 : s->n_nodes2;
    out->n_prims = s->use_kd ? s->host.kd_list.size() : s->n_prims;
    out->node_bytes = s->use_kd ? sizeof(KdNode) : wide ? sizeof(Node4) : sizeof(Node);
    out->tri_bytes = sizeof(PrimRec);
    out->device_bytes = s->device_bytes;
    out->bvh_depth = s->use_kd ? s->host.kd_depth : wide ? s->host.depth4 : s->host.depth;
    out->n_infinite = (uint32_t)s->host.inf.size();
    out->build_ms = s->host.build_ms;
    out->bvh_build_ms = s->host.bvh_ms;
    return PVX_OK;
}

// One closest-hit launch with an explicit traversal variant and grid.
// variant: PVX_TRAVERSAL_DEFAULT (by scene content and ray kind), _WIDE_WW,
// _WIDE_STEP (4-wide tree; triangle/sphere scenes) or _BINARY (binary tree,
// scenes without tori/CSG); every variant gives the same hits.
template <bool kCount>
static int launch_variant(pvx_scene *s, const pvx_ray *rays, const double *max_depth, uint32_t ray_kind,
                          const uint8_t *kinds, size_t n, pvx_hit *hits, uint32_t *nv, uint32_t *pt,
                          hipStream_t st, uint32_t variant, uint32_t grid) {
    int32_t *spill = nullptr;
    int r;
    if ((r = scene_spill(s, st, &spill))) return r;
    uint32_t nblocks = (uint32_t)std::min<size_t>(grid, (n + kBlock - 1) / kBlock);
    bool flags = s->view.any_flags != 0 && (ray_kind != PVX_RAY_PRIMARY || kinds);
    uint32_t nn = (uint32_t)n;
#define PVX_LAUNCH(L, F) \
    k_trace_closest<L, F, kCount><<<nblocks, kBlock, 0, st>>>(s->view, rays, max_depth, ray_kind, kinds, nn, hits, spill, nv, pt)
#define PVX_LAUNCH_WW(L, F) \
    k_trace_closest_ww<F, kCount><<<nblocks, kBlock, 0, st>>>(s->view, rays, max_depth, ray_kind, kinds, nn, hits, spill, nv, pt)
    const uint32_t level = kernel_level(s->view);
    if (s->use_kd) {
        if (flags)
            k_trace_closest_kd<true, kCount><<<nblocks, kBlock, 0, st>>>(s->view, s->kd, rays, max_depth, ray_kind,
                                                                        kinds, nn, hits, nv, pt);
        else
            k_trace_closest_kd<false, kCount><<<nblocks, kBlock, 0, st>>>(s->view, s->kd, rays, max_depth, ray_kind,
                                                                         kinds, nn, hits, nv, pt);
    } else if (level >= 2) {
        if (flags) PVX_LAUNCH(2, true); else PVX_LAUNCH(2, false);
    } else if (level == 1 || variant == PVX_TRAVERSAL_BINARY) {
        if (flags) PVX_LAUNCH(1, true); else PVX_LAUNCH(1, false);
    } else if (variant == PVX_TRAVERSAL_WIDE_WW ||
               (variant == PVX_TRAVERSAL_DEFAULT && (ray_kind & 0xFFu) != PVX_RAY_SHADOW)) {
        if (flags) PVX_LAUNCH_WW(0, true); else PVX_LAUNCH_WW(0, false);
    } else {
        if (flags) PVX_LAUNCH(0, true); else PVX_LAUNCH(0, false);
    }
#undef PVX_LAUNCH
#undef PVX_LAUNCH_WW
    HIP_TRY(hipGetLastError());
    return PVX_OK;
}

template <bool kCount>
static int launch_closest(pvx_scene *s, const pvx_ray *rays, const double *max_depth,
                          uint32_t ray_kind, const uint8_t *kinds, size_t n, pvx_hit *hits, uint32_t *nv,
                          uint32_t *pt, void *stream) {
    if (!s) return fail(PVX_EINVAL, "null scene");
    if (n == 0) return PVX_OK;
    if (!rays || !hits) return fail(PVX_EINVAL, "null ray or hit buffer");
    if (n > 0xFFFFFFFFull) return fail(PVX_EINVAL, "too many rays in one call");
    const int cls = (ray_kind & 0xFFu) == PVX_RAY_SHADOW ? 1 : 0; // the autotuner's choice for this ray class
    const uint32_t grid = s->tune_grid[cls] ? s->tune_grid[cls] : s->grid;
    return launch_variant<kCount>(s, rays, max_depth, ray_kind, kinds, n, hits, nv, pt,
                                  static_cast<hipStream_t>(stream), s->tune_variant[cls], grid);
}

// Variants a scene admits: the kd-tree and tori/CSG scenes have one walk;
// box/triangle scenes the binary tree; triangle/sphere scenes all three.
static std::vector<uint32_t> scene_variants(const pvx_scene *s) {
    const uint32_t level = kernel_level(s->view);
    if (s->use_kd || level >= 2) return {PVX_TRAVERSAL_DEFAULT};
    if (level == 1) return {PVX_TRAVERSAL_BINARY};
    return {PVX_TRAVERSAL_WIDE_WW, PVX_TRAVERSAL_WIDE_STEP, PVX_TRAVERSAL_BINARY};
}

extern "C" int pvx_scene_set_traversal(pvx_scene *s, uint32_t ray_kind, uint32_t variant, uint32_t grid_blocks) {
    if (!s) return fail(PVX_EINVAL, "null scene");
    if (grid_blocks > 2 * s->grid)
        return fail(PVX_EINVAL, "grid_blocks above twice the resident blocks (" + std::to_string(2 * s->grid) + ")");
    if (variant != PVX_TRAVERSAL_DEFAULT) {
        auto v = scene_variants(s);
        if (std::find(v.begin(), v.end(), variant) == v.end())
            return fail(PVX_EINVAL, "traversal variant " + std::to_string(variant) + " not available for this scene");
    }
    const int cls = (ray_kind & 0xFFu) == PVX_RAY_SHADOW ? 1 : 0;
    s->tune_variant[cls] = variant;
    s->tune_grid[cls] = grid_blocks;
    return PVX_OK;
}

// Times every traversal variant the scene admits, on persistent grids of
// one and two times the resident blocks, over the caller's rays (1 warm-up +
// 3 timed launches each), and keeps the fastest for rays of this class
// (light rays, or camera / bounce / radiosity rays).
int pvx::autotune_internal(pvx_scene *s, const pvx_ray *rays, const double *max_depth, size_t n,
                           uint32_t ray_kind, pvx_tune_result *out, hipStream_t st) {
    if (!s || !rays) return fail(PVX_EINVAL, "null argument");
    if (n == 0 || n > 0xFFFFFFFFull) return fail(PVX_EINVAL, "n must be in [1, 2^32)");
    const std::vector<uint32_t> variants = scene_variants(s);
    void *hits = nullptr;
    HIP_TRY(hipMalloc(&hits, n * sizeof(pvx_hit)));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int r = PVX_OK;
    double best = 1e30, dflt = 0.0;
    uint32_t best_v = PVX_TRAVERSAL_DEFAULT, best_g = s->grid;
    auto time_one = [&](uint32_t v, uint32_t g, double &ms) -> int {
        int rr;
        if ((rr = launch_variant<false>(s, rays, max_depth, ray_kind, nullptr, n, static_cast<pvx_hit *>(hits),
                                        nullptr, nullptr, st, v, g)))
            return rr;
        HIP_TRY(hipEventRecord(e0, st));
        for (int k = 0; k < 3; k++)
            if ((rr = launch_variant<false>(s, rays, max_depth, ray_kind, nullptr, n, static_cast<pvx_hit *>(hits),
                                            nullptr, nullptr, st, v, g)))
                return rr;
        HIP_TRY(hipEventRecord(e1, st));
        HIP_TRY(hipEventSynchronize(e1));
        float f = 0;
        HIP_TRY(hipEventElapsedTime(&f, e0, e1));
        ms = f / 3.0;
        return PVX_OK;
    };
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e != hipSuccess) {
        r = fail(PVX_EDEVICE, std::string("autotune events: ") + hipGetErrorString(e));
    } else {
        r = time_one(PVX_TRAVERSAL_DEFAULT, s->grid, dflt);
        for (uint32_t v : variants)
            for (uint32_t g : {s->grid, 2 * s->grid}) {
                if (r) break;
                double ms = 0;
                r = time_one(v, g, ms);
                if (!r && ms < best) {
                    best = ms;
                    best_v = v;
                    best_g = g;
                }
            }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(hits);
    if (r) return r;
    const int cls = (ray_kind & 0xFFu) == PVX_RAY_SHADOW ? 1 : 0;
    s->tune_variant[cls] = best_v;
    s->tune_grid[cls] = best_g;
    if (out) {
        out->variant = best_v;
        out->grid_blocks = best_g;
        out->best_ms = best;
        out->default_ms = dflt;
    }
    return PVX_OK;
}

extern "C" int pvx_scene_autotune(pvx_scene *s, const pvx_ray *rays, size_t n, uint32_t ray_kind,
                                  pvx_tune_result *out, void *stream) {
    return pvx::autotune_internal(s, rays, nullptr, n, ray_kind, out, static_cast<hipStream_t>(stream));
}

int pvx::trace_closest_internal(pvx_scene *s, const pvx_ray *rays, const double *max_depth,
                                uint32_t ray_kind, size_t n, pvx_hit *hits, hipStream_t stream,
                                const uint8_t *kinds) {
    return launch_closest<false>(s, rays, max_depth, ray_kind, kinds, n, hits, nullptr, nullptr, stream);
}

extern "C" int pvx_trace_closest(pvx_scene *s, const pvx_ray *rays, const double *max_depth,
                                 uint32_t ray_kind, size_t n, pvx_hit *hits, void *stream) {
    if (ray_kind & ~0xFFu) return fail(PVX_EINVAL, "unknown ray kind bits");
    return launch_closest<false>(s, rays, max_depth, ray_kind, nullptr, n, hits, nullptr, nullptr, stream);
}

extern "C" int pvx_trace_closest_counted(pvx_scene *s, const pvx_ray *rays,
                                         const double *max_depth, uint32_t ray_kind, size_t n,
                                         pvx_hit *hits, uint32_t *node_visits,
                                         uint32_t *prim_tests, void *stream) {
    if (!node_visits || !prim_tests) return fail(PVX_EINVAL, "null counter buffer");
    if (ray_kind & ~0xFFu) return fail(PVX_EINVAL, "unknown ray kind bits");
    return launch_closest<true>(s, rays, max_depth, ray_kind, nullptr, n, hits, node_visits, prim_tests,
                                stream);
}

// ----------------------------------------------------------------------------
// scene batches
// ----------------------------------------------------------------------------
extern "C" int pvx_batch_create(pvx_scene *const *scenes, uint32_t n_scenes, pvx_batch **out) {
    if (!out) return fail(PVX_EINVAL, "null argument");
    *out = nullptr;
    if (!scenes || n_scenes == 0) return fail(PVX_EINVAL, "empty scene list");
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    std::vector<SceneView> views(n_scenes);
    uint32_t level = 0, flags = 0;
    for (uint32_t k = 0; k < n_scenes; k++) {
        const pvx_scene *s = scenes[k];
        if (!s) return fail(PVX_EINVAL, "null scene in the batch");
        if (s->device != dev) return fail(PVX_EINVAL, "scene " + std::to_string(k) + " lives on another device");
        views[k] = s->view;
        level = std::max(level, kernel_level(s->view));
        flags |= s->view.any_flags;
    }
    pvx_batch *b = new pvx_batch();
    b->device = dev;
    b->scenes.assign(scenes, scenes + n_scenes);
    b->level = level;
    b->any_flags = flags;
    int r;
    if ((r = b->views.reserve(views.size() * sizeof(SceneView))) != PVX_OK) {
        delete b;
        return r;
    }
    hipError_t e = hipMemcpy(b->views.p, views.data(), views.size() * sizeof(SceneView), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->uploaded, hipEventDisableTiming);
    if (e != hipSuccess) {
        pvx_batch_destroy(b);
        return fail(PVX_EDEVICE, std::string("batch set-up: ") + hipGetErrorString(e));
    }
    // persistent grids of the two kernel families the batch can launch
    int cus = 0, per_ww = 0, per = 0;
    const void *kfn = level >= 2   ? (flags ? (const void *)k_trace_closest_multi<2, true, false>
                                            : (const void *)k_trace_closest_multi<2, false, false>)
                      : level == 1 ? (flags ? (const void *)k_trace_closest_multi<1, true, false>
                                            : (const void *)k_trace_closest_multi<1, false, false>)
                                   : (flags ? (const void *)k_trace_closest_multi<0, true, false>
                                            : (const void *)k_trace_closest_multi<0, false, false>);
    const void *kww = flags ? (const void *)k_trace_closest_ww_multi<true, false>
                            : (const void *)k_trace_closest_ww_multi<false, false>;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kfn, kBlock, 0);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_ww, kww, kBlock, 0);
    if (e != hipSuccess) {
        pvx_batch_destroy(b);
        return fail(PVX_EDEVICE, std::string("batch occupancy: ") + hipGetErrorString(e));
    }
    b->grid = (uint32_t)(cus * std::max(per, 1));
    b->grid_ww = (uint32_t)(cus * std::max(per_ww, 1));
    const size_t spill_bytes = (size_t)std::max(b->grid, b->grid_ww) * kBlock * kStackSpill * sizeof(int32_t);
    if ((r = b->spill.reserve(spill_bytes)) != PVX_OK) {
        pvx_batch_destroy(b);
        return r;
    }
    *out = b;
    return PVX_OK;
}

extern "C" int pvx_batch_destroy(pvx_batch *b) {
    if (!b) return PVX_OK;
    (void)hipDeviceSynchronize();
    if (b->uploaded) (void)hipEventDestroy(b->uploaded);
    delete b;
    return PVX_OK;
}

// Chunk table for the ray ranges `first` (host, n_scenes + 1 entries); kept
// on the device and rebuilt only when the ranges change.
static int batch_chunks(pvx_batch *b, const uint64_t *first, hipStream_t st) {
    const size_t ns = b->scenes.size();
    if (!first) return fail(PVX_EINVAL, "null ray_first");
    if (first[0] != 0) return fail(PVX_EINVAL, "ray_first[0] must be 0");
    for (size_t k = 0; k < ns; k++)
        if (first[k + 1] < first[k]) return fail(PVX_EINVAL, "ray_first must not decrease");
    if (first[ns] > 0xFFFFFFFFull) return fail(PVX_EINVAL, "too many rays in one call");
    if (b->last_first.size() == ns + 1 && std::equal(b->last_first.begin(), b->last_first.end(), first))
        return PVX_OK;
    uint64_t nc = 0;
    for (size_t k = 0; k < ns; k++) nc += (first[k + 1] - first[k] + 63) / 64;
    if (nc * 64 > 0xFFFFFFFFull) return fail(PVX_EINVAL, "too many rays in one call");
    b->last_first.clear();
    b->n_chunks = 0;
    if (nc == 0) {
        b->last_first.assign(first, first + ns + 1);
        return PVX_OK;
    }
    // the pinned table may still be feeding the previous upload
    HIP_TRY(hipEventSynchronize(b->uploaded));
    int r;
    if ((r = b->hchunks.reserve(nc * sizeof(uint4))) != PVX_OK) return r;
    if ((r = b->chunks.reserve(nc * sizeof(uint4))) != PVX_OK) return r;
    uint4 *h = static_cast<uint4 *>(b->hchunks.p);
    uint64_t c = 0;
    for (size_t k = 0; k < ns; k++)
        for (uint64_t f = first[k]; f < first[k + 1]; f += 64)
            h[c++] = make_uint4((uint32_t)k, (uint32_t)f, (uint32_t)std::min<uint64_t>(64, first[k + 1] - f), 0u);
    HIP_TRY(hipMemcpyAsync(b->chunks.p, h, nc * sizeof(uint4), hipMemcpyHostToDevice, st));
    HIP_TRY(hipEventRecord(b->uploaded, st));
    b->n_chunks = (uint32_t)nc;
    b->last_first.assign(first, first + ns + 1);
    return PVX_OK;
}

template <bool kCount>
static int launch_batch(pvx_batch *b, const pvx_ray *rays, const uint64_t *first, const double *max_depth,
                        uint32_t ray_kind, pvx_hit *hits, uint64_t *clocks, uint32_t *nv, uint32_t *pt,
                        void *stream) {
    if (!b) return fail(PVX_EINVAL, "null batch");
    hipStream_t st = static_cast<hipStream_t>(stream);
    int r;
    if ((r = batch_chunks(b, first, st)) != PVX_OK) return r;
    const uint32_t ns = (uint32_t)b->scenes.size();
    auto *clk = reinterpret_cast<unsigned long long *>(clocks);
    if (clk) {
        k_clock_init<<<(ns + 1 + 255) / 256, 256, 0, st>>>(clk, ns + 1);
        HIP_TRY(hipGetLastError());
    }
    if (b->n_chunks == 0) return PVX_OK;
    if (!rays || !hits) return fail(PVX_EINVAL, "null ray or hit buffer");
    const uint32_t n = b->n_chunks * 64u;
    const bool flags = b->any_flags != 0 && ray_kind != PVX_RAY_PRIMARY;
    const bool ww = b->level == 0 && ray_kind != PVX_RAY_SHADOW;
    const uint32_t nblocks = std::min<uint32_t>(ww ? b->grid_ww : b->grid, (n + kBlock - 1) / kBlock);
    const SceneView *views = static_cast<const SceneView *>(b->views.p);
    const uint4 *chunks = static_cast<const uint4 *>(b->chunks.p);
    int32_t *spill = static_cast<int32_t *>(b->spill.p);
#define PVX_LAUNCH(L, F)                                                                                     \
    k_trace_closest_multi<L, F, kCount><<<nblocks, kBlock, 0, st>>>(views, chunks, clk, rays, max_depth,     \
                                                                    ray_kind, nullptr, n, hits, spill, nv, pt)
#define PVX_LAUNCH_WW(F)                                                                                     \
    k_trace_closest_ww_multi<F, kCount><<<nblocks, kBlock, 0, st>>>(views, chunks, clk, rays, max_depth,     \
                                                                    ray_kind, nullptr, n, hits, spill, nv, pt)
    if (b->level >= 2) {
        if (flags) PVX_LAUNCH(2, true); else PVX_LAUNCH(2, false);
    } else if (b->level == 1) {
        if (flags) PVX_LAUNCH(1, true); else PVX_LAUNCH(1, false);
    } else if (ww) {
        if (flags) PVX_LAUNCH_WW(true); else PVX_LAUNCH_WW(false);
    } else {
        if (flags) PVX_LAUNCH(0, true); else PVX_LAUNCH(0, false);
    }
#undef PVX_LAUNCH
#undef PVX_LAUNCH_WW
    HIP_TRY(hipGetLastError());
    return PVX_OK;
}

extern "C" int pvx_batch_trace_closest(pvx_batch *b, const pvx_ray *rays, const uint64_t *ray_first,
                                       const double *max_depth, uint32_t ray_kind, pvx_hit *hits,
                                       uint64_t *scene_clock, void *stream) {
    if (ray_kind & ~0xFFu) return fail(PVX_EINVAL, "unknown ray kind bits");
    return launch_batch<false>(b, rays, ray_first, max_depth, ray_kind, hits, scene_clock, nullptr, nullptr,
                               stream);
}

extern "C" int pvx_batch_trace_closest_counted(pvx_batch *b, const pvx_ray *rays, const uint64_t *ray_first,
                                               const double *max_depth, uint32_t ray_kind, pvx_hit *hits,
                                               uint32_t *node_visits, uint32_t *prim_tests,
                                               uint64_t *scene_clock, void *stream) {
    if (!node_visits || !prim_tests) return fail(PVX_EINVAL, "null counter buffer");
    if (ray_kind & ~0xFFu) return fail(PVX_EINVAL, "unknown ray kind bits");
    return launch_batch<true>(b, rays, ray_first, max_depth, ray_kind, hits, scene_clock, node_visits, prim_tests,
                              stream);
}

extern "C" int pvx_clock_rate(double *hz) {
    if (!hz) return fail(PVX_EINVAL, "null argument");
    int dev = 0, khz = 0;
    HIP_TRY(hipGetDevice(&dev));
    HIP_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    if (khz <= 0) return fail(PVX_EDEVICE, "device reports no wall-clock rate");
    *hz = 1000.0 * (double)khz;
    return PVX_OK;
}

extern "C" int pvx_all_intersections(pvx_scene *s, const pvx_ray *rays, const uint32_t *obj,
                                     size_t n, uint32_t max_hits, double *depths, int32_t *prims,
                                     uint32_t *counts, void *stream) {
    if (!s) return fail(PVX_EINVAL, "null scene");
    if (n == 0) return PVX_OK;
    if (!rays || !obj || !depths || !prims || !counts || max_hits == 0)
        return fail(PVX_EINVAL, "null buffer or max_hits == 0");
    if (n > 0xFFFFFFFFull) return fail(PVX_EINVAL, "too many rays in one call");
    hipStream_t st = static_cast<hipStream_t>(stream);
    uint32_t nb = (uint32_t)((n + 127) / 128);
    k_all_intersections<<<nb, 128, 0, st>>>(s->view, rays, obj, (uint32_t)n, max_hits, depths,
                                            prims, counts, s->d_obj_type, s->d_obj_index);
    HIP_TRY(hipGetLastError());
    return PVX_OK;
}

extern "C" int pvx_inside(pvx_scene *s, const double *points, const uint32_t *obj, size_t n,
                          uint8_t *out, void *stream) {
    if (!s) return fail(PVX_EINVAL, "null scene");
    if (n == 0) return PVX_OK;
    if (!points || !obj || !out) return fail(PVX_EINVAL, "null buffer");
    if (n > 0xFFFFFFFFull) return fail(PVX_EINVAL, "too many points in one call");
    hipStream_t st = static_cast<hipStream_t>(stream);
    uint32_t nb = (uint32_t)((n + 255) / 256);
    k_inside<<<nb, 256, 0, st>>>(s->view, points, obj, (uint32_t)n, out, s->d_obj_type,
                                 s->d_obj_index, s->view.obj_flags);
    HIP_TRY(hipGetLastError());
    return PVX_OK;
}

extern "C" int pvx_camera_rays(const pvx_camera *cam, pvx_ray *rays, void *stream) {
    if (!cam || !rays) return fail(PVX_EINVAL, "null argument");
    if (cam->width == 0 || cam->height == 0 || cam->height > 65535)
        return fail(PVX_EINVAL, "bad image size");
    hipStream_t st = static_cast<hipStream_t>(stream);
    dim3 grid((cam->width + 255) / 256, cam->height);
    k_camera_rays<<<grid, 256, 0, st>>>(*cam, rays);
    HIP_TRY(hipGetLastError());
    return PVX_OK;
}

extern "C" int pvx_copy(void *dst, const void *src, size_t nbytes, void *stream) {
    if (nbytes == 0) return PVX_OK;
    if (!dst || !src) return fail(PVX_EINVAL, "null pointer");
    if ((nbytes & 15) || ((uintptr_t)dst & 15) || ((uintptr_t)src & 15))
        return fail(PVX_EINVAL, "pvx_copy needs 16 B multiples and alignment");
    size_t n16 = nbytes / 16;
    uint32_t blocks = (uint32_t)std::min<size_t>((n16 + 255) / 256, 65536);
    k_copy16<<<blocks, 256, 0, static_cast<hipStream_t>(stream)>>>(
        static_cast<uint4 *>(dst), static_cast<const uint4 *>(src), n16);
    HIP_TRY(hipGetLastError());
    return PVX_OK;
}
