// Closest-hit traversal body shared by the traversal kernels
// (trace_kernels.hip): FP32 slab tests over the BVH, the reference's FP64
// decision for every candidate primitive, the object-flag precondition and
// the depth filters of Trace::FindIntersection (core/render/trace.cpp:285-344).
#ifndef PVX_TRACE_BODY_H
#define PVX_TRACE_BODY_H

#include <hip/hip_runtime.h>

#include "../../include/pvx.h"
#include "pvx_types.h"
#include "shapes_device.h"
#include "trace_launch.h"

namespace pvx {
namespace {


struct RayF {
    float ix, iy, iz;    // 1 / direction (FP32, zero components clamped)
    float oix, oiy, oiz; // origin * inverse direction
};

__device__ inline float safe_inv(double d) {
    float f = (float)d;
    if (fabsf(f) < 1.0e-20f) f = f < 0.0f ? -1.0e-20f : 1.0e-20f;
    return 1.0f / f;
}

// Mesh::Intersect for one mesh2 triangle: for a placed mesh the ray moves
// into the mesh's own space with its direction normalised (mesh.cpp:179-230)
// and the depth returns to world units divided by that length (test_hit,
// mesh.cpp:1218-1243).  kind_axis: kind | axis << 4 | (placement + 1) << 8.
__device__ __forceinline__ bool mesh_tri_hit(const SceneView &sv, const PrimRec &p, uint32_t ka, D3 o, D3 d,
                                             double &dep) {
    const uint32_t axis = (ka >> 4) & 15u, xp = ka >> 8;
    if (xp == 0) return tri_hit64(p.tri, axis, o, d, dep);
    const Xform &x = sv.xforms[xp - 1];
    const D3 P = xf_point(x.inv, o);
    D3 D = xf_dir(x.inv, d);
    const double len = sqrt(D.x * D.x + D.y * D.y + D.z * D.z);
    D = D3{D.x / len, D.y / len, D.z / len};
    double t;
    if (!tri_hit64(p.tri, axis, P, D, t)) return false;
    dep = t / len;
    return true;
}

// The walks of scenes without placed meshes (kLevel < 2: the scene
// preparation sends scenes with placed meshes to the level-2 walk) take the
// unplaced test only, which keeps their register budget.
template <int kLevel>
__device__ __forceinline__ bool mesh_tri_hit_l(const SceneView &sv, const PrimRec &p, uint32_t ka, D3 o, D3 d,
                                               double &dep) {
    if (kLevel >= 2) return mesh_tri_hit(sv, p, ka, o, d, dep);
    return tri_hit64(p.tri, (ka >> 4) & 15u, o, d, dep);
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
// kRev: lathes and sors may be operands.
template <bool kRev = false>
__device__ inline bool csg_closest(const SceneView &sv, uint32_t root, D3 o, D3 d, bool shadow,
                                   double limit, double &dep_out, int32_t &slot_out) {
    uint32_t b = sv.csg_leaf_range[2 * root], e = sv.csg_leaf_range[2 * root + 1];
    bool found = false;
    double best = limit;
    for (uint32_t k = b; k < e; k++) {
        uint32_t slot = sv.csg_leaves[k];
        const CsgChild &c = sv.csg_children[slot];
        auto consider = [&](double dep) {
            if (!depth_ok(dep, shadow)) return;
            if (!(dep < best || (found && dep == best && (int32_t)slot < slot_out))) return;
            D3 ip = add3(o, mul3(d, dep));
            if (csg_survives<kRev>(sv, slot, ip)) {
                best = dep;
                slot_out = (int32_t)slot;
                found = true;
            }
        };
        if constexpr (kRev) {
            leaf_visit<true>(sv, c.type, c.index, o, d, true, [&](double dep, int, double) { consider(dep); });
        } else {
            double hs[kMaxLeafHits];
            int nh = leaf_hits(sv, c.type, c.index, o, d, hs);
            for (int j = 0; j < nh; j++) consider(hs[j]);
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
// (Find_Intersection over its IStack); with kRev also lathes and sors.
template <bool kRev = false>
__device__ inline bool simple_closest(const SceneView &sv, uint32_t kind, uint32_t index, D3 o,
                                      D3 d, bool shadow, double &dep_out) {
    if constexpr (kRev) {
        if (kind == kPrimRevolution) {
            bool found = false;
            double best = 0.0;
            rev_visit(sv, index, o, d, false, [&](double dep, int, double) {
                if (depth_ok(dep, shadow) && (!found || dep < best)) {
                    best = dep;
                    found = true;
                }
            });
            dep_out = best;
            return found;
        }
    }
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

// kLevel 0: triangles and spheres; 1: also boxes; 2: also tori and CSG; 3: also
// lathes and sors.
// kWW: while-while traversal (triangle scenes, camera and bounce rays).
// kMulti: rays of many scenes in one launch.  `n` is then 64 x the number of
// chunks; chunk c = {scene, first ray, ray count <= 64} covers the wavefront
// whose loop index is 64c..64c+63, so the scene (and its SceneView, read from
// `views`) is uniform across the wavefront.  With `clocks`, lane 0 of each
// wavefront records the wall clock at its start (minimum into clocks[0]) and
// at the end of each chunk (maximum into clocks[1 + scene]).
// kGen: light rays made here (ray k from gen_groups[k / gen_S].ip to grid
// point k mod gen_S of its group's light, rays past that light's grid
// skipped; max depth the distance) whose visibility byte goes to gen_vis[k]
// (1: nothing closer than the light minus SHADOW_TOLERANCE blocks it)
// instead of a hit record.
// kBin: a level-0 scene (triangles, spheres, planes) walked through the
// binary tree instead of the 4-wide one, with only level-0 leaf tests.
template <int kLevel, bool kFlags, bool kCount, bool kWW, bool kMulti = false, bool kGen = false, bool kBin = false>
__device__ __forceinline__ void trace_closest_body(const SceneView &sv0, const SceneView *__restrict__ views,
                                                   const uint4 *__restrict__ chunks,
                                                   unsigned long long *__restrict__ clocks,
                                                   const pvx_ray *__restrict__ rays,
                                                   const double *__restrict__ max_depth, uint32_t kind,
                                                   const uint8_t *__restrict__ kinds, uint32_t n,
                                                   pvx_hit *__restrict__ hits, int32_t *__restrict__ spill,
                                                   uint32_t *__restrict__ node_visits,
                                                   uint32_t *__restrict__ prim_tests, int32_t *s_stack,
                                                   const LightGroup *__restrict__ gen_groups = nullptr,
                                                   const LightRec *__restrict__ gen_lights = nullptr,
                                                   uint32_t gen_S = 1, uint8_t *__restrict__ gen_vis = nullptr) {
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
        // per-ray kind (precondition only; shadow rays come in launches of their own)
        const uint32_t rk = (kFlags && kinds) ? (uint32_t)kinds[ri] : kind;
        D3 o, d;
        double bt = kBoundHuge;
        if constexpr (kGen) {
            const uint32_t g = ri / gen_S, sidx = ri - g * gen_S;
            const LightGroup grp = gen_groups[g];
            const LightRec &lt = gen_lights[grp.light];
            if (sidx >= lt.n1 * lt.n2) continue; // past this light's grid
            o = D3{grp.ip[0], grp.ip[1], grp.ip[2]};
            double md;
            light_sample_ray(lt, sidx / lt.n2, sidx % lt.n2, o, d, md);
            if (md >= kEpsilon) bt = md; // trace.cpp:157-158
        } else {
            const pvx_ray r = rays[ri];
            o = d3(r.o[0], r.o[1], r.o[2]);
            d = d3(r.d[0], r.d[1], r.d[2]);
            if (max_depth) {
                double md = max_depth[ri];
                if (md >= kEpsilon) bt = md; // trace.cpp:157-158
            }
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
                                              : csg_closest<(kLevel >= 3)>(sv, inf.index, o, d, shadow, bt, dep, prim);
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

        int32_t node = (occl && bo >= 0 && bt < occl_lim) ? kEmpty : (kLevel == 0 && !kBin) ? sv.root4 : sv.root;
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
            auto push = [&](int32_t x) {
                if (sp < kStackLds)
                    s_stack[sp * kBlock + lane] = x;
                else
                    my_spill[sp - kStackLds] = x;
                sp++;
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
                    if (mesh_tri_hit_l<kLevel>(sv, p, ka, o, d, dep) && depth_ok(dep, shadow) &&
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
        if constexpr (kLevel == 0 && !kBin) {
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
        } else if constexpr (kBin && kWW) {
            // triangle / sphere scene, binary tree, while-while: the lanes of
            // a wavefront step through inner nodes together, then leaves
            auto bin_node = [&]() {
                if (kCount) cnt_nodes++;
                const float4 *np = reinterpret_cast<const float4 *>(sv.nodes + node);
                const float4 a = np[0], b = np[1], c = np[2];
                const int4 l = reinterpret_cast<const int4 *>(np)[3];
                float x0 = fmaf(a.x, rf.ix, -rf.oix), x1 = fmaf(a.y, rf.ix, -rf.oix);
                float y0 = fmaf(a.z, rf.iy, -rf.oiy), y1 = fmaf(a.w, rf.iy, -rf.oiy);
                float z0 = fmaf(c.x, rf.iz, -rf.oiz), z1 = fmaf(c.y, rf.iz, -rf.oiz);
                const float tn0 = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), 0.0f));
                const float tf0 = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fminf(fmaxf(z0, z1), tb));
                x0 = fmaf(b.x, rf.ix, -rf.oix);
                x1 = fmaf(b.y, rf.ix, -rf.oix);
                y0 = fmaf(b.z, rf.iy, -rf.oiy);
                y1 = fmaf(b.w, rf.iy, -rf.oiy);
                z0 = fmaf(c.z, rf.iz, -rf.oiz);
                z1 = fmaf(c.w, rf.iz, -rf.oiz);
                const float tn1 = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), 0.0f));
                const float tf1 = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fminf(fmaxf(z0, z1), tb));
                const bool h0 = tn0 <= tf0 * 1.000001f;
                const bool h1 = tn1 <= tf1 * 1.000001f;
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
                } else if (h0) {
                    node = l.x;
                } else if (h1) {
                    node = l.y;
                } else {
                    PVX_POP();
                }
            };
            auto bin_leaf = [&]() {
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
                        if (mesh_tri_hit_l<0>(sv, p, ka, o, d, dep) && depth_ok(dep, shadow) &&
                            better(dep, obj, p.tri.prim, bt, bo, bp)) {
                            bt = dep;
                            bo = obj;
                            bp = p.tri.prim;
                        }
                    } else if (pk == kPrimSphere) {
                        double hs[2];
                        const int nh = sphere_hits(d3(p.sph.c[0], p.sph.c[1], p.sph.c[2]), p.sph.r, o, d, hs);
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
            while (node != kEmpty) {
                while (node >= 0) bin_node();
                while (node < 0 && node != kEmpty) bin_leaf();
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
                            if (mesh_tri_hit_l<kLevel>(sv, p, ka, o, d, dep) && depth_ok(dep, shadow) &&
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
                                ok = csg_closest<(kLevel >= 3)>(sv, p.ref.index, o, d, shadow, bt, dep, prim);
                            else
                                ok = simple_closest<(kLevel >= 3)>(sv, pk, p.ref.index, o, d, shadow, dep);
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
        if constexpr (kGen) // TracePointLightShadowRay with an opaque blocker (trace.cpp:2021-2073)
            gen_vis[ri] = (bo >= 0 && bt < occl_lim && bt > kShadowTolerance) ? 0 : 1;
        else
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

// Closest valid hit among one primitive record of any kind (lathes and sors
// included), folded into
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
        if (mesh_tri_hit(sv, p, ka, o, d, dep) && depth_ok(dep, shadow) && better(dep, obj, p.tri.prim, bt, bo, bp)) {
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
            ok = csg_closest<true>(sv, p.ref.index, o, d, shadow, bt, dep, prim);
        else
            ok = simple_closest<true>(sv, pk, p.ref.index, o, d, shadow, dep);
        if (ok && better(dep, obj, prim, bt, bo, bp)) {
            bt = dep;
            bo = obj;
            bp = prim;
        }
    }
}

} // namespace
} // namespace pvx

#endif
