// Traversal kernels: closest hit over the BVH (one-lane-per-ray persistent
// grids), over the kd-tree, and the per-object All_Intersections / Inside
// queries.  This file is compiled once per part (-DPVX_TRACE_PART=n, see the
// Makefile); each part instantiates a few of the kernel templates, and
// trace_launch() in pvx.hip calls the part that holds the kernel asked for.
#include <hip/hip_runtime.h>

#include "../../include/pvx.h"
#include "trace_body.h"
#include "trace_launch.h"

#ifndef PVX_TRACE_PART
#error "build with -DPVX_TRACE_PART=<0..16>"
#endif

namespace pvx {
namespace {

template <int kLevel, bool kFlags, bool kCount>
__global__ __launch_bounds__(kBlock) void k_trace_closest(TraceArgs a) {
    __shared__ int32_t s_stack[kStackLds * kBlock];
    trace_closest_body<kLevel, kFlags, kCount, false>(a.sv, nullptr, nullptr, nullptr, a.rays, a.max_depth, a.kind,
                                                      a.kinds, a.n, a.hits, a.spill, a.node_visits, a.prim_tests,
                                                      s_stack);
}

// Level-0 scenes (triangles, spheres, planes) through the binary tree.
// 92 VGPRs, 5 wavefronts per SIMD (the level-1 binary kernel, with box
// leaves, needs 108 and spills; forcing 6 here spills and is slower).
// kWW: while-while passes (inner nodes, then leaves, lane by lane).
template <bool kWW, bool kFlags, bool kCount>
__global__ __launch_bounds__(kBlock) void k_trace_closest_bin0(TraceArgs a) {
    __shared__ int32_t s_stack[kStackLds * kBlock];
    trace_closest_body<0, kFlags, kCount, kWW, false, false, true>(a.sv, nullptr, nullptr, nullptr, a.rays,
                                                                    a.max_depth, a.kind, a.kinds, a.n, a.hits,
                                                                    a.spill, a.node_visits, a.prim_tests, s_stack);
}

// Triangle scenes, camera and bounce rays: while-while traversal (78 VGPRs
// and a 24 KB LDS stack per block: 6 wavefronts per SIMD; forcing 7 or 8
// costs more in register pressure than it gains).
template <bool kFlags, bool kCount>
__global__ __launch_bounds__(kBlock) void k_trace_closest_ww(TraceArgs a) {
    __shared__ int32_t s_stack[kStackLds * kBlock];
    trace_closest_body<0, kFlags, kCount, true>(a.sv, nullptr, nullptr, nullptr, a.rays, a.max_depth, a.kind,
                                                a.kinds, a.n, a.hits, a.spill, a.node_visits, a.prim_tests, s_stack);
}

// Light rays of pvx_render made in the kernel (triangle / sphere scenes,
// every material opaque): occlusion test, while-while walk, a visibility
// byte per ray.  5 wavefronts per SIMD: held to 80 VGPRs for 6 it was
// slower (18.7 against 18.3 ms of light passes on the 1.3M-triangle frame).
template <bool kFlags>
__global__ __launch_bounds__(kBlock, 5) void k_light_ww(TraceArgs a) {
    __shared__ int32_t s_stack[kStackLds * kBlock];
    const uint32_t n = a.gen_count ? a.gen_count[0] * a.gen_S : a.n;
    trace_closest_body<0, kFlags, false, true, false, true>(a.sv, nullptr, nullptr, nullptr, nullptr, nullptr,
                                                            PVX_RAY_SHADOW | kRayOcclusion, nullptr, n, nullptr,
                                                            a.spill, nullptr, nullptr, s_stack, a.gen_groups,
                                                            a.gen_lights, a.gen_S, a.gen_vis);
}

// The same light rays through the binary tree (level-0 leaf tests only).
template <bool kFlags>
__global__ __launch_bounds__(kBlock) void k_light_bin(TraceArgs a) {
    __shared__ int32_t s_stack[kStackLds * kBlock];
    const uint32_t n = a.gen_count ? a.gen_count[0] * a.gen_S : a.n;
    trace_closest_body<0, kFlags, false, false, false, true, true>(
        a.sv, nullptr, nullptr, nullptr, nullptr, nullptr, PVX_RAY_SHADOW | kRayOcclusion, nullptr, n, nullptr, a.spill,
        nullptr, nullptr, s_stack, a.gen_groups, a.gen_lights, a.gen_S, a.gen_vis);
}

// Light rays made in the kernel with refill: once no more than
// a.gen_refill_below lanes of a wavefront still hold rays, its free lanes
// take the next unstarted rays (one atomic per wavefront on a launch-wide
// counter), instead of waiting for the wavefront's slowest ray (Aila and
// Laine, "Understanding the efficiency of ray traversal on GPUs", HPG 2009:
// persistent threads with fetching and ray replacement).  Same rays, same
// tests and same visibility bytes as k_light_ww.  5 wavefronts per SIMD
// (92-96 VGPRs, no spills; held to 80 for 6 it spills and the camera-hit
// light rays of the 1.3M-triangle frame take 12.2 instead of 11.3 ms).
template <bool kFlags, bool kBin>
__global__ __launch_bounds__(kBlock, 5) void k_light_dyn(TraceArgs a) {
    __shared__ int32_t s_stack[kStackLds * kBlock];
    const SceneView &sv = a.sv;
    const uint32_t n = a.gen_count ? a.gen_count[0] * a.gen_S : a.n;
    const uint32_t lane = threadIdx.x;
    const uint32_t gtid = blockIdx.x * kBlock + lane;
    int32_t *my_spill = a.spill + (size_t)gtid * kStackSpill;
    const uint32_t wl = lane & 63u;
    bool have = false, done = false;
    uint32_t ri = 0;
    D3 o{0, 0, 0}, d{0, 0, 1};
    double bt = 0.0, occl_lim = 0.0;
    int32_t bo = -1, bp = 0;
    RayF rf{};
    float tb = 0.0f;
    int32_t node = kEmpty;
    int sp = 0;
    auto push = [&](int32_t x) {
        if (sp < kStackLds)
            s_stack[sp * kBlock + lane] = x;
        else
            my_spill[sp - kStackLds] = x;
        sp++;
    };
    auto pop = [&]() {
        if (sp == 0) {
            node = kEmpty;
        } else {
            sp--;
            node = sp < kStackLds ? s_stack[sp * kBlock + lane] : my_spill[sp - kStackLds];
        }
    };
    while (true) {
        // hand out new rays to the lanes that have none
        const uint64_t need = __ballot(!have && !done);
        if (need && (uint32_t)__popcll(__ballot(have)) <= a.gen_refill_below) {
            const int leader = __ffsll((unsigned long long)need) - 1;
            uint32_t base = 0;
            if ((int)wl == leader) base = atomicAdd(a.gen_next, (uint32_t)__popcll(need));
            base = (uint32_t)__shfl((int)base, leader);
            if (!have && !done) {
                // rank among the lanes that asked: those below this one in `need`
                ri = base + __builtin_amdgcn_mbcnt_hi((uint32_t)(need >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)need, 0u));
                if (ri >= n) {
                    done = true;
                } else {
                    const uint32_t g = ri / a.gen_S, sidx = ri - g * a.gen_S;
                    const LightGroup grp = a.gen_groups[g];
                    const LightRec &lt = a.gen_lights[grp.light];
                    if (sidx < lt.n1 * lt.n2) { // else: past this light's grid, take another
                        o = D3{grp.ip[0], grp.ip[1], grp.ip[2]};
                        double md;
                        light_sample_ray(lt, sidx / lt.n2, sidx % lt.n2, o, d, md);
                        bt = md >= kEpsilon ? md : kBoundHuge; // trace.cpp:157-158
                        occl_lim = bt - kShadowTolerance;
                        bo = -1;
                        bp = 0;
                        for (uint32_t q = 0; q < sv.n_inf; q++) { // planes
                            const InfRec inf = sv.inf[q];
                            if (kFlags && !precondition(sv.obj_flags[inf.obj], PVX_RAY_SHADOW)) continue;
                            double dep;
                            if (inf.kind == kPrimPlane && plane_hit(sv.planes[inf.index], o, d, dep) &&
                                depth_ok(dep, true) && better(dep, inf.obj, 0, bt, bo, bp)) {
                                bt = dep;
                                bo = inf.obj;
                                bp = 0;
                            }
                        }
                        rf.ix = safe_inv(d.x);
                        rf.iy = safe_inv(d.y);
                        rf.iz = safe_inv(d.z);
                        rf.oix = (float)o.x * rf.ix;
                        rf.oiy = (float)o.y * rf.iy;
                        rf.oiz = (float)o.z * rf.iz;
                        tb = (float)bt * 1.000001f;
                        node = (bo >= 0 && bt < occl_lim) ? kEmpty : kBin ? sv.root : sv.root4;
                        sp = 0;
                        have = true;
                    }
                }
            }
        }
        if (!__ballot(have)) {
            if (!__ballot(!done)) break; // every lane out of rays
            continue;
        }
        if (have) {
            // one while-while round: inner nodes, then leaves
            while (kBin && node >= 0) { // binary tree: near child first, the far one pushed
                const float4 *np = reinterpret_cast<const float4 *>(sv.nodes + node);
                const float4 na = np[0], nb = np[1], nc = np[2];
                const int4 nl = reinterpret_cast<const int4 *>(np)[3];
                float x0 = fmaf(na.x, rf.ix, -rf.oix), x1 = fmaf(na.y, rf.ix, -rf.oix);
                float y0 = fmaf(na.z, rf.iy, -rf.oiy), y1 = fmaf(na.w, rf.iy, -rf.oiy);
                float z0 = fmaf(nc.x, rf.iz, -rf.oiz), z1 = fmaf(nc.y, rf.iz, -rf.oiz);
                const float tn0 = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), 0.0f));
                const float tf0 = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fminf(fmaxf(z0, z1), tb));
                x0 = fmaf(nb.x, rf.ix, -rf.oix);
                x1 = fmaf(nb.y, rf.ix, -rf.oix);
                y0 = fmaf(nb.z, rf.iy, -rf.oiy);
                y1 = fmaf(nb.w, rf.iy, -rf.oiy);
                z0 = fmaf(nc.z, rf.iz, -rf.oiz);
                z1 = fmaf(nc.w, rf.iz, -rf.oiz);
                const float tn1 = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), 0.0f));
                const float tf1 = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fminf(fmaxf(z0, z1), tb));
                const bool h0 = tn0 <= tf0 * 1.000001f, h1 = tn1 <= tf1 * 1.000001f;
                if (h0 && h1) {
                    const bool swap = tn1 < tn0;
                    push(swap ? nl.x : nl.y);
                    node = swap ? nl.y : nl.x;
                } else if (h0) {
                    node = nl.x;
                } else if (h1) {
                    node = nl.y;
                } else {
                    pop();
                }
            }
            while (!kBin && node >= 0) {
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
                    pop();
                    continue;
                }
                if (d3 != INFINITY) push(l3);
                if (d2 != INFINITY) push(l2);
                if (d1 != INFINITY) push(l1);
                node = l0;
            }
            while (node < 0 && node != kEmpty) {
                const uint32_t leaf = (uint32_t)(~node);
                const uint32_t first = leaf >> 3, count = (leaf & 7u) + 1u;
                for (uint32_t k = 0; k < count; k++) {
                    const PrimRec &p = sv.prims[first + k];
                    const uint32_t ka = p.kind_axis;
                    const uint32_t pk = ka & 15u;
                    const int32_t obj = p.obj;
                    if (kFlags && !precondition(sv.obj_flags[obj], PVX_RAY_SHADOW)) continue;
                    if (pk == kPrimTriangle) {
                        double dep;
                        if (mesh_tri_hit_l<0>(sv, p, ka, o, d, dep) && depth_ok(dep, true) &&
                            better(dep, obj, p.tri.prim, bt, bo, bp)) {
                            bt = dep;
                            bo = obj;
                            bp = p.tri.prim;
                        }
                    } else if (pk == kPrimSphere) {
                        double hs[2];
                        const int nh = sphere_hits(d3(p.sph.c[0], p.sph.c[1], p.sph.c[2]), p.sph.r, o, d, hs);
#pragma unroll
                        for (int j = 0; j < 2; j++) // unrolled: hs stays in registers
                            if (j < nh && depth_ok(hs[j], true) && better(hs[j], obj, 0, bt, bo, bp)) {
                                bt = hs[j];
                                bo = obj;
                                bp = 0;
                            }
                    }
                }
                tb = (float)bt * 1.000001f;
                if (bo >= 0 && bt < occl_lim) {
                    node = kEmpty; // blocked: the occlusion test is decided
                    break;
                }
                pop();
            }
            if (node == kEmpty) { // TracePointLightShadowRay with an opaque blocker (trace.cpp:2021-2073)
                a.gen_vis[ri] = (bo >= 0 && bt < occl_lim && bt > kShadowTolerance) ? 0 : 1;
                have = false;
            }
        }
    }
}

// The same kernels over a batch of scenes (chunk table and SceneView table).
template <int kLevel, bool kFlags, bool kCount>
__global__ __launch_bounds__(kBlock) void k_trace_closest_multi(TraceArgs a) {
    __shared__ int32_t s_stack[kStackLds * kBlock];
    trace_closest_body<kLevel, kFlags, kCount, false, true>(a.views[0], a.views, a.chunks, a.clocks, a.rays,
                                                            a.max_depth, a.kind, a.kinds, a.n, a.hits, a.spill,
                                                            a.node_visits, a.prim_tests, s_stack);
}

template <bool kFlags, bool kCount>
__global__ __launch_bounds__(kBlock) void k_trace_closest_ww_multi(TraceArgs a) {
    __shared__ int32_t s_stack[kStackLds * kBlock];
    trace_closest_body<0, kFlags, kCount, true, true>(a.views[0], a.views, a.chunks, a.clocks, a.rays, a.max_depth,
                                                      a.kind, a.kinds, a.n, a.hits, a.spill, a.node_visits,
                                                      a.prim_tests, s_stack);
}

// Closest hit through the kd-tree (BSPTree::operator()(ray),
// core/bounding/bsptree.cpp:128-226): the ray's range inside the tree's box,
// the near child first with the far one remembered with its [entry, exit)
// range, every primitive of a reached leaf tested, and a stop once the next
// range starts beyond the closest hit.  Each primitive is decided as in
// k_trace_closest; one listed in several leaves is tested again (the
// reference's mailbox only saves that work).
template <bool kFlags, bool kCount>
__global__ __launch_bounds__(kBlock) void k_trace_closest_kd(TraceArgs a) {
    const SceneView &sv = a.sv;
    const KdView &kd = a.kd;
    const pvx_ray *__restrict__ rays = a.rays;
    const double *__restrict__ max_depth = a.max_depth;
    uint32_t kind = a.kind;
    const uint8_t *__restrict__ kinds = a.kinds;
    const uint32_t n = a.n;
    pvx_hit *__restrict__ hits = a.hits;
    uint32_t *__restrict__ node_visits = a.node_visits;
    uint32_t *__restrict__ prim_tests = a.prim_tests;
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
                                                 : csg_closest<true>(sv, inf.index, o, d, shadow, bt, dep, prim);
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

#if PVX_TRACE_PART == 12
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
            leaf_visit<true>(sv, c.type, c.index, o, d, true, [&](double dep, int, double) {
                if (csg_survives<true>(sv, slot, add3(o, mul3(d, dep)))) push(dep, (int32_t)slot);
            });
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
                    if (mesh_tri_hit(sv, p, p.kind_axis, o, d, dep)) push(dep, p.tri.prim);
                }
            }
            node = sp == 0 ? kEmpty : stack[--sp];
        }
    } else {
        leaf_visit<true>(sv, type, index, o, d, false, [&](double dep, int, double) { push(dep, 0); });
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
    out[i] = shape_inside<kCsgMaxDepth, true>(sv, obj_type[ob], obj_index[ob], obj_flags[ob], p) ? 1 : 0;
}
#endif

// One kernel: its occupancy, or a launch of nb blocks.
template <class F>
hipError_t run(F *kfn, const TraceArgs &a, uint32_t nb, hipStream_t st, int *occ) {
    if (occ) return hipOccupancyMaxActiveBlocksPerMultiprocessor(occ, reinterpret_cast<const void *>(kfn), kBlock, 0);
    hipLaunchKernelGGL(kfn, dim3(nb), dim3(kBlock), 0, st, a);
    return hipGetLastError();
}

} // namespace

#define PVX_FN2(n) trace_part##n
#define PVX_FN(n) PVX_FN2(n)
// flags x count selection of one template
#define PVX_FC(K, ...)                                                                                         \
    (k.flags ? (k.count ? run(K<__VA_ARGS__ true, true>, a, nb, st, occ) : run(K<__VA_ARGS__ true, false>, a, nb, st, occ)) \
             : (k.count ? run(K<__VA_ARGS__ false, true>, a, nb, st, occ) : run(K<__VA_ARGS__ false, false>, a, nb, st, occ)))

#if PVX_TRACE_PART == 12
hipError_t launch_all_intersections(const SceneView &sv, const pvx_ray *rays, const uint32_t *objs, uint32_t n,
                                    uint32_t max_hits, double *depths, int32_t *prims, uint32_t *counts,
                                    const uint32_t *obj_type, const uint32_t *obj_index, hipStream_t st) {
    const uint32_t nb = (n + 127) / 128;
    k_all_intersections<<<nb, 128, 0, st>>>(sv, rays, objs, n, max_hits, depths, prims, counts, obj_type, obj_index);
    return hipGetLastError();
}

hipError_t launch_inside(const SceneView &sv, const double *pts, const uint32_t *objs, uint32_t n, uint8_t *out,
                         const uint32_t *obj_type, const uint32_t *obj_index, const uint32_t *obj_flags,
                         hipStream_t st) {
    const uint32_t nb = (n + 255) / 256;
    k_inside<<<nb, 256, 0, st>>>(sv, pts, objs, n, out, obj_type, obj_index, obj_flags);
    return hipGetLastError();
}
#else
hipError_t PVX_FN(PVX_TRACE_PART)(const TraceKernel &k, const TraceArgs &a, uint32_t nb, hipStream_t st, int *occ) {
#if PVX_TRACE_PART == 0 // single scene, 4-wide walks
    return k.walk == kWalkWW4 ? PVX_FC(k_trace_closest_ww, ) : PVX_FC(k_trace_closest, 0, );
#elif PVX_TRACE_PART == 1 // single scene, binary walk
    return PVX_FC(k_trace_closest, 1, );
#elif PVX_TRACE_PART >= 2 && PVX_TRACE_PART <= 5 // single scene, every primitive kind
    constexpr bool kF = (PVX_TRACE_PART - 2) >= 2, kC = (PVX_TRACE_PART - 2) & 1;
    (void)k;
    return run(k_trace_closest<2, kF, kC>, a, nb, st, occ);
#elif PVX_TRACE_PART == 6 // kd-tree
    return PVX_FC(k_trace_closest_kd, );
#elif PVX_TRACE_PART == 7 // scene batches: 4-wide and binary walks
    return k.walk == kWalkWW4    ? PVX_FC(k_trace_closest_ww_multi, )
           : k.walk == kWalkStep4 ? PVX_FC(k_trace_closest_multi, 0, )
                                  : PVX_FC(k_trace_closest_multi, 1, );
#elif PVX_TRACE_PART >= 8 && PVX_TRACE_PART <= 11 // scene batches, every primitive kind
    constexpr bool kF = (PVX_TRACE_PART - 8) >= 2, kC = (PVX_TRACE_PART - 8) & 1;
    (void)k;
    return run(k_trace_closest_multi<2, kF, kC>, a, nb, st, occ);
#elif PVX_TRACE_PART == 13 // single scene with lathes or sors
    return PVX_FC(k_trace_closest, 3, );
#elif PVX_TRACE_PART == 14 // scene batches with lathes or sors
    return PVX_FC(k_trace_closest_multi, 3, );
#elif PVX_TRACE_PART == 16 // single scene, level-0 primitives on the binary tree
    return k.walk == kWalkBinary0WW ? PVX_FC(k_trace_closest_bin0, true, ) : PVX_FC(k_trace_closest_bin0, false, );
#elif PVX_TRACE_PART == 15 // light rays made in the kernel
    if (k.walk == kWalkLightDyn)
        return k.flags ? run(k_light_dyn<true, false>, a, nb, st, occ) : run(k_light_dyn<false, false>, a, nb, st, occ);
    if (k.walk == kWalkLightDynBin)
        return k.flags ? run(k_light_dyn<true, true>, a, nb, st, occ) : run(k_light_dyn<false, true>, a, nb, st, occ);
    if (k.walk == kWalkLightBin)
        return k.flags ? run(k_light_bin<true>, a, nb, st, occ) : run(k_light_bin<false>, a, nb, st, occ);
    return k.flags ? run(k_light_ww<true>, a, nb, st, occ) : run(k_light_ww<false>, a, nb, st, occ);
#else
#error "PVX_TRACE_PART out of range"
#endif
}
#endif

} // namespace pvx
