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
#include "trace_launch.h"

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

__global__ void k_clock_init(unsigned long long *clocks, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) clocks[i] = i == 0 ? ~0ull : 0ull;
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

// The same rays in tile order (tile_pixel), with each ray's pixel y * W + x.
__global__ void k_camera_rays_tiled(pvx_camera cam, uint32_t tw, uint32_t th, pvx_ray *rays, uint32_t *pixel) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cam.width * cam.height) return;
    uint32_t x, y;
    tile_pixel(cam.width, cam.height, tw, th, t, x, y);
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
    rays[t] = r;
    if (pixel) pixel[t] = y * cam.width + x;
}

__global__ void k_copy16(uint4 *__restrict__ dst, const uint4 *__restrict__ src, size_t n16) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16;
         i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

} // namespace

// The part of trace_kernels.hip that holds kernel `k`.
hipError_t pvx::trace_launch(const TraceKernel &k, const TraceArgs &a, uint32_t nb, hipStream_t st, int *occ) {
    if (!k.multi) {
        switch (k.walk) {
        case kWalkStep4:
        case kWalkWW4: return trace_part0(k, a, nb, st, occ);
        case kWalkBinary: return trace_part1(k, a, nb, st, occ);
        case kWalkBinary0:
        case kWalkBinary0WW: return trace_part16(k, a, nb, st, occ);
        case kWalkComplex:
            return (k.flags ? (k.count ? trace_part5 : trace_part4) : (k.count ? trace_part3 : trace_part2))(k, a, nb, st,
                                                                                                          occ);
        case kWalkKd: return trace_part6(k, a, nb, st, occ);
        case kWalkRevolution: return trace_part13(k, a, nb, st, occ);
        case kWalkLightWW:
        case kWalkLightDyn:
        case kWalkLightBin:
        case kWalkLightDynBin: return trace_part15(k, a, nb, st, occ);
        default: return hipErrorInvalidValue;
        }
    }
    switch (k.walk) {
    case kWalkStep4:
    case kWalkWW4:
    case kWalkBinary: return trace_part7(k, a, nb, st, occ);
    case kWalkRevolution: return trace_part14(k, a, nb, st, occ);
    case kWalkComplex:
        return (k.flags ? (k.count ? trace_part11 : trace_part10) : (k.count ? trace_part9 : trace_part8))(k, a, nb, st,
                                                                                                        occ);
    default: return hipErrorInvalidValue;
    }
}


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

// The no-allocation closest-hit entry point keeps its traversal-stack spill
// area in this fixed device array (one copy per device).
namespace {

constexpr size_t kNoAllocScratchBytes = (size_t)32 << 20; // 32 MiB per device
constexpr uint32_t kNoAllocMaxBlocks = (uint32_t)(kNoAllocScratchBytes / (kBlock * kSpillPerThread));
constexpr int kMaxDevices = 64;

static __device__ __align__(256) int32_t scratch[kNoAllocScratchBytes / sizeof(int32_t)];

std::mutex g_scratch_mu;
int32_t *g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device (each device holds
// its own copy), looked up once per device.
int scratch_address(int32_t **out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(PVX_EDEVICE, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(PVX_EDEVICE, "hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_ptr[dev] = static_cast<int32_t *>(p);
    }
    *out = g_scratch_ptr[dev];
    return PVX_OK;
}

} // namespace

extern "C" int pvx_init(int device) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) return fail(PVX_EDEVICE, "no HIP device available");
    if (device < 0 || device >= n) return fail(PVX_EINVAL, "device index out of range");
    HIP_TRY(hipSetDevice(device));
    int32_t *p = nullptr;
    return scratch_address(&p); // the no-allocation entry point's array on this device
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
    UP(h.obj_mesh_xf, obj_mesh_xf);
    UP(h.mesh_vtx, mesh_vtx);
    UP(h.mesh_nrm, mesh_nrm);
    UP(h.triangles, triangles);
    UP(h.revs, revs);
    UP(h.rev_segs, rev_segs);
    UP(h.rev_radius, rev_radius);
    UP(h.rev_height, rev_height);
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
    {
        const uint32_t level = kernel_level(v);
        const TraceKernel k{level >= 3   ? kWalkRevolution
                            : level == 2 ? kWalkComplex
                            : level == 1 ? kWalkBinary
                                         : kWalkWW4,
                            h.any_flags != 0, false, false};
        TraceArgs none{};
        HIP_TRY(trace_launch(k, none, 0, nullptr, &per_cu));
    }
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
                          hipStream_t st, uint32_t variant, uint32_t grid, int32_t *spill = nullptr) {
    int r;
    if (!spill && (r = scene_spill(s, st, &spill))) return r;
    uint32_t nblocks = (uint32_t)std::min<size_t>(grid, (n + kBlock - 1) / kBlock);
    const bool flags = s->view.any_flags != 0 && (ray_kind != PVX_RAY_PRIMARY || kinds);
    TraceKernel k{kWalkStep4, flags, kCount, false};
    const uint32_t level = kernel_level(s->view);
    if (s->use_kd)
        k.walk = kWalkKd;
    else if (level >= 3)
        k.walk = kWalkRevolution;
    else if (level == 2)
        k.walk = kWalkComplex;
    else if (level == 0 && variant == PVX_TRAVERSAL_BINARY_WW)
        k.walk = kWalkBinary0WW;
    else if (level == 1 || variant == PVX_TRAVERSAL_BINARY)
        k.walk = level == 0 ? kWalkBinary0 : kWalkBinary;
    else if (variant == PVX_TRAVERSAL_WIDE_WW ||
             (variant == PVX_TRAVERSAL_DEFAULT && (ray_kind & 0xFFu) != PVX_RAY_SHADOW))
        k.walk = kWalkWW4;
    TraceArgs ta{};
    ta.sv = s->view;
    ta.kd = s->kd;
    ta.rays = rays;
    ta.max_depth = max_depth;
    ta.kind = ray_kind;
    ta.kinds = kinds;
    ta.n = (uint32_t)n;
    ta.hits = hits;
    ta.spill = spill;
    ta.node_visits = nv;
    ta.prim_tests = pt;
    HIP_TRY(trace_launch(k, ta, nblocks, st));
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
    uint32_t variant = s->tune_variant[cls];
    if (is_light_walk(variant)) variant = PVX_TRAVERSAL_DEFAULT; // render-only light walks
    return launch_variant<kCount>(s, rays, max_depth, ray_kind, kinds, n, hits, nv, pt,
                                  static_cast<hipStream_t>(stream), variant, grid);
}

// Variants a scene admits: the kd-tree and tori/CSG scenes have one walk;
// box/triangle scenes the binary tree; triangle/sphere scenes all three.
static std::vector<uint32_t> scene_variants(const pvx_scene *s) {
    const uint32_t level = kernel_level(s->view);
    if (s->use_kd || level >= 2) return {PVX_TRAVERSAL_DEFAULT};
    if (level == 1) return {PVX_TRAVERSAL_BINARY};
    return {PVX_TRAVERSAL_WIDE_WW, PVX_TRAVERSAL_WIDE_STEP, PVX_TRAVERSAL_BINARY, PVX_TRAVERSAL_BINARY_WW};
}

extern "C" int pvx_scene_set_traversal(pvx_scene *s, uint32_t ray_kind, uint32_t variant, uint32_t grid_blocks) {
    if (!s) return fail(PVX_EINVAL, "null scene");
    if (grid_blocks > 2 * s->grid)
        return fail(PVX_EINVAL, "grid_blocks above twice the resident blocks (" + std::to_string(2 * s->grid) + ")");
    const int cls = (ray_kind & 0xFFu) == PVX_RAY_SHADOW ? 1 : 0;
    if (is_light_walk(variant)) {
        // pvx_render's light rays made in the kernel: triangle / sphere scenes
        if (cls != 1 || s->use_kd || kernel_level(s->view) != 0 || !s->view.wide)
            return fail(PVX_EINVAL,
                        "light-ray walks PACKET / FUSED / REFILL / LIGHT_BINARY / REFILL_BINARY need light rays of a "
                        "triangle / sphere scene");
    } else if (variant != PVX_TRAVERSAL_DEFAULT) {
        auto v = scene_variants(s);
        if (std::find(v.begin(), v.end(), variant) == v.end())
            return fail(PVX_EINVAL, "traversal variant " + std::to_string(variant) + " not available for this scene");
    }
    s->tune_variant[cls] = variant;
    s->tune_grid[cls] = grid_blocks;
    if (cls == 1) // pvx_render's light rays of every bounce
        for (int k = 0; k < 2; k++) {
            s->light_variant[k] = variant;
            s->light_grid[k] = grid_blocks;
        }
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

int pvx::light_rays_internal(pvx_scene *s, const LightGroup *groups, const LightRec *lights, uint32_t S,
                             uint32_t n_rays, uint8_t *vis, uint32_t *next, uint32_t grid_blocks,
                             const uint32_t *count, bool binary, hipStream_t st) {
    if (n_rays == 0) return PVX_OK;
    int r;
    int32_t *spill = nullptr;
    if ((r = scene_spill(s, st, &spill))) return r;
    // by default the fixed-stride walk on twice the resident blocks (the
    // light rays of one shaded point differ little in length, the tail of a
    // launch more: 19.3 ms against 21.8 ms on the resident grid for the
    // 1.3M-triangle frame); the refilling walk on the resident blocks, which
    // have no tail to fill
    const uint32_t grid = grid_blocks ? grid_blocks : next ? s->grid : 2 * s->grid;
    const uint32_t nb = (uint32_t)std::min<size_t>(grid, ((size_t)n_rays + kBlock - 1) / kBlock);
    TraceArgs ta{};
    ta.sv = s->view;
    ta.n = n_rays;
    ta.spill = spill;
    ta.gen_groups = groups;
    ta.gen_lights = lights;
    ta.gen_S = S;
    ta.gen_vis = vis;
    ta.gen_next = next;
    ta.gen_count = count;
    // refill a wavefront once no more than 16 of its lanes hold rays: new
    // rays then come in runs of at least 48 consecutive ones (rays of a few
    // shaded points), which keeps the lanes' walks alike; refilling every
    // free lane at once mixes rays of distant points and is slower (the
    // 1.3M-triangle frame: 27.5 ms against 18.0 ms at 16 and 18.5 ms at 0)
    ta.gen_refill_below = 16;
    HIP_TRY(trace_launch(TraceKernel{next ? (binary ? kWalkLightDynBin : kWalkLightDyn) : binary ? kWalkLightBin : kWalkLightWW,
                                     s->view.any_flags != 0, false, false},
                         ta, nb, st));
    return PVX_OK;
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
// closest hit without allocation: the traversal-stack spill area is a fixed
// device array instead of the per-stream area allocated on first use
// ----------------------------------------------------------------------------
extern "C" int pvx_trace_closest_noalloc(pvx_scene *s, const pvx_ray *rays, const double *max_depth,
                                         uint32_t ray_kind, size_t n, pvx_hit *hits, void *stream) {
    if (!s) return fail(PVX_EINVAL, "null scene");
    if (ray_kind & ~0xFFu) return fail(PVX_EINVAL, "unknown ray kind bits");
    if (n == 0) return PVX_OK;
    if (!rays || !hits) return fail(PVX_EINVAL, "null ray or hit buffer");
    if (n > 0xFFFFFFFFull) return fail(PVX_EINVAL, "too many rays in one call (at most 2^32 - 1)");
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev != s->device) return fail(PVX_EINVAL, "the scene lives on another device than the current one");
    int32_t *spill = nullptr;
    int r;
    if ((r = scratch_address(&spill))) return r;
    const int cls = (ray_kind & 0xFFu) == PVX_RAY_SHADOW ? 1 : 0;
    uint32_t grid = s->tune_grid[cls] ? s->tune_grid[cls] : s->grid;
    grid = std::min(grid, kNoAllocMaxBlocks); // the persistent grid is what the scratch array holds
    return launch_variant<false>(s, rays, max_depth, ray_kind, nullptr, n, hits, nullptr, nullptr,
                                 static_cast<hipStream_t>(stream), s->tune_variant[cls], grid, spill);
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
    const TraceKernel kb{level >= 3   ? kWalkRevolution
                         : level == 2 ? kWalkComplex
                         : level == 1 ? kWalkBinary
                                      : kWalkStep4,
                         flags != 0, false, true};
    const TraceKernel kw{kWalkWW4, flags != 0, false, true};
    TraceArgs none{};
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess) e = trace_launch(kb, none, 0, nullptr, &per);
    if (e == hipSuccess) e = trace_launch(kw, none, 0, nullptr, &per_ww);
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
    const TraceKernel k{b->level >= 3   ? kWalkRevolution
                        : b->level == 2 ? kWalkComplex
                        : b->level == 1 ? kWalkBinary
                        : ww            ? kWalkWW4
                                        : kWalkStep4,
                        flags, kCount, true};
    TraceArgs ta{};
    ta.views = views;
    ta.chunks = chunks;
    ta.clocks = clk;
    ta.rays = rays;
    ta.max_depth = max_depth;
    ta.kind = ray_kind;
    ta.n = n;
    ta.hits = hits;
    ta.spill = spill;
    ta.node_visits = nv;
    ta.prim_tests = pt;
    HIP_TRY(trace_launch(k, ta, nblocks, st));
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
    HIP_TRY(launch_all_intersections(s->view, rays, obj, (uint32_t)n, max_hits, depths, prims, counts,
                                     s->d_obj_type, s->d_obj_index, st));
    return PVX_OK;
}

extern "C" int pvx_inside(pvx_scene *s, const double *points, const uint32_t *obj, size_t n,
                          uint8_t *out, void *stream) {
    if (!s) return fail(PVX_EINVAL, "null scene");
    if (n == 0) return PVX_OK;
    if (!points || !obj || !out) return fail(PVX_EINVAL, "null buffer");
    if (n > 0xFFFFFFFFull) return fail(PVX_EINVAL, "too many points in one call");
    hipStream_t st = static_cast<hipStream_t>(stream);
    HIP_TRY(launch_inside(s->view, points, obj, (uint32_t)n, out, s->d_obj_type, s->d_obj_index,
                          s->view.obj_flags, st));
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

extern "C" int pvx_camera_rays_tiled(const pvx_camera *cam, uint32_t tile_w, uint32_t tile_h, pvx_ray *rays,
                                     uint32_t *pixel, void *stream) {
    if (!cam || !rays) return fail(PVX_EINVAL, "null argument");
    if (cam->width == 0 || cam->height == 0 || (uint64_t)cam->width * cam->height > 0xFFFFFFFFull)
        return fail(PVX_EINVAL, "bad image size");
    if (tile_w == 0 || tile_h == 0 || (uint64_t)tile_h * cam->width > 0xFFFFFFFFull)
        return fail(PVX_EINVAL, "bad tile size");
    const uint32_t n = cam->width * cam->height;
    k_camera_rays_tiled<<<(n + 255) / 256, 256, 0, static_cast<hipStream_t>(stream)>>>(*cam, tile_w, tile_h, rays,
                                                                                        pixel);
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
