// Internal declarations shared by the pvx translation units.
#ifndef PVX_INTERNAL_H
#define PVX_INTERNAL_H

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "../../include/pvx.h"
#include "pvx_types.h"
#include "scene_prep.h"

namespace pvx {

int fail(int code, const std::string &msg);

// pvx_render's light walks (light rays made in the traversal kernel).
inline bool is_light_walk(uint32_t v) {
    return v == PVX_TRAVERSAL_PACKET || v == PVX_TRAVERSAL_FUSED || v == PVX_TRAVERSAL_REFILL ||
           v == PVX_TRAVERSAL_LIGHT_BINARY || v == PVX_TRAVERSAL_REFILL_BINARY;
}

#define HIP_TRY(expr)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return ::pvx::fail(PVX_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Device buffer grown on demand (contents are not preserved).
struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf();
    int reserve(size_t bytes);
};

// Pinned host buffer grown on demand (contents are not preserved).
struct HostBuf {
    void *p = nullptr;
    size_t cap = 0;
    HostBuf() = default;
    HostBuf(const HostBuf &) = delete;
    HostBuf &operator=(const HostBuf &) = delete;
    ~HostBuf();
    int reserve(size_t bytes);
};

// Shading state attached by pvx_scene_set_shading.
struct ShadingState {
    bool set = false;
    std::vector<MatRec> mats;
    std::vector<LightRec> lights;
    float background[3] = {0, 0, 0};
    uint32_t max_level = 5;
    float bailout = 1.0f / 255.0f;
    uint32_t samples_per_hit = 0; // shadow samples over all lights
    bool all_opaque = false;      // no filter/transmit anywhere: light rays only test occlusion
    MatRec *d_mats = nullptr;
    LightRec *d_lights = nullptr;
};

// A stream created on first use (default priority) and destroyed with its owner.
struct StreamHolder {
    hipStream_t s = nullptr;
    StreamHolder() = default;
    StreamHolder(const StreamHolder &) = delete;
    StreamHolder &operator=(const StreamHolder &) = delete;
    ~StreamHolder() {
        if (s) (void)hipStreamDestroy(s);
    }
};

// Events kept between frames (created on demand, destroyed with the owner).
struct EventPool {
    std::vector<hipEvent_t> ev;
    EventPool() = default;
    EventPool(const EventPool &) = delete;
    EventPool &operator=(const EventPool &) = delete;
    ~EventPool() {
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    }
    hipError_t reserve(size_t n) {
        while (ev.size() < n) {
            hipEvent_t e;
            hipError_t r = hipEventCreate(&e);
            if (r != hipSuccess) return r;
            ev.push_back(e);
        }
        return hipSuccess;
    }
};

// Render work buffers, grown on demand and kept between frames.
struct RenderBuffers {
    DevBuf rays[2], state[2], kinds[2]; // ray queues (ping-pong)
    DevBuf hits, contrib, shade, lslot, lterm;
    DevBuf crays, cstate, nchild, offs, sums; // children of a bounce and their placement
    DevBuf srays[2], sdepth[2], sowner[2], shits, scol; // light rays
    DevBuf groups, vis; // packet light rays: shaded point per (hit, light), visibility per ray
    DevBuf counters, offsets;
    DevBuf pix[2]; // each hit's pixel (per bounce parity)
    // light passes overlapped with the next bounces (light rays made in the
    // kernel): the per-hit buffers of odd bounces, the refill counter, the
    // light stream and its events
    DevBuf contrib_b, shade_b, lslot_b, lterm_b, groups_b, refill;
    DevBuf gcount; // light groups of the bounces of each parity (two words)
    StreamHolder light_stream;
    EventPool events;
};

} // namespace pvx

struct pvx_scene {
    int device = 0;
    pvx::HostScene host;
    pvx::SceneView view{};
    uint32_t tune_variant[2] = {0, 0}; // autotuner's choice: [0] camera/bounce rays, [1] light rays
    uint32_t tune_grid[2] = {0, 0};    // (0: the default grid)
    // pvx_render's light rays made in the traversal kernel: walk and grid for
    // those of camera-ray hits [0] and of later bounces [1] (0: the default)
    uint32_t light_variant[2] = {0, 0};
    uint32_t light_grid[2] = {0, 0};
    // pixel tile of pvx_render's ray queue (tile_w 0: row-major)
    uint32_t render_tile[2] = {8, 8};
    pvx::KdView kd{};     // kd-tree of a PVX_ACCEL_BSP scene
    bool use_kd = false;
    std::vector<void *> allocs;
    uint32_t *d_obj_type = nullptr, *d_obj_index = nullptr;
    // Traversal-stack spill areas, one per stream that has traced this scene
    // (launches on different streams may overlap; each needs its own).
    std::mutex spill_mu;
    std::vector<std::pair<hipStream_t, int32_t *>> spills;
    size_t spill_bytes = 0;
    uint32_t grid = 0;
    uint64_t device_bytes = 0;
    uint64_t n_nodes2 = 0, n_nodes4 = 0, n_prims = 0; // tree sizes (host- or GPU-built)
    uint32_t n_objects = 0;
    pvx::ShadingState shading;
    pvx::RenderBuffers rb;
    void *rad_engine = nullptr;               // radiosity engine kept between frames
    void (*rad_engine_free)(void *) = nullptr;
};

// A batch of caller-owned scenes traced in one launch: a device table of
// their SceneViews and, per call, a table of 64-ray chunks {scene, first ray,
// count} that never straddle two scenes.
struct pvx_batch {
    int device = 0;
    std::vector<pvx_scene *> scenes;
    uint32_t level = 0;     // largest has_complex of the scenes
    uint32_t any_flags = 0; // union of the scenes' object flags
    uint32_t grid_ww = 0, grid = 0;
    pvx::DevBuf views, chunks, spill;
    pvx::HostBuf hchunks;
    hipEvent_t uploaded = nullptr; // last chunk-table upload (guards hchunks)
    std::vector<uint64_t> last_first;
    uint32_t n_chunks = 0;
};

namespace pvx {
// Closest-hit launch shared with the renderer (no counters).
// `kinds` (optional): one PVX_RAY_* kind per ray for the object-flag
// precondition of rays of mixed kinds (reflected and refracted).
int trace_closest_internal(pvx_scene *s, const pvx_ray *rays, const double *max_depth,
                           uint32_t ray_kind, size_t n, pvx_hit *hits, hipStream_t stream,
                           const uint8_t *kinds = nullptr);
// Light rays made in the traversal kernel from shaded points (LightGroup per
// lit point and light; S rays per group, the light's grid in the first ones)
// with the occlusion test; one visibility byte per ray.  Triangle / sphere
// scenes with a 4-wide tree.  `lights` is the device light array.  With
// `next` (a device word holding 0) lanes take new rays from that counter as
// their rays end; without it each lane walks a fixed stride of rays.  `grid`
// blocks (0: the walk's default).  With `count` (device), the kernel reads
// the number of groups there and n_rays only bounds the grid.  `binary`:
// through the binary tree (one node or leaf per pass per lane).
int light_rays_internal(pvx_scene *s, const LightGroup *groups, const LightRec *lights, uint32_t S,
                        uint32_t n_rays, uint8_t *vis, uint32_t *next, uint32_t grid, const uint32_t *count,
                        bool binary, hipStream_t stream);
// The autotuner over sample rays with optional per-ray max depths.
int autotune_internal(pvx_scene *s, const pvx_ray *rays, const double *max_depth, size_t n,
                      uint32_t ray_kind, pvx_tune_result *out, hipStream_t stream);
} // namespace pvx

#endif
