// Internal declarations shared by the pvx translation units.
#ifndef PVX_INTERNAL_H
#define PVX_INTERNAL_H

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/pvx.h"
#include "pvx_types.h"
#include "scene_prep.h"

namespace pvx {

int fail(int code, const std::string &msg);

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

// Shading state attached by pvx_scene_set_shading.
struct ShadingState {
    bool set = false;
    std::vector<MatRec> mats;
    std::vector<LightRec> lights;
    float background[3] = {0, 0, 0};
    uint32_t max_level = 5;
    float bailout = 1.0f / 255.0f;
    uint32_t samples_per_hit = 0; // shadow samples over all lights
    MatRec *d_mats = nullptr;
    LightRec *d_lights = nullptr;
};

// Render work buffers, grown on demand.
struct RenderBuffers {
    size_t cap_rays = 0, cap_shadow = 0, cap_shade = 0;
    void *rays[2] = {nullptr, nullptr};
    void *state[2] = {nullptr, nullptr};
    void *hits = nullptr;
    void *shade = nullptr;
    void *srays = nullptr, *sdepth = nullptr, *shits = nullptr;
    uint32_t *counters = nullptr; // [next rays, shadow rays]
};

} // namespace pvx

struct pvx_scene {
    int device = 0;
    pvx::HostScene host;
    pvx::SceneView view{};
    std::vector<void *> allocs;
    std::vector<void *> render_allocs;
    uint32_t *d_obj_type = nullptr, *d_obj_index = nullptr;
    int32_t *d_spill = nullptr;
    uint32_t grid = 0;
    uint64_t device_bytes = 0;
    uint32_t n_objects = 0;
    pvx::ShadingState shading;
    pvx::RenderBuffers rb;
};

namespace pvx {
// Closest-hit launch shared with the renderer (no counters).
int trace_closest_internal(pvx_scene *s, const pvx_ray *rays, const double *max_depth,
                           uint32_t ray_kind, size_t n, pvx_hit *hits, hipStream_t stream);
} // namespace pvx

#endif
