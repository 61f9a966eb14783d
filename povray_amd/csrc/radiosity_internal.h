// Radiosity engine state shared by the radiosity and surrogate translation
// units.
#ifndef PVX_RADIOSITY_INTERNAL_H
#define PVX_RADIOSITY_INTERNAL_H

#include <hip/hip_runtime.h>

#include <vector>

#include "internal.h"

namespace pvx {

constexpr int kMaxRadDepth = 4;
constexpr int kMaxResolveRounds = 32; // gather rounds per resolve before the rest goes at once

// One irradiance-cache record (ot_block fields that AverageNearBlock reads,
// core/lighting/radiosity.h:93-105): 64 bytes.
struct RadRec {
    float p[3];      // Point
    float n[3];      // S_Normal
    float E[3];      // Illuminance
    float tn[3];     // To_Nearest_Surface
    float r;         // Harmonic_Mean_Distance (clamped to the reuse range)
    float nearest;   // Nearest_Distance
    float pad[2];
};
static_assert(sizeof(RadRec) == 64, "RadRec is one 64-byte line");

// One irradiance query: a surface point, its (eye-facing) normal and the
// weight that widens the error bound (ComputeAmbient's arguments).
struct RadQuery {
    double p[3];
    float n[3];
    float w;
};
static_assert(sizeof(RadQuery) == 40, "RadQuery layout");

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

// Records of one recursion depth, with a uniform grid over their
// quick-check spheres (cell -> record ids, CSR).
struct RadCache {
    std::vector<RadRec> recs;
    std::vector<uint32_t> start, items;
    float lo[3] = {0, 0, 0};
    float cell = 1.0f;
    int dims[3] = {1, 1, 1};
    float bound_max = 1.0f; // largest error bound a lookup at this depth uses
    DevBuf d_recs, d_start, d_items;
    int rebuild(hipStream_t st);
};

struct RadiosityEngine {
    pvx_scene *scene;
    pvx_radiosity_settings rs;
    pvx_camera cam;
    hipStream_t stream;
    RadCache caches[kMaxRadDepth];
    uint32_t cursor[kMaxRadDepth];
    std::vector<double> pool;
    DevBuf d_pool;
    pvx_radiosity_stats stats{};
    uint64_t stats_rounds[kMaxRadDepth] = {0, 0, 0, 0};

    struct Scratch { // per gather depth
        DevBuf rays, hits, base, coef, amb, q, need, rec, rec_ok, sel, subq, subE;
    } scratch[kMaxRadDepth];
    struct Resolve { // per resolve depth
        DevBuf reuse, need, idx, Eg, subq, subE, subreuse, idx2;
    } rstate[kMaxRadDepth];
    Scratch prim; // primary rays of a pixel pass

    RadiosityEngine(pvx_scene *s, const pvx_radiosity_settings &set, const pvx_camera &c, hipStream_t st);
    int init();
    int rays_per_sample(int depth) const;
    float reuse_count(int depth) const;
    int lookup(int depth, const RadQuery *q, uint32_t n, bool final, float *E, float *reuse);
    int gather(int depth, const RadQuery *q, const uint32_t *idx, uint32_t m, bool final, float *Eg);
    int resolve(int depth, const RadQuery *q, uint32_t n, bool final, float *E);
    int pixel_pass(const std::vector<uint32_t> &px, bool final, float *rgb);
};

} // namespace pvx

#endif
