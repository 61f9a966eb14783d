// Radiosity engine state shared by the radiosity and surrogate translation
// units.
#ifndef PVX_RADIOSITY_INTERNAL_H
#define PVX_RADIOSITY_INTERNAL_H

#include <hip/hip_runtime.h>

#include <vector>

#include "internal.h"
#include "mlp_internal.h"

namespace pvx {

struct RadiosityEngine;

constexpr int kMaxRadDepth = 4;
constexpr int kMaxResolveRounds = 32; // gather rounds per resolve before the rest goes at once

// One irradiance-cache record (the ot_block fields that AverageNearBlock
// reads, core/support/octree.h:101-117): 64 bytes.  The reference keeps the
// point in FP64; FP32 here moves a point by < 1e-6 of the scene size, far
// below the smallest reuse radius.
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

// Running reuse state of one query: sum of weight * illuminance and sum of
// weights over the records seen so far (w < 0 marks a gathered query).
struct RadAcc {
    double e[3];
    double w;
};

// Uniform grid over the quick-check spheres of a range of records
// (cell -> record ids, CSR), host copy plus device copy.
struct RadGrid {
    float lo[3] = {0, 0, 0};
    float cell = 1.0f;
    int dims[3] = {1, 1, 1};
    uint32_t n = 0; // records indexed
    std::vector<uint32_t> start, items;
    DevBuf d_start, d_items;
    int build(const std::vector<RadRec> &recs, size_t first, size_t last, float reach, RadiosityEngine &eng);
};

// Records of one recursion depth.  `all` indexes every record; `fresh`
// indexes the records of the latest gather round only.
struct RadCache {
    std::vector<RadRec> recs;
    size_t uploaded = 0;  // records present in d_recs
    size_t indexed = 0;   // records covered by `all`
    float reach = 1.0f;   // largest quick-check radius / harmonic mean radius
    DevBuf d_recs;
    RadGrid all, fresh;
    int upload(RadiosityEngine &eng);
    int index_all(RadiosityEngine &eng);
};

struct RadiosityEngine {
    pvx_scene *scene;
    pvx_radiosity_settings rs{};
    pvx_camera cam{};
    hipStream_t stream = nullptr;
    RadCache caches[kMaxRadDepth];
    uint32_t cursor[kMaxRadDepth];
    std::vector<double> pool;
    DevBuf d_pool;
    pvx_radiosity_stats stats{};
    uint64_t stats_rounds[kMaxRadDepth] = {0, 0, 0, 0};

    struct Scratch { // per gather depth
        DevBuf rays, hits, base, coef, amb, q, need, rec, sel, subq, subE;
    } scratch[kMaxRadDepth];
    struct Resolve { // per resolve depth
        DevBuf acc, need, idx, Eg, tq;
    } rstate[kMaxRadDepth];
    Scratch prim; // primary rays of a pixel pass
    std::vector<int32_t> last_rec[kMaxRadDepth]; // cache index of each query of the last gather (-1: not filed)
    Surrogate net;  // learned mode: depth-0 irradiance
    Surrogate net1; // learned mode: depth-1 irradiance (inside depth-0 gathers)
    DevBuf lx, lt, ly, lidx, lg0, lg1, ax, ay, px6, py, pq;
    DevBuf lrole, lkeys, lfbflag, lfbtile, lcnt;      // learned final pass
    DevBuf psel, ppix, pqpix, cub_tmp, d_count;       // pixel-pass compaction
    // While set, depth-0 gathers take the irradiance of their hits from this
    // network instead of resolving depth 1 (inputs normalised by nlo / ninv).
    Surrogate *amb_net = nullptr;
    float nlo[3] = {0, 0, 0}, ninv = 1.0f;
    std::vector<RadRec> last_fetched[kMaxRadDepth]; // records of the last gather per depth
    std::vector<RadRec> preload[kMaxRadDepth];      // records loaded from a cache file (seed every frame)
    DevBuf amb_q, amb_flag; // pvx_radiosity_compute_ambient's queries and layer-normal flag
    bool ambient_open = false; // pvx_radiosity_begin has opened a per-query cache
    double train_ms = 0.0;
    double t_sync_ms = 0.0, t_grid_ms = 0.0; // host-side waits and grid builds
    uint64_t n_sync = 0;
    HostBuf bounce; // pinned staging for device -> host copies
    hipError_t sync();
    int fetch(void *host, const void *dev, size_t bytes);

    HostBuf up;            // pinned staging for host -> device copies (reused after each wait)
    size_t up_off = 0;
    int send(void *dev, const void *host, size_t bytes);

    // The engine lives with its scene so that device buffers survive from
    // frame to frame; begin() clears the per-frame state.
    explicit RadiosityEngine(pvx_scene *s);
    int begin(const pvx_radiosity_settings &set, const pvx_camera &c, hipStream_t st);
    int rays_per_sample(int depth) const;
    float reuse_count(int depth) const;
    int lookup(int depth, const RadGrid &g, const RadQuery *q, const uint32_t *idx, uint32_t n, bool final,
               bool accumulate, RadAcc *acc, const uint8_t *mask = nullptr);
    int compact(const uint8_t *flags, uint32_t n, uint32_t *out, uint32_t &count);
    int gather(int depth, const RadQuery *q, const uint32_t *idx, uint32_t m, bool final, float *Eg);
    int gather_trace(int depth, const RadQuery *q, const uint32_t *idx, uint32_t m);
    int gather_finish(int depth, const RadQuery *q, const uint32_t *idx, uint32_t m, bool final, float *Eg);
    int train_on(Surrogate &nt, const RadRec *recs, uint32_t n, uint64_t seed, float *loss);
    int probe_errors(Surrogate &nt, const RadQuery *q, const uint32_t *d_idx, uint32_t m, const RadRec *pr,
                     std::vector<double> &err);
    int resolve(int depth, const RadQuery *q, uint32_t n, bool final, float *E);
    int start_resolve(int depth, const RadQuery *q, uint32_t n, bool final, RadAcc *&acc,
                      std::vector<uint32_t> &todo, std::vector<uint32_t> &slot, std::vector<RadQuery> &hq);
    void spread_select(int depth, const std::vector<uint32_t> &todo, const std::vector<uint32_t> &slot,
                       const std::vector<RadQuery> &hq, std::vector<uint32_t> &sel, std::vector<uint32_t> &sel_slot,
                       std::vector<uint32_t> &rest, std::vector<uint32_t> &rest_slot) const;
    int gather_round(int depth, const RadQuery *q, const std::vector<uint32_t> &sel, bool final, RadAcc *acc,
                     float *E);
    int relookup(int depth, const RadQuery *q, size_t first_new, bool final, RadAcc *acc,
                 const std::vector<uint32_t> &rest, const std::vector<uint32_t> &rest_slot,
                 std::vector<uint32_t> &todo, std::vector<uint32_t> &slot);
    int gather_rounds(int depth, const RadQuery *q, bool final, std::vector<uint32_t> todo,
                      std::vector<uint32_t> slot, const std::vector<RadQuery> &hq, RadAcc *acc, float *E,
                      int max_rounds = kMaxResolveRounds);
    int resolve_learned(const RadQuery *q, uint32_t n, const uint32_t *d_pix, float *E);
    int pixel_pass(const std::vector<uint32_t> &px, bool final, float *rgb);
};

} // namespace pvx

#endif
