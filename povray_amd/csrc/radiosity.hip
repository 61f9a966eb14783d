// Radiosity on the GPU (gfx950): irradiance cache + batched gathers.
//
// The reference fills its Ward irradiance cache one gather at a time while
// pixels are traced in order (RadiosityFunction::ComputeAmbient /
// GatherLight, core/lighting/radiosity.cpp:385-743).  Here every query of a
// pass is looked up at once; the points the cache cannot serve are gathered
// in rounds — a spread-out subset first, its records inserted, the rest
// looked up again — so later points reuse earlier gathers as in the
// sequential algorithm, while each round is one wide launch:
//   gather rays (Halton(2,3) cosine-weighted pool, SampleDirectionGenerator
//   frame, radiosity.cpp:758-868) -> closest hit -> depth+1 irradiance
//   (recursively resolved the same way) -> per-query reduction (mean
//   radiance, harmonic mean distance, nearest hit) -> new cache records.
// Lookups follow AverageNearBlock (radiosity.cpp:1551-1769): quick-check
// radius, nearest-surface radius blend, error_reuse = dist/r + 2 sqrt(1-n.n'),
// in-front test, saw weight (pseudo-sigmoid for depth-0 records in the final
// pass).  A query keeps running sums (sum w*E, sum w), so after a round only
// the round's new records (their own small grid) are visited again.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_select.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <numeric>
#include <thread>
#include <vector>

#include "internal.h"
#include "normal_device.h"
#include "radiosity_internal.h"
#include "shapes_device.h"

using namespace pvx;

namespace {

constexpr double kInFrontLimit = -0.05;   // radiosity.cpp:100
constexpr double kAvgNearEps = 0.000001;  // radiosity.cpp:109
constexpr double kRadEps = 0.001;         // radiosity.cpp:110
constexpr float kWeightErrOffset = 0.25f; // radiosity.cpp:111
constexpr int kPoolSize = 1600;           // RADIOSITY_MAX_SAMPLE_DIRECTIONS

struct CacheView {
    const RadRec *recs;
    const uint32_t *cell_start; // ncells + 1
    const uint32_t *cell_items;
    float lo[3];
    float cell;
    int dims[3];
    uint32_t n;
};

__device__ inline float grey3(const float *c) { return 0.297f * c[0] + 0.589f * c[1] + 0.114f * c[2]; }

// FindReusableBlock + AverageNearBlock for a batch of queries: adds the
// weighted illuminance of every record of the grid that qualifies to the
// query's running sums (or replaces them when `accumulate` is 0).
// (Queries with mask[i] == 0 are unset: their sums are cleared.)
__global__ void k_lookup(const RadQuery *__restrict__ q, const uint32_t *__restrict__ idx, uint32_t n,
                         CacheView c, float eb, float factor, int sigmoid, int accumulate,
                         RadAcc *__restrict__ acc, const uint8_t *__restrict__ mask) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const uint32_t i = idx ? idx[t] : t;
    if (mask && !mask[i]) {
        RadAcc z;
        z.e[0] = z.e[1] = z.e[2] = 0.0;
        z.w = 0.0;
        acc[i] = z;
        return;
    }
    const RadQuery qq = q[i];
    double bound = (double)eb;
    if (qq.w < kWeightErrOffset) bound += (double)(kWeightErrOffset - qq.w);
    bound *= (double)factor;
    double W = 0.0, we[3] = {0, 0, 0};
    if (c.n > 0) {
        const float qx = (float)qq.p[0], qy = (float)qq.p[1], qz = (float)qq.p[2];
        int cx = (int)floorf((qx - c.lo[0]) / c.cell);
        int cy = (int)floorf((qy - c.lo[1]) / c.cell);
        int cz = (int)floorf((qz - c.lo[2]) / c.cell);
        if (cx >= 0 && cy >= 0 && cz >= 0 && cx < c.dims[0] && cy < c.dims[1] && cz < c.dims[2]) {
            const float fb = (float)bound * 1.0001f;
            uint32_t cid = ((uint32_t)cz * c.dims[1] + (uint32_t)cy) * c.dims[0] + (uint32_t)cx;
            for (uint32_t k = c.cell_start[cid]; k < c.cell_start[cid + 1]; k++) {
                const RadRec *bp = c.recs + c.cell_items[k];
                // FP32 pre-test of the quick-check sphere (conservative)
                float fx = qx - bp->p[0], fy = qy - bp->p[1], fz = qz - bp->p[2];
                float fq = bp->r * fb;
                if (fx * fx + fy * fy + fz * fz > fq * fq + 1e-12f) continue;
                const RadRec b = *bp;
                double dx = qq.p[0] - b.p[0], dy = qq.p[1] - b.p[1], dz = qq.p[2] - b.p[2];
                double sq = dx * dx + dy * dy + dz * dz;
                double quick = (double)b.r * bound;
                if (!(sq < quick * quick)) continue;
                double dist = sqrt(sq);
                double ri = b.r;
                bool far = dist > kAvgNearEps;
                double ux = 0, uy = 0, uz = 0;
                if (far) {
                    ux = dx / dist;
                    uy = dy / dist;
                    uz = dz / dist;
                    double cs = b.tn[0] * ux + b.tn[1] * uy + b.tn[2] * uz;
                    if (cs > 0.0) ri = cs * b.nearest + (1.0 - cs) * ri;
                }
                if (!(dist < ri * bound)) continue;
                double dd = (double)qq.n[0] * b.n[0] + (double)qq.n[1] * b.n[1] + (double)qq.n[2] * b.n[2];
                double err = dist / ri + 2.0 * sqrt(fabs(1.0 - dd));
                if (!(err < bound)) continue;
                double in_front = 1.0;
                if (far) {
                    double hx = qq.n[0] + b.n[0], hy = qq.n[1] + b.n[1], hz = qq.n[2] + b.n[2];
                    in_front = (ux * hx + uy * hy + uz * hz) / sqrt(hx * hx + hy * hy + hz * hz);
                }
                if (!(in_front > kInFrontLimit)) continue;
                double w;
                if (!sigmoid) {
                    w = 1.0 - err / bound;
                } else { // PSEUDO_SIGMOID_METHOD (final pass, bounce depth 0)
                    w = err / bound;
                    w = w < 0.5 ? 1.0 - (w * 2.0) * (w * 2.0) / 2.0 : ((1.0 - w) * 2.0) * ((1.0 - w) * 2.0) / 2.0;
                }
                if (in_front <= 0) w = w * (1 - in_front / kInFrontLimit);
                if (w > kRadEps) {
                    we[0] += b.E[0] * w;
                    we[1] += b.E[1] * w;
                    we[2] += b.E[2] * w;
                    W += w;
                }
            }
        }
    }
    RadAcc a;
    if (accumulate) {
        a = acc[i];
        a.e[0] += we[0];
        a.e[1] += we[1];
        a.e[2] += we[2];
        a.w += W;
    } else {
        a.e[0] = we[0];
        a.e[1] = we[1];
        a.e[2] = we[2];
        a.w = W;
    }
    acc[i] = a;
}

// Gather rays: m queries x N directions from the cosine-weighted pool,
// continuing a running cursor as SampleDirectionGenerator does.
__global__ void k_gather_rays(const RadQuery *__restrict__ q, const uint32_t *__restrict__ idx,
                              uint32_t m, uint32_t N, const double *__restrict__ pool,
                              uint32_t cursor0, pvx_ray *__restrict__ rays) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m * N) return;
    uint32_t s = k / N, j = k % N;
    const RadQuery qq = q[idx[s]];
    uint32_t pi = (uint32_t)(((uint64_t)cursor0 + (uint64_t)s * N + j) % kPoolSize);
    double vx = pool[3 * pi], vy = fabs(pool[3 * pi + 1]), vz = pool[3 * pi + 2];
    double fy[3] = {qq.n[0], qq.n[1], qq.n[2]};
    double off[3] = {0, 0, 1};
    if (fabs(fy[2]) > 0.9) {
        off[1] = 1;
        off[2] = 0;
    }
    // frameX = normalize(cross(frameY, offY)), frameZ = normalize(cross(frameX, frameY))
    double fx[3] = {fy[1] * off[2] - fy[2] * off[1], fy[2] * off[0] - fy[0] * off[2],
                    fy[0] * off[1] - fy[1] * off[0]};
    double lx = sqrt(fx[0] * fx[0] + fx[1] * fx[1] + fx[2] * fx[2]);
    for (int a = 0; a < 3; a++) fx[a] /= lx;
    double fz[3] = {fx[1] * fy[2] - fx[2] * fy[1], fx[2] * fy[0] - fx[0] * fy[2],
                    fx[0] * fy[1] - fx[1] * fy[0]};
    double lz = sqrt(fz[0] * fz[0] + fz[1] * fz[1] + fz[2] * fz[2]);
    for (int a = 0; a < 3; a++) fz[a] /= lz;
    double d[3];
    if (fy[1] > 1.0 - kRadEps) {
        d[0] = vx; d[1] = vy; d[2] = vz;
    } else if (fy[1] < -1.0 + kRadEps) {
        d[0] = -vx; d[1] = -vy; d[2] = -vz;
    } else {
        for (int a = 0; a < 3; a++) d[a] = fx[a] * vx + fy[a] * vy + fz[a] * vz;
    }
    pvx_ray r;
    for (int a = 0; a < 3; a++) {
        r.o[a] = qq.p[a];
        r.d[a] = d[a];
    }
    rays[k] = r;
}

struct ShadeParams : NormalParams {
    const MatRec *mats;
    float adc;
};

// Surface response at each hit: base = pigment*emission*opacity and
// coef = pigment*opacity*diffuse, plus a depth+1 irradiance query when
// radiosity is needed there (ComputeLightedTexture, trace.cpp:826-939).
// kRev: the scene has lathes / sors (their normal retraces the hit list).
template <bool kRev>
__global__ void k_hit_queries(ShadeParams p, const pvx_ray *__restrict__ rays,
                              const pvx_hit *__restrict__ hits, uint32_t n, float wnext,
                              int want_ambient, float *__restrict__ base, float *__restrict__ coef,
                              RadQuery *__restrict__ q, uint8_t *__restrict__ need) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const pvx_hit h = hits[i];
    need[i] = 0;
    if (h.obj < 0) {
        for (int a = 0; a < 3; a++) {
            base[3 * (size_t)i + a] = 0.0f;
            coef[3 * (size_t)i + a] = 0.0f;
        }
        return;
    }
    const pvx_ray r = rays[i];
    D3 o{r.o[0], r.o[1], r.o[2]}, d{r.d[0], r.d[1], r.d[2]};
    D3 ip = add3(o, mul3(d, h.t));
    D3 nrm = hit_normal<kRev>(p, h, o, d, ip);
    if (dot3(nrm, d) > 0.0) nrm = D3{-nrm.x, -nrm.y, -nrm.z};
    const MatRec &m = p.mats[h.obj];
    float cf[3];
    for (int a = 0; a < 3; a++) {
        base[3 * (size_t)i + a] = m.pigment[a] * m.ambient_only[a] * m.att;
        cf[a] = m.pigment[a] * (m.att * m.diffuse);
    }
    float maxc = grey3(m.pigment) * m.att * m.diffuse;
    bool amb = want_ambient && m.diffuse > 0.0f && maxc > p.adc;
    for (int a = 0; a < 3; a++) coef[3 * (size_t)i + a] = amb ? cf[a] : 0.0f;
    if (amb) {
        RadQuery qq;
        qq.p[0] = ip.x; qq.p[1] = ip.y; qq.p[2] = ip.z;
        qq.n[0] = (float)nrm.x; qq.n[1] = (float)nrm.y; qq.n[2] = (float)nrm.z;
        qq.w = wnext * maxc;
        q[i] = qq;
        need[i] = 1;
    }
}

// GatherLight's reduction (radiosity.cpp:585-700) for m queries of N rays
// each: one wavefront per query, lanes striding over the rays.
__global__ __launch_bounds__(256) void k_gather_reduce(
    const RadQuery *__restrict__ q, const uint32_t *__restrict__ idx, uint32_t m, uint32_t N,
    const pvx_ray *__restrict__ rays, const pvx_hit *__restrict__ hits, const float *__restrict__ base,
    const float *__restrict__ coef, const float *__restrict__ amb, double cx, double cy, double cz,
    float min_reuse, float max_reuse, float factor, float *__restrict__ Eg, RadRec *__restrict__ rec) {
    const uint32_t s = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t lane = threadIdx.x & 63;
    if (s >= m) return; // whole wavefronts leave together
    double sum[3] = {0, 0, 0}, inv = 0.0, smallest = kBoundHuge, sq = 0.0;
    uint32_t sj = 0xFFFFFFFFu;
    for (uint32_t j = lane; j < N; j += 64) {
        size_t k = (size_t)s * N + j;
        const pvx_hit h = hits[k];
        double depth = HUGE_VAL;
        if (h.obj >= 0) {
            depth = h.t;
            double c[3];
            for (int a = 0; a < 3; a++) {
                c[a] = (double)(base[3 * k + a] + coef[3 * k + a] * amb[3 * k + a]);
                sum[a] += c[a];
            }
            double gr = 0.297 * c[0] + 0.589 * c[1] + 0.114 * c[2];
            sq += gr * gr;
        }
        if (depth < smallest) { // first (lowest j) of the smallest depths
            smallest = depth;
            sj = j;
        }
        inv += 1.0 / depth;
    }
    for (int off = 32; off > 0; off >>= 1) {
        for (int a = 0; a < 3; a++) sum[a] += __shfl_xor(sum[a], off, 64);
        inv += __shfl_xor(inv, off, 64);
        sq += __shfl_xor(sq, off, 64);
        double os = __shfl_xor(smallest, off, 64);
        uint32_t oj = __shfl_xor(sj, off, 64);
        if (os < smallest || (os == smallest && oj < sj)) {
            smallest = os;
            sj = oj;
        }
    }
    if (lane != 0) return;
    const RadQuery qq = q[idx[s]];
    float e[3] = {(float)(sum[0] / N), (float)(sum[1] / N), (float)(sum[2] / N)};
    Eg[3 * (size_t)s] = e[0];
    Eg[3 * (size_t)s + 1] = e[1];
    Eg[3 * (size_t)s + 2] = e[2];
    double md[3] = {0, 0, 0};
    if (sj != 0xFFFFFFFFu) {
        const pvx_ray &r = rays[(size_t)s * N + sj];
        md[0] = r.d[0];
        md[1] = r.d[1];
        md[2] = r.d[2];
    }
    double ex = cx - qq.p[0], ey = cy - qq.p[1], ez = cz - qq.p[2];
    double to_eye = sqrt(ex * ex + ey * ey + ez * ez);
    double rmin = to_eye * (double)min_reuse * factor, rmax = to_eye * (double)max_reuse * factor;
    double mean = (double)N / inv;
    bool ok = smallest > rmax * 0.0001;
    if (mean < rmin) mean = rmin;
    if (mean > rmax) mean = rmax;
    RadRec r;
    for (int a = 0; a < 3; a++) {
        r.p[a] = (float)qq.p[a];
        r.n[a] = qq.n[a];
        r.E[a] = e[a];
        r.tn[a] = (float)md[a];
    }
    r.r = (float)mean;
    r.nearest = (float)smallest;
    // standard error of the grey mean over the N rays (the probe test's
    // allowance for the gather's own noise)
    double gm = (0.297 * sum[0] + 0.589 * sum[1] + 0.114 * sum[2]) / N;
    double var = N > 1 ? fmax(0.0, (sq - N * gm * gm) / (N - 1)) : 0.0;
    r.pad[0] = (float)sqrt(var / N);
    r.pad[1] = ok ? 1.0f : 0.0f; // filed in the cache (AddBlock condition)
    rec[s] = r;
}

// ComputeAmbient's combination of the partial reuse with the new gather
// (radiosity.cpp:430-440); marks the query as gathered.
__global__ void k_combine(const uint32_t *__restrict__ idx, uint32_t m, const float *__restrict__ Eg,
                          RadAcc *__restrict__ acc, float *__restrict__ E) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= m) return;
    uint32_t i = idx[s];
    RadAcc a = acc[i];
    for (int c = 0; c < 3; c++) {
        double g = Eg[3 * (size_t)s + c];
        E[3 * (size_t)i + c] = (float)(a.w > 0.0 ? (a.e[c] + g * 1.0) / (a.w + 1.0) : g);
    }
    a.w = -1.0;
    acc[i] = a;
}

// Reused irradiance of the queries that were not gathered.
__global__ void k_finish(const RadAcc *__restrict__ acc, uint32_t n, float brightness, float *__restrict__ E) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const RadAcc a = acc[i];
    for (int c = 0; c < 3; c++) {
        float v = E[3 * (size_t)i + c];
        if (a.w >= 0.0) v = a.w > 0.0 ? (float)(a.e[c] / a.w) : 0.0f;
        E[3 * (size_t)i + c] = v * brightness;
    }
}

// Needs-a-gather test of ComputeAmbient (radiosity.cpp:410-411; gray_threshold 0).
__global__ void k_need(const RadAcc *__restrict__ acc, const uint32_t *__restrict__ idx, uint32_t n,
                       float reuse_count, int final, uint8_t *__restrict__ need) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double w = acc[idx ? idx[t] : t].w;
    need[t] = (w * 4.0 >= (double)reuse_count || (final && w > 0.0)) ? 0 : 1;
}

// pixel colour = base + coef * E (the primary hit's ComputeLightedTexture)
__global__ void k_final_colour(const float *__restrict__ base, const float *__restrict__ coef,
                               const float *__restrict__ E, const uint32_t *__restrict__ qpix,
                               uint32_t n, const uint8_t *__restrict__ has_q, float *__restrict__ rgb) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int a = 0; a < 3; a++) {
        float c = base[3 * (size_t)i + a];
        if (has_q[i]) c = c + coef[3 * (size_t)i + a] * E[3 * (size_t)qpix[i] + a];
        rgb[3 * (size_t)i + a] = c;
    }
}

__global__ void k_pixel_rays(pvx_camera cam, const uint32_t *__restrict__ px, uint32_t n,
                             pvx_ray *__restrict__ rays) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t x = px[i] % cam.width, y = px[i] / cam.width;
    double x0 = ((double)x + cam.jitter_x) / (double)cam.width - 0.5;
    double y0 = 0.5 - ((double)y + cam.jitter_y) / (double)cam.height;
    double dx = cam.direction[0] + x0 * cam.right[0] + y0 * cam.up[0];
    double dy = cam.direction[1] + x0 * cam.right[1] + y0 * cam.up[1];
    double dz = cam.direction[2] + x0 * cam.right[2] + y0 * cam.up[2];
    double len = sqrt(dx * dx + dy * dy + dz * dz);
    pvx_ray r;
    r.o[0] = cam.location[0]; r.o[1] = cam.location[1]; r.o[2] = cam.location[2];
    r.d[0] = dx / len; r.d[1] = dy / len; r.d[2] = dz / len;
    rays[i] = r;
}

__global__ void k_gather_subq(const RadQuery *__restrict__ q, const uint32_t *__restrict__ idx, uint32_t m,
                              RadQuery *__restrict__ out) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < m) out[s] = q[idx[s]];
}

__global__ void k_scatter3(const float *__restrict__ src, const uint32_t *__restrict__ idx, uint32_t m,
                           float *__restrict__ dst) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= m) return;
    uint32_t i = idx[s];
    for (int a = 0; a < 3; a++) dst[3 * (size_t)i + a] = src[3 * (size_t)s + a];
}

// Surrogate inputs: position scaled into the unit cube, eye-facing normal.
__global__ void k_rec_to_sample(const RadRec *__restrict__ recs, const uint32_t *__restrict__ ridx, uint32_t n,
                                float lx, float ly, float lz, float inv, float *__restrict__ x6,
                                float *__restrict__ t) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const RadRec b = recs[ridx ? ridx[i] : i];
    x6[6 * (size_t)i + 0] = fminf(fmaxf((b.p[0] - lx) * inv, 0.0f), 1.0f);
    x6[6 * (size_t)i + 1] = fminf(fmaxf((b.p[1] - ly) * inv, 0.0f), 1.0f);
    x6[6 * (size_t)i + 2] = fminf(fmaxf((b.p[2] - lz) * inv, 0.0f), 1.0f);
    for (int a = 0; a < 3; a++) {
        x6[6 * (size_t)i + 3 + a] = b.n[a];
        t[3 * (size_t)i + a] = b.E[a];
    }
}

// (idx may be null: queries 0..m-1; where mask[s] is 0 the query is unset
// and the inputs are zero.)
__global__ void k_query_to_x6(const RadQuery *__restrict__ q, const uint32_t *__restrict__ idx,
                              const uint8_t *__restrict__ mask, uint32_t m, float lx, float ly, float lz,
                              float inv, float *__restrict__ x6) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= m) return;
    if (mask && !mask[s]) {
        for (int c = 0; c < 6; c++) x6[6 * (size_t)s + c] = 0.0f;
        return;
    }
    const RadQuery a = q[idx ? idx[s] : s];
    x6[6 * (size_t)s + 0] = fminf(fmaxf(((float)a.p[0] - lx) * inv, 0.0f), 1.0f);
    x6[6 * (size_t)s + 1] = fminf(fmaxf(((float)a.p[1] - ly) * inv, 0.0f), 1.0f);
    x6[6 * (size_t)s + 2] = fminf(fmaxf(((float)a.p[2] - lz) * inv, 0.0f), 1.0f);
    for (int c = 0; c < 3; c++) x6[6 * (size_t)s + 3 + c] = a.n[c];
}

// Surrogate answer for the queries idx[*]: irradiance max(y, 0); marks them done.
__global__ void k_apply_pred(const uint32_t *__restrict__ idx, uint32_t m, const float *__restrict__ y,
                             RadAcc *__restrict__ acc, float *__restrict__ E) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= m) return;
    uint32_t i = idx[s];
    for (int c = 0; c < 3; c++) E[3 * (size_t)i + c] = fmaxf(y[3 * (size_t)s + c], 0.0f);
    acc[i].w = -1.0;
}

// ComputeAmbient's arguments -> queries (radiosity.cpp:385-400): the point,
// the raw normal (the effective normal with `normal off`), the weight that
// widens the error bound.  A layer normal that is not `similar` to the raw
// one (vector.h:591, summed component difference < EPSILON) would make
// GatherLight sample around the perturbed normal: flagged, not served.
__global__ void k_ambient_queries(const double *__restrict__ ip, const double *__restrict__ rn,
                                  const double *__restrict__ ln, const double *__restrict__ wt, size_t n,
                                  RadQuery *__restrict__ q, uint32_t *__restrict__ bad) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    RadQuery qq;
    for (int a = 0; a < 3; a++) {
        qq.p[a] = ip[3 * i + a];
        qq.n[a] = (float)rn[3 * i + a];
    }
    qq.w = wt ? (float)wt[i] : 1.0f;
    if (ln) {
        double s = fabs(ln[3 * i] - rn[3 * i]) + fabs(ln[3 * i + 1] - rn[3 * i + 1]) +
                   fabs(ln[3 * i + 2] - rn[3 * i + 2]);
        if (!(s < 1.0e-10)) atomicOr(bad, 1u);
    }
    q[i] = qq;
}

__global__ void k_fill_f32(float *__restrict__ p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

__global__ void k_acc_zero(RadAcc *__restrict__ acc, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    RadAcc a;
    a.e[0] = a.e[1] = a.e[2] = 0.0;
    a.w = 0.0;
    acc[i] = a;
}

// depth+1 irradiance of the gather hits that want it, from the network
// (what resolve(depth + 1) would return: max(y, 0) x brightness).
__global__ void k_amb_from_pred(const uint8_t *__restrict__ need, const float *__restrict__ y, uint32_t n,
                                float brightness, float *__restrict__ amb) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int c = 0; c < 3; c++)
        amb[3 * (size_t)i + c] = need[i] ? fmaxf(y[3 * (size_t)i + c], 0.0f) * brightness : 0.0f;
}

__device__ inline uint32_t dmix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

__global__ void k_fill_u64(unsigned long long *__restrict__ p, size_t n, unsigned long long v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_u8(uint8_t *__restrict__ p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Training points: per cell of cpx x cpx pixels, the query with the lowest
// pixel hash (key = hash << 32 | query; the minimum does not depend on the
// order of the atomics).
// atomicMin of (slot, key) per lane, combined first across the lanes of the
// wavefront that share lane 0's slot (consecutive queries mostly share a
// cell or tile), so a hot slot takes one atomic per wavefront.
__device__ inline void wave_min_into(unsigned long long *keys, uint32_t slot, unsigned long long key, bool active) {
    const uint32_t s0 = __shfl(slot, 0, 64);
    const bool act0 = __shfl((int)active, 0, 64) != 0;
    const bool same = active && act0 && slot == s0;
    unsigned long long v = same ? key : ~0ull;
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(v, off, 64);
        v = o < v ? o : v;
    }
    if ((threadIdx.x & 63) == 0 && v != ~0ull) atomicMin(&keys[s0], v);
    if (active && !same) atomicMin(&keys[slot], key);
}

__global__ void k_cell_keys(const uint32_t *__restrict__ pix, uint32_t n, uint32_t W, uint32_t cpx, uint32_t ncx,
                            unsigned long long *__restrict__ keys) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = k < n;
    uint32_t c = 0;
    unsigned long long key = ~0ull;
    if (active) {
        const uint32_t p = pix[k];
        c = (p % W) / cpx + ((p / W) / cpx) * ncx;
        key = ((unsigned long long)dmix32(p) << 32) | k;
    }
    wave_min_into(keys, c, key, active);
}

// Probes: per probe tile, the not-yet-chosen query with the lowest salted hash.
__global__ void k_tile_keys(const uint32_t *__restrict__ pix, const uint8_t *__restrict__ role, uint32_t n,
                            uint32_t W, uint32_t tsz, uint32_t tx, uint32_t salt,
                            unsigned long long *__restrict__ keys) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = k < n && !role[k];
    uint32_t t = 0;
    unsigned long long key = ~0ull;
    if (active) {
        const uint32_t p = pix[k];
        t = (p % W) / tsz + ((p / W) / tsz) * tx;
        key = ((unsigned long long)dmix32(p ^ salt) << 32) | k;
    }
    wave_min_into(keys, t, key, active);
}

__global__ void k_mark_role(const uint32_t *__restrict__ idx, uint32_t m, uint8_t v, uint8_t *__restrict__ role) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < m) role[idx[s]] = v;
}

__global__ void k_acc_reset(const uint32_t *__restrict__ idx, uint32_t m, RadAcc *__restrict__ acc) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= m) return;
    RadAcc z;
    z.e[0] = z.e[1] = z.e[2] = 0.0;
    z.w = 0.0;
    acc[idx[s]] = z;
}

// Final-pass misses (no record reaches them, not gathered): the network's
// answer, or a fallback flag in tiles whose probes failed.  counters[0] /
// [1] count the two.
__global__ void k_learned_apply(RadAcc *__restrict__ acc, const uint32_t *__restrict__ pix, uint32_t n, uint32_t W,
                                uint32_t tsz, uint32_t tx, const uint8_t *__restrict__ fb_tile,
                                const float *__restrict__ y, float *__restrict__ E, uint8_t *__restrict__ fb,
                                uint32_t *__restrict__ counters) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    int kind = 0; // 1: network, 2: fallback
    if (k < n) {
        fb[k] = 0;
        if (acc[k].w == 0.0) { // neither reused (> 0) nor gathered (< 0)
            const uint32_t p = pix[k];
            const size_t t = (p % W) / tsz + (size_t)((p / W) / tsz) * tx;
            if (fb_tile[t]) {
                fb[k] = 1;
                kind = 2;
            } else {
                for (int c = 0; c < 3; c++) E[3 * (size_t)k + c] = fmaxf(y[3 * (size_t)k + c], 0.0f);
                acc[k].w = -1.0;
                kind = 1;
            }
        }
    }
    const unsigned long long b1 = __ballot(kind == 1), b2 = __ballot(kind == 2);
    if ((threadIdx.x & 63) == 0) {
        if (b1) atomicAdd(&counters[0], (uint32_t)__popcll(b1));
        if (b2) atomicAdd(&counters[1], (uint32_t)__popcll(b2));
    }
}

__global__ void k_gather_u32(const uint32_t *__restrict__ src, const uint32_t *__restrict__ idx, uint32_t m,
                             uint32_t *__restrict__ out) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < m) out[s] = src[idx[s]];
}

// rank[sel[s]] = s: each pixel's query slot
__global__ void k_scatter_rank(const uint32_t *__restrict__ sel, uint32_t m, uint32_t *__restrict__ rank) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < m) rank[sel[s]] = s;
}

__global__ void k_fill_u32(uint32_t *__restrict__ p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

inline uint32_t blocks(size_t n, uint32_t b = 256) { return (uint32_t)((n + b - 1) / b); }

} // namespace

// ----------------------------------------------------------------------------
// host side
// ----------------------------------------------------------------------------
namespace pvx {

DevBuf::~DevBuf() {
    if (p) (void)hipFree(p);
}
int DevBuf::reserve(size_t bytes) {
    if (bytes <= cap) return PVX_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    size_t want = std::max(bytes + bytes / 4, (size_t)4096); // slack: fewer regrowths
    HIP_TRY(hipMalloc(&p, want));
    cap = want;
    return PVX_OK;
}

int RadGrid::build(const std::vector<RadRec> &recs, size_t first, size_t last, float reach, RadiosityEngine &eng) {
    n = (uint32_t)(last - first);
    if (n == 0) return PVX_OK;
    float lo_[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    std::vector<float> rad(n), rs(n);
    for (size_t i = 0; i < n; i++) {
        const RadRec &b = recs[first + i];
        rad[i] = b.r * reach * 1.001f + 1e-6f; // >= any quick-check radius of the record
        rs[i] = rad[i];
        for (int a = 0; a < 3; a++) {
            lo_[a] = std::min(lo_[a], b.p[a] - rad[i]);
            hi[a] = std::max(hi[a], b.p[a] + rad[i]);
        }
    }
    std::nth_element(rs.begin(), rs.begin() + rs.size() / 2, rs.end());
    float ext = std::max(hi[0] - lo_[0], std::max(hi[1] - lo_[1], hi[2] - lo_[2]));
    // cells about the median sphere size, at most 64 per axis and about four
    // cells per record, grown until the record->cell lists stay within a few
    // entries per record
    const float vol = std::max(hi[0] - lo_[0], 1e-6f) * std::max(hi[1] - lo_[1], 1e-6f) *
                      std::max(hi[2] - lo_[2], 1e-6f);
    cell = std::max(std::max(rs[rs.size() / 2], ext / 64.0f), std::cbrt(vol / (4.0f * n + 64.0f)));
    cell = std::max(cell, 1e-6f);
    const double budget = 8.0 * (double)n + 65536.0;
    for (int tries = 0; tries < 64; tries++) {
        double total = 0;
        for (size_t i = 0; i < n && total <= budget; i++) {
            double k = 2.0 * rad[i] / cell + 2.0;
            total += k * k * k;
        }
        if (total <= budget) break;
        cell *= 1.5f;
    }
    for (int a = 0; a < 3; a++) {
        lo[a] = lo_[a];
        dims[a] = std::max(1, (int)std::ceil((hi[a] - lo_[a]) / cell) + 1);
    }
    size_t ncell = (size_t)dims[0] * dims[1] * dims[2];
    std::vector<uint32_t> count(ncell + 1, 0);
    auto range = [&](size_t i, int a, int &c0, int &c1) {
        const RadRec &b = recs[first + i];
        c0 = std::max(0, (int)std::floor((b.p[a] - rad[i] - lo[a]) / cell));
        c1 = std::min(dims[a] - 1, (int)std::floor((b.p[a] + rad[i] - lo[a]) / cell));
    };
    for (int pass = 0; pass < 2; pass++) {
        if (pass == 1) {
            uint32_t acc = 0;
            for (size_t c = 0; c <= ncell; c++) {
                uint32_t v = count[c];
                count[c] = acc;
                acc += v;
            }
            items.assign(count[ncell], 0);
            start = count;
        }
        for (size_t i = 0; i < n; i++) {
            int x0, x1, y0, y1, z0, z1;
            range(i, 0, x0, x1);
            range(i, 1, y0, y1);
            range(i, 2, z0, z1);
            for (int z = z0; z <= z1; z++)
                for (int y = y0; y <= y1; y++)
                    for (int x = x0; x <= x1; x++) {
                        size_t c = ((size_t)z * dims[1] + y) * dims[0] + x;
                        if (pass == 0)
                            count[c]++;
                        else
                            items[count[c]++] = (uint32_t)(first + i);
                    }
        }
    }
    // count[] now holds the cell ends; start[] the cell starts
    int r;
    if ((r = d_start.reserve(start.size() * 4)) || (r = d_items.reserve(std::max<size_t>(1, items.size()) * 4)))
        return r;
    if ((r = eng.send(d_start.p, start.data(), start.size() * 4)) ||
        (r = eng.send(d_items.p, items.data(), items.size() * 4)))
        return r;
    return PVX_OK;
}

int RadCache::upload(RadiosityEngine &eng) {
    if (uploaded == recs.size()) return PVX_OK;
    size_t bytes = recs.size() * sizeof(RadRec);
    if (bytes > d_recs.cap) { // grow (doubling) and send everything
        int r;
        if ((r = d_recs.reserve(std::max(bytes, 2 * d_recs.cap)))) return r;
        uploaded = 0;
    }
    int r2;
    if ((r2 = eng.send(static_cast<RadRec *>(d_recs.p) + uploaded, recs.data() + uploaded,
                       (recs.size() - uploaded) * sizeof(RadRec))))
        return r2;
    uploaded = recs.size();
    return PVX_OK;
}

int RadCache::index_all(RadiosityEngine &eng) {
    if (indexed == recs.size()) return PVX_OK;
    int r;
    if ((r = upload(eng)) || (r = all.build(recs, 0, recs.size(), reach, eng))) return r;
    indexed = recs.size();
    return PVX_OK;
}

RadiosityEngine::RadiosityEngine(pvx_scene *s) : scene(s) {
    pool.resize(3 * kPoolSize);
    for (int i = 0; i < kPoolSize; i++) { // HaltonCosWeightedDirectionGenerator (bases 2, 3)
        auto halton = [](size_t index, unsigned base) {
            size_t k = 1 + index;
            double h = 0, q = 1.0 / base;
            while (k > 0) {
                h = h + (unsigned)(k % base) * q;
                k /= base;
                q /= base;
            }
            return h;
        };
        double r = std::sqrt(halton(i, 2));
        double theta = halton(i, 3) * (2 * M_PI);
        double x = r * std::cos(theta), z = r * std::sin(theta);
        pool[3 * i] = x;
        pool[3 * i + 1] = std::sqrt(1 - x * x - z * z);
        pool[3 * i + 2] = z;
    }
}

int RadiosityEngine::begin(const pvx_radiosity_settings &set, const pvx_camera &c, hipStream_t st) {
    rs = set;
    cam = c;
    stream = st;
    for (int d = 0; d < kMaxRadDepth; d++) {
        RadCache &cc = caches[d];
        cc.recs = preload[d]; // a loaded cache file seeds the frame (RadiosityCache::Load)
        cc.uploaded = cc.indexed = 0;
        cc.all.n = cc.fresh.n = 0;
        cc.reach = (float)(rs.error_bound + kWeightErrOffset) * std::pow(2.0f, (float)d);
        cursor[d] = 0;
        stats_rounds[d] = 0;
    }
    stats = pvx_radiosity_stats{};
    amb_net = nullptr;
    train_ms = t_sync_ms = t_grid_ms = 0.0;
    n_sync = 0;
    up_off = 0;
    if (!d_pool.p) {
        int r;
        if ((r = d_pool.reserve(pool.size() * 8))) return r;
        HIP_TRY(hipMemcpy(d_pool.p, pool.data(), pool.size() * 8, hipMemcpyHostToDevice));
    }
    return PVX_OK;
}

// Host -> device through the pinned staging buffer; regions are reused only
// after the next wait on the stream, when every earlier copy has finished.
int RadiosityEngine::send(void *dev, const void *host, size_t bytes) {
    if (bytes == 0) return PVX_OK;
    const size_t need = (bytes + 255) & ~(size_t)255;
    if (up_off + need > up.cap) {
        HIP_TRY(sync()); // every copy from the buffer has finished
        if (need > up.cap) {
            int r;
            if ((r = up.reserve(std::max(need, 2 * up.cap)))) return r;
        }
    }
    char *p = static_cast<char *>(up.p) + up_off;
    std::memcpy(p, host, bytes);
    HIP_TRY(hipMemcpyAsync(dev, p, bytes, hipMemcpyHostToDevice, stream));
    up_off += need;
    return PVX_OK;
}


// Device -> host through a pinned bounce buffer (an asynchronous copy and
// one wait, instead of a pageable copy that blocks inside the runtime).
int RadiosityEngine::fetch(void *host, const void *dev, size_t bytes) {
    if (bytes == 0) return PVX_OK;
    int r;
    if ((r = bounce.reserve(bytes))) return r;
    HIP_TRY(hipMemcpyAsync(bounce.p, dev, bytes, hipMemcpyDeviceToHost, stream));
    HIP_TRY(sync());
    std::memcpy(host, bounce.p, bytes);
    return PVX_OK;
}

HostBuf::~HostBuf() {
    if (p) (void)hipHostFree(p);
}
int HostBuf::reserve(size_t bytes) {
    if (bytes <= cap) return PVX_OK;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
    size_t want = std::max(bytes, (size_t)1 << 20);
    HIP_TRY(hipHostMalloc(&p, want, hipHostMallocDefault));
    cap = want;
    return PVX_OK;
}

hipError_t RadiosityEngine::sync() {
    auto t0 = std::chrono::steady_clock::now();
    hipError_t e = hipStreamSynchronize(stream);
    if (e == hipSuccess) up_off = 0;
    t_sync_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    n_sync++;
    return e;
}

int RadiosityEngine::rays_per_sample(int d) const {
    return std::max(5, (int)(rs.count * std::pow(0.5, (double)d)));
}
float RadiosityEngine::reuse_count(int d) const {
    return (float)(d == 0 ? (int)rs.nearest_count : (d == 1 ? std::min(2, (int)rs.nearest_count) : 1));
}

int RadiosityEngine::lookup(int depth, const RadGrid &g, const RadQuery *q, const uint32_t *idx, uint32_t n,
                            bool final, bool accumulate, RadAcc *acc, const uint8_t *mask) {
    if (n == 0) return PVX_OK;
    if (accumulate && g.n == 0) return PVX_OK;
    RadCache &c = caches[depth];
    CacheView v;
    v.recs = static_cast<const RadRec *>(c.d_recs.p);
    v.cell_start = static_cast<const uint32_t *>(g.d_start.p);
    v.cell_items = static_cast<const uint32_t *>(g.d_items.p);
    for (int a = 0; a < 3; a++) {
        v.lo[a] = g.lo[a];
        v.dims[a] = g.dims[a];
    }
    v.cell = g.cell;
    v.n = g.n;
    k_lookup<<<blocks(n), 256, 0, stream>>>(q, idx, n, v, rs.error_bound, std::pow(2.0f, (float)depth),
                                            (final && depth == 0) ? 1 : 0, accumulate ? 1 : 0, acc, mask);
    HIP_TRY(hipGetLastError());
    return PVX_OK;
}

// Gather at `depth` for the m queries q[idx[*]]; writes their irradiance to Eg
// (3m floats, device) and appends the new records to the depth cache.
int RadiosityEngine::gather(int depth, const RadQuery *q, const uint32_t *idx, uint32_t m, bool final,
                            float *Eg) {
    if (m == 0) return PVX_OK;
    int r;
    if ((r = gather_trace(depth, q, idx, m))) return r;
    return gather_finish(depth, q, idx, m, final, Eg);
}

// First half of a gather: the rays, their closest hits and what each hit
// contributes (scratch[depth]: base, coef and the depth+1 queries).
int RadiosityEngine::gather_trace(int depth, const RadQuery *q, const uint32_t *idx, uint32_t m) {
    const uint32_t N = (uint32_t)rays_per_sample(depth);
    const size_t nr = (size_t)m * N;
    Scratch &S = scratch[depth];
    int r;
    if ((r = S.rays.reserve(nr * sizeof(pvx_ray))) || (r = S.hits.reserve(nr * sizeof(pvx_hit))) ||
        (r = S.base.reserve(nr * 12)) || (r = S.coef.reserve(nr * 12)) || (r = S.amb.reserve(nr * 12)) ||
        (r = S.q.reserve(nr * sizeof(RadQuery))) || (r = S.need.reserve(nr)) ||
        (r = S.rec.reserve(m * sizeof(RadRec))))
        return r;
    k_gather_rays<<<blocks(nr), 256, 0, stream>>>(q, idx, m, N, static_cast<const double *>(d_pool.p),
                                                  cursor[depth], static_cast<pvx_ray *>(S.rays.p));
    HIP_TRY(hipGetLastError());
    cursor[depth] = (uint32_t)((cursor[depth] + (uint64_t)m * N) % kPoolSize);
    if ((r = trace_closest_internal(scene, static_cast<const pvx_ray *>(S.rays.p), nullptr,
                                    PVX_RAY_RADIOSITY, nr, static_cast<pvx_hit *>(S.hits.p), stream)))
        return r;
    stats.gathers[std::min(depth, 1)] += m;
    stats.gather_rays += nr;
    bool deeper = depth + 1 < (int)rs.recursion_limit;
    float wnext = std::max(rs.adc_bailout + 1e-10f, (float)std::pow(0.5 * rs.brightness, depth + 1.0));
    ShadeParams sp;
    sp.sv = scene->view;
    sp.obj_type = scene->d_obj_type;
    sp.obj_index = scene->d_obj_index;
    sp.mats = scene->shading.d_mats;
    sp.adc = rs.adc_bailout;
    (scene->view.has_complex >= 3 ? k_hit_queries<true> : k_hit_queries<false>)<<<blocks(nr), 256, 0, stream>>>(sp, static_cast<const pvx_ray *>(S.rays.p),
                                                  static_cast<const pvx_hit *>(S.hits.p), (uint32_t)nr,
                                                  wnext, deeper ? 1 : 0, static_cast<float *>(S.base.p),
                                                  static_cast<float *>(S.coef.p),
                                                  static_cast<RadQuery *>(S.q.p),
                                                  static_cast<uint8_t *>(S.need.p));
    HIP_TRY(hipGetLastError());
    return PVX_OK;
}

// Second half: the depth+1 irradiance of the hits (resolved with the
// depth+1 cache, or from `amb_net` at depth 0), GatherLight's reduction,
// and the new records (AddBlock).
int RadiosityEngine::gather_finish(int depth, const RadQuery *q, const uint32_t *idx, uint32_t m, bool final,
                                   float *Eg) {
    const uint32_t N = (uint32_t)rays_per_sample(depth);
    const size_t nr = (size_t)m * N;
    Scratch &S = scratch[depth];
    int r;
    bool deeper = depth + 1 < (int)rs.recursion_limit;
    k_fill_f32<<<blocks(nr * 3), 256, 0, stream>>>(static_cast<float *>(S.amb.p), nr * 3);
    HIP_TRY(hipGetLastError());
    if (deeper && depth == 0 && amb_net) {
        // depth+1 irradiance of every hit from the network (reusing this
        // frame's sparse depth+1 records as well measured no better, at the
        // cost of a lookup over scattered hits)
        if ((r = ax.reserve(nr * 24)) || (r = ay.reserve(nr * 12))) return r;
        k_query_to_x6<<<blocks(nr), 256, 0, stream>>>(static_cast<const RadQuery *>(S.q.p), nullptr,
                                                      static_cast<const uint8_t *>(S.need.p), (uint32_t)nr,
                                                      nlo[0], nlo[1], nlo[2], ninv, static_cast<float *>(ax.p));
        HIP_TRY(hipGetLastError());
        if ((r = amb_net->infer(static_cast<const float *>(ax.p), (uint32_t)nr, static_cast<float *>(ay.p),
                                stream)))
            return r;
        k_amb_from_pred<<<blocks(nr), 256, 0, stream>>>(static_cast<const uint8_t *>(S.need.p),
                                                        static_cast<const float *>(ay.p), (uint32_t)nr,
                                                        rs.brightness, static_cast<float *>(S.amb.p));
        HIP_TRY(hipGetLastError());
    } else if (deeper) {
        // compact the hits that need irradiance at depth + 1 (deterministic order)
        std::vector<uint8_t> need(nr);
        if ((r = fetch(need.data(), S.need.p, nr))) return r;
        std::vector<uint32_t> sel;
        sel.reserve(nr);
        for (size_t k = 0; k < nr; k++)
            if (need[k]) sel.push_back((uint32_t)k);
        if (!sel.empty()) {
            DevBuf &di = S.sel;
            if ((r = di.reserve(sel.size() * 4)) || (r = S.subq.reserve(sel.size() * sizeof(RadQuery))) ||
                (r = S.subE.reserve(sel.size() * 12)))
                return r;
            if ((r = send(di.p, sel.data(), sel.size() * 4))) return r;
            k_gather_subq<<<blocks(sel.size()), 256, 0, stream>>>(
                static_cast<const RadQuery *>(S.q.p), static_cast<const uint32_t *>(di.p),
                (uint32_t)sel.size(), static_cast<RadQuery *>(S.subq.p));
            if ((r = resolve(depth + 1, static_cast<const RadQuery *>(S.subq.p), (uint32_t)sel.size(), final,
                             static_cast<float *>(S.subE.p))))
                return r;
            k_scatter3<<<blocks(sel.size()), 256, 0, stream>>>(
                static_cast<const float *>(S.subE.p), static_cast<const uint32_t *>(di.p),
                (uint32_t)sel.size(), static_cast<float *>(S.amb.p));
            HIP_TRY(hipGetLastError());
        }
    }
    float factor = std::pow(2.0f, (float)depth);
    k_gather_reduce<<<blocks((size_t)m * 64), 256, 0, stream>>>(
        q, idx, m, N, static_cast<const pvx_ray *>(S.rays.p), static_cast<const pvx_hit *>(S.hits.p),
        static_cast<const float *>(S.base.p), static_cast<const float *>(S.coef.p),
        static_cast<const float *>(S.amb.p), cam.location[0], cam.location[1], cam.location[2],
        rs.minimum_reuse, rs.maximum_reuse, factor, Eg, static_cast<RadRec *>(S.rec.p));
    HIP_TRY(hipGetLastError());
    // append the records to the cache (AddBlock)
    std::vector<RadRec> &rec = last_fetched[depth];
    rec.resize(m);
    if ((r = fetch(rec.data(), S.rec.p, m * sizeof(RadRec)))) return r;
    RadCache &c = caches[depth];
    std::vector<int32_t> &lr = last_rec[depth];
    lr.assign(m, -1);
    for (uint32_t k = 0; k < m; k++)
        if (rec[k].pad[1] != 0.0f) {
            lr[k] = (int32_t)c.recs.size();
            c.recs.push_back(rec[k]);
        }
    return c.upload(*this);
}

// Spread-out subset of the pending queries: the first pending point of
// every grid cell whose edge is the typical reach of a record (median
// harmonic-mean radius of the records so far times the error bound), so one
// round's gathers do not crowd each other and the rest of each cell can
// reuse them, as with the sequential order.
void RadiosityEngine::spread_select(int depth, const std::vector<uint32_t> &todo, const std::vector<uint32_t> &slot,
                                    const std::vector<RadQuery> &hq, std::vector<uint32_t> &sel,
                                    std::vector<uint32_t> &sel_slot, std::vector<uint32_t> &rest,
                                    std::vector<uint32_t> &rest_slot) const {
    const RadCache &c = caches[depth];
    const double factor = std::pow(2.0, (double)depth);
    std::vector<double> rad;
    if (!c.recs.empty()) {
        size_t stride = std::max<size_t>(1, c.recs.size() / 4096);
        for (size_t k = 0; k < c.recs.size(); k += stride) rad.push_back(c.recs[k].r);
    } else { // nothing gathered yet at this depth: half the largest reach
        size_t stride = std::max<size_t>(1, todo.size() / 4096);
        for (size_t k = 0; k < todo.size(); k += stride) {
            const RadQuery &a = hq[slot[k]];
            double ex = cam.location[0] - a.p[0], ey = cam.location[1] - a.p[1], ez = cam.location[2] - a.p[2];
            rad.push_back(0.5 * std::sqrt(ex * ex + ey * ey + ez * ez) * rs.maximum_reuse * factor);
        }
    }
    std::nth_element(rad.begin(), rad.begin() + rad.size() / 2, rad.end());
    const double D = std::max(rad[rad.size() / 2] * rs.error_bound * factor, 1e-9);
    // open-addressing set of occupied cells (21-bit cell coordinates)
    size_t cap = 16;
    while (cap < 2 * todo.size()) cap <<= 1;
    std::vector<uint64_t> table(cap, 0);
    double pmax = 0;
    for (size_t k = 0; k < todo.size(); k += std::max<size_t>(1, todo.size() / 256))
        for (int d = 0; d < 3; d++) pmax = std::max(pmax, std::fabs(hq[slot[k]].p[d]));
    const double Dc = std::max(D, (pmax + 1.0) * 2e-6);
    for (size_t k = 0; k < todo.size(); k++) {
        const RadQuery &a = hq[slot[k]];
        uint64_t key = 1; // never 0: 0 marks an empty slot
        for (int d = 0; d < 3; d++) {
            int64_t cidx = (int64_t)std::floor(a.p[d] / Dc) + (1 << 20);
            key = (key << 21) | ((uint64_t)cidx & 0x1FFFFFu);
        }
        uint64_t h = key * 0x9E3779B97F4A7C15ull;
        size_t pos = (size_t)(h >> 20) & (cap - 1);
        bool fresh = true;
        while (table[pos]) {
            if (table[pos] == key) {
                fresh = false;
                break;
            }
            pos = (pos + 1) & (cap - 1);
        }
        if (fresh) table[pos] = key;
        if (fresh) {
            sel.push_back(todo[k]);
            sel_slot.push_back(slot[k]);
        } else {
            rest.push_back(todo[k]);
            rest_slot.push_back(slot[k]);
        }
    }
}

// One round: gather at the `sel` queries, combine with their partial reuse
// (ComputeAmbient, radiosity.cpp:430-440) and file the new records.
int RadiosityEngine::gather_round(int depth, const RadQuery *q, const std::vector<uint32_t> &sel, bool final,
                                  RadAcc *acc, float *E) {
    if (sel.empty()) return PVX_OK;
    Resolve &R = rstate[depth];
    uint32_t *d_idx = static_cast<uint32_t *>(R.idx.p);
    stats_rounds[depth]++;
    int r;
    if ((r = send(d_idx, sel.data(), sel.size() * 4))) return r;
    if ((r = gather(depth, q, d_idx, (uint32_t)sel.size(), final, static_cast<float *>(R.Eg.p)))) return r;
    k_combine<<<blocks(sel.size()), 256, 0, stream>>>(d_idx, (uint32_t)sel.size(),
                                                      static_cast<const float *>(R.Eg.p), acc, E);
    HIP_TRY(hipGetLastError());
    return PVX_OK;
}

// The `rest` queries meet the records filed since `first_new`; those that
// still need a gather go back to todo (with their hq slots).
int RadiosityEngine::relookup(int depth, const RadQuery *q, size_t first_new, bool final, RadAcc *acc,
                              const std::vector<uint32_t> &rest, const std::vector<uint32_t> &rest_slot,
                              std::vector<uint32_t> &todo, std::vector<uint32_t> &slot) {
    todo.clear();
    slot.clear();
    if (rest.empty()) return PVX_OK;
    Resolve &R = rstate[depth];
    RadCache &c = caches[depth];
    uint32_t *d_idx = static_cast<uint32_t *>(R.idx.p);
    uint8_t *d_need = static_cast<uint8_t *>(R.need.p);
    int r;
    auto tg = std::chrono::steady_clock::now();
    if ((r = c.fresh.build(c.recs, first_new, c.recs.size(), c.reach, *this))) return r;
    t_grid_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tg).count();
    if ((r = send(d_idx, rest.data(), rest.size() * 4))) return r;
    if ((r = lookup(depth, c.fresh, q, d_idx, (uint32_t)rest.size(), final, true, acc))) return r;
    k_need<<<blocks(rest.size()), 256, 0, stream>>>(acc, d_idx, (uint32_t)rest.size(), reuse_count(depth),
                                                    final ? 1 : 0, d_need);
    std::vector<uint8_t> need(rest.size());
    if ((r = fetch(need.data(), d_need, rest.size()))) return r;
    for (size_t k = 0; k < rest.size(); k++)
        if (need[k]) {
            todo.push_back(rest[k]);
            slot.push_back(rest_slot[k]);
        }
    return PVX_OK;
}

// Gathers rounds until every pending query has its irradiance.
int RadiosityEngine::gather_rounds(int depth, const RadQuery *q, bool final, std::vector<uint32_t> todo,
                                   std::vector<uint32_t> slot, const std::vector<RadQuery> &hq, RadAcc *acc,
                                   float *E, int max_rounds) {
    RadCache &c = caches[depth];
    int r;
    for (int round = 0; !todo.empty(); round++) {
        std::vector<uint32_t> sel, sel_slot, rest, rest_slot;
        if (round < max_rounds - 1 && todo.size() > 64)
            spread_select(depth, todo, slot, hq, sel, sel_slot, rest, rest_slot);
        else
            sel.swap(todo);
        const size_t first_new = c.recs.size();
        if ((r = gather_round(depth, q, sel, final, acc, E))) return r;
        if ((r = relookup(depth, q, first_new, final, acc, rest, rest_slot, todo, slot))) return r;
    }
    return PVX_OK;
}

// Initial lookup of n queries against the whole depth cache; returns the
// queries that need a gather and their positions (hq, indexed by slot).
int RadiosityEngine::start_resolve(int depth, const RadQuery *q, uint32_t n, bool final, RadAcc *&acc,
                                   std::vector<uint32_t> &todo, std::vector<uint32_t> &slot,
                                   std::vector<RadQuery> &hq) {
    Resolve &R = rstate[depth];
    RadCache &c = caches[depth];
    int r;
    if ((r = R.acc.reserve((size_t)n * sizeof(RadAcc))) || (r = R.need.reserve(n)) ||
        (r = R.idx.reserve((size_t)n * 4)) || (r = R.Eg.reserve((size_t)n * 12)) ||
        (r = R.tq.reserve((size_t)n * sizeof(RadQuery))))
        return r;
    acc = static_cast<RadAcc *>(R.acc.p);
    uint32_t *d_idx = static_cast<uint32_t *>(R.idx.p);
    uint8_t *d_need = static_cast<uint8_t *>(R.need.p);
    auto tg = std::chrono::steady_clock::now();
    if ((r = c.index_all(*this))) return r;
    t_grid_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tg).count();
    if ((r = lookup(depth, c.all, q, nullptr, n, final, false, acc))) return r;
    k_need<<<blocks(n), 256, 0, stream>>>(acc, nullptr, n, reuse_count(depth), final ? 1 : 0, d_need);
    std::vector<uint8_t> need(n);
    if ((r = fetch(need.data(), d_need, n))) return r;
    todo.clear();
    for (uint32_t i = 0; i < n; i++)
        if (need[i]) todo.push_back(i);
    if (depth == 0 && final) stats.final_misses += todo.size();
    slot.resize(todo.size());
    for (size_t k = 0; k < todo.size(); k++) slot[k] = (uint32_t)k;
    hq.clear();
    if (!todo.empty()) {
        if ((r = send(d_idx, todo.data(), todo.size() * 4))) return r;
        k_gather_subq<<<blocks(todo.size()), 256, 0, stream>>>(q, d_idx, (uint32_t)todo.size(),
                                                               static_cast<RadQuery *>(R.tq.p));
        hq.resize(todo.size());
        if ((r = fetch(hq.data(), R.tq.p, todo.size() * sizeof(RadQuery)))) return r;
    }
    return PVX_OK;
}

// ComputeAmbient for n queries at `depth`: reuse the cache where it has
// enough weight, gather the rest in rounds (spread-out subsets first).
int RadiosityEngine::resolve(int depth, const RadQuery *q, uint32_t n, bool final, float *E) {
    if (n == 0) return PVX_OK;
    RadAcc *acc = nullptr;
    std::vector<uint32_t> todo, slot;
    std::vector<RadQuery> hq;
    int r;
    if ((r = start_resolve(depth, q, n, final, acc, todo, slot, hq)) ||
        (r = gather_rounds(depth, q, final, std::move(todo), std::move(slot), hq, acc, E)))
        return r;
    k_finish<<<blocks(n), 256, 0, stream>>>(acc, n, rs.brightness, E);
    HIP_TRY(hipGetLastError());
    return PVX_OK;
}

// Trains `nt` from scratch on n records (device), inputs normalised by
// (nlo, ninv); adds the time to train_ms.
int RadiosityEngine::train_on(Surrogate &nt, const RadRec *recs, uint32_t n, uint64_t seed, float *loss) {
    if (n == 0) return PVX_OK;
    int r;
    if ((r = lx.reserve((size_t)n * 24)) || (r = lt.reserve((size_t)n * 12))) return r;
    k_rec_to_sample<<<blocks(n), 256, 0, stream>>>(recs, nullptr, n, nlo[0], nlo[1], nlo[2], ninv,
                                                    static_cast<float *>(lx.p), static_cast<float *>(lt.p));
    HIP_TRY(hipGetLastError());
    hipEvent_t t0, t1;
    HIP_TRY(hipEventCreate(&t0));
    HIP_TRY(hipEventCreate(&t1));
    HIP_TRY(hipEventRecord(t0, stream));
    // samples drawn with replacement from the n records per step; the rate
    // grows with the square root of the batch
    const uint32_t batch = rs.train_batch ? rs.train_batch : 2048;
    const float lr = 5e-3f * std::sqrt((float)batch / 2048.0f);
    if ((r = nt.init(nullptr, seed, stream)) ||
        (r = nt.train(static_cast<const float *>(lx.p), static_cast<const float *>(lt.p), n, rs.train_steps, batch,
                      lr, loss, stream))) {
        (void)hipEventDestroy(t0);
        (void)hipEventDestroy(t1);
        return r;
    }
    HIP_TRY(hipEventRecord(t1, stream));
    HIP_TRY(hipEventSynchronize(t1));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, t0, t1));
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    train_ms += ms;
    return PVX_OK;
}

// Probe test: for the m queries q[d_idx[*]] whose plain gathers gave the
// records pr[*] (standard error of the gather in pad[0]), the grey
// difference between the network and the gather beyond twice the gather's
// standard error, times the query weight (the surface's diffuse
// reflectance, times 1/2 per bounce), i.e. in displayed colour.
int RadiosityEngine::probe_errors(Surrogate &nt, const RadQuery *q, const uint32_t *d_idx, uint32_t m,
                                  const RadRec *pr, std::vector<double> &err) {
    err.assign(m, 0.0);
    if (m == 0) return PVX_OK;
    int r;
    if ((r = px6.reserve((size_t)m * 24)) || (r = py.reserve((size_t)m * 12)) ||
        (r = pq.reserve((size_t)m * sizeof(RadQuery))))
        return r;
    k_query_to_x6<<<blocks(m), 256, 0, stream>>>(q, d_idx, nullptr, m, nlo[0], nlo[1], nlo[2], ninv,
                                                 static_cast<float *>(px6.p));
    k_gather_subq<<<blocks(m), 256, 0, stream>>>(q, d_idx, m, static_cast<RadQuery *>(pq.p));
    HIP_TRY(hipGetLastError());
    if ((r = nt.infer(static_cast<const float *>(px6.p), m, static_cast<float *>(py.p), stream))) return r;
    std::vector<float> y(3 * (size_t)m);
    std::vector<RadQuery> hq(m);
    if ((r = fetch(y.data(), py.p, y.size() * 4)) || (r = fetch(hq.data(), pq.p, m * sizeof(RadQuery)))) return r;
    const double gw[3] = {0.297, 0.589, 0.114};
    for (uint32_t k = 0; k < m; k++) {
        double dg = 0;
        for (int a = 0; a < 3; a++) dg += gw[a] * ((double)std::max(y[3 * (size_t)k + a], 0.0f) - pr[k].E[a]);
        err[k] = std::max(0.0, std::fabs(dg) - 2.0 * pr[k].pad[0]) * hq[k].w * rs.brightness;
    }
    return PVX_OK;
}

// Indices i < n with flags[i] != 0, ascending (device `out`); count to the host.
int RadiosityEngine::compact(const uint8_t *flags, uint32_t n, uint32_t *out, uint32_t &count) {
    count = 0;
    if (n == 0) return PVX_OK;
    int r;
    if ((r = d_count.reserve(4))) return r;
    rocprim::counting_iterator<uint32_t> it(0u);
    size_t tb = 0;
    HIP_TRY(rocprim::select(nullptr, tb, it, flags, out, static_cast<uint32_t *>(d_count.p), (size_t)n, stream));
    if ((r = cub_tmp.reserve(std::max<size_t>(tb, 16)))) return r;
    HIP_TRY(rocprim::select(cub_tmp.p, tb, it, flags, out, static_cast<uint32_t *>(d_count.p), (size_t)n, stream));
    return fetch(&count, d_count.p, 4);
}

// Final-pass irradiance with the surrogate (learned mode; no pretrace).
//  1. Plain depth-0 gathers at a spread-out training set (per cell of about
//     W*H / train_samples pixels, the point with the lowest pixel hash) and
//     at probes (up to probes_per_tile per probe_tile^2 image tile), all in
//     one batch.
//  2. Their rays' hits want depth-1 irradiance: plain depth-1 gathers at a
//     spread-out subset of those hits (every 17th held out as a probe)
//     train a depth-1 network, which then answers every hit of the batch —
//     unless its probes disagree, in which case depth 1 is resolved the
//     plain way (irradiance cache and gather rounds).
//  3. The depth-0 training gathers train the depth-0 network.  Every query
//     first reuses this frame's depth-0 records as the plain final pass
//     would; the misses take the network's answer, except in tiles whose
//     probes disagree by more than probe_threshold, which fall back to the
//     plain method (gather rounds over the same cache).
// Each network is scored only on plain gathers it was not trained on.
int RadiosityEngine::resolve_learned(const RadQuery *q, uint32_t n, const uint32_t *d_pix, float *E) {
    if (n == 0) return PVX_OK;
    int r;
    Resolve &R = rstate[0];
    const uint32_t W = cam.width, H = cam.height;
    const uint32_t T = std::max<uint32_t>(1, rs.train_samples);
    const uint32_t cpx = std::max<uint32_t>(1, (uint32_t)std::floor(std::sqrt((double)W * H / T)));
    const uint32_t ncx = (W + cpx - 1) / cpx, ncy = (H + cpx - 1) / cpx;
    const size_t ncell = (size_t)ncx * ncy;
    const uint32_t tsz = std::max<uint32_t>(1, rs.probe_tile);
    const uint32_t tx = (W + tsz - 1) / tsz, ty = (H + tsz - 1) / tsz;
    const size_t ntile = (size_t)tx * ty;
    if ((r = R.acc.reserve((size_t)n * sizeof(RadAcc))) || (r = R.need.reserve(n)) ||
        (r = R.idx.reserve((size_t)n * 4)) || (r = R.Eg.reserve((size_t)n * 12)) ||
        (r = R.tq.reserve((size_t)n * sizeof(RadQuery))) || (r = lrole.reserve(n)) ||
        (r = lkeys.reserve(std::max(ncell, ntile) * 8)) || (r = lfbflag.reserve(n)) ||
        (r = lfbtile.reserve(ntile)) || (r = lcnt.reserve(8)) || (r = lx.reserve((size_t)n * 24)) ||
        (r = ly.reserve((size_t)n * 12)))
        return r;
    RadAcc *acc = static_cast<RadAcc *>(R.acc.p);
    uint8_t *role = static_cast<uint8_t *>(lrole.p);
    auto *keys = static_cast<unsigned long long *>(lkeys.p);

    // (1) training points and probes
    k_fill_u64<<<blocks(ncell), 256, 0, stream>>>(keys, ncell, ~0ull);
    k_cell_keys<<<blocks(n), 256, 0, stream>>>(d_pix, n, W, cpx, ncx, keys);
    k_fill_u8<<<blocks(n), 256, 0, stream>>>(role, n);
    HIP_TRY(hipGetLastError());
    std::vector<unsigned long long> hk(std::max(ncell, ntile));
    if ((r = fetch(hk.data(), keys, ncell * 8))) return r;
    std::vector<uint32_t> g0;
    for (size_t c = 0; c < ncell; c++)
        if (hk[c] != ~0ull) g0.push_back((uint32_t)hk[c]);
    const uint32_t n_train = (uint32_t)g0.size();
    if ((r = lg0.reserve(((size_t)n_train + ntile * 16 + 1) * 4))) return r;
    uint32_t *d_g0 = static_cast<uint32_t *>(lg0.p);
    if ((r = send(d_g0, g0.data(), (size_t)n_train * 4))) return r;
    k_mark_role<<<blocks(n_train), 256, 0, stream>>>(d_g0, n_train, 1, role);
    std::vector<uint32_t> probe_tile;
    const uint32_t P = std::min<uint32_t>(rs.probes_per_tile, 16);
    for (uint32_t j = 0; j < P; j++) {
        k_fill_u64<<<blocks(ntile), 256, 0, stream>>>(keys, ntile, ~0ull);
        k_tile_keys<<<blocks(n), 256, 0, stream>>>(d_pix, role, n, W, tsz, tx, 0x5bd1e995u * (j + 1), keys);
        HIP_TRY(hipGetLastError());
        if ((r = fetch(hk.data(), keys, ntile * 8))) return r;
        const size_t first = g0.size();
        for (size_t t = 0; t < ntile; t++)
            if (hk[t] != ~0ull) {
                g0.push_back((uint32_t)hk[t]);
                probe_tile.push_back((uint32_t)t);
            }
        const uint32_t add = (uint32_t)(g0.size() - first);
        if (add == 0) break;
        if ((r = send(d_g0 + first, g0.data() + first, (size_t)add * 4))) return r;
        k_mark_role<<<blocks(add), 256, 0, stream>>>(d_g0 + first, add, 2, role);
    }
    const uint32_t m0 = (uint32_t)g0.size();
    const uint32_t n_probe = m0 - n_train;
    stats.probes += n_probe;

    // (2) the depth-0 gathers' rays; normalisation from the gathered points
    if ((r = gather_trace(0, q, d_g0, m0))) return r;
    std::vector<RadQuery> hq0(m0);
    k_gather_subq<<<blocks(m0), 256, 0, stream>>>(q, d_g0, m0, static_cast<RadQuery *>(R.tq.p));
    HIP_TRY(hipGetLastError());
    if ((r = fetch(hq0.data(), R.tq.p, (size_t)m0 * sizeof(RadQuery)))) return r;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    auto grow = [&](const float *pp) {
        for (int a = 0; a < 3; a++) {
            lo[a] = std::min(lo[a], pp[a]);
            hi[a] = std::max(hi[a], pp[a]);
        }
    };
    for (const RadQuery &a : hq0) {
        float pp[3] = {(float)a.p[0], (float)a.p[1], (float)a.p[2]};
        grow(pp);
    }
    auto set_norm = [&]() {
        float ext = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
        ext = std::max(ext * 1.04f, 1e-6f);
        for (int a = 0; a < 3; a++) nlo[a] = lo[a] - 0.02f * ext;
        ninv = 1.0f / (ext * 1.02f);
    };
    set_norm();
    amb_net = nullptr;
    Scratch &S0 = scratch[0];
    const size_t nr0 = (size_t)m0 * rays_per_sample(0);
    if (rs.recursion_limit > 1) {
        std::vector<uint8_t> need(nr0);
        if ((r = fetch(need.data(), S0.need.p, nr0))) return r;
        size_t cntn = 0;
        for (size_t k = 0; k < nr0; k++) cntn += need[k];
        const uint32_t want = T + (P > 0 ? std::max<uint32_t>(1, T / 16) : 0);
        std::vector<uint32_t> tr1, pr1;
        if (cntn) {
            const uint32_t cnt = (uint32_t)std::min<size_t>(want, cntn);
            const double stride = (double)cntn / cnt;
            size_t seen = 0, next = 0;
            uint32_t j = 0;
            for (size_t k = 0; k < nr0 && j < cnt; k++) {
                if (!need[k]) continue;
                if (seen == next) {
                    (P > 0 && j % 17 == 16 ? pr1 : tr1).push_back((uint32_t)k);
                    j++;
                    next = (size_t)(j * stride);
                }
                seen++;
            }
        }
        if (!tr1.empty()) {
            std::vector<uint32_t> g1 = tr1;
            g1.insert(g1.end(), pr1.begin(), pr1.end());
            const uint32_t m1 = (uint32_t)g1.size();
            Resolve &R1 = rstate[1];
            if ((r = lg1.reserve((size_t)m1 * 4)) || (r = R1.Eg.reserve((size_t)m1 * 12))) return r;
            uint32_t *d_g1 = static_cast<uint32_t *>(lg1.p);
            if ((r = send(d_g1, g1.data(), (size_t)m1 * 4))) return r;
            const RadQuery *q1 = static_cast<const RadQuery *>(S0.q.p);
            if ((r = gather(1, q1, d_g1, m1, true, static_cast<float *>(R1.Eg.p)))) return r;
            const std::vector<RadRec> rec1 = last_fetched[1];
            for (const RadRec &b : rec1) grow(b.p);
            set_norm();
            float loss1 = 0.0f;
            if ((r = train_on(net1, static_cast<const RadRec *>(scratch[1].rec.p), (uint32_t)tr1.size(),
                              0xC0FFEEull + 1, &loss1)))
                return r;
            bool ok = true;
            if (!pr1.empty()) {
                std::vector<double> e1;
                if ((r = probe_errors(net1, q1, d_g1 + tr1.size(), (uint32_t)pr1.size(), rec1.data() + tr1.size(),
                                      e1)))
                    return r;
                double tot = 0;
                for (double e : e1) tot += e;
                ok = tot / e1.size() <= rs.probe_threshold;
                stats.probes += pr1.size();
            }
            if (ok) amb_net = &net1;
        }
    }
    if ((r = gather_finish(0, q, d_g0, m0, true, static_cast<float *>(R.Eg.p)))) return r;
    const std::vector<RadRec> rec0 = last_fetched[0];

    // (3) the depth-0 network and its probes
    float loss = 0.0f;
    if ((r = train_on(net, static_cast<const RadRec *>(S0.rec.p), n_train, 0xC0FFEEull, &loss))) return r;
    stats.train_loss = loss;
    std::vector<uint8_t> fb_tile(ntile, 0);
    if (n_probe) {
        std::vector<double> e0;
        if ((r = probe_errors(net, q, d_g0 + n_train, n_probe, rec0.data() + n_train, e0))) return r;
        std::vector<double> err(ntile, 0.0);
        std::vector<uint32_t> cnt(ntile, 0);
        double tot = 0;
        for (uint32_t k = 0; k < n_probe; k++) {
            err[probe_tile[k]] += e0[k];
            cnt[probe_tile[k]]++;
            tot += e0[k];
        }
        stats.probe_error = (float)(tot / n_probe);
        for (size_t t = 0; t < ntile; t++) fb_tile[t] = cnt[t] && err[t] / cnt[t] > rs.probe_threshold;
    }

    // (4) reuse of this frame's records, the gathers themselves, the network
    RadCache &c = caches[0];
    if ((r = c.index_all(*this))) return r;
    if ((r = lookup(0, c.all, q, nullptr, n, true, false, acc))) return r;
    k_acc_reset<<<blocks(m0), 256, 0, stream>>>(d_g0, m0, acc);
    k_combine<<<blocks(m0), 256, 0, stream>>>(d_g0, m0, static_cast<const float *>(R.Eg.p), acc, E);
    k_query_to_x6<<<blocks(n), 256, 0, stream>>>(q, nullptr, nullptr, n, nlo[0], nlo[1], nlo[2], ninv,
                                                 static_cast<float *>(lx.p));
    HIP_TRY(hipGetLastError());
    if ((r = net.infer(static_cast<const float *>(lx.p), n, static_cast<float *>(ly.p), stream))) return r;
    if ((r = send(lfbtile.p, fb_tile.data(), ntile))) return r;
    uint32_t *cnt = static_cast<uint32_t *>(lcnt.p);
    k_fill_u32<<<1, 64, 0, stream>>>(cnt, 2);
    k_learned_apply<<<blocks(n), 256, 0, stream>>>(acc, d_pix, n, W, tsz, tx, static_cast<const uint8_t *>(lfbtile.p),
                                                   static_cast<const float *>(ly.p), E,
                                                   static_cast<uint8_t *>(lfbflag.p), cnt);
    HIP_TRY(hipGetLastError());
    uint32_t hc[2];
    if ((r = fetch(hc, cnt, 8))) return r;
    stats.surrogate_queries += hc[0];
    stats.fallbacks += hc[1];
    stats.final_misses += (uint64_t)hc[0] + hc[1];
    if (hc[1]) {
        uint32_t mf = 0;
        if ((r = compact(static_cast<const uint8_t *>(lfbflag.p), n, static_cast<uint32_t *>(R.idx.p), mf))) return r;
        std::vector<uint32_t> todo(mf), slot(mf);
        std::vector<RadQuery> hq(mf);
        if ((r = fetch(todo.data(), R.idx.p, (size_t)mf * 4))) return r;
        k_gather_subq<<<blocks(mf), 256, 0, stream>>>(q, static_cast<const uint32_t *>(R.idx.p), mf,
                                                      static_cast<RadQuery *>(R.tq.p));
        HIP_TRY(hipGetLastError());
        if ((r = fetch(hq.data(), R.tq.p, (size_t)mf * sizeof(RadQuery)))) return r;
        for (uint32_t k = 0; k < mf; k++) slot[k] = k;
        // one spread-out round with reuse, then every point still unreached
        if ((r = gather_rounds(0, q, true, std::move(todo), std::move(slot), hq, acc, E, 2))) return r;
    }
    amb_net = nullptr;
    k_finish<<<blocks(n), 256, 0, stream>>>(acc, n, rs.brightness, E);
    HIP_TRY(hipGetLastError());
    return PVX_OK;
}

} // namespace pvx

// ----------------------------------------------------------------------------
// entry point
// ----------------------------------------------------------------------------
namespace pvx {

// One pass over a pixel set: primary rays, depth-0 irradiance for every hit
// that wants it, and (final pass) the pixel colours.
int RadiosityEngine::pixel_pass(const std::vector<uint32_t> &px, bool final, float *rgb) {
    const uint32_t n = (uint32_t)px.size();
    if (n == 0) return PVX_OK;
    Scratch &S = prim;
    int r;
    if ((r = S.sel.reserve(n * 4)) || (r = S.rays.reserve(n * sizeof(pvx_ray))) ||
        (r = S.hits.reserve(n * sizeof(pvx_hit))) || (r = S.base.reserve(n * 12)) ||
        (r = S.coef.reserve(n * 12)) || (r = S.q.reserve(n * sizeof(RadQuery))) || (r = S.need.reserve(n)) ||
        (r = S.subq.reserve(n * sizeof(RadQuery))) || (r = S.subE.reserve(n * 12)) ||
        (r = psel.reserve(n * 4)) || (r = ppix.reserve(n * 4)) || (r = pqpix.reserve(n * 4)))
        return r;
    if ((r = send(S.sel.p, px.data(), n * 4))) return r;
    k_pixel_rays<<<blocks(n), 256, 0, stream>>>(cam, static_cast<const uint32_t *>(S.sel.p), n,
                                                static_cast<pvx_ray *>(S.rays.p));
    HIP_TRY(hipGetLastError());
    if ((r = trace_closest_internal(scene, static_cast<const pvx_ray *>(S.rays.p), nullptr, PVX_RAY_IMAGE, n,
                                    static_cast<pvx_hit *>(S.hits.p), stream)))
        return r;
    ShadeParams sp;
    sp.sv = scene->view;
    sp.obj_type = scene->d_obj_type;
    sp.obj_index = scene->d_obj_index;
    sp.mats = scene->shading.d_mats;
    sp.adc = rs.adc_bailout;
    (scene->view.has_complex >= 3 ? k_hit_queries<true> : k_hit_queries<false>)<<<blocks(n), 256, 0, stream>>>(sp, static_cast<const pvx_ray *>(S.rays.p),
                                                 static_cast<const pvx_hit *>(S.hits.p), n, 1.0f,
                                                 rs.recursion_limit > 0 ? 1 : 0, static_cast<float *>(S.base.p),
                                                 static_cast<float *>(S.coef.p), static_cast<RadQuery *>(S.q.p),
                                                 static_cast<uint8_t *>(S.need.p));
    HIP_TRY(hipGetLastError());
    // the hits that want irradiance, in pixel-list order
    uint32_t m = 0;
    uint32_t *sel = static_cast<uint32_t *>(psel.p);
    if ((r = compact(static_cast<const uint8_t *>(S.need.p), n, sel, m))) return r;
    if (m) {
        k_gather_subq<<<blocks(m), 256, 0, stream>>>(static_cast<const RadQuery *>(S.q.p), sel, m,
                                                     static_cast<RadQuery *>(S.subq.p));
        HIP_TRY(hipGetLastError());
        if (final && rs.mode == PVX_RADIOSITY_LEARNED) {
            k_gather_u32<<<blocks(m), 256, 0, stream>>>(static_cast<const uint32_t *>(S.sel.p), sel, m,
                                                        static_cast<uint32_t *>(ppix.p));
            HIP_TRY(hipGetLastError());
            if ((r = resolve_learned(static_cast<const RadQuery *>(S.subq.p), m,
                                     static_cast<const uint32_t *>(ppix.p), static_cast<float *>(S.subE.p))))
                return r;
        } else if ((r = resolve(0, static_cast<const RadQuery *>(S.subq.p), m, final,
                                static_cast<float *>(S.subE.p)))) {
            return r;
        }
    }
    if (final) {
        if (m) k_scatter_rank<<<blocks(m), 256, 0, stream>>>(sel, m, static_cast<uint32_t *>(pqpix.p));
        k_final_colour<<<blocks(n), 256, 0, stream>>>(
            static_cast<const float *>(S.base.p), static_cast<const float *>(S.coef.p),
            static_cast<const float *>(S.subE.p), static_cast<const uint32_t *>(pqpix.p), n,
            static_cast<const uint8_t *>(S.need.p), rgb);
        HIP_TRY(hipGetLastError());
        HIP_TRY(sync());
    }
    return PVX_OK;
}

} // namespace pvx

static RadiosityEngine *engine_of(pvx_scene *s) {
    if (!s->rad_engine) {
        RadiosityEngine *e = new (std::nothrow) RadiosityEngine(s);
        if (!e) return nullptr;
        s->rad_engine = e;
        s->rad_engine_free = [](void *p) { delete static_cast<RadiosityEngine *>(p); };
    }
    return static_cast<RadiosityEngine *>(s->rad_engine);
}

// The packed unit vector of a cache file: three 2-digit hex bytes,
// (v + 1) / 2 * 254 rounded (octree.cpp:1345-1370); read back as
// byte / 254 * 2 - 1, normalised (radiosity.cpp:970-985).
static void pack_dir(const float *v, char *out) {
    int b[3];
    for (int k = 0; k < 3; k++) b[k] = (int)(((double)v[k] + 1.) * .5 * 254. + .499999);
    std::snprintf(out, 8, "%02x%02x%02x", b[0] & 0xFF, b[1] & 0xFF, b[2] & 0xFF);
}
static bool unpack_dir(const char *s, float *v) {
    unsigned tx, ty, tz;
    if (std::sscanf(s, "%02x%02x%02x", &tx, &ty, &tz) != 3) return false;
    double d[3] = {((double)tx * (1. / 254.)) * 2. - 1., ((double)ty * (1. / 254.)) * 2. - 1.,
                   ((double)tz * (1. / 254.)) * 2. - 1.};
    const double l = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    for (int k = 0; k < 3; k++) v[k] = (float)(l != 0 ? d[k] / l : d[k]);
    return true;
}

// Writes the last frame's irradiance cache as a POV-Ray radiosity cache file
// (.rca): one "C" line per record with 1-based bounce depth, point, packed
// normal, illuminance, harmonic mean distance, nearest distance and packed
// direction to the nearest surface (ot_write_block, octree.cpp:1345-1370).
extern "C" int pvx_radiosity_save(pvx_scene *s, const char *path, uint64_t *n_written) {
    if (!s || !path) return fail(PVX_EINVAL, "null argument");
    if (!s->rad_engine) return fail(PVX_ESTATE, "no radiosity frame rendered on this scene");
    const RadiosityEngine &eng = *static_cast<RadiosityEngine *>(s->rad_engine);
    FILE *f = std::fopen(path, "w");
    if (!f) return fail(PVX_EINVAL, std::string("cannot open ") + path + " for writing");
    uint64_t n = 0;
    for (int d = 0; d < kMaxRadDepth; d++)
        for (const RadRec &r : eng.caches[d].recs) {
            char ns[8], ts[8];
            pack_dir(r.n, ns);
            pack_dir(r.tn, ts);
            std::fprintf(f, "C%d\t%g\t%g\t%g\t%s\t%.4f\t%.4f\t%.4f\t%g\t%g\t%s\n", d + 1, (double)r.p[0],
                         (double)r.p[1], (double)r.p[2], ns, (double)r.E[0], (double)r.E[1], (double)r.E[2],
                         (double)r.r, (double)r.nearest, ts);
            n++;
        }
    const bool ok = std::fclose(f) == 0;
    if (!ok) return fail(PVX_EINVAL, std::string("error writing ") + path);
    if (n_written) *n_written = n;
    return PVX_OK;
}

// Reads a radiosity cache file into the records every later frame on this
// scene starts from (RadiosityCache::Load, radiosity.cpp:909-1000): "C" lines
// of 11 fields are taken, other lines are skipped.  path NULL drops them.
extern "C" int pvx_radiosity_load(pvx_scene *s, const char *path, uint64_t *n_read) {
    if (!s) return fail(PVX_EINVAL, "null scene");
    RadiosityEngine *e = engine_of(s);
    if (!e) return fail(PVX_ENOMEM, "host allocation failed");
    for (int d = 0; d < kMaxRadDepth; d++) e->preload[d].clear();
    if (n_read) *n_read = 0;
    if (!path) return PVX_OK;
    FILE *f = std::fopen(path, "r");
    if (!f) return fail(PVX_EINVAL, std::string("cannot open ") + path);
    char line[256];
    uint64_t n = 0;
    while (std::fgets(line, sizeof line, f)) {
        if (line[0] != 'C') continue;
        int depth;
        double p[3], hm, nearest;
        float E[3];
        char ns[32], ts[32];
        if (std::sscanf(line, "C%d %lf %lf %lf %31s %f %f %f %lf %lf %31s", &depth, &p[0], &p[1], &p[2], ns, &E[0],
                        &E[1], &E[2], &hm, &nearest, ts) != 11)
            continue;
        depth -= 1; // 1-based in the file
        if (depth < 0 || depth >= kMaxRadDepth) continue;
        RadRec r{};
        for (int k = 0; k < 3; k++) {
            r.p[k] = (float)p[k];
            r.E[k] = E[k];
        }
        if (!unpack_dir(ns, r.n) || !unpack_dir(ts, r.tn)) continue;
        r.r = (float)hm;
        r.nearest = (float)nearest;
        e->preload[depth].push_back(r);
        n++;
    }
    std::fclose(f);
    if (n_read) *n_read = n;
    return PVX_OK;
}

static int check_settings(pvx_scene *s, const pvx_radiosity_settings *set) {
    if (!s->shading.set) return fail(PVX_ESTATE, "pvx_scene_set_shading was not called");
    if (set->count == 0 || set->count > 1600) return fail(PVX_EINVAL, "count must be in [1, 1600]");
    if (set->nearest_count == 0 || set->nearest_count > 20)
        return fail(PVX_EINVAL, "nearest_count must be in [1, 20]");
    if (set->recursion_limit > (uint32_t)kMaxRadDepth)
        return fail(PVX_EUNSUPPORTED, "recursion_limit above " + std::to_string(kMaxRadDepth));
    if (!(set->error_bound > 0.0f) || !(set->minimum_reuse > 0.0f) || !(set->maximum_reuse >= set->minimum_reuse) ||
        !(set->pretrace_start > 0.0f) || !(set->pretrace_end > 0.0f) || !(set->pretrace_end <= set->pretrace_start))
        return fail(PVX_EINVAL, "bad radiosity settings");
    if (set->mode != PVX_RADIOSITY_PLAIN && set->mode != PVX_RADIOSITY_LEARNED)
        return fail(PVX_EINVAL, "unknown radiosity mode");
    if (set->mode == PVX_RADIOSITY_LEARNED &&
        (set->probe_tile == 0 || set->train_steps == 0 || set->train_samples == 0 || !(set->probe_threshold >= 0.0f)))
        return fail(PVX_EINVAL, "learned radiosity needs probe_tile, train_steps, train_samples > 0");
    if (set->mode == PVX_RADIOSITY_LEARNED && set->train_batch % 256)
        return fail(PVX_EINVAL, "train_batch must be a multiple of 256");
    for (const MatRec &m : s->shading.mats)
        if (m.reflective)
            return fail(PVX_EUNSUPPORTED, "radiosity frames take plain diffuse / emissive finishes only");
    return PVX_OK;
}

extern "C" int pvx_radiosity_begin(pvx_scene *s, const pvx_camera *cam, const pvx_radiosity_settings *set,
                                   void *stream) {
    if (!s || !cam || !set) return fail(PVX_EINVAL, "null argument");
    int r;
    if ((r = check_settings(s, set))) return r;
    if (set->mode != PVX_RADIOSITY_PLAIN)
        return fail(PVX_EUNSUPPORTED, "per-query ambient takes the plain gather (the surrogate works per frame)");
    if (!engine_of(s)) return fail(PVX_ENOMEM, "host allocation failed");
    RadiosityEngine &eng = *static_cast<RadiosityEngine *>(s->rad_engine);
    if ((r = eng.begin(*set, *cam, static_cast<hipStream_t>(stream)))) return r;
    eng.ambient_open = true;
    return PVX_OK;
}

extern "C" int pvx_radiosity_compute_ambient(pvx_scene *s, size_t n, const double *ipoint, const double *raw_normal,
                                             const double *layer_normal, const double *weight, uint32_t depth,
                                             uint32_t final_trace, float *ambient_rgb, pvx_radiosity_stats *stats,
                                             void *stream) {
    if (!s) return fail(PVX_EINVAL, "null scene");
    RadiosityEngine *e = static_cast<RadiosityEngine *>(s->rad_engine);
    if (!e || !e->ambient_open) return fail(PVX_ESTATE, "pvx_radiosity_begin was not called");
    if (n == 0) return PVX_OK;
    if (!ipoint || !raw_normal || !ambient_rgb) return fail(PVX_EINVAL, "null argument");
    if (n > 0x3FFFFFFFu) return fail(PVX_EINVAL, "too many queries");
    RadiosityEngine &eng = *e;
    if (depth >= eng.rs.recursion_limit)
        return fail(PVX_EINVAL, "depth must be below recursion_limit (CheckRadiosityTraceLevel)");
    eng.stream = static_cast<hipStream_t>(stream);
    int r;
    if ((r = eng.amb_q.reserve(n * sizeof(RadQuery))) || (r = eng.amb_flag.reserve(4))) return r;
    uint32_t *flag = static_cast<uint32_t *>(eng.amb_flag.p);
    k_fill_u32<<<1, 64, 0, eng.stream>>>(flag, 1);
    RadQuery *q = static_cast<RadQuery *>(eng.amb_q.p);
    k_ambient_queries<<<blocks(n), 256, 0, eng.stream>>>(ipoint, raw_normal, layer_normal, weight, n, q, flag);
    HIP_TRY(hipGetLastError());
    uint32_t bad = 0;
    if ((r = eng.fetch(&bad, flag, 4))) return r;
    if (bad) return fail(PVX_EUNSUPPORTED, "a layer normal differs from its raw normal (perturbed normals)");
    const pvx_radiosity_stats before = eng.stats;
    if ((r = eng.resolve((int)depth, q, (uint32_t)n, final_trace != 0, ambient_rgb))) return r;
    HIP_TRY(eng.sync());
    if (stats) {
        *stats = eng.stats;
        for (int d = 0; d < 2; d++) {
            stats->gathers[d] -= before.gathers[d];
            stats->records[d] = eng.caches[d].recs.size();
        }
        stats->gather_rays -= before.gather_rays;
        stats->final_misses -= before.final_misses;
    }
    return PVX_OK;
}

extern "C" int pvx_radiosity_render(pvx_scene *s, const pvx_camera *cam, const pvx_radiosity_settings *set,
                                    float *rgb, pvx_radiosity_stats *stats, void *stream) {
    if (!s || !cam || !set || !rgb) return fail(PVX_EINVAL, "null argument");
    if (cam->width == 0 || cam->height == 0) return fail(PVX_EINVAL, "bad image size");
    size_t npix = (size_t)cam->width * cam->height;
    if (npix > 0x3FFFFFFFu) return fail(PVX_EINVAL, "image too large");
    {
        int rc = check_settings(s, set);
        if (rc) return rc;
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (!engine_of(s)) return fail(PVX_ENOMEM, "host allocation failed");
    RadiosityEngine &eng = *static_cast<RadiosityEngine *>(s->rad_engine);
    int r;
    if ((r = eng.begin(*set, *cam, st))) return r;
    eng.ambient_open = false; // a frame leaves the per-query cache closed
    hipEvent_t e0, e1, e2;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventCreate(&e2));
    HIP_TRY(hipEventRecord(e0, st));
    const uint32_t W = cam->width, H = cam->height;
    // pretrace: pixel step from pretrace_start to pretrace_end of the larger
    // image side, halving (View::StartRender pretrace steps,
    // backend/scene/view.cpp:1056-1106; integer pixel steps here)
    double step = std::max(1.0, std::floor((double)set->pretrace_start * std::max(W, H)));
    double end = std::max(1.0, std::floor((double)set->pretrace_end * std::max(W, H)));
    // the learned mode has no pretrace: its networks stand in for the cache
    for (; set->mode != PVX_RADIOSITY_LEARNED;) {
        std::vector<uint32_t> px;
        uint32_t k = (uint32_t)step;
        for (uint32_t y = 0; y < H; y += k)
            for (uint32_t x = 0; x < W; x += k) px.push_back(y * W + x);
        if ((r = eng.pixel_pass(px, false, nullptr))) return r;
        if (step <= end) break;
        step = std::max(end, std::floor(step / 2));
    }
    HIP_TRY(hipEventRecord(e1, st));
    std::vector<uint32_t> px(npix);
    for (size_t i = 0; i < npix; i++) px[i] = (uint32_t)i;
    if ((r = eng.pixel_pass(px, true, rgb))) return r;
    HIP_TRY(hipEventRecord(e2, st));
    HIP_TRY(hipEventSynchronize(e2));
    if (stats) {
        float a = 0, b = 0;
        HIP_TRY(hipEventElapsedTime(&a, e0, e1));
        HIP_TRY(hipEventElapsedTime(&b, e1, e2));
        *stats = eng.stats;
        stats->pretrace_ms = a;
        stats->final_ms = b;
        stats->frame_ms = (double)a + b;
        stats->records[0] = eng.caches[0].recs.size();
        stats->records[1] = eng.caches[1].recs.size();
        stats->train_ms = eng.train_ms;
    }
    if (std::getenv("PVX_RADIOSITY_VERBOSE"))
        std::fprintf(stderr,
                     "[pvx radiosity] rounds d0=%llu d1=%llu gathers d0=%llu d1=%llu records d0=%zu d1=%zu "
                     "syncs=%llu wait=%.1fms grids=%.1fms\n",
                     (unsigned long long)eng.stats_rounds[0], (unsigned long long)eng.stats_rounds[1],
                     (unsigned long long)eng.stats.gathers[0], (unsigned long long)eng.stats.gathers[1],
                     eng.caches[0].recs.size(), eng.caches[1].recs.size(), (unsigned long long)eng.n_sync,
                     eng.t_sync_ms, eng.t_grid_ms);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(e2);
    return PVX_OK;
}


// ----------------------------------------------------------------------------
// many scenes: their radiosity frames on several streams at once
// ----------------------------------------------------------------------------
extern "C" int pvx_batch_radiosity_render(pvx_batch *b, const pvx_camera *cams,
                                          const pvx_radiosity_settings *settings, float *const *rgb,
                                          uint32_t n_streams, pvx_radiosity_stats *scene_stats,
                                          double *scene_done_ms, pvx_batch_radiosity_stats *stats) {
    if (!b || !cams || !settings || !rgb) return fail(PVX_EINVAL, "null argument");
    const uint32_t n = (uint32_t)b->scenes.size();
    if (n_streams == 0 || n_streams > 16) return fail(PVX_EINVAL, "n_streams must be in [1, 16]");
    for (uint32_t k = 0; k < n; k++) {
        if (!rgb[k]) return fail(PVX_EINVAL, "null image buffer for scene " + std::to_string(k));
        const uint64_t px = (uint64_t)cams[k].width * cams[k].height;
        if (px == 0 || px > 0x3FFFFFFFu) return fail(PVX_EINVAL, "bad image size for scene " + std::to_string(k));
    }
    // largest frames first, so that the streams finish together
    std::vector<uint32_t> order(n);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t c) {
        return (uint64_t)cams[a].width * cams[a].height > (uint64_t)cams[c].width * cams[c].height;
    });
    const uint32_t W = std::min(n_streams, n);
    std::vector<hipStream_t> streams(W, nullptr);
    for (uint32_t w = 0; w < W; w++) HIP_TRY(hipStreamCreate(&streams[w]));
    std::vector<pvx_radiosity_stats> st(n);
    std::vector<double> done(n, 0.0);
    std::atomic<uint32_t> next{0};
    std::mutex err_mu;
    int err_rc = PVX_OK;
    std::string err_msg;
    const auto t0 = std::chrono::steady_clock::now();
    auto worker = [&](uint32_t w) {
        if (hipSetDevice(b->device) != hipSuccess) {
            std::lock_guard<std::mutex> lock(err_mu);
            if (!err_rc) err_rc = PVX_EDEVICE, err_msg = "hipSetDevice failed in a batch worker";
            return;
        }
        for (;;) {
            const uint32_t i = next.fetch_add(1);
            if (i >= n) return;
            {
                std::lock_guard<std::mutex> lock(err_mu);
                if (err_rc) return;
            }
            const uint32_t k = order[i];
            const int rc = pvx_radiosity_render(b->scenes[k], &cams[k], &settings[k], rgb[k], &st[k], streams[w]);
            if (rc) {
                std::lock_guard<std::mutex> lock(err_mu);
                if (!err_rc) err_rc = rc, err_msg = "scene " + std::to_string(k) + ": " + pvx_last_error();
                return;
            }
            done[k] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        }
    };
    std::vector<std::thread> pool;
    for (uint32_t w = 1; w < W; w++) pool.emplace_back(worker, w);
    worker(0);
    for (auto &t : pool) t.join();
    const double wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    for (hipStream_t sm : streams) (void)hipStreamDestroy(sm);
    if (err_rc) return fail(err_rc, err_msg);
    if (scene_stats) std::copy(st.begin(), st.end(), scene_stats);
    if (scene_done_ms) std::copy(done.begin(), done.end(), scene_done_ms);
    if (stats) {
        *stats = pvx_batch_radiosity_stats{};
        stats->batch_ms = wall;
        stats->streams = W;
        for (uint32_t k = 0; k < n; k++) {
            stats->pixels += (uint64_t)cams[k].width * cams[k].height;
            stats->gather_rays += st[k].gather_rays;
            stats->surrogate_queries += st[k].surrogate_queries;
            stats->fallbacks += st[k].fallbacks;
            stats->final_misses += st[k].final_misses;
            stats->gpu_frame_ms += st[k].frame_ms;
        }
    }
    return PVX_OK;
}
