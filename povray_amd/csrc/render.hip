// Wavefront shading of opaque plain-finish surfaces (gfx950).
//
// Per bounce: one closest-hit launch over the live rays, one shade launch
// (normal, classic ambient, per-light diffuse/Phong terms, shadow-ray and
// reflection-ray emission through atomic queues), one shadow-ray launch and
// one finish launch (light visibility -> colour).  A path has at most one
// live ray per pixel (reflection only), so pixels are accumulated without
// atomics and the frame is deterministic.
//
// Follows Trace::TraceRay / ComputeLightedTexture / ComputeOneDiffuseLight /
// TracePointLightShadowRay / TraceAreaLightSubsetShadowRay /
// ComputeDiffuseColour / ComputePhongColour / ComputeReflectivity /
// ComputeReflection (core/render/trace.cpp:135-228, 739-1179, 1264-1321,
// 1637-1729, 1951-2272, 2441-2558, 2627-2656).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "internal.h"
#include "normal_device.h"
#include "shapes_device.h"

using namespace pvx;

namespace {

constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr double kShadowTolerance = 1.0e-3; // trace.cpp:81

struct PathState {
    uint32_t pixel;
    uint32_t level; // trace level before this ray's TraceRay
    float thr[3];   // product of reflectivities along the path
    float weight;   // ADC weight
};

struct ShadeOut {
    uint32_t pixel;
    uint32_t obj;
    float thr[3];
    float pad;
};

struct RenderParams : NormalParams {
    const MatRec *mats;
    const LightRec *lights;
    uint32_t n_lights;
    uint32_t max_level;
    float bailout;
    float bg[3];
};

__global__ void k_fill(float *p, size_t n, float v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_zero_counters(uint32_t *c) {
    if (threadIdx.x < 4) c[threadIdx.x] = 0;
}

__global__ void k_init_paths(const pvx_ray *__restrict__ cam_rays, uint32_t n, PathState *st) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    st[i] = PathState{i, 0u, {1.0f, 1.0f, 1.0f}, 1.0f};
}

// Shade the hits of one bounce: ambient now, per-light terms and shadow rays
// for the finish pass, and the reflected ray of the next bounce.
__global__ void k_shade(RenderParams p, const pvx_ray *__restrict__ rays,
                        const PathState *__restrict__ st, const pvx_hit *__restrict__ hits,
                        uint32_t n, uint32_t ray_kind, float *__restrict__ rgb,
                        ShadeOut *__restrict__ shade, uint32_t *__restrict__ lslot,
                        double *__restrict__ lterm, pvx_ray *__restrict__ srays,
                        double *__restrict__ sdepth, pvx_ray *__restrict__ next_rays,
                        PathState *__restrict__ next_st, uint32_t *__restrict__ counters) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const pvx_ray r = rays[i];
    const PathState ps = st[i];
    const pvx_hit h = hits[i];
    float *px = rgb + 3 * (size_t)ps.pixel;
    ShadeOut so;
    so.pixel = kNone;
    if (h.obj < 0) { // ComputeSky: background colour
        px[0] += ps.thr[0] * p.bg[0];
        px[1] += ps.thr[1] * p.bg[1];
        px[2] += ps.thr[2] * p.bg[2];
        shade[i] = so;
        return;
    }
    D3 o{r.o[0], r.o[1], r.o[2]}, d{r.d[0], r.d[1], r.d[2]};
    D3 ip = add3(o, mul3(d, h.t));
    D3 nrm = hit_normal(p, h, o, d, ip);
    if (dot3(nrm, d) > 0.0) nrm = D3{-nrm.x, -nrm.y, -nrm.z};
    const MatRec m = p.mats[h.obj];
    // classic ambient: pigment * (ambient * ambient_light) * opacity
    {
        float a0 = (float)((double)m.att * (double)(m.pigment[0] * m.emission[0]));
        float a1 = (float)((double)m.att * (double)(m.pigment[1] * m.emission[1]));
        float a2 = (float)((double)m.att * (double)(m.pigment[2] * m.emission[2]));
        px[0] += ps.thr[0] * a0;
        px[1] += ps.thr[1] * a1;
        px[2] += ps.thr[2] * a2;
    }
    for (uint32_t L = 0; L < p.n_lights; L++) {
        const LightRec &lt = p.lights[L];
        D3 C{lt.pos[0], lt.pos[1], lt.pos[2]};
        D3 dir = sub3(C, ip);
        double lsd = sqrt(dot3(dir, dir));
        dir = D3{dir.x / lsd, dir.y / lsd, dir.z / lsd};
        size_t li = (size_t)i * p.n_lights + L;
        double cosi = dot3(nrm, dir);
        if (cosi < kEpsilon) { // light behind the surface (no diffuse_back)
            lslot[li] = kNone;
            continue;
        }
        double inten = (m.brilliance != 1.0f) ? pow(fabs(cosi), (double)m.brilliance) : fabs(cosi);
        inten *= (double)m.diffuse * (double)m.att;
        double ph = 0.0;
        if (m.phong > 0.0f && ray_kind != PVX_RAY_RADIOSITY) {
            double c2 = -2.0 * dot3(d, nrm);
            D3 rd = add3(d, mul3(nrm, c2));
            double ca = dot3(rd, dir);
            if (ca > 0.0 && (m.phong_size < 60 || ca > 0.0008))
                ph = (double)m.phong * pow(ca, (double)m.phong_size);
        }
        lterm[2 * li] = m.diffuse > 0.0f ? inten : 0.0;
        lterm[2 * li + 1] = ph;
        uint32_t ns = lt.n1 * lt.n2;
        uint32_t base = atomicAdd(&counters[1], ns);
        lslot[li] = base;
        for (uint32_t u = 0; u < lt.n1; u++)
            for (uint32_t v = 0; v < lt.n2; v++) {
                D3 tgt = C;
                if (lt.n1 > 1) {
                    double f = (double)u / (double)(lt.n1 - 1) - 0.5;
                    tgt = add3(tgt, D3{lt.axis1[0] * f, lt.axis1[1] * f, lt.axis1[2] * f});
                }
                if (lt.n2 > 1) {
                    double f = (double)v / (double)(lt.n2 - 1) - 0.5;
                    tgt = add3(tgt, D3{lt.axis2[0] * f, lt.axis2[1] * f, lt.axis2[2] * f});
                }
                D3 sd = sub3(tgt, ip);
                double sl = sqrt(dot3(sd, sd));
                pvx_ray sr;
                sr.o[0] = ip.x; sr.o[1] = ip.y; sr.o[2] = ip.z;
                sr.d[0] = sd.x / sl; sr.d[1] = sd.y / sl; sr.d[2] = sd.z / sl;
                uint32_t k = base + u * lt.n2 + v;
                srays[k] = sr;
                sdepth[k] = sl;
            }
    }
    so.pixel = ps.pixel;
    so.obj = (uint32_t)h.obj;
    so.thr[0] = ps.thr[0];
    so.thr[1] = ps.thr[1];
    so.thr[2] = ps.thr[2];
    shade[i] = so;
    // reflection (ComputeReflectivity without fresnel, ComputeReflection)
    if (m.reflective & 1u) {
        double cos_inc = -dot3(d, nrm);
        double w = (double)ps.weight * (double)m.rweight;
        double frac = fabs((double)m.falloff - 1.0) > kEpsilon ? pow(1.0 - cos_inc, (double)m.falloff)
                                                              : 1.0 - cos_inc;
        float rf[3];
        for (int k = 0; k < 3; k++) {
            if (fabs(frac) < kEpsilon)
                rf[k] = m.rmin[k];
            else if (fabs(frac - 1.0) < kEpsilon)
                rf[k] = m.rmax[k];
            else
                rf[k] = (float)(frac * m.rmax[k]) + (float)((1.0 - frac) * m.rmin[k]);
        }
        bool nonzero = rf[0] != 0.0f || rf[1] != 0.0f || rf[2] != 0.0f;
        if (nonzero && ps.level + 1 < p.max_level && (float)w >= p.bailout) {
            double k2 = -2.0 * dot3(d, nrm);
            D3 nd = normalize3(add3(d, mul3(nrm, k2)));
            uint32_t slot = atomicAdd(&counters[0], 1u);
            pvx_ray nr;
            nr.o[0] = ip.x; nr.o[1] = ip.y; nr.o[2] = ip.z;
            nr.d[0] = nd.x; nr.d[1] = nd.y; nr.d[2] = nd.z;
            next_rays[slot] = nr;
            PathState ns;
            ns.pixel = ps.pixel;
            ns.level = ps.level + 1;
            ns.thr[0] = ps.thr[0] * rf[0];
            ns.thr[1] = ps.thr[1] * rf[1];
            ns.thr[2] = ps.thr[2] * rf[2];
            ns.weight = (float)w;
            next_st[slot] = ns;
        }
    }
}

// TraceAreaLightSubsetShadowRay's quadtree average with adaptive 100 (every
// region subdivides until it spans <= 2x2 grid points).
template <int D>
__device__ float area_average(const float *vis, int n2, int u1, int v1, int u2, int v2) {
    float c0 = vis[u1 * n2 + v1], c1 = vis[u2 * n2 + v1], c2 = vis[u1 * n2 + v2], c3 = vis[u2 * n2 + v2];
    if constexpr (D > 0) {
        if ((u2 - u1 > 1) || (v2 - v1 > 1)) {
            int mu0 = (u1 + u2) / 2, mu1 = (u1 + u2 + 1) / 2;
            int mv0 = (v1 + v2) / 2, mv1 = (v1 + v2 + 1) / 2;
            c0 = area_average<D - 1>(vis, n2, u1, v1, mu0, mv0);
            c1 = area_average<D - 1>(vis, n2, mu1, v1, u2, mv0);
            c2 = area_average<D - 1>(vis, n2, u1, mv1, mu0, v2);
            c3 = area_average<D - 1>(vis, n2, mu1, mv1, u2, v2);
        }
    }
    return (c0 + c1 + c2 + c3) * 0.25f;
}

constexpr int kMaxAreaGrid = 16; // per axis

__global__ void k_finish(RenderParams p, const ShadeOut *__restrict__ shade, uint32_t n,
                         const uint32_t *__restrict__ lslot, const double *__restrict__ lterm,
                         const pvx_hit *__restrict__ shits, const double *__restrict__ sdepth,
                         float *__restrict__ rgb) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const ShadeOut so = shade[i];
    if (so.pixel == kNone) return;
    const MatRec &m = p.mats[so.obj];
    float col[3] = {0.0f, 0.0f, 0.0f};
    for (uint32_t L = 0; L < p.n_lights; L++) {
        size_t li = (size_t)i * p.n_lights + L;
        uint32_t base = lslot[li];
        if (base == kNone) continue;
        const LightRec &lt = p.lights[L];
        float f;
        uint32_t ns = lt.n1 * lt.n2;
        if (ns == 1) {
            const pvx_hit sh = shits[base];
            bool blocked = sh.obj >= 0 && sh.t < sdepth[base] - kShadowTolerance && sh.t > kShadowTolerance;
            f = blocked ? 0.0f : 1.0f;
        } else {
            float vis[kMaxAreaGrid * kMaxAreaGrid];
            for (uint32_t k = 0; k < ns; k++) {
                const pvx_hit sh = shits[base + k];
                bool blocked = sh.obj >= 0 && sh.t < sdepth[base + k] - kShadowTolerance &&
                               sh.t > kShadowTolerance;
                vis[k] = blocked ? 0.0f : 1.0f;
            }
            f = area_average<5>(vis, (int)lt.n2, 0, 0, (int)lt.n1 - 1, (int)lt.n2 - 1);
        }
        float lc[3] = {lt.colour[0] * f, lt.colour[1] * f, lt.colour[2] * f};
        if (fabsf(lc[0]) < 1e-10f && fabsf(lc[1]) < 1e-10f && fabsf(lc[2]) < 1e-10f) continue;
        double inten = lterm[2 * li], ph = lterm[2 * li + 1];
        for (int k = 0; k < 3; k++) {
            if (m.diffuse > 0.0f) col[k] = col[k] + (float)(inten * (double)(m.pigment[k] * lc[k]));
        }
        if (ph != 0.0)
            for (int k = 0; k < 3; k++) col[k] = col[k] + (float)(ph * (double)lc[k]);
    }
    float *px = rgb + 3 * (size_t)so.pixel;
    px[0] += so.thr[0] * col[0];
    px[1] += so.thr[1] * col[1];
    px[2] += so.thr[2] * col[2];
}

template <class T>
int grow(pvx_scene *s, void *&ptr, size_t bytes) {
    if (ptr) {
        for (auto &a : s->render_allocs)
            if (a == ptr) a = nullptr;
        (void)hipFree(ptr);
        ptr = nullptr;
    }
    HIP_TRY(hipMalloc(&ptr, bytes));
    s->render_allocs.push_back(ptr);
    return PVX_OK;
}

} // namespace

extern "C" int pvx_scene_set_shading(pvx_scene *s, const pvx_shading_desc *d) {
    if (!s || !d) return fail(PVX_EINVAL, "null argument");
    if (d->n_materials != s->n_objects)
        return fail(PVX_EINVAL, "one material per scene object is required");
    if (d->n_lights && !d->lights) return fail(PVX_EINVAL, "null light array");
    if (!d->materials) return fail(PVX_EINVAL, "null material array");
    ShadingState &sh = s->shading;
    sh.mats.resize(d->n_materials);
    for (uint32_t i = 0; i < d->n_materials; i++) {
        const pvx_material &a = d->materials[i];
        MatRec &m = sh.mats[i];
        for (int k = 0; k < 3; k++) {
            m.pigment[k] = a.pigment[k];
            m.emission[k] = a.emission[k] + a.ambient[k] * d->ambient_light[k];
            m.ambient_only[k] = a.emission[k];
            m.rmin[k] = a.reflection_min[k];
            m.rmax[k] = a.reflection_max[k];
        }
        m.att = (float)(1.0 - (double)a.filter - (double)a.transmit);
        m.diffuse = a.diffuse;
        m.brilliance = a.brilliance;
        m.phong = a.phong;
        m.phong_size = a.phong_size;
        m.falloff = a.reflection_falloff;
        float wmax = std::max(a.reflection_max[0], std::max(a.reflection_max[1], a.reflection_max[2]));
        float wmin = std::max(a.reflection_min[0], std::max(a.reflection_min[1], a.reflection_min[2]));
        m.rweight = std::max(wmax, wmin);
        m.reflective = (wmax != 0.0f || wmin != 0.0f ||
                        a.reflection_min[0] != 0.0f || a.reflection_max[0] != 0.0f) ? 1u : 0u;
        if (a.filter != 0.0f || a.transmit != 0.0f || a.ior != 0.0f) m.reflective |= 2u;
    }
    sh.lights.resize(d->n_lights);
    sh.samples_per_hit = 0;
    for (uint32_t i = 0; i < d->n_lights; i++) {
        const pvx_light &a = d->lights[i];
        LightRec &l = sh.lights[i];
        for (int k = 0; k < 3; k++) {
            l.pos[k] = a.position[k];
            l.axis1[k] = a.axis1[k];
            l.axis2[k] = a.axis2[k];
            l.colour[k] = a.colour[k];
        }
        bool area = a.kind == PVX_LIGHT_AREA;
        l.n1 = area ? std::max(1u, a.size1) : 1u;
        l.n2 = area ? std::max(1u, a.size2) : 1u;
        if (l.n1 > kMaxAreaGrid || l.n2 > kMaxAreaGrid)
            return fail(PVX_EUNSUPPORTED, "area light grid larger than 16x16");
        l.first = sh.samples_per_hit;
        sh.samples_per_hit += l.n1 * l.n2;
    }
    for (int k = 0; k < 3; k++) sh.background[k] = d->background[k];
    sh.max_level = d->max_trace_level;
    sh.bailout = d->adc_bailout;
    if (sh.d_mats) (void)hipFree(sh.d_mats);
    if (sh.d_lights) (void)hipFree(sh.d_lights);
    sh.d_mats = nullptr;
    sh.d_lights = nullptr;
    HIP_TRY(hipMalloc((void **)&sh.d_mats, sizeof(MatRec) * std::max<size_t>(1, sh.mats.size())));
    HIP_TRY(hipMemcpy(sh.d_mats, sh.mats.data(), sizeof(MatRec) * sh.mats.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMalloc((void **)&sh.d_lights, sizeof(LightRec) * std::max<size_t>(1, sh.lights.size())));
    if (!sh.lights.empty())
        HIP_TRY(hipMemcpy(sh.d_lights, sh.lights.data(), sizeof(LightRec) * sh.lights.size(),
                          hipMemcpyHostToDevice));
    sh.set = true;
    return PVX_OK;
}

extern "C" int pvx_render(pvx_scene *s, const pvx_camera *cam, float *rgb, pvx_render_stats *stats,
                          void *stream) {
    if (!s || !cam || !rgb) return fail(PVX_EINVAL, "null argument");
    ShadingState &sh = s->shading;
    if (!sh.set) return fail(PVX_ESTATE, "pvx_scene_set_shading was not called");
    for (const MatRec &m : sh.mats)
        if (m.reflective & 2u)
            return fail(PVX_EUNSUPPORTED, "transparent / refractive materials are not rendered yet");
    if (cam->width == 0 || cam->height == 0 || cam->height > 65535)
        return fail(PVX_EINVAL, "bad image size");
    hipStream_t st = static_cast<hipStream_t>(stream);
    size_t npix = (size_t)cam->width * cam->height;
    if (npix > 0xFFFFFFF0u) return fail(PVX_EINVAL, "image too large");
    RenderBuffers &b = s->rb;
    size_t nl = std::max<size_t>(1, sh.lights.size());
    size_t nshadow = npix * std::max<uint32_t>(1, sh.samples_per_hit);
    if (b.cap_rays < npix) {
        int r;
        if ((r = grow<char>(s, b.rays[0], npix * sizeof(pvx_ray))) ||
            (r = grow<char>(s, b.rays[1], npix * sizeof(pvx_ray))) ||
            (r = grow<char>(s, b.state[0], npix * sizeof(PathState))) ||
            (r = grow<char>(s, b.state[1], npix * sizeof(PathState))) ||
            (r = grow<char>(s, b.hits, npix * sizeof(pvx_hit))))
            return r;
        b.cap_rays = npix;
    }
    size_t shade_bytes = npix * (sizeof(ShadeOut) + nl * (sizeof(uint32_t) + 2 * sizeof(double))) + 64;
    if (b.cap_shade < shade_bytes) {
        int r;
        if ((r = grow<char>(s, b.shade, shade_bytes))) return r;
        b.cap_shade = shade_bytes;
    }
    if (b.cap_shadow < nshadow) {
        int r;
        if ((r = grow<char>(s, b.srays, nshadow * sizeof(pvx_ray))) ||
            (r = grow<char>(s, b.sdepth, nshadow * sizeof(double))) ||
            (r = grow<char>(s, b.shits, nshadow * sizeof(pvx_hit))))
            return r;
        b.cap_shadow = nshadow;
    }
    if (!b.counters) {
        void *c = nullptr;
        HIP_TRY(hipMalloc(&c, 16));
        s->render_allocs.push_back(c);
        b.counters = static_cast<uint32_t *>(c);
    }
    // the per-light arrays live behind the ShadeOut records (re-derived each
    // frame because the light count may change between frames)
    ShadeOut *shade = static_cast<ShadeOut *>(b.shade);
    uint32_t *lslot = reinterpret_cast<uint32_t *>(shade + npix);
    double *lterm = reinterpret_cast<double *>(
        reinterpret_cast<char *>(lslot) + ((npix * nl * sizeof(uint32_t) + 7) & ~size_t(7)));

    RenderParams p;
    p.sv = s->view;
    p.obj_type = s->d_obj_type;
    p.obj_index = s->d_obj_index;
    p.mats = sh.d_mats;
    p.lights = sh.d_lights;
    p.n_lights = (uint32_t)sh.lights.size();
    p.max_level = sh.max_level;
    p.bailout = sh.bailout;
    for (int k = 0; k < 3; k++) p.bg[k] = sh.background[k];

    hipEvent_t e0, e1, t0, t1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventCreate(&t0));
    HIP_TRY(hipEventCreate(&t1));
    double trace_ms = 0.0, shadow_ms = 0.0;
    uint64_t n_closest = 0, n_shadow = 0;
    uint32_t bounces = 0;
    HIP_TRY(hipEventRecord(e0, st));
    {
        size_t n3 = 3 * npix;
        k_fill<<<(unsigned)((n3 + 255) / 256), 256, 0, st>>>(rgb, n3, 0.0f);
    }
    int rc = pvx_camera_rays(cam, static_cast<pvx_ray *>(b.rays[0]), st);
    if (rc) return rc;
    uint32_t n = (uint32_t)npix;
    k_init_paths<<<(n + 255) / 256, 256, 0, st>>>(static_cast<pvx_ray *>(b.rays[0]), n,
                                                   static_cast<PathState *>(b.state[0]));
    int cur = 0;
    uint32_t kind = PVX_RAY_IMAGE;
    while (n > 0 && bounces < sh.max_level) {
        const pvx_ray *rays = static_cast<const pvx_ray *>(b.rays[cur]);
        HIP_TRY(hipEventRecord(t0, st));
        rc = trace_closest_internal(s, rays, nullptr, kind, n, static_cast<pvx_hit *>(b.hits), st);
        if (rc) return rc;
        HIP_TRY(hipEventRecord(t1, st));
        n_closest += n;
        k_zero_counters<<<1, 64, 0, st>>>(b.counters);
        k_shade<<<(n + 127) / 128, 128, 0, st>>>(
            p, rays, static_cast<const PathState *>(b.state[cur]), static_cast<const pvx_hit *>(b.hits),
            n, kind, rgb, shade, lslot, lterm, static_cast<pvx_ray *>(b.srays),
            static_cast<double *>(b.sdepth), static_cast<pvx_ray *>(b.rays[1 - cur]),
            static_cast<PathState *>(b.state[1 - cur]), b.counters);
        HIP_TRY(hipGetLastError());
        uint32_t cnt[2];
        HIP_TRY(hipMemcpyAsync(cnt, b.counters, sizeof cnt, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, t0, t1));
        trace_ms += ms;
        if (cnt[1] > 0) {
            HIP_TRY(hipEventRecord(t0, st));
            rc = trace_closest_internal(s, static_cast<const pvx_ray *>(b.srays),
                                        static_cast<const double *>(b.sdepth), PVX_RAY_SHADOW, cnt[1],
                                        static_cast<pvx_hit *>(b.shits), st);
            if (rc) return rc;
            HIP_TRY(hipEventRecord(t1, st));
            n_shadow += cnt[1];
        }
        k_finish<<<(n + 127) / 128, 128, 0, st>>>(p, shade, n, lslot, lterm,
                                                  static_cast<const pvx_hit *>(b.shits),
                                                  static_cast<const double *>(b.sdepth), rgb);
        HIP_TRY(hipGetLastError());
        if (cnt[1] > 0) {
            HIP_TRY(hipEventSynchronize(t1));
            HIP_TRY(hipEventElapsedTime(&ms, t0, t1));
            shadow_ms += ms;
        }
        n = cnt[0];
        cur = 1 - cur;
        kind = PVX_RAY_REFLECTION;
        bounces++;
    }
    HIP_TRY(hipEventRecord(e1, st));
    HIP_TRY(hipEventSynchronize(e1));
    float total = 0.0f;
    HIP_TRY(hipEventElapsedTime(&total, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    if (stats) {
        stats->closest_rays = n_closest;
        stats->shadow_rays = n_shadow;
        stats->bounces = bounces;
        stats->pad = 0;
        stats->trace_ms = trace_ms;
        stats->shadow_ms = shadow_ms;
        stats->total_ms = total;
    }
    return PVX_OK;
}
