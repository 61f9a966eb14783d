// Wavefront shading of plain-finish surfaces (gfx950).
//
// Per bounce: one closest-hit launch over the live rays, one shade launch
// (normal, classic ambient, per-light diffuse/Phong terms and shadow rays,
// the refracted / reflected rays of the next bounce), shadow launches that
// follow each light ray through transparent objects, one finish launch
// (light colour -> surface colour) and one accumulation launch.  Every ray
// writes its own contribution; the queue stays ordered by pixel (children
// are placed by a prefix sum in parent order), so each pixel's rays are
// contiguous and are summed in a fixed order: the frame is deterministic.
//
// Follows Trace::TraceRay / ComputeLightedTexture / ComputeOneDiffuseLight /
// TracePointLightShadowRay / TraceAreaLightSubsetShadowRay /
// ComputeShadowColour / ComputeDiffuseColour / ComputePhongColour /
// ComputeReflectivity / ComputeReflection / ComputeRefraction /
// TraceRefractionRay (core/render/trace.cpp:135-228, 739-1179, 1181-1262,
// 1264-1483, 1637-1729, 1951-2420, 2441-2558, 2627-2656).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "internal.h"
#include "normal_device.h"
#include "shapes_device.h"
#include "trace_body.h"

using namespace pvx;

namespace {

constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr int kMaxInteriors = 10;           // objects a ray can be inside at once
constexpr int kMaxShadowSteps = 64;         // transparent layers followed per light ray

// Ray type bits, as Ray::SetFlags(type, parent) sets them (ray.cpp:155-166).
enum : uint8_t { kFPrimary = 1, kFRefl = 2, kFRefr = 4, kFContinued = 8 };

__host__ __device__ inline uint8_t ray_kind(uint8_t f) { // Ray::IsImageRay etc. (ray.h:99-106)
    uint8_t k = 0;
    if ((f & kFPrimary) || ((f & kFRefr) && !(f & kFRefl))) k |= PVX_RAY_IMAGE;
    if (f & kFRefl) k |= PVX_RAY_REFLECTION;
    return k;
}

struct PathState {
    uint32_t pixel;
    uint16_t level; // trace level when this ray's TraceRay starts
    uint8_t flags;  // kF* bits
    uint8_t n_in;   // interiors entered
    float thr[3];   // product of reflectivities / filter colours along the path
    float weight;   // ADC weight
    int32_t ins[kMaxInteriors];
};
static_assert(sizeof(PathState) == 64, "PathState is one 64-byte line");

struct ShadeOut {
    uint32_t obj; // kNone: nothing for the finish pass
    float thr[3];
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

// counters: [0] light rays, [1] light-ray slots, [2] next queue size,
// [3] lit (hit, light) groups whose light rays the traversal kernel makes
constexpr uint32_t kGroupCounter = 3;
__global__ void k_zero_counters(uint32_t *c) {
    if (threadIdx.x < 64) c[threadIdx.x] = 0;
}

__global__ void k_zero_word(uint32_t *c) {
    if (threadIdx.x == 0) *c = 0;
}

// The queue holds the pixels in tile order (tile_pixel with tw x th tiles,
// 8 x 8 unless pvx_render_autotune or pvx_scene_set_render_tile chose
// others; tw 0 is row-major), a pixel's samples in a run: camera ray of
// sample s of the pixel at tile position p sits at queue position p * S + s.
__device__ inline void queue_pixel(uint32_t W, uint32_t H, uint32_t tw, uint32_t th, uint32_t p, uint32_t &x,
                                   uint32_t &y) {
    if (tw == 0) {
        x = p % W;
        y = p / W;
    } else {
        tile_pixel(W, H, tw, th, p, x, y);
    }
}

// (TracePixel::CreateCameraRay perspective branch, tracepixel.cpp:357-372).
__global__ void k_sample_rays(pvx_camera cam, const double *__restrict__ off, uint32_t S, uint32_t n, uint32_t tw,
                              uint32_t th, pvx_ray *__restrict__ rays) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t s = i % S;
    uint32_t x, y;
    queue_pixel(cam.width, cam.height, tw, th, i / S, x, y);
    double x0 = ((double)x + off[2 * s]) / (double)cam.width - 0.5;
    double y0 = 0.5 - ((double)y + off[2 * s + 1]) / (double)cam.height;
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
    rays[i] = r;
}

__global__ void k_scale3(float *rgb, size_t n3, float k) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n3) rgb[i] *= k;
}

__global__ void k_init_paths(uint32_t n, uint32_t S, uint32_t W, uint32_t H, uint32_t tw, uint32_t th, PathState *st,
                             uint8_t *kinds) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    PathState ps;
    uint32_t x, y;
    queue_pixel(W, H, tw, th, i / S, x, y);
    ps.pixel = y * W + x;
    ps.level = 0;
    ps.flags = kFPrimary;
    ps.n_in = 0;
    ps.thr[0] = ps.thr[1] = ps.thr[2] = 1.0f;
    ps.weight = 1.0f;
    for (int k = 0; k < kMaxInteriors; k++) ps.ins[k] = -1;
    st[i] = ps;
    kinds[i] = ray_kind(kFPrimary);
}

__device__ inline void write_child(pvx_ray *crays, PathState *cst, size_t slot, D3 o, D3 d, const PathState &ps,
                                   const float *thr, float weight, uint8_t flags, uint16_t level) {
    pvx_ray r;
    r.o[0] = o.x; r.o[1] = o.y; r.o[2] = o.z;
    r.d[0] = d.x; r.d[1] = d.y; r.d[2] = d.z;
    crays[slot] = r;
    PathState c = ps;
    c.thr[0] = thr[0];
    c.thr[1] = thr[1];
    c.thr[2] = thr[2];
    c.weight = weight;
    c.flags = flags;
    c.level = level;
    cst[slot] = c;
}

// Shade the hits of one bounce: ambient now, per-light terms and shadow rays
// for the finish pass, and up to two rays of the next bounce (refracted or
// totally reflected first, reflected second) in slots 2i, 2i+1.
// kPacket: light rays are not written; each lit (hit, light) appends its
// shaded point and light to the group list and the traversal kernel makes
// the rays.
// kRev: the scene has lathes / sors (their normal retraces the hit list).
template <bool kPacket, bool kRev>
__global__ void k_shade(RenderParams p, const pvx_ray *__restrict__ rays, const PathState *__restrict__ st,
                        const pvx_hit *__restrict__ hits, uint32_t n, float *__restrict__ contrib,
                        ShadeOut *__restrict__ shade, uint32_t *__restrict__ lslot, double *__restrict__ lterm,
                        pvx_ray *__restrict__ srays, double *__restrict__ sdepth, uint32_t *__restrict__ sowner,
                        float *__restrict__ scol, pvx_ray *__restrict__ crays, PathState *__restrict__ cst,
                        uint32_t *__restrict__ nchild, uint32_t *__restrict__ counters,
                        LightGroup *__restrict__ groups, uint32_t *__restrict__ pix, uint32_t *__restrict__ gcount) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const pvx_ray r = rays[i];
    const PathState ps = st[i];
    pix[i] = ps.pixel;
    const pvx_hit h = hits[i];
    float *ct = contrib + 3 * (size_t)i;
    ShadeOut so;
    so.obj = kNone;
    nchild[i] = 0;
    if (h.obj < 0) { // ComputeSky: background colour
        ct[0] = ps.thr[0] * p.bg[0];
        ct[1] = ps.thr[1] * p.bg[1];
        ct[2] = ps.thr[2] * p.bg[2];
        shade[i] = so;
        return;
    }
    // TraceRay raises the trace level unless the ray continues a transmission
    const uint16_t lev = (uint16_t)(ps.level + ((ps.flags & kFContinued) ? 0 : 1));
    D3 o{r.o[0], r.o[1], r.o[2]}, d{r.d[0], r.d[1], r.d[2]};
    D3 ip = add3(o, mul3(d, h.t));
    D3 nrm = hit_normal<kRev>(p, h, o, d, ip);
    if (dot3(nrm, d) > 0.0) nrm = D3{-nrm.x, -nrm.y, -nrm.z};
    const MatRec m = p.mats[h.obj];
    // classic ambient: pigment * (ambient * ambient_light) * opacity
    ct[0] = ps.thr[0] * (float)((double)m.att * (double)(m.pigment[0] * m.emission[0]));
    ct[1] = ps.thr[1] * (float)((double)m.att * (double)(m.pigment[1] * m.emission[1]));
    ct[2] = ps.thr[2] * (float)((double)m.att * (double)(m.pigment[2] * m.emission[2]));
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
        if (m.phong > 0.0f) {
            double c2 = -2.0 * dot3(d, nrm);
            D3 rd = add3(d, mul3(nrm, c2));
            double ca = dot3(rd, dir);
            if (ca > 0.0 && (m.phong_size < 60 || ca > 0.0008))
                ph = (double)m.phong * pow(ca, (double)m.phong_size);
        }
        lterm[2 * li] = m.diffuse > 0.0f ? inten : 0.0;
        lterm[2 * li + 1] = ph;
        if constexpr (kPacket) {
            atomicAdd(&counters[1], lt.n1 * lt.n2); // light rays (statistics)
            const uint32_t g = atomicAdd(gcount, 1u);
            lslot[li] = g;
            groups[g] = LightGroup{{ip.x, ip.y, ip.z}, i, L};
            continue;
        }
        uint32_t ns = lt.n1 * lt.n2;
        uint32_t base = atomicAdd(&counters[1], ns);
        lslot[li] = base;
        for (uint32_t u = 0; u < lt.n1; u++)
            for (uint32_t v = 0; v < lt.n2; v++) {
                D3 sd;
                double sl;
                light_sample_ray(lt, u, v, ip, sd, sl);
                pvx_ray sr;
                sr.o[0] = ip.x; sr.o[1] = ip.y; sr.o[2] = ip.z;
                sr.d[0] = sd.x; sr.d[1] = sd.y; sr.d[2] = sd.z;
                uint32_t k = base + u * lt.n2 + v;
                srays[k] = sr;
                sdepth[k] = sl;
                sowner[k] = k;
                scol[3 * (size_t)k] = scol[3 * (size_t)k + 1] = scol[3 * (size_t)k + 2] = 1.0f;
            }
    }
    so.obj = (uint32_t)h.obj;
    so.thr[0] = ps.thr[0];
    so.thr[1] = ps.thr[1];
    so.thr[2] = ps.thr[2];
    shade[i] = so;

    uint32_t nc = 0;
    bool tir = false;
    // transmitted component (ComputeLightedTexture, trace.cpp:1075-1131);
    // none when every material is opaque (kPacket)
    if (!kPacket && m.trans > p.bailout) {
        const float nw = (float)((double)ps.weight * (double)m.fmax);
        // ComputeRefraction (trace.cpp:1323-1452): ior ratio from the interiors
        PathState c = ps;
        double ior;
        const double own = (double)m.ior;
        if (c.n_in == 0) {
            c.ins[c.n_in++] = h.obj;
            ior = 1.0 / own;
        } else if (c.ins[c.n_in - 1] == h.obj) {
            c.n_in--;
            ior = c.n_in == 0 ? own : own / (double)p.mats[c.ins[c.n_in - 1]].ior;
        } else {
            int at = -1;
            for (int k = 0; k < c.n_in; k++)
                if (c.ins[k] == h.obj) {
                    at = k;
                    break;
                }
            if (at >= 0) { // leaving an overlap
                for (int k = at; k + 1 < c.n_in; k++) c.ins[k] = c.ins[k + 1];
                c.n_in--;
                ior = 1.0;
            } else {
                ior = (double)p.mats[c.ins[c.n_in - 1]].ior / own;
                if (c.n_in < kMaxInteriors) c.ins[c.n_in++] = h.obj;
            }
        }
        const uint8_t rf = (uint8_t)(kFRefr | (ps.flags & kFRefl));
        const size_t slot = 2 * (size_t)i;
        if (fabs(ior - 1.0) < kEpsilon) { // only transmit the ray (a continued ray)
            float thr[3] = {ps.thr[0] * m.fil[0], ps.thr[1] * m.fil[1], ps.thr[2] * m.fil[2]};
            if (lev < p.max_level && nw >= p.bailout) {
                write_child(crays, cst, slot, ip, d, c, thr, nw, (uint8_t)(rf | kFContinued), lev);
                nc = 1;
            }
        } else {
            double nn = dot3(d, nrm);
            D3 ln = nrm;
            if (nn <= 0.0)
                nn = -nn;
            else
                ln = D3{-nrm.x, -nrm.y, -nrm.z};
            double t = 1.0 + ior * ior * (nn * nn - 1.0);
            if (t < 0.0) { // total internal reflection: a reflected ray, unfiltered
                tir = true;
                double k2 = -2.0 * dot3(d, nrm);
                D3 nd = normalize3(add3(d, mul3(nrm, k2)));
                if (lev < p.max_level && nw >= p.bailout) {
                    write_child(crays, cst, slot, ip, nd, ps, ps.thr, nw,
                                (uint8_t)(kFRefl | (ps.flags & kFRefr)), lev);
                    nc = 1;
                }
            } else {
                t = ior * nn - sqrt(t);
                D3 nd = add3(mul3(d, ior), mul3(ln, t));
                float thr[3] = {ps.thr[0] * m.fil[0], ps.thr[1] * m.fil[1], ps.thr[2] * m.fil[2]};
                if (lev < p.max_level && nw >= p.bailout) {
                    write_child(crays, cst, slot, ip, nd, c, thr, nw, rf, lev);
                    nc = 1;
                }
            }
        }
    }
    // reflection (ComputeReflectivity without fresnel, ComputeReflection);
    // skipped after total internal reflection (trace.cpp:1141-1147)
    if ((m.reflective & 1u) && !tir) {
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
        if (nonzero && lev < p.max_level && (float)w >= p.bailout) {
            double k2 = -2.0 * dot3(d, nrm);
            D3 nd = normalize3(add3(d, mul3(nrm, k2)));
            float thr[3] = {ps.thr[0] * rf[0], ps.thr[1] * rf[1], ps.thr[2] * rf[2]};
            write_child(crays, cst, 2 * (size_t)i + nc, ip, nd, ps, thr, (float)w,
                        (uint8_t)(kFRefl | (ps.flags & kFRefr)), lev);
            nc++;
        }
    }
    nchild[i] = nc;
}

// One step of TracePointLightShadowRay's loop (trace.cpp:2021-2073): a light
// ray blocked by a transparent object is filtered by it and continues from
// the hit point; an opaque blocker ends it.
__global__ void k_shadow_step(RenderParams p, const pvx_ray *__restrict__ srays, const double *__restrict__ sdepth,
                              const uint32_t *__restrict__ sowner, const pvx_hit *__restrict__ shits, uint32_t n,
                              float *__restrict__ scol, pvx_ray *__restrict__ nrays, double *__restrict__ ndepth,
                              uint32_t *__restrict__ nowner, uint32_t *__restrict__ counter) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const pvx_hit h = shits[k];
    const double lsd = sdepth[k];
    if (!(h.obj >= 0 && h.t < lsd - kShadowTolerance && h.t > kShadowTolerance)) return;
    const uint32_t ow = sowner[k];
    const MatRec &m = p.mats[h.obj];
    float *c = scol + 3 * (size_t)ow;
    if (m.opaque) {
        c[0] = c[1] = c[2] = 0.0f;
        return;
    }
    c[0] *= m.fil[0];
    c[1] *= m.fil[1];
    c[2] *= m.fil[2];
    const pvx_ray r = srays[k];
    pvx_ray nr = r;
    for (int a = 0; a < 3; a++) nr.o[a] = r.o[a] + r.d[a] * h.t;
    uint32_t slot = atomicAdd(counter, 1u);
    nrays[slot] = nr;
    ndepth[slot] = lsd - h.t;
    nowner[slot] = ow;
}

// Light rays in packets (triangle / sphere scenes, every material opaque):
// each wavefront takes 64 consecutive rays (ray k = group k / S, grid point
// k % S of the group's light; consecutive groups are neighbouring hits) and
// walks the 4-wide tree as one packet: one node sequence for the whole
// wavefront with one stack in LDS, a child visited when any live lane's ray
// enters its box within that ray's light distance (ballot), children in the
// entry order of the first live lane.  A lane's ray is blocked by the first
// valid hit closer than the light minus SHADOW_TOLERANCE (the occlusion test
// of TracePointLightShadowRay, trace.cpp:2021-2073, with an opaque blocker);
// a lane with a blocked ray takes no further part, and the walk ends when no
// lane is live.  Every hit a lane's own walk could find lies in a box that
// lane enters, so the visibility is the per-ray walk's: vis[k] = 1 when
// nothing blocks.  The stack holds at most 3 entries per tree level (the
// scene's 4-wide tree qualifies only when 3 * depth + 1 < kStackMax).
constexpr int kPacketWaves = 4; // wavefronts per block of 256
template <bool kFlags>
__global__ __launch_bounds__(256) void k_light_packet(SceneView sv, const LightRec *__restrict__ lights,
                                                      const LightGroup *__restrict__ groups, uint32_t S,
                                                      uint32_t n_bound, const uint32_t *__restrict__ count,
                                                      uint8_t *__restrict__ vis) {
    // the group count on the device when given (n_bound then only sized the grid)
    const uint32_t n_rays = count ? count[0] * S : n_bound;
    __shared__ int32_t wstack[kPacketWaves][kStackMax];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    int32_t *stk = wstack[wave];
    const uint32_t stride = gridDim.x * kPacketWaves * 64u;
    for (uint32_t base = (blockIdx.x * kPacketWaves + wave) * 64u; base < n_rays; base += stride) {
        const uint32_t k = base + lane;
        bool valid = k < n_rays;
        D3 o{0.0, 0.0, 0.0}, d{0.0, 0.0, 1.0};
        double sl = 1.0;
        if (valid) {
            const uint32_t g = k / S, sidx = k - g * S;
            const LightGroup grp = groups[g];
            const LightRec &lt = lights[grp.light];
            if (sidx < lt.n1 * lt.n2) {
                o = D3{grp.ip[0], grp.ip[1], grp.ip[2]};
                light_sample_ray(lt, sidx / lt.n2, sidx % lt.n2, o, d, sl);
            } else {
                valid = false; // past this light's grid
            }
        }
        const double occl_lim = sl - kShadowTolerance;
        bool blocked = false;
        // objects outside the tree (planes)
        if (valid)
            for (uint32_t q = 0; q < sv.n_inf; q++) {
                const InfRec inf = sv.inf[q];
                if (kFlags && !precondition(sv.obj_flags[inf.obj], PVX_RAY_SHADOW)) continue;
                double dep;
                if (inf.kind == kPrimPlane && plane_hit(sv.planes[inf.index], o, d, dep) && depth_ok(dep, true) &&
                    dep < occl_lim)
                    blocked = true;
            }
        RayF rf;
        rf.ix = safe_inv(d.x);
        rf.iy = safe_inv(d.y);
        rf.iz = safe_inv(d.z);
        rf.oix = (float)o.x * rf.ix;
        rf.oiy = (float)o.y * rf.iy;
        rf.oiz = (float)o.z * rf.iz;
        const float tb = (float)sl * 1.000001f;
        uint64_t live = __ballot(valid && !blocked);
        int32_t node = live ? sv.root4 : kEmpty;
        int sp = 0;
        while (node != kEmpty) {
            node = __builtin_amdgcn_readfirstlane(node);
            if (node >= 0) {
                const float4 *np = reinterpret_cast<const float4 *>(sv.nodes4 + node);
                const float4 xl = np[0], xh = np[1], yl = np[2], yh = np[3], zl = np[4], zh = np[5];
                const int4 cl = reinterpret_cast<const int4 *>(np)[6];
                const bool me = valid && !blocked;
                auto slab = [&](float lx, float hx, float ly, float hy, float lz, float hz, int32_t link) {
                    const float x0 = fmaf(lx, rf.ix, -rf.oix), x1 = fmaf(hx, rf.ix, -rf.oix);
                    const float y0 = fmaf(ly, rf.iy, -rf.oiy), y1 = fmaf(hy, rf.iy, -rf.oiy);
                    const float z0 = fmaf(lz, rf.iz, -rf.oiz), z1 = fmaf(hz, rf.iz, -rf.oiz);
                    const float tn = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), 0.0f));
                    const float tf = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fminf(fmaxf(z0, z1), tb));
                    return (me && link != kEmpty && tn <= tf * 1.000001f) ? tn : INFINITY;
                };
                const float e[4] = {slab(xl.x, xh.x, yl.x, yh.x, zl.x, zh.x, cl.x),
                                    slab(xl.y, xh.y, yl.y, yh.y, zl.y, zh.y, cl.y),
                                    slab(xl.z, xh.z, yl.z, yh.z, zl.z, zh.z, cl.z),
                                    slab(xl.w, xh.w, yl.w, yh.w, zl.w, zh.w, cl.w)};
                const int32_t links[4] = {cl.x, cl.y, cl.z, cl.w};
                const int leader = __builtin_amdgcn_readfirstlane(__ffsll((unsigned long long)live) - 1);
                // wave-uniform order keys: the leader's entry distance, after
                // it the children only other lanes enter (in slot order)
                float key[4];
                int32_t lk[4];
                int cnt = 0;
                for (int c = 0; c < 4; c++) {
                    const uint64_t m = __ballot(e[c] != INFINITY);
                    if (!m) continue;
                    const float kl = __builtin_bit_cast(
                        float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, e[c]), leader));
                    key[cnt] = kl != INFINITY ? kl : 3.0e38f; // the sort is stable: slot order
                    lk[cnt] = links[c];
                    cnt++;
                }
                if (cnt == 0) {
                    node = sp == 0 ? kEmpty : stk[--sp];
                    continue;
                }
                for (int a = 1; a < cnt; a++) // insertion sort of <= 4 uniform entries
                    for (int b = a; b > 0 && key[b] < key[b - 1]; b--) {
                        const float tk = key[b];
                        key[b] = key[b - 1];
                        key[b - 1] = tk;
                        const int32_t tl = lk[b];
                        lk[b] = lk[b - 1];
                        lk[b - 1] = tl;
                    }
                for (int a = cnt - 1; a >= 1; a--) stk[sp++] = lk[a]; // farthest first
                node = lk[0];
            } else {
                const uint32_t leaf = (uint32_t)(~node);
                const uint32_t first = leaf >> 3, count = (leaf & 7u) + 1u;
                for (uint32_t q = 0; q < count; q++) {
                    const PrimRec &pr = sv.prims[first + q];
                    const uint32_t ka = pr.kind_axis;
                    const uint32_t pk = ka & 15u;
                    if (!(valid && !blocked)) continue;
                    if (kFlags && !precondition(sv.obj_flags[pr.obj], PVX_RAY_SHADOW)) continue;
                    if (pk == kPrimTriangle) {
                        double dep;
                        if (mesh_tri_hit_l<0>(sv, pr, ka, o, d, dep) && depth_ok(dep, true) && dep < occl_lim)
                            blocked = true;
                    } else if (pk == kPrimSphere) {
                        double hs[2];
                        const int nh = sphere_hits(d3(pr.sph.c[0], pr.sph.c[1], pr.sph.c[2]), pr.sph.r, o, d, hs);
                        for (int j = 0; j < nh; j++)
                            if (depth_ok(hs[j], true) && hs[j] < occl_lim) blocked = true;
                    }
                }
                live = __ballot(valid && !blocked);
                if (!live) break;
                node = sp == 0 ? kEmpty : stk[--sp];
            }
        }
        if (valid) vis[k] = blocked ? 0 : 1;
    }
}

// TraceAreaLightSubsetShadowRay's quadtree average with adaptive 100 (every
// region subdivides until it spans <= 2x2 grid points), the three colour
// channels together; get(k) is the light colour that reached sample k.
struct F3 {
    float x, y, z;
};
template <int D, class Get>
__device__ F3 area_average(const Get &get, int n2, int u1, int v1, int u2, int v2) {
    F3 c0, c1, c2, c3;
    if (D > 0 && ((u2 - u1 > 1) || (v2 - v1 > 1))) {
        if constexpr (D > 0) {
            const int mu0 = (u1 + u2) / 2, mu1 = (u1 + u2 + 1) / 2;
            const int mv0 = (v1 + v2) / 2, mv1 = (v1 + v2 + 1) / 2;
            c0 = area_average<D - 1>(get, n2, u1, v1, mu0, mv0);
            c1 = area_average<D - 1>(get, n2, mu1, v1, u2, mv0);
            c2 = area_average<D - 1>(get, n2, u1, mv1, mu0, v2);
            c3 = area_average<D - 1>(get, n2, mu1, mv1, u2, v2);
        }
    } else {
        c0 = get(u1 * n2 + v1);
        c1 = get(u2 * n2 + v1);
        c2 = get(u1 * n2 + v2);
        c3 = get(u2 * n2 + v2);
    }
    return F3{(c0.x + c1.x + c2.x + c3.x) * 0.25f, (c0.y + c1.y + c2.y + c3.y) * 0.25f,
              (c0.z + c1.z + c2.z + c3.z) * 0.25f};
}

constexpr int kMaxAreaGrid = 16; // per axis
// A region spanning w grid steps splits into two of floor(w / 2) steps until
// w <= 1: 15 -> 7 -> 3 -> 1, so three levels of splits cover 16 points.
constexpr int kAreaSplits = 3;

// Light colour -> surface colour for each shaded hit, added to its contribution.
// kPacket: the light colour of sample k of group g is the light's colour
// times the visibility byte vis[g * S + k].
template <bool kPacket>
__global__ void k_finish(RenderParams p, const ShadeOut *__restrict__ shade, uint32_t n,
                         const uint32_t *__restrict__ lslot, const double *__restrict__ lterm,
                         const float *__restrict__ scol, float *__restrict__ contrib,
                         const uint8_t *__restrict__ vis, uint32_t S) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const ShadeOut so = shade[i];
    if (so.obj == kNone) return;
    const MatRec &m = p.mats[so.obj];
    float col[3] = {0.0f, 0.0f, 0.0f};
    for (uint32_t L = 0; L < p.n_lights; L++) {
        size_t li = (size_t)i * p.n_lights + L;
        uint32_t base = lslot[li];
        if (base == kNone) continue;
        const LightRec &lt = p.lights[L];
        float lc[3];
        const uint32_t ns = lt.n1 * lt.n2;
        F3 avg;
        if constexpr (kPacket) {
            // visibility bits of the group's samples (at most 16 x 16)
            uint64_t m0 = 0, m1 = 0, m2 = 0, m3 = 0;
            const uint8_t *vb = vis + (size_t)base * S;
            for (uint32_t k = 0; k < ns; k++) {
                const uint64_t b = (uint64_t)(vb[k] != 0) << (k & 63u);
                if (k < 64) m0 |= b;
                else if (k < 128) m1 |= b;
                else if (k < 192) m2 |= b;
                else m3 |= b;
            }
            auto get = [&](int k) -> F3 {
                const uint64_t w = k < 128 ? (k < 64 ? m0 : m1) : (k < 192 ? m2 : m3);
                const float v = ((w >> (k & 63)) & 1u) ? 1.0f : 0.0f;
                return F3{lt.colour[0] * v, lt.colour[1] * v, lt.colour[2] * v};
            };
            avg = ns == 1 ? get(0) : area_average<kAreaSplits>(get, (int)lt.n2, 0, 0, (int)lt.n1 - 1, (int)lt.n2 - 1);
        } else {
            // the filter colour that reached the light from sample k
            auto get = [&](int k) -> F3 {
                const float *c = scol + 3 * ((size_t)base + (size_t)k);
                return F3{lt.colour[0] * c[0], lt.colour[1] * c[1], lt.colour[2] * c[2]};
            };
            avg = ns == 1 ? get(0) : area_average<kAreaSplits>(get, (int)lt.n2, 0, 0, (int)lt.n1 - 1, (int)lt.n2 - 1);
        }
        lc[0] = avg.x;
        lc[1] = avg.y;
        lc[2] = avg.z;
        if (fabsf(lc[0]) < 1e-10f && fabsf(lc[1]) < 1e-10f && fabsf(lc[2]) < 1e-10f) continue;
        double inten = lterm[2 * li], ph = lterm[2 * li + 1];
        for (int k = 0; k < 3; k++) {
            if (m.diffuse > 0.0f) col[k] = col[k] + (float)(inten * (double)(m.pigment[k] * lc[k]));
        }
        if (ph != 0.0)
            for (int k = 0; k < 3; k++) col[k] = col[k] + (float)(ph * (double)lc[k]);
    }
    float *ct = contrib + 3 * (size_t)i;
    ct[0] += so.thr[0] * col[0];
    ct[1] += so.thr[1] * col[1];
    ct[2] += so.thr[2] * col[2];
}

// Adds each pixel's contributions in queue order (the queue is ordered by
// pixel, so a pixel's rays form one run).
__global__ void k_accumulate(const uint32_t *__restrict__ pix, const float *__restrict__ contrib, uint32_t n,
                             float *__restrict__ rgb) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t px = pix[i];
    if (i > 0 && pix[i - 1] == px) return;
    float c[3] = {0.0f, 0.0f, 0.0f};
    for (uint32_t j = i; j < n && pix[j] == px; j++)
        for (int a = 0; a < 3; a++) c[a] += contrib[3 * (size_t)j + a];
    float *o = rgb + 3 * (size_t)px;
    o[0] += c[0];
    o[1] += c[1];
    o[2] += c[2];
}

// Exclusive prefix sum of child counts: per-block scan (1024 per block) ...
constexpr int kScanBlock = 1024;
__global__ __launch_bounds__(kScanBlock) void k_scan_block(const uint32_t *__restrict__ in, uint32_t n,
                                                           uint32_t *__restrict__ out, uint32_t *__restrict__ sums) {
    __shared__ uint32_t s[kScanBlock];
    const uint32_t t = threadIdx.x, i = blockIdx.x * kScanBlock + t;
    uint32_t v = i < n ? in[i] : 0u;
    s[t] = v;
    __syncthreads();
    for (int off = 1; off < kScanBlock; off <<= 1) {
        uint32_t a = t >= (uint32_t)off ? s[t - off] : 0u;
        __syncthreads();
        s[t] += a;
        __syncthreads();
    }
    if (i < n) out[i] = s[t] - v;
    if (t == kScanBlock - 1) sums[blockIdx.x] = s[t];
}
// ... scan of the block sums (one block, up to 1024 x 8 sums) ...
__global__ __launch_bounds__(kScanBlock) void k_scan_sums(uint32_t *__restrict__ sums, uint32_t nb,
                                                          uint32_t *__restrict__ total) {
    __shared__ uint32_t s[kScanBlock];
    const uint32_t t = threadIdx.x;
    constexpr int kPer = 8;
    uint32_t loc[kPer], acc = 0;
    for (int k = 0; k < kPer; k++) {
        uint32_t idx = t * kPer + k;
        loc[k] = acc;
        acc += idx < nb ? sums[idx] : 0u;
    }
    s[t] = acc;
    __syncthreads();
    for (int off = 1; off < kScanBlock; off <<= 1) {
        uint32_t a = t >= (uint32_t)off ? s[t - off] : 0u;
        __syncthreads();
        s[t] += a;
        __syncthreads();
    }
    const uint32_t before = s[t] - acc;
    for (int k = 0; k < kPer; k++) {
        uint32_t idx = t * kPer + k;
        if (idx < nb) sums[idx] = before + loc[k];
    }
    if (t == kScanBlock - 1) *total = s[t];
}
// ... then every child is copied to its place in the next queue.
__global__ void k_emit(const uint32_t *__restrict__ nchild, const uint32_t *__restrict__ offs,
                       const uint32_t *__restrict__ sums, uint32_t n, const pvx_ray *__restrict__ crays,
                       const PathState *__restrict__ cst, pvx_ray *__restrict__ next_rays,
                       PathState *__restrict__ next_st, uint8_t *__restrict__ next_kinds) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t c = nchild[i];
    if (!c) return;
    const uint32_t o = offs[i] + sums[i / kScanBlock];
    for (uint32_t j = 0; j < c; j++) {
        next_rays[o + j] = crays[2 * (size_t)i + j];
        const PathState ps = cst[2 * (size_t)i + j];
        next_st[o + j] = ps;
        next_kinds[o + j] = ray_kind(ps.flags);
    }
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
    bool all_opaque = true;
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
        // TransColour::TransmittedColour (base/colour.h:1912-1915) and the
        // remaining translucency of ComputeLightedTexture (trace.cpp:1068)
        for (int k = 0; k < 3; k++) m.fil[k] = a.pigment[k] * a.filter + a.transmit;
        m.trans = std::min(1.0f, std::fabs(0.297f * m.fil[0] + 0.589f * m.fil[1] + 0.114f * m.fil[2]));
        m.fmax = std::max(std::fabs(m.fil[0]), std::max(std::fabs(m.fil[1]), std::fabs(m.fil[2])));
        m.ior = a.ior > 0.0f ? a.ior : 1.0f; // every object has an interior (parser.cpp:9226-9230)
        m.opaque = (a.filter == 0.0f && a.transmit == 0.0f) ? 1u : 0u;
        if (m.trans > 0.0f) m.reflective |= 2u;
        if (!m.opaque) all_opaque = false;
    }
    sh.all_opaque = all_opaque;
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
    if (!cam) return fail(PVX_EINVAL, "null argument");
    const double off[2] = {cam->jitter_x, cam->jitter_y};
    return pvx_render_samples(s, cam, 1, off, rgb, stats, stream);
}

static int render_frame(pvx_scene *s, const pvx_camera *cam, uint32_t S, const double *offsets, float *rgb,
                        pvx_render_stats *stats, void *stream, pvx_tune_result *tune);

extern "C" int pvx_render_samples(pvx_scene *s, const pvx_camera *cam, uint32_t S, const double *offsets,
                                  float *rgb, pvx_render_stats *stats, void *stream) {
    return render_frame(s, cam, S, offsets, rgb, stats, stream, nullptr);
}

extern "C" int pvx_scene_set_render_tile(pvx_scene *s, uint32_t tile_w, uint32_t tile_h) {
    if (!s) return fail(PVX_EINVAL, "null scene");
    if (tile_w == 0 ? tile_h != 1 : (tile_h == 0 || tile_w > 4096 || tile_h > 4096))
        return fail(PVX_EINVAL, "tile must be w x h in [1, 4096], or 0 x 1 (row-major)");
    s->render_tile[0] = tile_w;
    s->render_tile[1] = tile_h;
    return PVX_OK;
}

// One frame that first tunes the queue's pixel tile and the closest-hit walk
// on its own camera rays, and the light walks on its light rays.
extern "C" int pvx_render_autotune(pvx_scene *s, const pvx_camera *cam, float *rgb, pvx_tune_result *tune,
                                   pvx_render_stats *stats, void *stream) {
    if (!cam) return fail(PVX_EINVAL, "null argument");
    pvx_tune_result local[2];
    const double off[2] = {cam->jitter_x, cam->jitter_y};
    return render_frame(s, cam, 1, off, rgb, stats, stream, tune ? tune : local);
}

static int render_frame(pvx_scene *s, const pvx_camera *cam, uint32_t S, const double *offsets, float *rgb,
                        pvx_render_stats *stats, void *stream, pvx_tune_result *tune) {
    if (!s || !cam || !rgb || !offsets) return fail(PVX_EINVAL, "null argument");
    if (S == 0 || S > 256) return fail(PVX_EINVAL, "n_samples must be in [1, 256]");
    ShadingState &sh = s->shading;
    if (!sh.set) return fail(PVX_ESTATE, "pvx_scene_set_shading was not called");
    if (cam->width == 0 || cam->height == 0 || cam->height > 65535)
        return fail(PVX_EINVAL, "bad image size");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const size_t npix = (size_t)cam->width * cam->height;
    const size_t nray = npix * S;
    if (nray > 0x3FFFFFFFu) return fail(PVX_EINVAL, "image too large");
    RenderBuffers &b = s->rb;
    const size_t nl = std::max<size_t>(1, sh.lights.size());
    const size_t spp = std::max<uint32_t>(1, sh.samples_per_hit);
    int r;
    // queue buffers sized for `cap` rays; grown when a bounce needs more
    auto ensure_queue = [&](size_t cap) -> int {
        int rr;
        if ((rr = b.rays[0].reserve(cap * sizeof(pvx_ray))) || (rr = b.rays[1].reserve(cap * sizeof(pvx_ray))) ||
            (rr = b.state[0].reserve(cap * sizeof(PathState))) || (rr = b.state[1].reserve(cap * sizeof(PathState))) ||
            (rr = b.kinds[0].reserve(cap)) || (rr = b.kinds[1].reserve(cap)) ||
            (rr = b.hits.reserve(cap * sizeof(pvx_hit))) || (rr = b.contrib.reserve(cap * 12)) ||
            (rr = b.shade.reserve(cap * sizeof(ShadeOut))) || (rr = b.lslot.reserve(cap * nl * 4)) ||
            (rr = b.lterm.reserve(cap * nl * 16)) || (rr = b.crays.reserve(2 * cap * sizeof(pvx_ray))) ||
            (rr = b.cstate.reserve(2 * cap * sizeof(PathState))) || (rr = b.nchild.reserve(cap * 4)) ||
            (rr = b.offs.reserve(cap * 4)) || (rr = b.sums.reserve((cap / kScanBlock + 2) * 4)) ||
            (rr = b.counters.reserve(64)) || (rr = b.pix[0].reserve(cap * 4)) || (rr = b.pix[1].reserve(cap * 4)) ||
            (rr = b.contrib_b.reserve(cap * 12)) || (rr = b.shade_b.reserve(cap * sizeof(ShadeOut))) ||
            (rr = b.lslot_b.reserve(cap * nl * 4)) || (rr = b.lterm_b.reserve(cap * nl * 16)))
            return rr;
        return PVX_OK;
    };
    auto ensure_shadow = [&](size_t ns) -> int {
        int rr;
        for (int k = 0; k < 2; k++)
            if ((rr = b.srays[k].reserve(ns * sizeof(pvx_ray))) || (rr = b.sdepth[k].reserve(ns * 8)) ||
                (rr = b.sowner[k].reserve(ns * 4)))
                return rr;
        if ((rr = b.shits.reserve(ns * sizeof(pvx_hit))) || (rr = b.scol.reserve(ns * 12))) return rr;
        return PVX_OK;
    };
    // light rays in packets: triangle / sphere scenes walked through the
    // 4-wide tree, every material opaque, unless the caller picked another
    // walk for light rays
    const bool packet = sh.all_opaque && s->view.has_complex == 0 && s->view.wide && !s->use_kd &&
                        (s->tune_variant[1] == PVX_TRAVERSAL_DEFAULT ||
                         is_light_walk(s->tune_variant[1]));
    uint32_t smax = 1; // rays per group: the largest light grid
    for (const LightRec &lt : sh.lights) smax = std::max(smax, lt.n1 * lt.n2);
    auto ensure_packet = [&](size_t nh) -> int {
        int rr;
        if ((rr = b.groups.reserve(nh * nl * sizeof(LightGroup))) ||
            (rr = b.groups_b.reserve(nh * nl * sizeof(LightGroup))) || (rr = b.vis.reserve(nh * nl * smax)))
            return rr;
        return PVX_OK;
    };
    // With light rays made in the kernel, a bounce's light pass, k_finish and
    // k_accumulate run on a second stream while the next bounces trace and
    // shade: odd bounces keep their per-hit buffers in a second set, the
    // light stream keeps the bounce order of the accumulation into rgb
    // (so the image is the sequential one, bit for bit), and bounce j waits
    // for the light pass of bounce j - 2 before reusing its buffers.
    const bool overlap = packet;
    if ((r = ensure_queue(nray)) || (r = b.offsets.reserve(16 * S)) || (r = b.counters.reserve(256)) ||
        (r = b.refill.reserve(256)) || (r = b.gcount.reserve(256)))
        return r;
    if ((r = packet ? ensure_packet(nray) : ensure_shadow(nray * spp))) return r;
    HIP_TRY(hipMemcpyAsync(b.offsets.p, offsets, 16 * S, hipMemcpyHostToDevice, st));
    hipStream_t ls = st;
    if (overlap) {
        if (!b.light_stream.s) HIP_TRY(hipStreamCreateWithFlags(&b.light_stream.s, hipStreamNonBlocking));
        ls = b.light_stream.s;
    }
    // per-hit buffers of bounce parity `par` (the second set only when overlapping)
    struct HitBufs {
        float *contrib;
        ShadeOut *shade;
        uint32_t *lslot;
        double *lterm;
        LightGroup *groups;
        uint32_t *pix;
    };
    auto hit_bufs = [&](int par) -> HitBufs {
        const bool second = overlap && par == 1;
        return HitBufs{static_cast<float *>(second ? b.contrib_b.p : b.contrib.p),
                       static_cast<ShadeOut *>(second ? b.shade_b.p : b.shade.p),
                       static_cast<uint32_t *>(second ? b.lslot_b.p : b.lslot.p),
                       static_cast<double *>(second ? b.lterm_b.p : b.lterm.p),
                       static_cast<LightGroup *>(second ? b.groups_b.p : b.groups.p),
                       static_cast<uint32_t *>(b.pix[overlap ? par : 0].p)};
    };

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
    // events per bounce k: [4k] after k_shade (st), [4k+1], [4k+2] around the
    // light pass (ls), [4k+3] after k_accumulate (ls)
    std::vector<int> timed_light; // bounces whose light pass is timed at the end
    auto ev = [&](uint32_t k, int j) { return b.events.ev[4 * (size_t)k + j]; };
    HIP_TRY(hipEventRecord(e0, st));
    {
        size_t n3 = 3 * npix;
        k_fill<<<(unsigned)((n3 + 255) / 256), 256, 0, st>>>(rgb, n3, 0.0f);
    }
    uint32_t n = (uint32_t)nray;
    auto sample_rays = [&](uint32_t tw, uint32_t th) {
        k_sample_rays<<<(n + 255) / 256, 256, 0, st>>>(*cam, static_cast<const double *>(b.offsets.p), S, n, tw, th,
                                                       static_cast<pvx_ray *>(b.rays[0].p));
    };
    if (tune) {
        // the queue's pixel order: each candidate tile's camera rays are
        // timed with every walk and grid the scene admits; the fastest
        // {tile, walk, grid} is kept for this and later frames
        const uint32_t cands[4][2] = {{8, 8}, {16, 4}, {32, 2}, {0, 1}};
        pvx_tune_result best{};
        for (const auto &c : cands) {
            sample_rays(c[0], c[1]);
            HIP_TRY(hipGetLastError());
            pvx_tune_result tr{};
            if ((r = autotune_internal(s, static_cast<const pvx_ray *>(b.rays[0].p), nullptr, n, PVX_RAY_IMAGE, &tr,
                                       st)))
                return r;
            if (best.best_ms == 0.0 || tr.best_ms < best.best_ms) {
                best = tr;
                best.tile_w = c[0];
                best.tile_h = c[1];
            }
        }
        s->render_tile[0] = best.tile_w;
        s->render_tile[1] = best.tile_h;
        s->tune_variant[0] = best.variant;
        s->tune_grid[0] = best.grid_blocks;
        tune[0] = best;
    }
    const uint32_t tw = s->render_tile[0], th = s->render_tile[1];
    sample_rays(tw, th);
    k_init_paths<<<(n + 255) / 256, 256, 0, st>>>(n, S, cam->width, cam->height, tw, th,
                                                  static_cast<PathState *>(b.state[0].p),
                                                  static_cast<uint8_t *>(b.kinds[0].p));
    int rc = PVX_OK;
    int cur = 0;
    // with every material opaque a light ray only needs to know whether
    // something blocks it: the traversal may stop at the first blocking hit
    const uint32_t light_kind = PVX_RAY_SHADOW | (sh.all_opaque ? kRayOcclusion : 0u);
    // every bounce raises the trace level of all but continued rays; the cap
    // only guards against pathological chains of continued rays
    const uint32_t max_bounces = 4 * std::max<uint32_t>(1, sh.max_level) + 16;
    // buffers the light stream may still be reading are reallocated only
    // after it has drained
    auto drain_light = [&]() -> int {
        if (overlap) HIP_TRY(hipStreamSynchronize(ls));
        return PVX_OK;
    };
    auto packet_fits = [&](size_t nh) {
        return nh * nl * sizeof(LightGroup) <= b.groups.cap && nh * nl * sizeof(LightGroup) <= b.groups_b.cap &&
               nh * nl * smax <= b.vis.cap;
    };
    // a light pass on the light stream: walk, grid (0: the walk's default),
    // rays to launch for (with `count`, the kernel reads the number of groups
    // there and n_launch only sizes the grid)
    auto light_pass = [&](const LightGroup *groups, uint32_t variant, uint32_t grid, uint64_t n_launch,
                          const uint32_t *count) -> int {
        if (n_launch == 0) return PVX_OK;
        uint8_t *vl = static_cast<uint8_t *>(b.vis.p);
        if (variant == PVX_TRAVERSAL_PACKET) {
            const uint32_t nb = (uint32_t)std::min<uint64_t>(grid ? grid : s->grid, (n_launch + 255) / 256);
            if (s->view.any_flags)
                k_light_packet<true><<<nb, 256, 0, ls>>>(s->view, sh.d_lights, groups, smax, (uint32_t)n_launch, count,
                                                         vl);
            else
                k_light_packet<false><<<nb, 256, 0, ls>>>(s->view, sh.d_lights, groups, smax, (uint32_t)n_launch, count,
                                                          vl);
            HIP_TRY(hipGetLastError());
            return PVX_OK;
        }
        uint32_t *next = nullptr;
        if (variant == PVX_TRAVERSAL_REFILL || variant == PVX_TRAVERSAL_REFILL_BINARY) {
            next = static_cast<uint32_t *>(b.refill.p);
            k_zero_word<<<1, 64, 0, ls>>>(next);
        }
        return light_rays_internal(s, groups, sh.d_lights, smax, (uint32_t)n_launch, vl, next, grid, count,
                                   variant == PVX_TRAVERSAL_LIGHT_BINARY || variant == PVX_TRAVERSAL_REFILL_BINARY, ls);
    };
    while (n > 0 && bounces < max_bounces) {
        const int par = (int)(bounces & 1u);
        if (overlap) HIP_TRY(b.events.reserve(4 * ((size_t)bounces + 1)));
        const pvx_ray *rays = static_cast<const pvx_ray *>(b.rays[cur].p);
        const PathState *pst = static_cast<const PathState *>(b.state[cur].p);
        HIP_TRY(hipEventRecord(t0, st));
        rc = trace_closest_internal(s, rays, nullptr, bounces == 0 ? (uint32_t)PVX_RAY_IMAGE : (uint32_t)PVX_RAY_REFLECTION,
                                    n, static_cast<pvx_hit *>(b.hits.p), st,
                                    bounces == 0 ? nullptr : static_cast<const uint8_t *>(b.kinds[cur].p));
        if (rc) return rc;
        HIP_TRY(hipEventRecord(t1, st));
        n_closest += n;
        uint32_t *counters = static_cast<uint32_t *>(b.counters.p);
        if (packet) { // light groups of this bounce
            if (!packet_fits(n) && (r = drain_light())) return r;
            if ((r = ensure_packet(n))) return r;
        } else if ((r = ensure_shadow((size_t)n * spp))) {
            return r;
        }
        // the light pass of bounce - 2 used this parity's buffers
        if (overlap && bounces >= 2) HIP_TRY(hipStreamWaitEvent(st, ev(bounces - 2, 3), 0));
        const HitBufs hb = hit_bufs(par);
        uint32_t *gcount = static_cast<uint32_t *>(b.gcount.p) + par;
        k_zero_counters<<<1, 64, 0, st>>>(counters);
        k_zero_word<<<1, 64, 0, st>>>(gcount);
        const bool rev = s->view.has_complex >= 3; // lathes / sors in the scene
        (packet ? (rev ? k_shade<true, true> : k_shade<true, false>)
                : (rev ? k_shade<false, true> : k_shade<false, false>))<<<(n + 127) / 128, 128, 0, st>>>(
            p, rays, pst, static_cast<const pvx_hit *>(b.hits.p), n, hb.contrib, hb.shade, hb.lslot, hb.lterm,
            static_cast<pvx_ray *>(b.srays[0].p), static_cast<double *>(b.sdepth[0].p),
            static_cast<uint32_t *>(b.sowner[0].p), static_cast<float *>(b.scol.p),
            static_cast<pvx_ray *>(b.crays.p), static_cast<PathState *>(b.cstate.p),
            static_cast<uint32_t *>(b.nchild.p), counters, hb.groups, hb.pix, gcount);
        HIP_TRY(hipGetLastError());
        if (overlap) HIP_TRY(hipEventRecord(ev(bounces, 0), st));
        // The light rays of camera-ray hits come in pixel order and walk
        // alike in runs; those of later bounces less so, and the walk that
        // suits them may differ: each keeps its own choice.
        const int lcls = bounces == 0 ? 0 : 1;
        uint32_t lvariant = s->light_variant[lcls], lgrid = s->light_grid[lcls];
        if (!is_light_walk(lvariant)) lvariant = PVX_TRAVERSAL_FUSED;
        // outside the tuning bounces the light pass starts at once, with the
        // group count read on the device (no wait for the host's read-back)
        const uint64_t nr_bound = (uint64_t)n * nl * smax;
        const bool early = packet && !(tune && bounces < 2) && nr_bound <= 0xFFFFFFFFull;
        if (early) {
            HIP_TRY(hipStreamWaitEvent(ls, ev(bounces, 0), 0));
            HIP_TRY(hipEventRecord(ev(bounces, 1), ls));
            if ((r = light_pass(hb.groups, lvariant, lgrid, nr_bound, gcount))) return r;
            HIP_TRY(hipEventRecord(ev(bounces, 2), ls));
            timed_light.push_back((int)bounces);
        }
        // next queue size: exclusive scan of the child counts
        const uint32_t nb = (n + kScanBlock - 1) / kScanBlock;
        if (nb > (uint32_t)kScanBlock * 8) return fail(PVX_EUNSUPPORTED, "ray queue above 8M rays");
        k_scan_block<<<nb, kScanBlock, 0, st>>>(static_cast<const uint32_t *>(b.nchild.p), n,
                                                static_cast<uint32_t *>(b.offs.p), static_cast<uint32_t *>(b.sums.p));
        k_scan_sums<<<1, kScanBlock, 0, st>>>(static_cast<uint32_t *>(b.sums.p), nb, counters + 2);
        HIP_TRY(hipGetLastError());
        uint32_t cnt[kGroupCounter + 1];
        HIP_TRY(hipMemcpyAsync(cnt, counters, sizeof cnt, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(&cnt[kGroupCounter], gcount, 4, hipMemcpyDeviceToHost, st)); // light groups
        HIP_TRY(hipStreamSynchronize(st));
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, t0, t1));
        trace_ms += ms;
        if (packet) { // light rays made in the kernel and tested for occlusion, on the light stream
            const uint64_t nr = (uint64_t)cnt[kGroupCounter] * smax;
            if (nr > 0xFFFFFFFFull) return fail(PVX_EUNSUPPORTED, "more than 2^32 light rays in one bounce");
            if (!early) {
                HIP_TRY(hipStreamWaitEvent(ls, ev(bounces, 0), 0));
                auto timed_pass = [&](uint32_t variant, uint32_t grid, float &out_ms) -> int {
                    HIP_TRY(hipEventRecord(t0, ls));
                    int rr = light_pass(hb.groups, variant, grid, nr, nullptr);
                    if (rr) return rr;
                    HIP_TRY(hipEventRecord(t1, ls));
                    HIP_TRY(hipEventSynchronize(t1));
                    HIP_TRY(hipEventElapsedTime(&out_ms, t0, t1));
                    return PVX_OK;
                };
                if (tune && bounces < 2 && nr > 0) { // the walks and grids on this frame's light rays; the fastest is kept
                    struct Cand {
                        uint32_t variant, grid;
                        float ms;
                    } cands[7] = {{PVX_TRAVERSAL_FUSED, 2 * s->grid, 0.0f},
                                  {PVX_TRAVERSAL_FUSED, s->grid, 0.0f},
                                  {PVX_TRAVERSAL_REFILL, s->grid, 0.0f},
                                  {PVX_TRAVERSAL_REFILL_BINARY, s->grid, 0.0f},
                                  {PVX_TRAVERSAL_LIGHT_BINARY, 2 * s->grid, 0.0f},
                                  {PVX_TRAVERSAL_LIGHT_BINARY, s->grid, 0.0f},
                                  {PVX_TRAVERSAL_PACKET, s->grid, 0.0f}};
                    // warm-up: the first launch is slower
                    if ((rc = light_pass(hb.groups, PVX_TRAVERSAL_FUSED, 0, nr, nullptr))) return rc;
                    for (Cand &c : cands)
                        if ((rc = timed_pass(c.variant, c.grid, c.ms))) return rc;
                    const Cand *best = &cands[0];
                    for (const Cand &c : cands)
                        if (c.ms < best->ms) best = &c;
                    lvariant = best->variant;
                    lgrid = best->grid;
                    s->light_variant[lcls] = lvariant;
                    s->light_grid[lcls] = lgrid;
                    if (bounces == 0) {
                        s->tune_variant[1] = lvariant;
                        s->tune_grid[1] = lgrid;
                        tune[1] = pvx_tune_result{lvariant, lgrid, best->ms, cands[0].ms, 0, 0};
                    }
                }
                // the pass itself, timed by events read at the end of the frame
                HIP_TRY(hipEventRecord(ev(bounces, 1), ls));
                if ((rc = light_pass(hb.groups, lvariant, lgrid, nr, nullptr))) return rc;
                HIP_TRY(hipEventRecord(ev(bounces, 2), ls));
                timed_light.push_back((int)bounces);
            }
            n_shadow += cnt[1]; // the light rays made (k_shade counts them as it did when writing them)
        }
        // light rays, followed through transparent objects
        uint32_t ns = packet ? 0u : cnt[1];
        int sq = 0;
        if (tune && bounces == 0 && !packet) {
            if (ns > 0) {
                if ((rc = autotune_internal(s, static_cast<const pvx_ray *>(b.srays[0].p),
                                            static_cast<const double *>(b.sdepth[0].p), ns, light_kind, &tune[1],
                                            st)))
                    return rc;
            } else {
                tune[1] = pvx_tune_result{0, 0, 0.0, 0.0, 0, 0}; // no light rays: nothing to tune
            }
        }
        for (int step = 0; ns > 0 && step < kMaxShadowSteps; step++) {
            HIP_TRY(hipEventRecord(t0, st));
            rc = trace_closest_internal(s, static_cast<const pvx_ray *>(b.srays[sq].p),
                                        static_cast<const double *>(b.sdepth[sq].p), light_kind, ns,
                                        static_cast<pvx_hit *>(b.shits.p), st);
            if (rc) return rc;
            HIP_TRY(hipEventRecord(t1, st));
            n_shadow += ns;
            k_zero_counters<<<1, 64, 0, st>>>(counters);
            k_shadow_step<<<(ns + 255) / 256, 256, 0, st>>>(
                p, static_cast<const pvx_ray *>(b.srays[sq].p), static_cast<const double *>(b.sdepth[sq].p),
                static_cast<const uint32_t *>(b.sowner[sq].p), static_cast<const pvx_hit *>(b.shits.p), ns,
                static_cast<float *>(b.scol.p), static_cast<pvx_ray *>(b.srays[1 - sq].p),
                static_cast<double *>(b.sdepth[1 - sq].p), static_cast<uint32_t *>(b.sowner[1 - sq].p), counters);
            HIP_TRY(hipGetLastError());
            uint32_t c2;
            HIP_TRY(hipMemcpyAsync(&c2, counters, 4, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipStreamSynchronize(st));
            HIP_TRY(hipEventElapsedTime(&ms, t0, t1));
            shadow_ms += ms;
            ns = c2;
            sq = 1 - sq;
        }
        (packet ? k_finish<true> : k_finish<false>)<<<(n + 127) / 128, 128, 0, ls>>>(
            p, hb.shade, n, hb.lslot, hb.lterm, static_cast<const float *>(b.scol.p), hb.contrib,
            static_cast<const uint8_t *>(b.vis.p), smax);
        k_accumulate<<<(n + 255) / 256, 256, 0, ls>>>(hb.pix, hb.contrib, n, rgb);
        HIP_TRY(hipGetLastError());
        if (overlap) HIP_TRY(hipEventRecord(ev(bounces, 3), ls));
        const uint32_t next = cnt[2];
        if (next > 0) {
            if (next > b.rays[1 - cur].cap / sizeof(pvx_ray) || next > b.nchild.cap / 4) {
                // grow the queue (keeping this bounce's children, which live in crays/cstate)
                HIP_TRY(hipStreamSynchronize(st));
                if ((r = drain_light())) return r;
                DevBuf keep_r, keep_s;
                std::swap(keep_r.p, b.crays.p);
                std::swap(keep_r.cap, b.crays.cap);
                std::swap(keep_s.p, b.cstate.p);
                std::swap(keep_s.cap, b.cstate.cap);
                DevBuf keep_n, keep_o, keep_u;
                std::swap(keep_n.p, b.nchild.p);
                std::swap(keep_n.cap, b.nchild.cap);
                std::swap(keep_o.p, b.offs.p);
                std::swap(keep_o.cap, b.offs.cap);
                std::swap(keep_u.p, b.sums.p);
                std::swap(keep_u.cap, b.sums.cap);
                // the current queue must survive too
                DevBuf cur_r, cur_s, cur_k;
                std::swap(cur_r.p, b.rays[cur].p);
                std::swap(cur_r.cap, b.rays[cur].cap);
                std::swap(cur_s.p, b.state[cur].p);
                std::swap(cur_s.cap, b.state[cur].cap);
                std::swap(cur_k.p, b.kinds[cur].p);
                std::swap(cur_k.cap, b.kinds[cur].cap);
                if ((r = ensure_queue((size_t)next + next / 2)) ||
                    (r = packet ? ensure_packet((size_t)next + next / 2) : ensure_shadow(((size_t)next + next / 2) * spp)))
                    return r;
                k_emit<<<(n + 255) / 256, 256, 0, st>>>(
                    static_cast<const uint32_t *>(keep_n.p), static_cast<const uint32_t *>(keep_o.p),
                    static_cast<const uint32_t *>(keep_u.p), n, static_cast<const pvx_ray *>(keep_r.p),
                    static_cast<const PathState *>(keep_s.p), static_cast<pvx_ray *>(b.rays[1 - cur].p),
                    static_cast<PathState *>(b.state[1 - cur].p), static_cast<uint8_t *>(b.kinds[1 - cur].p));
                HIP_TRY(hipGetLastError());
                HIP_TRY(hipStreamSynchronize(st));
            } else {
                k_emit<<<(n + 255) / 256, 256, 0, st>>>(
                    static_cast<const uint32_t *>(b.nchild.p), static_cast<const uint32_t *>(b.offs.p),
                    static_cast<const uint32_t *>(b.sums.p), n, static_cast<const pvx_ray *>(b.crays.p),
                    static_cast<const PathState *>(b.cstate.p), static_cast<pvx_ray *>(b.rays[1 - cur].p),
                    static_cast<PathState *>(b.state[1 - cur].p), static_cast<uint8_t *>(b.kinds[1 - cur].p));
                HIP_TRY(hipGetLastError());
            }
        }
        n = next;
        cur = 1 - cur;
        bounces++;
    }
    // the light stream's last accumulation, then the caller's stream goes on
    if (overlap && bounces > 0) HIP_TRY(hipStreamWaitEvent(st, ev(bounces - 1, 3), 0));
    if (S > 1) // the pixel is the mean of its samples (tracetask.cpp:889)
        k_scale3<<<(unsigned)((3 * npix + 255) / 256), 256, 0, st>>>(rgb, 3 * npix, 1.0f / (float)S);
    HIP_TRY(hipEventRecord(e1, st));
    HIP_TRY(hipEventSynchronize(e1));
    float total = 0.0f;
    HIP_TRY(hipEventElapsedTime(&total, e0, e1));
    for (int k : timed_light) {
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, ev((uint32_t)k, 1), ev((uint32_t)k, 2)));
        shadow_ms += ms;
    }
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

