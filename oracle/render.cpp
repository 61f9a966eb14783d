// FP64 CPU ORACLE — TEST INFRASTRUCTURE ONLY (see oracle.hpp).
// Restatement of the plain-finish shading path:
// Trace::TraceRay (core/render/trace.cpp:135-228), ComputeTextureColour
// (457-586), ComputeLightedTexture (739-1179, single layer, no radiosity,
// no media, no photons), ComputeOneDiffuseLight (1637-1729),
// ComputeOneWhiteLightRay (2710-2767), TracePointLightShadowRay /
// TraceAreaLightSubsetShadowRay (1951-2272, without the order-dependent
// shadow cache), ComputeShadowColour / ComputeShadowTexture (2274-2420,
// 1181-1262: filter * pigment + transmit, no caustics or fade),
// ComputeDiffuseColour (2441-2484), ComputePhongColour (2520-2558),
// ComputeReflectivity (2627-2656), ComputeReflection (1264-1321),
// ComputeRefraction / TraceRefractionRay (1323-1483: interior list,
// Heckbert refraction, total internal reflection; no dispersion).
// Colours are FP32 (MathColour), geometry FP64.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <thread>

#include "oracle.hpp"

namespace orc {

bool scene_find_intersection(const Scene &s, const Ray &r, double max_depth, uint32_t kind,
                             Isect &best, int32_t &best_obj, int method);

namespace {

const double kShadowTolerance = 1.0e-3; // trace.cpp:81

struct RGB {
    float r = 0, g = 0, b = 0;
};
inline RGB operator+(RGB a, RGB b) { return {a.r + b.r, a.g + b.g, a.b + b.b}; }
inline RGB operator*(RGB a, RGB b) { return {a.r * b.r, a.g * b.g, a.b * b.b}; }
inline RGB scale(RGB a, double s) {
    return {(float)(s * a.r), (float)(s * a.g), (float)(s * a.b)};
}
inline RGB rgb3(const float *p) { return {p[0], p[1], p[2]}; }
inline bool near_zero(RGB c) {
    return std::fabs(c.r) < kEpsilon && std::fabs(c.g) < kEpsilon && std::fabs(c.b) < kEpsilon;
}
inline float max3f(const float *p) { return std::max(p[0], std::max(p[1], p[2])); }

struct Shading {
    int method = 1;
    std::vector<pvx_material> mats;
    std::vector<pvx_light> lights;
    float amb[3], bg[3];
    uint32_t max_level;
    float bailout;
};

struct Counters {
    std::atomic<uint64_t> closest{0}, shadow{0};
};

V3 xf_normal_out(const Transform *tr, V3 n) { // MTransNormal + normalize
    if (!tr) return n;
    return normalized(minv_trans_normal(tr->inv, n));
}

// Object::Normal for a hit (raw, not flipped).
V3 shape_normal(const Scene &s, uint32_t type, uint32_t index, const Isect &h, V3 ip) {
    switch (type) {
    case PVX_SPHERE: // Sphere::Normal (sphere.cpp), spherical mode
        return (ip - s.sph_c[index]) / s.sph_r[index];
    case PVX_PLANE:
        return s.pl_n[index];
    case PVX_MESH:
        return mesh_normal_world(s.meshes[index], s.tr(s.mesh_tr[index]), h.prim, ip);
    case PVX_TRIANGLE:
    case PVX_SMOOTH_TRIANGLE:
        return tri_normal(s.tris[index], ip);
    case PVX_BOX: { // Box::Normal (box.cpp:600-625)
        static const double tbl[7][3] = {{0, 0, 0}, {-1, 0, 0}, {1, 0, 0}, {0, -1, 0},
                                         {0, 1, 0},  {0, 0, -1}, {0, 0, 1}};
        int side = h.side >= 1 && h.side <= 6 ? h.side : 0;
        V3 n(tbl[side][0], tbl[side][1], tbl[side][2]);
        return xf_normal_out(s.tr(s.box_tr[index]), n);
    }
    case PVX_TORUS: { // Torus::Normal (torus.cpp:458-490)
        const Transform *tr = s.tr(s.tor_tr[index]);
        static const Transform id = {{{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}},
                                     {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}}};
        if (!tr) tr = &id;
        V3 P = mtrans_point(tr->inv, ip);
        double dist = std::sqrt(P.x * P.x + P.z * P.z);
        V3 M(0, 0, 0);
        if (dist > kEpsilon) M = V3(s.tor_R[index] * P.x / dist, 0.0, s.tor_R[index] * P.z / dist);
        return normalized(minv_trans_normal(tr->inv, P - M));
    }
    case PVX_POLYGON: // Polygon::Normal (polygon.cpp:308): S_Normal
        return s.polygons[index].normal;
    case PVX_QUADRIC: // Quadric::Normal (quadric.cpp)
        return quadric_normal(s.quadrics[index], ip);
    case PVX_DISC: // Disc::Normal (disc.cpp): the stored normal
        return s.discs[index].normal;
    case PVX_CONE: // Cone::Normal (cone.cpp:440-480)
        return cone_normal(s.cones[index], h.side, ip);
    case PVX_LATHE:
    case PVX_SOR:
        return revolution_normal(s.revs[index], h, ip);
    case PVX_CSG: { // the leaf that produced the hit
        const pvx_object &c = s.csg_children[h.leaf];
        Isect leaf = h;
        return shape_normal(s, c.type, c.index, leaf, ip);
    }
    }
    return V3(0, 1, 0);
}

// TransColour::TransmittedColour (base/colour.h:1912-1915): filter * pigment + transmit.
inline RGB transmitted(const pvx_material &m) {
    return {m.pigment[0] * m.filter + m.transmit, m.pigment[1] * m.filter + m.transmit,
            m.pigment[2] * m.filter + m.transmit};
}
inline bool opaque(const pvx_material &m) { return m.filter == 0.0f && m.transmit == 0.0f; }
inline double interior_ior(const pvx_material &m) { return m.ior > 0.0f ? (double)m.ior : 1.0; }

// Light reaching ip from target: TracePointLightShadowRay's loop, which
// filters the light through every transparent object in between and stops
// at the first opaque one.
RGB light_through(const Scene &s, const Shading &sh, Counters &cnt, V3 ip, V3 target, RGB lc) {
    Ray lr;
    lr.o = ip;
    V3 dir = target - ip;
    double lsd = length(dir);
    lr.d = dir / lsd;
    for (int guard = 0; guard < 64; guard++) {
        Isect h;
        int32_t obj;
        cnt.shadow++;
        bool found = scene_find_intersection(s, lr, lsd, PVX_RAY_SHADOW, h, obj, sh.method);
        if (!(found && h.depth < lsd - kShadowTolerance && h.depth > kShadowTolerance)) return lc;
        const pvx_material &m = sh.mats[obj];
        if (opaque(m)) return RGB{};
        lc = lc * transmitted(m);
        lsd -= h.depth;
        lr.o = lr.eval(h.depth);
    }
    return lc;
}

// Interiors a ray is inside (object ids), as Ray::GetInteriors.
using Interiors = std::vector<int32_t>;

// Ray type bits as Ray::SetFlags(type, parent) sets them (ray.cpp:155-166).
struct RayFlags {
    bool primary = false, refl = false, refr = false;
};
inline RayFlags reflected(RayFlags p) { return RayFlags{false, true, p.refr}; }
inline RayFlags refracted(RayFlags p) { return RayFlags{false, p.refl, true}; }
// The ray kind NoSomethingFlagRayObjectCondition tests (trace.cpp:84-95,
// Ray::IsImageRay ray.h:100).
inline uint32_t kind_of(RayFlags f) {
    uint32_t k = 0;
    if (f.primary || (f.refr && !f.refl)) k |= PVX_RAY_IMAGE;
    if (f.refl) k |= PVX_RAY_REFLECTION;
    return k;
}

RGB trace_ray(const Scene &s, const Shading &sh, Counters &cnt, const Ray &ray, float weight,
              uint32_t level, RayFlags rf, const Interiors &ins, bool continued) {
    const uint32_t kind = kind_of(rf);
    if (level >= sh.max_level || weight < sh.bailout) return RGB{};
    Isect best;
    int32_t obj;
    cnt.closest++;
    if (!scene_find_intersection(s, ray, 0.0, kind, best, obj, sh.method)) return rgb3(sh.bg);
    if (!continued) level++;
    const pvx_material &m = sh.mats[obj];
    V3 ip = ray.eval(best.depth);
    const Object &o = s.objects[obj];
    V3 n = shape_normal(s, o.type, o.index, best, ip);
    if (dot(n, ray.d) > 0.0) n = n * -1.0;
    RGB pig = rgb3(m.pigment);
    double att = 1.0 - m.filter - m.transmit;
    // emission + classic ambient (Ambient * ambient_light), times opacity
    RGB emission = rgb3(m.emission) + rgb3(m.ambient) * rgb3(sh.amb);
    RGB col = scale(pig * emission, att);
    for (const pvx_light &L : sh.lights) {
        V3 C(L.position[0], L.position[1], L.position[2]);
        V3 dir = C - ip;
        double lsd = length(dir);
        dir = dir / lsd;
        RGB lc = rgb3(L.colour); // Attenuate_Light = 1 without fade
        if (dot(n, dir) < kEpsilon) continue;
        if (L.kind == PVX_LIGHT_AREA && (L.size1 > 1 || L.size2 > 1)) {
            V3 a1(L.axis1[0], L.axis1[1], L.axis1[2]), a2(L.axis2[0], L.axis2[1], L.axis2[2]);
            // default adaptive 100: every grid point is sampled; the
            // quadtree average equals the mean over the grid for 2^k+1 and
            // 4x4 grids alike, computed here as nested corner averages.
            int n1 = (int)L.size1, n2 = (int)L.size2;
            std::vector<RGB> vis((size_t)n1 * n2);
            for (int u = 0; u < n1; u++)
                for (int v = 0; v < n2; v++) {
                    V3 j1 = n1 > 1 ? a1 * ((double)u / (double)(n1 - 1) - 0.5) : V3(0, 0, 0);
                    V3 j2 = n2 > 1 ? a2 * ((double)v / (double)(n2 - 1) - 0.5) : V3(0, 0, 0);
                    vis[(size_t)u * n2 + v] = light_through(s, sh, cnt, ip, C + (j1 + j2), lc);
                }
            // TraceAreaLightSubsetShadowRay recursion (trace.cpp:2131-2272)
            struct R {
                static RGB sub(const std::vector<RGB> &g, int n2, int u1, int v1, int u2, int v2) {
                    RGB c[4] = {g[(size_t)u1 * n2 + v1], g[(size_t)u2 * n2 + v1], g[(size_t)u1 * n2 + v2],
                                g[(size_t)u2 * n2 + v2]};
                    if ((u2 - u1 > 1) || (v2 - v1 > 1)) {
                        int mu0 = (int)std::floor((u1 + u2) / 2.0), mu1 = (int)std::ceil((u1 + u2) / 2.0);
                        int mv0 = (int)std::floor((v1 + v2) / 2.0), mv1 = (int)std::ceil((v1 + v2) / 2.0);
                        c[0] = sub(g, n2, u1, v1, mu0, mv0);
                        c[1] = sub(g, n2, mu1, v1, u2, mv0);
                        c[2] = sub(g, n2, u1, mv1, mu0, v2);
                        c[3] = sub(g, n2, mu1, mv1, u2, v2);
                    }
                    RGB a = c[0] + c[1] + c[2] + c[3];
                    return RGB{a.r * 0.25f, a.g * 0.25f, a.b * 0.25f};
                }
            };
            lc = R::sub(vis, n2, 0, 0, n1 - 1, n2 - 1);
        } else {
            lc = light_through(s, sh, cnt, ip, C, lc);
        }
        if (near_zero(lc)) continue;
        double cosi = dot(n, dir);
        double intensity = (m.brilliance != 1.0f) ? std::pow(std::fabs(cosi), (double)m.brilliance)
                                                  : std::fabs(cosi);
        intensity *= (double)m.diffuse * att;
        if (m.diffuse > 0.0f) col = col + scale(pig * lc, intensity);
        if (m.phong > 0.0f && kind != PVX_RAY_RADIOSITY) {
            double c2 = -2.0 * dot(ray.d, n);
            V3 rd = ray.d + n * c2;
            double ca = dot(rd, dir);
            if (ca > 0.0 && (m.phong_size < 60 || ca > 0.0008))
                col = col + scale(lc, (double)m.phong * std::pow(ca, (double)m.phong_size));
        }
    }
    // transmitted component (ComputeLightedTexture, trace.cpp:1075-1131)
    const RGB fil = transmitted(m);
    const double trans = std::min(1.0, (double)std::fabs(0.297f * fil.r + 0.589f * fil.g + 0.114f * fil.b));
    bool tir = false;
    if (trans > sh.bailout) {
        const float nw = (float)((double)weight *
                                 std::max(std::fabs(fil.r), std::max(std::fabs(fil.g), std::fabs(fil.b))));
        // ComputeRefraction: ior ratio from the interior list
        Interiors ni = ins;
        const double own = interior_ior(m);
        double ior;
        if (ni.empty()) {
            ni.push_back(obj);
            ior = 1.0 / own;
        } else if (ni.back() == obj) {
            ni.pop_back();
            ior = ni.empty() ? own : own / interior_ior(sh.mats[ni.back()]);
        } else {
            auto it = std::find(ni.begin(), ni.end(), obj);
            if (it != ni.end()) {
                ni.erase(it);
                ior = 1.0;
            } else {
                ior = interior_ior(sh.mats[ni.back()]) / own;
                ni.push_back(obj);
            }
        }
        RGB rfr;
        if (std::fabs(ior - 1.0) < kEpsilon) {
            rfr = trace_ray(s, sh, cnt, Ray{ip, ray.d}, nw, level, refracted(rf), ni, true);
        } else {
            double nn = dot(ray.d, n);
            V3 ln = n;
            if (nn <= 0.0)
                nn = -nn;
            else
                ln = n * -1.0;
            double t = 1.0 + ior * ior * (nn * nn - 1.0);
            if (t < 0.0) { // total internal reflection: a reflected ray, unfiltered
                tir = true;
                double k = -2.0 * dot(ray.d, n);
                Ray nr{ip, normalized(ray.d + n * k)};
                rfr = trace_ray(s, sh, cnt, nr, nw, level, reflected(rf), ins, false);
            } else {
                t = ior * nn - std::sqrt(t);
                Ray nr{ip, ray.d * ior + ln * t};
                rfr = trace_ray(s, sh, cnt, nr, nw, level, refracted(rf), ni, false);
            }
        }
        col = col + (tir ? rfr : rfr * fil);
    }
    // reflection: ComputeReflectivity (no fresnel) + ComputeReflection;
    // skipped after total internal reflection (trace.cpp:1141-1147)
    double cos_inc = -dot(ray.d, n);
    double w = (double)weight * std::max(max3f(m.reflection_max), max3f(m.reflection_min));
    double frac = std::fabs((double)m.reflection_falloff - 1.0) > kEpsilon
                      ? std::pow(1.0 - cos_inc, (double)m.reflection_falloff)
                      : 1.0 - cos_inc;
    RGB refl;
    if (std::fabs(frac) < kEpsilon)
        refl = rgb3(m.reflection_min);
    else if (std::fabs(frac - 1.0) < kEpsilon)
        refl = rgb3(m.reflection_max);
    else
        refl = scale(rgb3(m.reflection_max), frac) + scale(rgb3(m.reflection_min), 1.0 - frac);
    if (!tir && !(refl.r == 0.0f && refl.g == 0.0f && refl.b == 0.0f)) {
        Ray nr;
        double k = -2.0 * dot(ray.d, n);
        nr.d = normalized(ray.d + n * k);
        nr.o = ip;
        RGB rc = trace_ray(s, sh, cnt, nr, (float)w, level, reflected(rf), ins, false);
        col = col + refl * rc;
    }
    return col;
}

} // namespace

V3 oracle_shape_normal(const Scene &s, uint32_t obj, const Isect &h, V3 ip) {
    const Object &o = s.objects[obj];
    return shape_normal(s, o.type, o.index, h, ip);
}

} // namespace orc

using namespace orc;

// n_samples rays per pixel at `offsets` (from the pixel corner); the pixel is
// their sum divided by the count (SupersampleOnePixel, tracetask.cpp:861-890).
extern "C" int orc_render(void *sp, const pvx_shading_desc *sd, const pvx_camera *cam, float *rgb,
                          int threads, uint64_t *counters, int method, int n_samples, const double *offsets) {
    const Scene &s = *static_cast<Scene *>(sp);
    Shading sh;
    sh.method = method;
    sh.mats.assign(sd->materials, sd->materials + sd->n_materials);
    sh.lights.assign(sd->lights, sd->lights + sd->n_lights);
    std::memcpy(sh.amb, sd->ambient_light, sizeof sh.amb);
    std::memcpy(sh.bg, sd->background, sizeof sh.bg);
    sh.max_level = sd->max_trace_level;
    sh.bailout = sd->adc_bailout;
    Counters cnt;
    size_t npix = (size_t)cam->width * cam->height;
    std::atomic<size_t> next(0);
    auto work = [&]() {
        for (;;) {
            size_t b = next.fetch_add(64);
            if (b >= npix) break;
            for (size_t i = b; i < std::min(npix, b + 64); i++) {
                uint32_t x = (uint32_t)(i % cam->width), y = (uint32_t)(i / cam->width);
                RGB c;
                const int S = n_samples > 0 ? n_samples : 1;
                for (int smp = 0; smp < S; smp++) {
                    double ox = n_samples > 0 ? offsets[2 * smp] : cam->jitter_x;
                    double oy = n_samples > 0 ? offsets[2 * smp + 1] : cam->jitter_y;
                    double x0 = ((double)x + ox) / (double)cam->width - 0.5;
                    double y0 = 0.5 - ((double)y + oy) / (double)cam->height;
                    V3 d;
                    for (int k = 0; k < 3; k++)
                        d[k] = cam->direction[k] + x0 * cam->right[k] + y0 * cam->up[k];
                    Ray r{V3(cam->location[0], cam->location[1], cam->location[2]), normalized(d)};
                    c = c + trace_ray(s, sh, cnt, r, 1.0f, 0, RayFlags{true, false, false}, Interiors{}, false);
                }
                if (S > 1) c = RGB{c.r / (float)S, c.g / (float)S, c.b / (float)S};
                rgb[3 * i] = c.r;
                rgb[3 * i + 1] = c.g;
                rgb[3 * i + 2] = c.b;
            }
        }
    };
    std::vector<std::thread> pool;
    for (int t = 0; t < std::max(1, threads); t++) pool.emplace_back(work);
    for (auto &t : pool) t.join();
    if (counters) {
        counters[0] = cnt.closest;
        counters[1] = cnt.shadow;
    }
    return 0;
}

// TraceRay for caller-given camera rays (n x 6: origin, unit direction),
// with the primary-ray flags of orc_render; rgb receives n x 3.
extern "C" int orc_trace_rays(void *sp, const pvx_shading_desc *sd, const double *rays, size_t n, float *rgb,
                              int method) {
    const Scene &s = *static_cast<Scene *>(sp);
    Shading sh;
    sh.method = method;
    sh.mats.assign(sd->materials, sd->materials + sd->n_materials);
    sh.lights.assign(sd->lights, sd->lights + sd->n_lights);
    std::memcpy(sh.amb, sd->ambient_light, sizeof sh.amb);
    std::memcpy(sh.bg, sd->background, sizeof sh.bg);
    sh.max_level = sd->max_trace_level;
    sh.bailout = sd->adc_bailout;
    Counters cnt;
    for (size_t i = 0; i < n; i++) {
        const double *q = rays + 6 * i;
        Ray r{V3(q[0], q[1], q[2]), V3(q[3], q[4], q[5])};
        const RGB c = trace_ray(s, sh, cnt, r, 1.0f, 0, RayFlags{true, false, false}, Interiors{}, false);
        rgb[3 * i] = c.r;
        rgb[3 * i + 1] = c.g;
        rgb[3 * i + 2] = c.b;
    }
    return 0;
}

// Closest hit per ray (method 1) and the raw Object::Normal there (zeros on
// a miss): 3 doubles per ray.
extern "C" int orc_hit_normals(void *sp, const double *rays, size_t n, double *out) {
    const Scene &s = *static_cast<Scene *>(sp);
    for (size_t i = 0; i < n; i++) {
        Ray r;
        r.o = V3(rays[6 * i], rays[6 * i + 1], rays[6 * i + 2]);
        r.d = V3(rays[6 * i + 3], rays[6 * i + 4], rays[6 * i + 5]);
        Isect h;
        int32_t obj;
        V3 nrm(0, 0, 0);
        if (scene_find_intersection(s, r, 0.0, PVX_RAY_IMAGE, h, obj, 1)) {
            const Object &o = s.objects[obj];
            nrm = shape_normal(s, o.type, o.index, h, r.eval(h.depth));
        }
        out[3 * i] = nrm.x, out[3 * i + 1] = nrm.y, out[3 * i + 2] = nrm.z;
    }
    return 0;
}
