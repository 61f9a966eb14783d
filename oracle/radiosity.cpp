// FP64 CPU ORACLE — TEST INFRASTRUCTURE ONLY (see oracle.hpp).
// Sequential restatement of POV-Ray's radiosity for plain diffuse surfaces:
// RadiosityFunction::ComputeAmbient (core/lighting/radiosity.cpp:385-461),
// GatherLight (509-743), SampleDirectionGenerator (758-868) over the
// Halton(2,3) cosine-weighted direction pool (core/math/randomsequence.cpp:
// 730-747, 907-914, 1019-1038), RadiosityCache::FindReusableBlock /
// AverageNearBlock (1458-1769, PSEUDO_SIGMOID_METHOD, no RAD_GRADIENT) and
// the per-depth settings of GetRecursionSettings (177-218).  The cache is
// kept in a uniform grid instead of the octree (same set of records is
// visited: every record whose quick-check radius covers the query point).
// Pretrace runs single-threaded over a pixel grid whose step halves from
// pretrace_start to pretrace_end; the final pass visits every pixel.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <unordered_map>

#include "oracle.hpp"

namespace orc {

bool scene_find_intersection(const Scene &s, const Ray &r, double max_depth, uint32_t kind,
                             Isect &best, int32_t &best_obj, int method);
V3 oracle_shape_normal(const Scene &s, uint32_t obj, const Isect &h, V3 ip);

namespace {

const double kInFrontLimit = -0.05;  // radiosity.cpp:100
const double kAvgNearEps = 0.000001; // radiosity.cpp:109
const double kRadEps = 0.001;        // radiosity.cpp:110
const double kWeightErrOffset = 0.25;// radiosity.cpp:111
const int kPoolSize = 1600;          // RADIOSITY_MAX_SAMPLE_DIRECTIONS

struct Col {
    double r = 0, g = 0, b = 0;
};
inline Col operator+(Col a, Col b) { return {a.r + b.r, a.g + b.g, a.b + b.b}; }
inline Col operator*(Col a, double s) { return {a.r * s, a.g * s, a.b * s}; }
inline Col mulc(Col a, Col b) { return {a.r * b.r, a.g * b.g, a.b * b.b}; }
inline double grey(Col c) { return 0.297 * c.r + 0.589 * c.g + 0.114 * c.b; }

struct Record {
    V3 p, n;
    Col E;
    double r;        // Harmonic_Mean_Distance (clamped)
    double nearest;  // Nearest_Distance
    V3 to_nearest;   // To_Nearest_Surface
    int depth;
};

// Halton generator with the reference's index offset (index + 1).
double halton(size_t index, unsigned base) {
    size_t i = 1 + index;
    double h = 0, q = 1.0 / base;
    while (i > 0) {
        h = h + (unsigned)(i % base) * q;
        i /= base;
        q /= base;
    }
    return h;
}

struct Rad {
    const Scene &s;
    const std::vector<pvx_material> &mats;
    pvx_radiosity_settings st;
    V3 cam;
    std::vector<V3> pool;
    size_t cursor[8] = {0};
    std::vector<Record> recs[8];
    double cell = 0.25; // lookup grid cell (an index only: results do not depend on it)
    std::unordered_map<uint64_t, std::vector<int>> grid[8];
    uint64_t gathers[8] = {0}, rays = 0;

    Rad(const Scene &sc, const std::vector<pvx_material> &m, const pvx_radiosity_settings &set, V3 c)
        : s(sc), mats(m), st(set), cam(c) {
        // reuse radii grow with the distance to the eye: scale the cells with it
        cell = 0.25 * std::max(1.0, length(c) / 5.0);
        pool.resize(kPoolSize);
        for (int i = 0; i < kPoolSize; i++) { // HaltonCosWeightedDirectionGenerator::operator[]
            double r = std::sqrt(halton(i, 2));
            double theta = halton(i, 3) * (2 * M_PI);
            double x = r * std::cos(theta), z = r * std::sin(theta);
            pool[i] = V3(x, std::sqrt(1 - x * x - z * z), z);
        }
    }

    int rays_per_sample(int d) const { return std::max(5, (int)(st.count * std::pow(0.5, (double)d))); }
    int reuse_count(int d) const {
        return d == 0 ? (int)st.nearest_count : (d == 1 ? std::min(2, (int)st.nearest_count) : 1);
    }
    double bound_factor(int d) const { return std::pow(2.0, (double)d); }

    static uint64_t key(int x, int y, int z) {
        return ((uint64_t)(uint32_t)(x + (1 << 20)) << 42) | ((uint64_t)(uint32_t)(y + (1 << 20)) << 21) |
               (uint64_t)(uint32_t)(z + (1 << 20));
    }

    void add(const Record &rc) {
        int d = rc.depth;
        int id = (int)recs[d].size();
        recs[d].push_back(rc);
        double R = rc.r * st.error_bound * bound_factor(d) * 1.5 + 1e-9; // >= quick-check radius
        int x0 = (int)std::floor((rc.p.x - R) / cell), x1 = (int)std::floor((rc.p.x + R) / cell);
        int y0 = (int)std::floor((rc.p.y - R) / cell), y1 = (int)std::floor((rc.p.y + R) / cell);
        int z0 = (int)std::floor((rc.p.z - R) / cell), z1 = (int)std::floor((rc.p.z + R) / cell);
        for (int x = x0; x <= x1; x++)
            for (int y = y0; y <= y1; y++)
                for (int z = z0; z <= z1; z++) grid[d][key(x, y, z)].push_back(id);
    }

    // FindReusableBlock + AverageNearBlock
    double lookup(V3 p, V3 n, double bound, int depth, bool final, Col &E) {
        auto it = grid[depth].find(key((int)std::floor(p.x / cell), (int)std::floor(p.y / cell),
                                       (int)std::floor(p.z / cell)));
        if (it == grid[depth].end()) return 0.0;
        double W = 0;
        Col WE;
        for (int id : it->second) {
            const Record &b = recs[depth][id];
            V3 delta = p - b.p;
            double sq = length_sqr(delta);
            double quick = b.r * bound;
            if (!(sq < quick * quick)) continue;
            double dist = std::sqrt(sq);
            double ri = b.r;
            bool far = dist > kAvgNearEps;
            V3 du;
            if (far) {
                du = delta / dist;
                double c = dot(b.to_nearest, du);
                if (c > 0.0) ri = c * b.nearest + (1.0 - c) * ri;
            }
            if (!(dist < ri * bound)) continue;
            double dd = dot(n, b.n);
            double err = dist / ri + 2.0 * std::sqrt(std::fabs(1.0 - dd));
            if (!(err < bound)) continue;
            double in_front = 1.0;
            if (far) {
                V3 half = n + b.n;
                in_front = dot(du, half) / length(half);
            }
            if (!(in_front > kInFrontLimit)) continue;
            double w;
            if (!final || b.depth > 0) {
                w = 1.0 - err / bound;
            } else { // PSEUDO_SIGMOID_METHOD
                w = err / bound;
                w = w < 0.5 ? 1.0 - (w * 2.0) * (w * 2.0) / 2.0 : ((1.0 - w) * 2.0) * ((1.0 - w) * 2.0) / 2.0;
            }
            if (in_front <= 0) w = w * (1 - in_front / kInFrontLimit);
            if (w > kRadEps) {
                WE = WE + b.E * w;
                W += w;
            }
        }
        if (W > 0) E = WE * (1.0 / W);
        return W;
    }

    // Surface colour seen by a radiosity ray (ComputeLightedTexture with
    // radiosity ambient at recursion depth `depth`).
    Col shade(const Ray &ray, const Isect &h, int obj, int depth, double weight, bool final) {
        const pvx_material &m = mats[obj];
        V3 ip = ray.eval(h.depth);
        V3 nrm = oracle_shape_normal(s, (uint32_t)obj, h, ip);
        if (dot(nrm, ray.d) > 0.0) nrm = nrm * -1.0;
        Col pig{m.pigment[0], m.pigment[1], m.pigment[2]};
        double att = 1.0 - m.filter - m.transmit;
        Col col = mulc(pig, Col{m.emission[0], m.emission[1], m.emission[2]}) * att;
        if (depth < (int)st.recursion_limit && m.diffuse > 0.0f) {
            double maxc = grey(pig) * att * m.diffuse;
            if (maxc > st.adc_bailout) {
                Col E = ambient(ip, nrm, weight * maxc, depth, final);
                col = col + mulc(pig, E) * (att * m.diffuse);
            }
        }
        return col;
    }

    Col ambient(V3 p, V3 n, double weight, int depth, bool final) { // ComputeAmbient
        double bound = st.error_bound;
        if (weight < kWeightErrOffset) bound += kWeightErrOffset - weight;
        Col E;
        double reuse = lookup(p, n, bound * bound_factor(depth), depth, final, E);
        if (!(reuse * 4 >= reuse_count(depth) || (final && reuse > 0))) {
            Col g;
            double q = gather(p, n, depth, final, g);
            E = reuse > 0 ? (E * reuse + g * q) * (1.0 / (reuse + q)) : g;
        }
        return E * st.brightness;
    }

    double gather(V3 p, V3 n, int depth, bool final, Col &out) { // GatherLight
        int N = rays_per_sample(depth);
        double to_eye = length(cam - p);
        double reuse_min = to_eye * (double)st.minimum_reuse * bound_factor(depth);
        double reuse_max = to_eye * (double)st.maximum_reuse * bound_factor(depth);
        double wnext = std::max((double)st.adc_bailout + kEpsilon, std::pow(0.5 * (double)st.brightness, depth + 1.0));
        V3 fy = n, off = std::fabs(fy.z) > 0.9 ? V3(0, 1, 0) : V3(0, 0, 1);
        V3 fx = normalized(cross(fy, off)), fz = normalized(cross(fx, fy));
        Col sum;
        double inv_sum = 0, smallest = kBoundHuge;
        V3 min_dir;
        gathers[depth]++;
        for (int i = 0; i < N; i++) {
            V3 v = pool[cursor[depth]];
            cursor[depth] = (cursor[depth] + 1) % kPoolSize;
            v.y = std::fabs(v.y);
            V3 dir;
            if (fy.y > 1.0 - kRadEps)
                dir = v;
            else if (fy.y < -1.0 + kRadEps)
                dir = v * -1.0;
            else
                dir = fx * v.x + fy * v.y + fz * v.z;
            Ray r{p, dir};
            Isect h;
            int32_t obj;
            rays++;
            double depth_hit = HUGE_VAL;
            if (scene_find_intersection(s, r, 0.0, PVX_RAY_RADIOSITY, h, obj, 1)) {
                depth_hit = h.depth;
                sum = sum + shade(r, h, obj, depth + 1, wnext, final);
            }
            if (depth_hit < smallest) {
                smallest = depth_hit;
                min_dir = dir;
            }
            inv_sum += 1.0 / depth_hit;
        }
        out = sum * (1.0 / N);
        double mean = N / inv_sum;
        if (smallest > reuse_max * 0.0001) {
            if (mean < reuse_min) mean = reuse_min;
            if (mean > reuse_max) mean = reuse_max;
            add(Record{p, n, out, mean, smallest, min_dir, depth});
        }
        return 1.0;
    }
};

} // namespace

// Object normal helper for the radiosity oracle (render.cpp's shape normals).
} // namespace orc

using namespace orc;

extern "C" int orc_radiosity_render(void *sp, const pvx_shading_desc *sd, const pvx_camera *cam,
                                    const pvx_radiosity_settings *rs, float *rgb, uint64_t *counters) {
    const Scene &s = *static_cast<Scene *>(sp);
    std::vector<pvx_material> mats(sd->materials, sd->materials + sd->n_materials);
    Rad rad(s, mats, *rs, V3(cam->location[0], cam->location[1], cam->location[2]));
    uint32_t W = cam->width, H = cam->height;
    auto primary = [&](double px, double py, Ray &r) {
        double x0 = px / W - 0.5, y0 = 0.5 - py / H;
        V3 d;
        for (int k = 0; k < 3; k++) d[k] = cam->direction[k] + x0 * cam->right[k] + y0 * cam->up[k];
        r = Ray{V3(cam->location[0], cam->location[1], cam->location[2]), normalized(d)};
    };
    // pretrace: pixel step from pretrace_start to pretrace_end of the width
    double step = std::max(1.0, std::floor((double)rs->pretrace_start * std::max(W, H)));
    double end = std::max(1.0, std::floor((double)rs->pretrace_end * std::max(W, H)));
    for (;;) {
        for (double y = 0; y < H; y += step)
            for (double x = 0; x < W; x += step) {
                Ray r;
                primary(x + 0.5, y + 0.5, r);
                Isect h;
                int32_t obj;
                if (scene_find_intersection(s, r, 0.0, PVX_RAY_IMAGE, h, obj, 1))
                    (void)rad.shade(r, h, obj, 0, 1.0, false);
            }
        if (step <= end) break;
        step = std::max(end, std::floor(step / 2));
    }
    // final pass
    for (uint32_t y = 0; y < H; y++)
        for (uint32_t x = 0; x < W; x++) {
            Ray r;
            primary(x + 0.5, y + 0.5, r);
            Isect h;
            int32_t obj;
            Col c;
            if (scene_find_intersection(s, r, 0.0, PVX_RAY_IMAGE, h, obj, 1))
                c = rad.shade(r, h, obj, 0, 1.0, true);
            size_t i = (size_t)y * W + x;
            rgb[3 * i] = (float)c.r;
            rgb[3 * i + 1] = (float)c.g;
            rgb[3 * i + 2] = (float)c.b;
        }
    if (counters) {
        counters[0] = rad.gathers[0];
        counters[1] = rad.gathers[1];
        counters[2] = rad.rays;
        counters[3] = rad.recs[0].size();
        counters[4] = rad.recs[1].size();
    }
    return 0;
}

// The direction pool the gathers draw from (kPoolSize x 3).
// RadiosityFunction over one cache, queried point by point in the caller's
// order (ComputeAmbient, radiosity.cpp:385-461): open, query, close.
struct OrcRadiosity {
    std::vector<pvx_material> mats;
    Rad rad;
    OrcRadiosity(const Scene &s, const pvx_shading_desc *sd, const pvx_radiosity_settings *rs, V3 cam)
        : mats(sd->materials, sd->materials + sd->n_materials), rad(s, mats, *rs, cam) {}
};

extern "C" void *orc_rad_open(void *sp, const pvx_shading_desc *sd, const pvx_camera *cam,
                              const pvx_radiosity_settings *rs) {
    return new OrcRadiosity(*static_cast<Scene *>(sp), sd, rs,
                            V3(cam->location[0], cam->location[1], cam->location[2]));
}

extern "C" void orc_rad_close(void *h) { delete static_cast<OrcRadiosity *>(h); }

// ambient (n x 3) for points p, normals nrm (n x 3) and weights w (n, or
// NULL = 1); counters: gathers at depth 0 / 1, rays, records at depth 0 / 1.
extern "C" int orc_rad_ambient(void *h, size_t n, const double *p, const double *nrm, const double *w, int depth,
                               int final, float *out, uint64_t *counters) {
    Rad &rad = static_cast<OrcRadiosity *>(h)->rad;
    for (size_t i = 0; i < n; i++) {
        Col c = rad.ambient(V3(p[3 * i], p[3 * i + 1], p[3 * i + 2]),
                            V3(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]), w ? w[i] : 1.0, depth, final != 0);
        out[3 * i] = (float)c.r;
        out[3 * i + 1] = (float)c.g;
        out[3 * i + 2] = (float)c.b;
    }
    if (counters) {
        counters[0] = rad.gathers[0];
        counters[1] = rad.gathers[1];
        counters[2] = rad.rays;
        counters[3] = rad.recs[0].size();
        counters[4] = rad.recs[1].size();
    }
    return 0;
}

extern "C" int orc_halton_pool(double *out) {
    static const Scene empty;
    static const std::vector<pvx_material> no_mats;
    pvx_radiosity_settings st{};
    Rad rad(empty, no_mats, st, V3(0, 0, 0));
    for (int i = 0; i < kPoolSize; i++) {
        out[3 * i] = rad.pool[i].x;
        out[3 * i + 1] = rad.pool[i].y;
        out[3 * i + 2] = rad.pool[i].z;
    }
    return kPoolSize;
}

// FindReusableBlock + AverageNearBlock of one query against hand-built
// records (15 doubles each: point 3, normal 3, irradiance 3, harmonic-mean
// distance, nearest distance, direction to the nearest surface 3, depth).
// `error_bound` places the records in the lookup grid like the cache's own
// insertions; `bound` is the query's error bound.  out4 = averaged
// irradiance (r, g, b) and the weight sum.
extern "C" int orc_rad_average(const double *recs, size_t n, double error_bound, const double *p3, const double *n3,
                               double bound, int depth, int final_pass, double *out4) {
    static const Scene empty;
    static const std::vector<pvx_material> no_mats;
    pvx_radiosity_settings st{};
    st.error_bound = (float)error_bound;
    Rad rad(empty, no_mats, st, V3(0, 0, 0));
    for (size_t i = 0; i < n; i++) {
        const double *q = recs + 15 * i;
        Record rc;
        rc.p = V3(q[0], q[1], q[2]);
        rc.n = V3(q[3], q[4], q[5]);
        rc.E = Col{q[6], q[7], q[8]};
        rc.r = q[9];
        rc.nearest = q[10];
        rc.to_nearest = V3(q[11], q[12], q[13]);
        rc.depth = (int)q[14];
        rad.add(rc);
    }
    Col E;
    const double W = rad.lookup(V3(p3[0], p3[1], p3[2]), V3(n3[0], n3[1], n3[2]), bound, depth, final_pass != 0, E);
    out4[0] = E.r;
    out4[1] = E.g;
    out4[2] = E.b;
    out4[3] = W;
    return 0;
}
