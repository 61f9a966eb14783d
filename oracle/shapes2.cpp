// FP64 CPU ORACLE — TEST INFRASTRUCTURE ONLY (see oracle.hpp).
// Box, torus and CSG restatements (core/shape/box.cpp, core/shape/torus.cpp,
// core/shape/csg.cpp, core/bounding/boundingbox.cpp Recompute_BBox).
#include <algorithm>

#include "oracle.hpp"

namespace orc {

namespace {
const double kBoxDepthTol = 1.0e-6;   // box.cpp:67
const double kCloseTol = 1.0e-6;      // box.cpp:71 CLOSE_TOLERANCE
const double kTorusDepthTol = 1.0e-4; // torus.cpp:78
const double kRootTol = 1.0e-4;       // torus.cpp:82
enum { X0 = 1, X1 = 2, Y0 = 3, Y1 = 4, Z0 = 5, Z1 = 6 }; // box.h:81-87
} // namespace

// Box::Intersect (box.cpp:172-505).
bool box_intersect(const Ray &r, const Transform *tr, V3 c1, V3 c2, double *d1, double *d2,
                   int *s1, int *s2) {
    int smin = 0, smax = 0;
    double t, tmin = 0.0, tmax = kBoundHuge;
    V3 P = r.o, D = r.d;
    if (tr) {
        P = mtrans_point(tr->inv, r.o);
        D = mtrans_direction(tr->inv, r.d);
    }
    // X sides: no tie handling.
    if (D.x < -kEpsilon) {
        t = (c1.x - P.x) / D.x;
        if (t < tmin) return false;
        if (t <= tmax) { smax = X0; tmax = t; }
        t = (c2.x - P.x) / D.x;
        if (t >= tmin) {
            if (t > tmax) return false;
            smin = X1;
            tmin = t;
        }
    } else if (D.x > kEpsilon) {
        t = (c2.x - P.x) / D.x;
        if (t < tmin) return false;
        if (t <= tmax) { smax = X1; tmax = t; }
        t = (c1.x - P.x) / D.x;
        if (t >= tmin) {
            if (t > tmax) return false;
            smin = X0;
            tmin = t;
        }
    } else if (P.x < c1.x || P.x > c2.x) {
        return false;
    }
    // Y sides: ties resolved against the X component.
    if (D.y < -kEpsilon) {
        t = (c1.y - P.y) / D.y;
        if (t < tmin) return false;
        if (t <= tmax - kCloseTol) { smax = Y0; tmax = t; }
        else if (t <= tmax + kCloseTol) { if (-D.y > std::fabs(D.x)) smax = Y0; }
        t = (c2.y - P.y) / D.y;
        if (t >= tmin + kCloseTol) {
            if (t > tmax) return false;
            smin = Y1;
            tmin = t;
        } else if (t >= tmin - kCloseTol) {
            if (-D.y > std::fabs(D.x)) smin = Y1;
        }
    } else if (D.y > kEpsilon) {
        t = (c2.y - P.y) / D.y;
        if (t < tmin) return false;
        if (t <= tmax - kCloseTol) { smax = Y1; tmax = t; }
        else if (t <= tmax + kCloseTol) { if (D.y > std::fabs(D.x)) smax = Y1; }
        t = (c1.y - P.y) / D.y;
        if (t >= tmin + kCloseTol) {
            if (t > tmax) return false;
            smin = Y0;
            tmin = t;
        } else if (t >= tmin - kCloseTol) {
            if (D.y > std::fabs(D.x)) smin = Y0;
        }
    } else if (P.y < c1.y || P.y > c2.y) {
        return false;
    }
    // Z sides: ties resolved against the axis of the side chosen so far.
    auto tie = [&](int side, double dz) {
        if (side == X0 || side == X1) return dz > std::fabs(D.x);
        if (side == Y0 || side == Y1) return dz > std::fabs(D.y);
        return false;
    };
    if (D.z < -kEpsilon) {
        t = (c1.z - P.z) / D.z;
        if (t < tmin) return false;
        if (t <= tmax - kCloseTol) { smax = Z0; tmax = t; }
        else if (t <= tmax + kCloseTol) { if (tie(smax, -D.z)) smax = Z0; }
        t = (c2.z - P.z) / D.z;
        if (t >= tmin + kCloseTol) {
            if (t > tmax) return false;
            smin = Z1;
            tmin = t;
        } else if (t >= tmin - kCloseTol) {
            if (tie(smin, -D.z)) smin = Z1;
        }
    } else if (D.z > kEpsilon) {
        t = (c2.z - P.z) / D.z;
        if (t < tmin) return false;
        if (t <= tmax - kCloseTol) { smax = Z1; tmax = t; }
        else if (t <= tmax + kCloseTol) { if (tie(smax, D.z)) smax = Z1; }
        t = (c1.z - P.z) / D.z;
        if (t >= tmin + kCloseTol) {
            if (t > tmax) return false;
            smin = Z0;
            tmin = t;
        } else if (t >= tmin - kCloseTol) {
            if (tie(smin, D.z)) smin = Z0;
        }
    } else if (P.z < c1.z || P.z > c2.z) {
        return false;
    }
    if (tmax < kBoxDepthTol) return false;
    *d1 = tmin;
    *d2 = tmax;
    *s1 = smin;
    *s2 = smax;
    return true;
}

// Box::Inside (box.cpp:535-570)
bool box_inside(V3 p, const Transform *tr, V3 c1, V3 c2, bool inverted) {
    V3 q = tr ? mtrans_point(tr->inv, p) : p;
    if (q.x < c1.x || q.x > c2.x) return inverted;
    if (q.y < c1.y || q.y > c2.y) return inverted;
    if (q.z < c1.z || q.z > c2.z) return inverted;
    return !inverted;
}

// Torus::Test_Thick_Cylinder (torus.cpp:932-1085)
static bool test_thick_cylinder(V3 P, V3 D, double h1, double h2, double r1, double r2) {
    double k, u, v, r, h;
    if (std::fabs(D.y) < kEpsilon) {
        if (P.y < h1 || P.y > h2) return false;
    } else {
        k = (h2 - P.y) / D.y;
        u = P.x + k * D.x;
        v = P.z + k * D.z;
        if (k > kEpsilon && k < kMaxDistance) {
            r = u * u + v * v;
            if (r >= r1 && r <= r2) return true;
        }
        k = (h1 - P.y) / D.y;
        u = P.x + k * D.x;
        v = P.z + k * D.z;
        if (k > kEpsilon && k < kMaxDistance) {
            r = u * u + v * v;
            if (r >= r1 && r <= r2) return true;
        }
    }
    double a = D.x * D.x + D.z * D.z;
    if (a > kEpsilon) {
        double b = P.x * D.x + P.z * D.z;
        for (double rr : {r2, r1}) { // outer, then inner cylinder
            double c = P.x * P.x + P.z * P.z - rr;
            double d = b * b - a * c;
            if (d >= 0.0) {
                d = std::sqrt(d);
                k = (-b + d) / a;
                if (k > kEpsilon && k < kMaxDistance) {
                    h = P.y + k * D.y;
                    if (h >= h1 && h <= h2) return true;
                }
                k = (-b - d) / a;
                if (k > kEpsilon && k < kMaxDistance) {
                    h = P.y + k * D.y;
                    if (h >= h1 && h <= h2) return true;
                }
            }
        }
    }
    return false;
}

static const Transform kIdentity = {
    {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}},
    {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}}};

// Torus::Intersect (torus.cpp:241-329) + All_Intersections (133-160).
bool torus_all_intersections(const Ray &ray, const Transform *tr, double R, double rr, bool sturm,
                             std::vector<Isect> &stack, int32_t slot) {
    if (!tr) tr = &kIdentity;
    V3 P = mtrans_point(tr->inv, ray.o);
    V3 D = mtrans_direction(tr->inv, ray.d);
    double len = length(D);
    D = D / len;
    double depth[4];
    int i = 0;
    double y1 = -rr, y2 = rr;
    double r1 = (R - rr) * (R - rr);
    if (R < rr) r1 = 0;
    double r2 = (R + rr) * (R + rr);
    if (test_thick_cylinder(P, D, y1, y2, r1, r2)) {
        double bsr = R + rr + rr;
        double dist = length_sqr(P);
        double closer = 0.0;
        if (dist > bsr * bsr) {
            dist = std::sqrt(dist);
            closer = dist - bsr;
            P = P + closer * D;
        }
        double R2 = R * R, rr2 = rr * rr;
        double Py2 = P.y * P.y, Dy2 = D.y * D.y, PDy2 = P.y * D.y;
        double k1 = P.x * P.x + P.z * P.z + Py2 - R2 - rr2;
        double k2 = P.x * D.x + P.z * D.z + PDy2;
        double c[5];
        c[0] = 1.0;
        c[1] = 4.0 * k2;
        c[2] = 2.0 * (k1 + 2.0 * (k2 * k2 + R2 * Dy2));
        c[3] = 4.0 * (k2 * k1 + 2.0 * R2 * PDy2);
        c[4] = k1 * k1 + 4.0 * R2 * (Py2 - rr2);
        double roots[4];
        int n = solve_polynomial(4, c, roots, sturm, kRootTol);
        while (n--) depth[i++] = (roots[n] + closer) / len;
    }
    bool found = false;
    for (int k = 0; k < i; k++) {
        if (depth[k] > kTorusDepthTol && depth[k] < kMaxDistance) {
            stack.push_back(Isect{depth[k], slot, 0, slot});
            found = true;
        }
    }
    return found;
}

// Torus::Inside (torus.cpp:364-385)
bool torus_inside(V3 p, const Transform *tr, double R, double rr, bool inverted) {
    if (!tr) tr = &kIdentity;
    V3 P = mtrans_point(tr->inv, p);
    double r = std::sqrt(P.x * P.x + P.z * P.z);
    double r2 = P.y * P.y + (r - R) * (r - R);
    return (r2 <= rr * rr) ? !inverted : inverted;
}

// Recompute_BBox (boundingbox.cpp:168-210)
FBox recompute_bbox(const FBox &b, const Transform *tr) {
    if (!tr) return b;
    V3 ll(b.ll[0], b.ll[1], b.ll[2]), len(b.size[0], b.size[1], b.size[2]);
    V3 mins(kBoundHuge, kBoundHuge, kBoundHuge), maxs(-kBoundHuge, -kBoundHuge, -kBoundHuge);
    for (int i = 1; i <= 8; i++) {
        V3 c = ll;
        c.x += (i & 1) ? len.x : 0.0;
        c.y += (i & 2) ? len.y : 0.0;
        c.z += (i & 4) ? len.z : 0.0;
        c = mtrans_point(tr->m, c);
        for (int k = 0; k < 3; k++) {
            if (c[k] < mins[k]) mins[k] = c[k];
            if (c[k] > maxs[k]) maxs[k] = c[k];
        }
    }
    for (int k = 0; k < 3; k++) {
        if (mins[k] < -kBoundHuge / 2) mins[k] = -kBoundHuge / 2;
        if (maxs[k] > kBoundHuge / 2) maxs[k] = kBoundHuge / 2;
    }
    return make_bbox_from_min_max(mins, maxs);
}

// CSG::Compute_BBox for boxes/tori lives with the scene (scene.cpp).
FBox box_bbox(V3 c1, V3 c2, const Transform *tr) { // Box::Compute_BBox (box.cpp:947-960)
    FBox b;
    for (int k = 0; k < 3; k++) {
        b.ll[k] = (float)c1[k];
        b.size[k] = (float)(c2[k] - c1[k]);
    }
    return recompute_bbox(b, tr);
}

FBox torus_bbox(double R, double rr, const Transform *tr) { // Torus::Compute_BBox (torus.cpp)
    double r1 = rr, r2 = R + rr;
    FBox b;
    b.ll[0] = (float)-r2; b.ll[1] = (float)-r1; b.ll[2] = (float)-r2;
    b.size[0] = (float)(2.0 * r2); b.size[1] = (float)(2.0 * r1); b.size[2] = (float)(2.0 * r2);
    return recompute_bbox(b, tr ? tr : &kIdentity);
}

} // namespace orc
