// FP64 CPU ORACLE — TEST INFRASTRUCTURE ONLY (see oracle.hpp).
// Stand-alone triangles: Parse_Triangle / Parse_Smooth_Triangle
// (parser/parser.cpp:5539-5590, 6023-6047), Triangle::Compute_Triangle /
// SmoothTriangle::Compute_Triangle / Compute_Smooth_Triangle
// (core/shape/triangle.cpp:95-420), Triangle::Intersect (479-572),
// Triangle::Normal / SmoothTriangle::Normal (641-644, 712-738) and
// Triangle::Compute_BBox (1296-1311).
#include "oracle.hpp"

namespace orc {

namespace {
const double kTriDepthTol = 1.0e-6; // triangle.cpp:62 DEPTH_TOLERANCE

// max3_coordinate (triangle.cpp:64-65)
int max3_coordinate(double x, double y, double z) { return (x > y) ? ((x > z) ? 0 : 2) : ((y > z) ? 1 : 2); }

V3 normalized_or_same(V3 a) { // Vector3d::normalized (vector.h:523-529)
    double l = length(a);
    return l != 0 ? a / l : a;
}
} // namespace

Tri make_tri(const double *P, const double *N, bool smooth) {
    Tri t;
    t.smooth = smooth;
    t.p1 = V3(P[0], P[1], P[2]);
    t.p2 = V3(P[3], P[4], P[5]);
    t.p3 = V3(P[6], P[7], P[8]);
    if (smooth) { // Parse_Smooth_Triangle normalises each normal (a zero one: no Compute_Triangle)
        V3 *ns[3] = {&t.n1, &t.n2, &t.n3};
        for (int j = 0; j < 3; j++) {
            V3 v(N[3 * j], N[3 * j + 1], N[3 * j + 2]);
            double l = length(v);
            if (l == 0.0)
                t.degenerate = true;
            else
                v = V3(v.x / l, v.y / l, v.z / l);
            *ns[j] = v;
        }
        if (t.degenerate) return t;
    }
    // Compute_Triangle
    V3 v1 = t.p1 - t.p2, v2 = t.p3 - t.p2;
    V3 nv = cross(v1, v2);
    double len = length(nv);
    if (len == 0.0) {
        t.degenerate = true;
        return t;
    }
    nv = V3(nv.x / len, nv.y / len, nv.z / len);
    t.normal = nv;
    t.distance = -dot(nv, t.p1);
    t.axis = max3_coordinate(std::fabs(nv.x), std::fabs(nv.y), std::fabs(nv.z));
    bool swap = false;
    const V3 &P1 = t.p1, &P2 = t.p2, &P3 = t.p3;
    if (t.axis == 0)
        swap = (P2.y - P3.y) * (P2.z - P1.z) < (P2.z - P3.z) * (P2.y - P1.y);
    else if (t.axis == 1)
        swap = (P2.x - P3.x) * (P2.z - P1.z) < (P2.z - P3.z) * (P2.x - P1.x);
    else
        swap = (P2.x - P3.x) * (P2.y - P1.y) < (P2.y - P3.y) * (P2.x - P1.x);
    if (swap) {
        std::swap(t.p1, t.p2);
        std::swap(t.n1, t.n2);
    }
    if (smooth) { // Compute_Smooth_Triangle
        V3 d32 = t.p3 - t.p2;
        t.vaxis = max3_coordinate(std::fabs(d32.x), std::fabs(d32.y), std::fabs(d32.z));
        V3 a = normalized_or_same(t.p2 - t.p3);
        V3 b = t.p1 - t.p3;
        double proj = dot(b, a);
        a = a * proj;
        V3 perp = normalized_or_same(a - b);
        double uden = dot(b, perp);
        t.perp = perp / -uden;
        double x = dot(t.normal, t.n1), y = dot(t.normal, t.n2), z = dot(t.normal, t.n3);
        if (!(((x < 0.0) && (y < 0.0) && (z < 0.0)) || ((x > 0.0) && (y > 0.0) && (z > 0.0))))
            t.degenerate = true;
    }
    return t;
}

// Triangle::Intersect: plane depth, then the edge tests in the projection
// that drops the dominant axis.
bool tri_intersect(const Tri &t, const Ray &r, double *depth) {
    if (t.degenerate) return false;
    double ndd = dot(t.normal, r.d);
    if (std::fabs(ndd) < kEpsilon) return false;
    double ndo = dot(t.normal, r.o);
    *depth = -(t.distance + ndo) / ndd;
    if ((*depth < kTriDepthTol) || (*depth > kMaxDistance)) return false;
    const int a = t.axis == 0 ? 1 : 0, b = t.axis == 2 ? 1 : 2;
    double s = r.o[a] + *depth * r.d[a];
    double u = r.o[b] + *depth * r.d[b];
    const V3 &P1 = t.p1, &P2 = t.p2, &P3 = t.p3;
    if ((P2[a] - s) * (P2[b] - P1[b]) < (P2[b] - u) * (P2[a] - P1[a])) return false;
    if ((P3[a] - s) * (P3[b] - P2[b]) < (P3[b] - u) * (P3[a] - P2[a])) return false;
    if ((P1[a] - s) * (P1[b] - P3[b]) < (P1[b] - u) * (P1[a] - P3[a])) return false;
    return true;
}

V3 tri_normal(const Tri &t, V3 ip) {
    if (!t.smooth) return t.normal;
    V3 pi = ip - t.p1;
    double u = dot(pi, t.perp);
    if (u < kEpsilon) return t.n1;
    int ax = t.vaxis;
    double v = (pi[ax] / u + t.p1[ax] - t.p2[ax]) / (t.p3[ax] - t.p2[ax]);
    V3 n = t.n1 + u * (t.n2 - t.n1 + v * (t.n3 - t.n2));
    return normalized_or_same(n);
}

FBox tri_bbox(const double *P) {
    V3 mn(std::min(P[0], std::min(P[3], P[6])), std::min(P[1], std::min(P[4], P[7])),
          std::min(P[2], std::min(P[5], P[8])));
    V3 mx(std::max(P[0], std::max(P[3], P[6])), std::max(P[1], std::max(P[4], P[7])),
          std::max(P[2], std::max(P[5], P[8])));
    V3 e(kEpsilon, kEpsilon, kEpsilon);
    return make_bbox_from_min_max(mn - e, mx + e);
}

} // namespace orc
