// Host-side scene preparation for the pvx engine.
//
// * mesh2 triangles are prepared exactly as the reference's parser leaves
//   them (Parse_Mesh2 -> Mesh::Compute_Mesh_Triangle, parser/parser.cpp:
//   4512-4565, core/shape/mesh.cpp:838-945): FP64 face normal from the FP32
//   vertices, stored as FP32 together with an FP32 Distance; vertex order
//   swapped by the dominant-axis orientation test.  Degenerate triangles keep
//   a zero normal in the reference and can never be hit, so they are dropped.
// * every bounded primitive gets a conservative FP32 box (FP64 bounds
//   rounded outward, then padded) and goes into one BVH.
// * planes and unbounded CSG objects are tested outside the BVH, like the
//   reference's infinite objects (boundingbox.cpp:288-306).
#include "scene_prep.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>

#include "bvh_build.h"
#include "kd_build.h"

namespace pvx {

namespace {

struct DBox {
    double lo[3], hi[3];
    bool bounded;
    static DBox empty() {
        DBox b;
        for (int k = 0; k < 3; k++) {
            b.lo[k] = INFINITY;
            b.hi[k] = -INFINITY;
        }
        b.bounded = true;
        return b;
    }
    static DBox unbounded() {
        DBox b;
        for (int k = 0; k < 3; k++) {
            b.lo[k] = -INFINITY;
            b.hi[k] = INFINITY;
        }
        b.bounded = false;
        return b;
    }
    void grow(const double *p) {
        for (int k = 0; k < 3; k++) {
            lo[k] = std::min(lo[k], p[k]);
            hi[k] = std::max(hi[k], p[k]);
        }
    }
};

inline float round_down(double v) {
    float f = (float)v;
    if ((double)f > v) f = std::nextafter(f, -INFINITY);
    return f;
}
inline float round_up(double v) {
    float f = (float)v;
    if ((double)f < v) f = std::nextafter(f, INFINITY);
    return f;
}

// Conservative FP32 box with a small relative pad (absorbs FP32 slab-test
// rounding in traversal).
void to_fbox(const DBox &b, float *lo, float *hi) {
    for (int k = 0; k < 3; k++) {
        double ext = b.hi[k] - b.lo[k];
        double mag = std::max(std::fabs(b.lo[k]), std::fabs(b.hi[k]));
        double pad = ext * 1.0e-6 + mag * 4.8e-7 + 1.0e-30;
        lo[k] = round_down(b.lo[k] - pad);
        hi[k] = round_up(b.hi[k] + pad);
    }
}

void xf_point(const double *m, const double *p, double *r) {
    for (int i = 0; i < 3; i++) r[i] = p[0] * m[i] + p[1] * m[4 + i] + p[2] * m[8 + i] + m[12 + i];
}

DBox transformed_box(const double *lo, const double *hi, const double *m) {
    DBox b = DBox::empty();
    for (int i = 0; i < 8; i++) {
        double c[3] = {(i & 1) ? hi[0] : lo[0], (i & 2) ? hi[1] : lo[1], (i & 4) ? hi[2] : lo[2]};
        double r[3];
        if (m)
            xf_point(m, c, r);
        else
            std::memcpy(r, c, sizeof r);
        b.grow(r);
    }
    return b;
}

// TRANSFORM builders in the reference's arithmetic (core/math/matrix.cpp),
// on the flat row-major Xform layout: m[4 * row + col].
void xf_identity(Xform &x) {
    for (int k = 0; k < 16; k++) x.m[k] = x.inv[k] = (k % 5 == 0) ? 1.0 : 0.0;
}
// Compose_Transforms (matrix.cpp:788-793): m = m * a.m (MTimesA),
// inv = a.inv * inv (MTimesB), each entry summed over k = 0..3 in order.
void xf_compose(Xform &x, const Xform &a) {
    double r[16];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            r[4 * i + j] = x.m[4 * i] * a.m[j] + x.m[4 * i + 1] * a.m[4 + j] + x.m[4 * i + 2] * a.m[8 + j] +
                           x.m[4 * i + 3] * a.m[12 + j];
    std::memcpy(x.m, r, sizeof r);
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            r[4 * i + j] = a.inv[4 * i] * x.inv[j] + a.inv[4 * i + 1] * x.inv[4 + j] +
                           a.inv[4 * i + 2] * x.inv[8 + j] + a.inv[4 * i + 3] * x.inv[12 + j];
    std::memcpy(x.inv, r, sizeof r);
}
// Compute_Coordinate_Transform (matrix.cpp:880-906): scale (radius, radius,
// length), rotate z onto `up` about (-up.y, up.x, 0) (Compute_Axis_Rotation_
// Transform, matrix.cpp:825-850), translate to `origin`.
void xf_coordinate(Xform &x, const double *origin, double *up, double radius, double length) {
    xf_identity(x);
    x.m[0] = radius;
    x.m[5] = radius;
    x.m[10] = length;
    x.inv[0] = 1.0 / radius;
    x.inv[5] = 1.0 / radius;
    x.inv[10] = 1.0 / length;
    double ax[3];
    if (std::fabs(up[2]) > 1.0 - 1.0e-10) {
        ax[0] = 1.0, ax[1] = 0.0, ax[2] = 0.0;
        up[2] = up[2] < 0.0 ? -1.0 : 1.0;
    } else {
        ax[0] = -up[1], ax[1] = up[0], ax[2] = 0.0;
    }
    const double angle = std::acos(up[2]);
    const double l = std::sqrt(ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2]);
    const double v[3] = {l != 0 ? ax[0] / l : ax[0], l != 0 ? ax[1] / l : ax[1], l != 0 ? ax[2] / l : ax[2]};
    const double c = std::cos(angle), s = std::sin(angle);
    Xform rot;
    xf_identity(rot);
    rot.m[0] = v[0] * v[0] + c * (1.0 - v[0] * v[0]);
    rot.m[1] = v[0] * v[1] * (1.0 - c) + v[2] * s;
    rot.m[2] = v[0] * v[2] * (1.0 - c) - v[1] * s;
    rot.m[4] = v[0] * v[1] * (1.0 - c) - v[2] * s;
    rot.m[5] = v[1] * v[1] + c * (1.0 - v[1] * v[1]);
    rot.m[6] = v[1] * v[2] * (1.0 - c) + v[0] * s;
    rot.m[8] = v[0] * v[2] * (1.0 - c) + v[1] * s;
    rot.m[9] = v[1] * v[2] * (1.0 - c) - v[0] * s;
    rot.m[10] = v[2] * v[2] + c * (1.0 - v[2] * v[2]);
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) rot.inv[4 * i + j] = rot.m[4 * j + i];
    xf_compose(x, rot);
    Xform tr;
    xf_identity(tr);
    for (int k = 0; k < 3; k++) {
        tr.m[12 + k] = origin[k];
        tr.inv[12 + k] = -origin[k];
    }
    xf_compose(x, tr);
}

// A cone / cylinder statement as Parse_Cone / Parse_Cylinder leave it
// (parser.cpp:2255-2315) after Compute_Cone_Data / Compute_Cylinder_Data
// (cone.cpp:715-840).  Returns false for a degenerate shape (the reference
// throws "Degenerate cone/cylinder").
bool make_cone(const double *apex_in, const double *base_in, const double *radius, uint8_t flags, ConeRec &c,
               Xform &x) {
    double apex[3], base[3];
    std::memcpy(apex, apex_in, sizeof apex);
    std::memcpy(base, base_in, sizeof base);
    const bool cyl_stmt = (flags & PVX_CONE_CYLINDER) != 0;
    double ar = radius[0], br = cyl_stmt ? radius[0] : radius[1];
    c.flags = (flags & PVX_CONE_OPEN) ? 0u : kConeClosed;
    c.dist = 0.0;
    auto cylinder_data = [&]() -> bool {
        double axis[3] = {apex[0] - base[0], apex[1] - base[1], apex[2] - base[2]};
        const double l = std::sqrt(axis[0] * axis[0] + axis[1] * axis[1] + axis[2] * axis[2]);
        if (l < kEpsilon) return false;
        for (int k = 0; k < 3; k++) axis[k] = axis[k] / l;
        xf_coordinate(x, base, axis, ar, l);
        c.flags |= kConeCylinder;
        c.dist = 0.0;
        return true;
    };
    if (cyl_stmt) return cylinder_data();
    double axis[3] = {apex[0] - base[0], apex[1] - base[1], apex[2] - base[2]};
    const double len = std::sqrt(axis[0] * axis[0] + axis[1] * axis[1] + axis[2] * axis[2]);
    if (len < kEpsilon) return false;
    for (int k = 0; k < 3; k++) axis[k] = axis[k] / len;
    if (std::fabs(ar - br) < kEpsilon) return cylinder_data();
    if (ar < br) { // the bigger end at the top
        for (int k = 0; k < 3; k++) {
            std::swap(apex[k], base[k]);
            axis[k] = -axis[k];
        }
        std::swap(ar, br);
    }
    const double tmpf = br * len / (ar - br);
    const double origin[3] = {base[0] - axis[0] * tmpf, base[1] - axis[1] * tmpf, base[2] - axis[2] * tmpf};
    const double tlen = tmpf + len;
    if (((ar - br) * len / tlen) < kEpsilon) return cylinder_data();
    c.dist = tmpf / tlen;
    xf_coordinate(x, origin, axis, ar, tlen);
    return true;
}

// MTransNormal (matrix.cpp) with a transform's inverse: row i of the inverse.
void xf_normal(const Xform &x, const double *n, double *out) {
    for (int i = 0; i < 3; i++) out[i] = n[0] * x.inv[4 * i] + n[1] * x.inv[4 * i + 1] + n[2] * x.inv[4 * i + 2];
}

// Parse_Disc (parser.cpp) + Compute_Disc (disc.cpp): the normal normalised,
// the placement Compute_Coordinate_Transform(center, normal, 1, 1), which
// may snap the stored normal's z to +-1.
void make_disc(const double *center, const double *normal, const double *radius, DiscRec &c, Xform &x) {
    double n[3] = {normal[0], normal[1], normal[2]};
    const double l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (l != 0)
        for (int k = 0; k < 3; k++) n[k] = n[k] / l;
    c.or2 = radius[0] * radius[0];
    c.ir2 = radius[1] * radius[1];
    xf_coordinate(x, center, n, 1.0, 1.0);
    std::memcpy(c.n, n, sizeof n);
    c.pad = 0;
}

inline int max3_axis(double x, double y, double z) { // mesh.cpp:94 max3_coordinate
    return (x > y) ? ((x > z) ? 0 : 2) : ((y > z) ? 1 : 2);
}

// Mesh::compute_smooth_triangle (mesh.cpp:983-1008) on post-swap P1, P2, P3.
void smooth_triangle(const double *P1, const double *P2, const double *P3, ShadeTri &t) {
    double d32[3] = {P3[0] - P2[0], P3[1] - P2[1], P3[2] - P2[2]};
    int vaxis = max3_axis(std::fabs(d32[0]), std::fabs(d32[1]), std::fabs(d32[2]));
    double v1[3] = {P2[0] - P3[0], P2[1] - P3[1], P2[2] - P3[2]};
    double l1 = std::sqrt(v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2]);
    if (l1 != 0.0)
        for (int k = 0; k < 3; k++) v1[k] = v1[k] / l1;
    double v2[3] = {P1[0] - P3[0], P1[1] - P3[1], P1[2] - P3[2]};
    double proj = (v2[0] * v1[0]) + (v2[1] * v1[1]) + (v2[2] * v1[2]);
    double pp[3];
    for (int k = 0; k < 3; k++) pp[k] = v1[k] * proj - v2[k];
    double lp = std::sqrt(pp[0] * pp[0] + pp[1] * pp[1] + pp[2] * pp[2]);
    float pf[3];
    for (int k = 0; k < 3; k++) pf[k] = (float)(lp != 0.0 ? pp[k] / lp : pp[k]);
    double uden = -((v2[0] * (double)pf[0]) + (v2[1] * (double)pf[1]) + (v2[2] * (double)pf[2]));
    float udf = (float)uden;
    for (int k = 0; k < 3; k++) t.perp[k] = pf[k] / udf;
    t.vaxis_smooth = (uint32_t)vaxis;
}

// Triangle::Compute_Triangle / SmoothTriangle::Compute_Triangle and
// Compute_Smooth_Triangle (core/shape/triangle.cpp:95-420) on the parsed
// points; smooth normals are first normalised as Parse_Smooth_Triangle
// does (parser/parser.cpp:5555-5583).  Same operations in the same order,
// so every stored value matches the reference's bits.
TriangleRec make_triangle(const double *P, const double *N, bool smooth) {
    TriangleRec t;
    std::memset(&t, 0, sizeof t);
    double P1[3], P2[3], P3[3], N1[3] = {0, 0, 0}, N2[3] = {0, 0, 0}, N3[3] = {0, 0, 0};
    for (int k = 0; k < 3; k++) {
        P1[k] = P[k];
        P2[k] = P[3 + k];
        P3[k] = P[6 + k];
    }
    auto length = [](const double *v) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); };
    auto dot = [](const double *a, const double *b) { return ((a[0] * b[0]) + (a[1] * b[1])) + (a[2] * b[2]); };
    auto store = [&]() {
        for (int k = 0; k < 3; k++) {
            t.p1[k] = P1[k];
            t.p2[k] = P2[k];
            t.p3[k] = P3[k];
            t.n1[k] = N1[k];
            t.n2[k] = N2[k];
            t.n3[k] = N3[k];
        }
    };
    t.smooth = smooth ? 1 : 0;
    if (smooth) {
        double *Ns[3] = {N1, N2, N3};
        bool degen = false;
        for (int j = 0; j < 3; j++) {
            for (int k = 0; k < 3; k++) Ns[j][k] = N[3 * j + k];
            const double l = length(Ns[j]);
            if (l == 0.0)
                degen = true;
            else
                for (int k = 0; k < 3; k++) Ns[j][k] /= l;
        }
        if (degen) { // a zero normal: no triangle is computed
            t.degenerate = 1;
            store();
            return t;
        }
    }
    double V1[3], V2[3], Nv[3];
    for (int k = 0; k < 3; k++) {
        V1[k] = P1[k] - P2[k];
        V2[k] = P3[k] - P2[k];
    }
    Nv[0] = (V1[1] * V2[2]) - (V1[2] * V2[1]);
    Nv[1] = (V1[2] * V2[0]) - (V1[0] * V2[2]);
    Nv[2] = (V1[0] * V2[1]) - (V1[1] * V2[0]);
    const double len = length(Nv);
    if (len == 0.0) { // DEGENERATE_FLAG
        t.degenerate = 1;
        store();
        return t;
    }
    for (int k = 0; k < 3; k++) Nv[k] /= len;
    const double dist = -dot(Nv, P1);
    const double ax = std::fabs(Nv[0]), ay = std::fabs(Nv[1]), az = std::fabs(Nv[2]);
    const int axis = (ax > ay) ? ((ax > az) ? 0 : 2) : ((ay > az) ? 1 : 2);
    bool swap = false;
    switch (axis) {
    case 0: swap = (P2[1] - P3[1]) * (P2[2] - P1[2]) < (P2[2] - P3[2]) * (P2[1] - P1[1]); break;
    case 1: swap = (P2[0] - P3[0]) * (P2[2] - P1[2]) < (P2[2] - P3[2]) * (P2[0] - P1[0]); break;
    default: swap = (P2[0] - P3[0]) * (P2[1] - P1[1]) < (P2[1] - P3[1]) * (P2[0] - P1[0]); break;
    }
    if (swap) {
        for (int k = 0; k < 3; k++) {
            std::swap(P1[k], P2[k]);
            std::swap(N1[k], N2[k]);
        }
    }
    for (int k = 0; k < 3; k++) t.n[k] = Nv[k];
    t.dist = dist;
    t.axis = axis;
    if (smooth) { // Compute_Smooth_Triangle
        double d32[3], v1[3], v2[3], perp[3];
        for (int k = 0; k < 3; k++) d32[k] = P3[k] - P2[k];
        const double x = std::fabs(d32[0]), y = std::fabs(d32[1]), z = std::fabs(d32[2]);
        t.vaxis = (x > y) ? ((x > z) ? 0 : 2) : ((y > z) ? 1 : 2);
        for (int k = 0; k < 3; k++) v1[k] = P2[k] - P3[k];
        const double l1 = length(v1);
        if (l1 != 0)
            for (int k = 0; k < 3; k++) v1[k] = v1[k] / l1;
        for (int k = 0; k < 3; k++) v2[k] = P1[k] - P3[k];
        const double proj = dot(v2, v1);
        for (int k = 0; k < 3; k++) v1[k] *= proj;
        for (int k = 0; k < 3; k++) perp[k] = v1[k] - v2[k];
        const double lp = length(perp);
        if (lp != 0)
            for (int k = 0; k < 3; k++) perp[k] = perp[k] / lp;
        const double uden = dot(v2, perp);
        for (int k = 0; k < 3; k++) perp[k] /= -uden;
        for (int k = 0; k < 3; k++) t.perp[k] = perp[k];
        // smooth normals more than 90 degrees from the face normal (or its
        // inverse) on some vertex: DEGENERATE_FLAG
        const double dx = dot(Nv, N1), dy = dot(Nv, N2), dz = dot(Nv, N3);
        if (!(((dx < 0.0) && (dy < 0.0) && (dz < 0.0)) || ((dx > 0.0) && (dy > 0.0) && (dz > 0.0))))
            t.degenerate = 1;
    }
    store();
    return t;
}

// Solve_Polynomial for a quadratic with epsilon 0 (polynomialsolver.cpp:
// 1585-1625 and solve_quadratic 810-860), c descending; the extremes of a
// spline segment are the roots of its derivative.
int quadratic_roots(const double *c0, double *r) {
    int i = 0, n = 2;
    while (i < n && std::fabs(c0[i]) < 1.0e-10) i++;
    n -= i;
    const double *c = c0 + i;
    if (n == 0) return 0;
    if (n == 1) {
        if (c[0] == 0.0) return 0;
        r[0] = -c[1] / c[0];
        return 1;
    }
    double a = c[0], b = -c[1], cc = c[2];
    if (a == 0.0) {
        if (b == 0.0) return 0;
        r[0] = cc / b;
        return 1;
    }
    b /= a; // monic form
    cc /= a;
    a = 1.0;
    double d = b * b - 4.0 * a * cc;
    if (d > -1.0e-10 && d < 1.0e-10) { // a double root
        r[0] = 0.5 * b / a;
        return 1;
    }
    if (d < 0.0) return 0;
    d = std::sqrt(d);
    const double t = 2.0 * a;
    r[0] = (b + d) / t;
    r[1] = (b - d) / t;
    return 2;
}

// Inverse of a 4x4 matrix (MInvers, matrix.cpp:1154-1202): the adjugate
// over the determinant, each cofactor the rule-of-Sarrus determinant of its
// minor with the terms in the reference's order (the sor coefficients then
// round as the reference's do); false when singular (the reference throws).
bool invert4(const double (&m)[4][4], double (&r)[4][4]) {
    double cof[4][4];
    for (int row = 0; row < 4; row++)
        for (int col = 0; col < 4; col++) {
            double a[3][3];
            for (int i = 0, ii = 0; i < 4; i++) {
                if (i == row) continue;
                for (int j = 0, jj = 0; j < 4; j++)
                    if (j != col) a[ii][jj++] = m[i][j];
                ii++;
            }
            cof[row][col] = a[0][0] * a[1][1] * a[2][2] + a[0][1] * a[1][2] * a[2][0] + a[0][2] * a[1][0] * a[2][1] -
                            a[2][0] * a[1][1] * a[0][2] - a[2][1] * a[1][2] * a[0][0] - a[2][2] * a[1][0] * a[0][1];
        }
    const double det = m[0][0] * cof[0][0] - m[0][1] * cof[0][1] + m[0][2] * cof[0][2] - m[0][3] * cof[0][3];
    if (det == 0.0) return false;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) r[i][j] = (((i + j) & 1) ? -cof[j][i] : cof[j][i]) / det;
    return true;
}

// A lathe or sor as Parse_Lathe / Parse_Sor leave it (parser/parser.cpp:
// 2812-2995, 5628-5714): the segments of Compute_Lathe (lathe.cpp:488-760)
// or Compute_Sor (sor.cpp:920-1180) with each segment's bounding cylinder,
// then Create_BCyl's lists of distinct radii (squared) and heights
// (boundingcylinder.cpp:408-560) appended to the scene-wide lists.  Returns
// an error message (empty on success).
std::string make_revolution(const double *pts, int n, uint32_t kind, uint32_t flags, RevRec &rec,
                            std::vector<RevSeg> &segs, std::vector<double> &rad_all,
                            std::vector<double> &hgt_all) {
    auto X = [&](int i) { return pts[2 * i]; };
    auto Y = [&](int i) { return pts[2 * i + 1]; };
    int nseg = 0;
    switch (kind) {
    case PVX_LATHE_LINEAR: nseg = n - 1; break;
    case PVX_LATHE_QUADRATIC: nseg = n - 2; break;
    case PVX_LATHE_CUBIC: nseg = n - 3; break;
    case PVX_LATHE_BEZIER: nseg = (n & 3) ? 0 : n / 4; break;
    case PVX_SOR_CUBIC: nseg = n - 3; break;
    default: return "unknown lathe or sor spline kind " + std::to_string(kind);
    }
    if (nseg < 1) return "too few points for the lathe or sor spline kind";
    if (nseg > kRevMaxSegments) return "more than " + std::to_string(kRevMaxSegments) + " segments in a lathe or sor";
    for (int i = 0; i < n; i++) { // the parser's point checks (v3.8 rules)
        const double x = X(i);
        bool bad;
        switch (kind) {
        case PVX_LATHE_LINEAR: bad = x < 0.0; break;
        case PVX_LATHE_QUADRATIC: bad = i > 0 && x < 0.0; break;
        case PVX_LATHE_CUBIC: bad = i > 0 && i < n - 1 && x < 0.0; break;
        case PVX_LATHE_BEZIER: bad = (i % 4 == 0 || i % 4 == 3) && x < 0.0; break;
        default: bad = x < 0.0 || (i > 1 && i < n - 1 && Y(i) <= Y(i - 1)); break;
        }
        if (bad) return kind == PVX_SOR_CUBIC ? "incorrect point in surface of revolution"
                                              : "lathe point with an x value < 0";
    }
    std::vector<double> r1(nseg), r2(nseg), h1(nseg), h2(nseg);
    const size_t seg0 = segs.size();
    segs.resize(seg0 + nseg);
    for (int g = 0; g < nseg; g++) {
        RevSeg &sg = segs[seg0 + g];
        std::memset(&sg, 0, sizeof sg);
        double xs[4], ys[4];
        if (kind == PVX_SOR_CUBIC) {
            const int i = g;
            if (std::fabs(Y(i + 2) - Y(i)) < kEpsilon || std::fabs(Y(i + 3) - Y(i + 1)) < kEpsilon)
                return "incorrect point in surface of revolution";
            // Hermite cubic in the height through r^2 at points i+1, i+2 with
            // the slopes 2 r dr/dh of the neighbouring chords
            const double k[4] = {X(i + 1) * X(i + 1), X(i + 2) * X(i + 2),
                                 (X(i + 2) - X(i)) / (Y(i + 2) - Y(i)) * (2.0 * X(i + 1)),
                                 (X(i + 3) - X(i + 1)) / (Y(i + 3) - Y(i + 1)) * (2.0 * X(i + 2))};
            const double w1 = Y(i + 1), w2 = Y(i + 2);
            const double M[4][4] = {{w1 * w1 * w1, w1 * w1, w1, 1.0},
                                    {w2 * w2 * w2, w2 * w2, w2, 1.0},
                                    {3.0 * w1 * w1, 2.0 * w1, 1.0, 0.0},
                                    {3.0 * w2 * w2, 2.0 * w2, 1.0, 0.0}};
            double Mi[4][4];
            if (!invert4(M, Mi)) return "singular matrix in a surface of revolution";
            double c[4];
            for (int r = 0; r < 4; r++) {
                c[r] = k[0] * Mi[r][0] + k[1] * Mi[r][1] + k[2] * Mi[r][2] + k[3] * Mi[r][3];
                if (std::fabs(c[r]) < kEpsilon) c[r] = 0.0;
            }
            sg.a[0] = c[0], sg.b[0] = c[1], sg.c[0] = c[2], sg.d[0] = c[3];
            xs[0] = xs[2] = X(i + 1);
            xs[1] = xs[3] = X(i + 2);
            const double dq[3] = {3.0 * c[0], 2.0 * c[1], c[2]};
            double rr[2];
            int nr = quadratic_roots(dq, rr);
            while (nr--)
                if (rr[nr] >= w1 && rr[nr] <= w2)
                    xs[nr] = std::sqrt(rr[nr] * (rr[nr] * (rr[nr] * c[0] + c[1]) + c[2]) + c[3]);
            r1[g] = std::min(std::min(xs[0], xs[1]), std::min(xs[2], xs[3]));
            r2[g] = std::max(std::max(xs[0], xs[1]), std::max(xs[2], xs[3]));
            h1[g] = w1;
            h2[g] = w2;
            continue;
        }
        const int i = kind == PVX_LATHE_BEZIER ? 4 * g : g;
        for (int ax = 0; ax < 2; ax++) {
            auto P = [&](int j) { return ax == 0 ? X(j) : Y(j); };
            double A = 0.0, B = 0.0, C, D;
            if (kind == PVX_LATHE_LINEAR) {
                C = -1.0 * P(i) + 1.0 * P(i + 1);
                D = 1.0 * P(i);
            } else if (kind == PVX_LATHE_QUADRATIC) {
                B = 0.5 * P(i) - 1.0 * P(i + 1) + 0.5 * P(i + 2);
                C = -0.5 * P(i) + 0.5 * P(i + 2);
                D = 1.0 * P(i + 1);
            } else if (kind == PVX_LATHE_CUBIC) { // Catmull-Rom through points i+1, i+2
                A = -0.5 * P(i) + 1.5 * P(i + 1) - 1.5 * P(i + 2) + 0.5 * P(i + 3);
                B = P(i) - 2.5 * P(i + 1) + 2.0 * P(i + 2) - 0.5 * P(i + 3);
                C = -0.5 * P(i) + 0.5 * P(i + 2);
                D = P(i + 1);
            } else { // Bezier, Bernstein form
                A = P(i + 3) - 3.0 * P(i + 2) + 3.0 * P(i + 1) - P(i);
                B = 3.0 * P(i + 2) - 6.0 * P(i + 1) + 3.0 * P(i);
                C = 3.0 * P(i + 1) - 3.0 * P(i);
                D = P(i);
            }
            sg.a[ax] = A, sg.b[ax] = B, sg.c[ax] = C, sg.d[ax] = D;
            double *v = ax == 0 ? xs : ys;
            if (kind == PVX_LATHE_BEZIER) {
                for (int j = 0; j < 4; j++) v[j] = P(i + j);
            } else {
                const int j0 = kind == PVX_LATHE_LINEAR ? i : i + 1;
                v[0] = v[2] = P(j0);
                v[1] = v[3] = P(j0 + 1);
            }
            if (kind == PVX_LATHE_QUADRATIC || kind == PVX_LATHE_CUBIC) { // extremes inside [0, 1]
                const double dq[3] = {3.0 * A, 2.0 * B, C};
                double rr[2];
                int nr = quadratic_roots(dq, rr);
                while (nr--)
                    if (rr[nr] >= 0.0 && rr[nr] <= 1.0) v[nr] = rr[nr] * (rr[nr] * (rr[nr] * A + B) + C) + D;
            }
        }
        r1[g] = std::min(std::min(xs[0], xs[1]), std::min(xs[2], xs[3]));
        r2[g] = std::max(std::max(xs[0], xs[1]), std::max(xs[2], xs[3]));
        h1[g] = std::min(std::min(ys[0], ys[1]), std::min(ys[2], ys[3]));
        h2[g] = std::max(std::max(ys[0], ys[1]), std::max(ys[2], ys[3]));
    }
    // the overall cylinder, and the distinct radii / heights in first-use order
    double xmin = kBoundHuge, xmax = -kBoundHuge, ymin = kBoundHuge, ymax = -kBoundHuge;
    std::vector<double> rad, hgt;
    auto index_of = [](std::vector<double> &list, double v) {
        size_t j = 0;
        while (j < list.size() && list[j] != v) j++;
        if (j == list.size()) list.push_back(v);
        return (uint32_t)j;
    };
    const uint32_t rbase = (uint32_t)rad_all.size(), hbase = (uint32_t)hgt_all.size();
    for (int g = 0; g < nseg; g++) {
        xmin = std::min(xmin, r1[g]);
        xmax = std::max(xmax, r2[g]);
        ymin = std::min(ymin, h1[g]);
        ymax = std::max(ymax, h2[g]);
        RevSeg &sg = segs[seg0 + g];
        sg.r1 = rbase + index_of(rad, r1[g]);
        sg.r2 = rbase + index_of(rad, r2[g]);
        sg.h1 = hbase + index_of(hgt, h1[g]);
        sg.h2 = hbase + index_of(hgt, h2[g]);
    }
    for (double v : rad) rad_all.push_back(v * v);
    hgt_all.insert(hgt_all.end(), hgt.begin(), hgt.end());
    rec.radius1 = xmin;
    rec.radius2 = xmax;
    rec.height1 = ymin;
    rec.height2 = ymax;
    rec.base_r2 = rec.cap_r2 = 0.0;
    if (kind == PVX_SOR_CUBIC) { // cap and base radii squared, clamped at 0
        const RevSeg &top = segs[seg0 + nseg - 1], &bot = segs[seg0];
        const double wt = h2[nseg - 1], wb = h1[0];
        rec.cap_r2 = std::max(0.0, wt * (wt * (top.a[0] * wt + top.b[0]) + top.c[0]) + top.d[0]);
        rec.base_r2 = std::max(0.0, wb * (wb * (bot.a[0] * wb + bot.b[0]) + bot.c[0]) + bot.d[0]);
    }
    rec.seg0 = (uint32_t)seg0;
    rec.nseg = (uint32_t)nseg;
    rec.flags = kind | ((kind == PVX_SOR_CUBIC && !(flags & PVX_REV_OPEN)) ? kRevClosed : 0u) |
                ((flags & PVX_REV_STURM) ? kRevSturm : 0u);
    return std::string();
}

} // namespace

int prepare_scene(const pvx_scene_desc &d, const pvx_build_opts *opts, HostScene &out,
                  std::string &err) {
    auto t0 = std::chrono::steady_clock::now();
    if (d.n_objects > 0 && !d.objects) { // an empty scene is valid: every ray misses
        err = "null object table";
        return PVX_EINVAL;
    }
    // --- validate arrays -----------------------------------------------------
    auto need = [&](bool ok, const char *what) {
        if (!ok) err = std::string("missing or invalid array: ") + what;
        return ok;
    };
    if (!need(d.n_spheres == 0 || (d.sphere_center && d.sphere_radius), "spheres")) return PVX_EINVAL;
    if (!need(d.n_planes == 0 || (d.plane_normal && d.plane_distance), "planes")) return PVX_EINVAL;
    if (!need(d.n_boxes == 0 || (d.box_corner1 && d.box_corner2), "boxes")) return PVX_EINVAL;
    if (!need(d.n_tori == 0 || (d.torus_major && d.torus_minor), "tori")) return PVX_EINVAL;
    if (!need(d.n_transforms == 0 || d.transforms, "transforms")) return PVX_EINVAL;
    if (!need(d.n_meshes == 0 || d.meshes, "meshes")) return PVX_EINVAL;
    if (!need(d.n_csg == 0 || (d.csg && d.csg_children), "csg")) return PVX_EINVAL;
    if (!need(d.n_triangles == 0 || d.triangle_vertices, "triangles")) return PVX_EINVAL;
    if (!need(d.n_cones == 0 || (d.cone_apex && d.cone_base && d.cone_radius), "cones")) return PVX_EINVAL;
    if (!need(d.n_discs == 0 || (d.disc_center && d.disc_normal && d.disc_radius), "discs")) return PVX_EINVAL;
    if (!need(d.n_quadrics == 0 || d.quadric_coef, "quadrics")) return PVX_EINVAL;
    if (!need(d.n_polygons == 0 || (d.polygon_first && d.polygon_points), "polygons")) return PVX_EINVAL;
    if (!need(d.n_revolutions == 0 || (d.revolution_first && d.revolution_points && d.revolution_kind),
              "lathes and sors"))
        return PVX_EINVAL;

    bool uses_rev = false; // a lathe or sor is referred to (the level-3 walks)
    auto check_ref = [&](uint32_t type, uint32_t index) -> bool {
        uint32_t lim = 0;
        switch (type) {
        case PVX_SPHERE: lim = d.n_spheres; break;
        case PVX_BOX: lim = d.n_boxes; break;
        case PVX_TORUS: lim = d.n_tori; break;
        case PVX_PLANE: lim = d.n_planes; break;
        case PVX_MESH: lim = d.n_meshes; break;
        case PVX_CSG: lim = d.n_csg; break;
        case PVX_TRIANGLE:
        case PVX_SMOOTH_TRIANGLE: lim = d.n_triangles; break;
        case PVX_CONE: lim = d.n_cones; break;
        case PVX_DISC: lim = d.n_discs; break;
        case PVX_QUADRIC: lim = d.n_quadrics; break;
        case PVX_POLYGON: lim = d.n_polygons; break;
        case PVX_LATHE:
        case PVX_SOR:
            if (index < d.n_revolutions && (d.revolution_kind[index] == PVX_SOR_CUBIC) != (type == PVX_SOR)) {
                err = "a lathe or sor object refers to a shape of the other kind (index " + std::to_string(index) + ")";
                return false;
            }
            uses_rev = true;
            lim = d.n_revolutions;
            break;
        default:
            err = "unknown object type " + std::to_string(type);
            return false;
        }
        if (index >= lim) {
            err = "object index out of range (type " + std::to_string(type) + ", index " +
                  std::to_string(index) + ")";
            return false;
        }
        return true;
    };

    // --- shared shape arrays -------------------------------------------------
    out.xforms.resize(d.n_transforms);
    for (uint32_t i = 0; i < d.n_transforms; i++) {
        std::memcpy(out.xforms[i].m, d.transforms + 32 * (size_t)i, 16 * sizeof(double));
        std::memcpy(out.xforms[i].inv, d.transforms + 32 * (size_t)i + 16, 16 * sizeof(double));
    }
    int32_t identity = -1;
    auto get_identity = [&]() {
        if (identity < 0) {
            Xform x{};
            for (int k = 0; k < 4; k++) x.m[5 * k] = x.inv[5 * k] = 1.0;
            identity = (int32_t)out.xforms.size();
            out.xforms.push_back(x);
        }
        return identity;
    };
    out.sph.resize(4 * (size_t)d.n_spheres);
    for (uint32_t i = 0; i < d.n_spheres; i++) {
        for (int k = 0; k < 3; k++) out.sph[4 * i + k] = d.sphere_center[3 * i + k];
        out.sph[4 * i + 3] = d.sphere_radius[i];
    }
    out.planes.resize(d.n_planes);
    for (uint32_t i = 0; i < d.n_planes; i++) {
        for (int k = 0; k < 3; k++) out.planes[i].n[k] = d.plane_normal[3 * i + k];
        out.planes[i].d = d.plane_distance[i];
    }
    out.boxes.resize(d.n_boxes);
    for (uint32_t i = 0; i < d.n_boxes; i++) {
        BoxRec &b = out.boxes[i];
        for (int k = 0; k < 3; k++) {
            b.c1[k] = d.box_corner1[3 * i + k];
            b.c2[k] = d.box_corner2[3 * i + k];
        }
        b.xf = d.box_transform ? d.box_transform[i] : -1;
        if (b.xf >= (int32_t)d.n_transforms) {
            err = "box transform index out of range";
            return PVX_EINVAL;
        }
        b.flags = 0;
    }
    out.tori.resize(d.n_tori);
    for (uint32_t i = 0; i < d.n_tori; i++) {
        TorusRec &t = out.tori[i];
        t.R = d.torus_major[i];
        t.r = d.torus_minor[i];
        t.xf = d.torus_transform ? d.torus_transform[i] : -1;
        if (t.xf >= (int32_t)d.n_transforms) {
            err = "torus transform index out of range";
            return PVX_EINVAL;
        }
        if (t.xf < 0) t.xf = get_identity(); // the reference always gives a torus a TRANSFORM
        t.sturm = d.torus_sturm ? d.torus_sturm[i] : 0;
        t.flags = 0;
        t.pad = 0;
    }

    // cones / cylinders: the shape's own placement, then the object's transform
    out.cones.resize(d.n_cones);
    for (uint32_t i = 0; i < d.n_cones; i++) {
        Xform x;
        if (!make_cone(d.cone_apex + 3 * (size_t)i, d.cone_base + 3 * (size_t)i, d.cone_radius + 2 * (size_t)i,
                       d.cone_flags ? d.cone_flags[i] : 0, out.cones[i], x)) {
            err = "degenerate cone/cylinder " + std::to_string(i) + " (base point = apex point)";
            return PVX_EINVAL;
        }
        const int32_t t = d.cone_transform ? d.cone_transform[i] : -1;
        if (t >= (int32_t)d.n_transforms) {
            err = "cone transform index out of range";
            return PVX_EINVAL;
        }
        if (t >= 0) {
            Xform u;
            std::memcpy(u.m, d.transforms + 32 * (size_t)t, 16 * sizeof(double));
            std::memcpy(u.inv, d.transforms + 32 * (size_t)t + 16, 16 * sizeof(double));
            xf_compose(x, u);
        }
        out.cones[i].xf = (int32_t)out.xforms.size();
        out.xforms.push_back(x);
    }

    // discs: placement, then the object's transform (Disc::Transform carries
    // the normal with MTransNormal and normalises it)
    out.discs.resize(d.n_discs);
    for (uint32_t i = 0; i < d.n_discs; i++) {
        Xform x;
        DiscRec &c = out.discs[i];
        make_disc(d.disc_center + 3 * (size_t)i, d.disc_normal + 3 * (size_t)i, d.disc_radius + 2 * (size_t)i, c, x);
        const int32_t t = d.disc_transform ? d.disc_transform[i] : -1;
        if (t >= (int32_t)d.n_transforms) {
            err = "disc transform index out of range";
            return PVX_EINVAL;
        }
        if (t >= 0) {
            Xform u;
            std::memcpy(u.m, d.transforms + 32 * (size_t)t, 16 * sizeof(double));
            std::memcpy(u.inv, d.transforms + 32 * (size_t)t + 16, 16 * sizeof(double));
            double nn[3];
            xf_normal(u, c.n, nn);
            const double l = std::sqrt(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]);
            for (int k = 0; k < 3; k++) c.n[k] = l != 0 ? nn[k] / l : nn[k];
            xf_compose(x, u);
        }
        c.xf = (int32_t)out.xforms.size();
        out.xforms.push_back(x);
    }

    // quadrics: Quadric::Transform (quadric.cpp) on the coefficient matrix
    // (Quadric_To_Matrix; M = inverse * M (MTimesB); M = M * inverse^T
    // (MTimesA); Matrix_To_Quadric)
    out.quadrics.resize(d.n_quadrics);
    for (uint32_t i = 0; i < d.n_quadrics; i++) {
        const double *c = d.quadric_coef + 10 * (size_t)i;
        QuadricRec &q = out.quadrics[i];
        q = QuadricRec{c[0], c[1], c[2], c[3], c[4], c[5], c[6], c[7], c[8], c[9]};
        const int32_t t = d.quadric_transform ? d.quadric_transform[i] : -1;
        if (t >= (int32_t)d.n_transforms) {
            err = "quadric transform index out of range";
            return PVX_EINVAL;
        }
        if (t < 0) continue;
        const double *inv = d.transforms + 32 * (size_t)t + 16;
        double M[16] = {q.qa, q.qb, q.qc, q.qd, 0.0, q.qe, q.qf, q.qg, 0.0, 0.0, q.qh, q.qi, 0.0, 0.0, 0.0, q.qj};
        double R[16];
        for (int j = 0; j < 4; j++) // MTimesB(inverse, M): column j of inverse * M
            for (int r = 0; r < 4; r++)
                R[4 * r + j] = inv[4 * r] * M[j] + inv[4 * r + 1] * M[4 + j] + inv[4 * r + 2] * M[8 + j] +
                               inv[4 * r + 3] * M[12 + j];
        for (int r = 0; r < 4; r++) // MTimesA(R, inverse^T): row r of R * inverse^T
            for (int j = 0; j < 4; j++)
                M[4 * r + j] = R[4 * r] * inv[4 * j] + R[4 * r + 1] * inv[4 * j + 1] + R[4 * r + 2] * inv[4 * j + 2] +
                               R[4 * r + 3] * inv[4 * j + 3];
        q = QuadricRec{M[0], M[5], M[10], M[1] + M[4], M[2] + M[8], M[6] + M[9], M[3] + M[12], M[7] + M[13],
                       M[11] + M[14], M[15]};
    }

    std::vector<DBox> poly_boxes;
    // polygons: Parse_Polygon's closing of the point list, then Compute_Polygon
    // (polygon.cpp:646-805) and Polygon::Transform
    out.polygons.resize(d.n_polygons);
    for (uint32_t i = 0; i < d.n_polygons; i++) {
        const uint32_t b = d.polygon_first[i], e = d.polygon_first[i + 1];
        if (e < b + 3) {
            err = "polygon " + std::to_string(i) + " needs at least three points";
            return PVX_EINVAL;
        }
        std::vector<double> P(d.polygon_points + 3 * (size_t)b, d.polygon_points + 3 * (size_t)e);
        int number = (int)(e - b);
        bool closed = false;
        double Q[3] = {P[0], P[1], P[2]};
        for (int k = 1; k < number; k++) {
            closed = false;
            if (std::fabs(Q[0] - P[3 * k]) < kEpsilon && std::fabs(Q[1] - P[3 * k + 1]) < kEpsilon &&
                std::fabs(Q[2] - P[3 * k + 2]) < kEpsilon) {
                for (int a = 0; a < 3; a++) P[3 * k + a] = Q[a];
                k++;
                if (k < number)
                    for (int a = 0; a < 3; a++) Q[a] = P[3 * k + a];
                closed = true;
            }
        }
        if (!closed) {
            P.insert(P.end(), Q, Q + 3);
            number++;
        }
        PolygonRec &pg = out.polygons[i];
        pg.degenerate = 0;
        pg.pad = 0;
        auto sub = [&](int k, double *r) {
            for (int a = 0; a < 3; a++) r[a] = P[3 * k + a] - P[a];
        };
        auto cross = [](const double *x, const double *y, double *r) {
            r[0] = x[1] * y[2] - x[2] * y[1];
            r[1] = x[2] * y[0] - x[0] * y[2];
            r[2] = x[0] * y[1] - x[1] * y[0];
        };
        auto len2 = [](const double *x) { return x[0] * x[0] + x[1] * x[1] + x[2] * x[2]; };
        double u[3] = {0, 0, 0}, v[3] = {0, 0, 0}, w[3] = {0, 0, 0};
        int k = 1;
        for (; k < number; k++) {
            sub(k, u);
            if (len2(u) > kEpsilon) break;
        }
        if (k == number) pg.degenerate = 1;
        for (k++; k < number; k++) {
            sub(k, v);
            cross(u, v, w);
            if (len2(v) > kEpsilon && len2(w) > kEpsilon) break;
        }
        if (k == number) pg.degenerate = 1;
        double t[3];
        cross(v, w, t);
        std::memcpy(u, t, sizeof t);
        cross(w, u, t);
        std::memcpy(v, t, sizeof t);
        for (double *x : {u, v, w}) {
            const double l = std::sqrt(len2(x));
            if (l != 0)
                for (int a = 0; a < 3; a++) x[a] = x[a] / l;
        }
        // inverse = A * B (MTimesC): A translates by -o, B's columns are u, v, w
        Xform x;
        xf_identity(x);
        double A_[16], B_[16];
        for (int q = 0; q < 16; q++) A_[q] = B_[q] = (q % 5 == 0) ? 1.0 : 0.0;
        for (int a = 0; a < 3; a++) {
            A_[12 + a] = -P[a];
            B_[4 * a + 0] = u[a];
            B_[4 * a + 1] = v[a];
            B_[4 * a + 2] = w[a];
        }
        for (int r = 0; r < 4; r++)
            for (int c = 0; c < 4; c++)
                x.inv[4 * r + c] = A_[4 * r] * B_[c] + A_[4 * r + 1] * B_[4 + c] + A_[4 * r + 2] * B_[8 + c] +
                                   A_[4 * r + 3] * B_[12 + c];
        // (x.m is not used: hits need the inverse only; the box comes from the points)
        pg.first = (uint32_t)(out.poly_pts.size() / 2);
        pg.count = (uint32_t)number;
        for (int q = 0; q < number; q++) {
            const double px = P[3 * q] - P[0], py = P[3 * q + 1] - P[1], pz = P[3 * q + 2] - P[2];
            const double dd = px * w[0] + py * w[1] + pz * w[2];
            if (std::fabs(dd) > 1.0e-10) pg.degenerate = 1; // not co-planar (ZERO_TOLERANCE)
            out.poly_pts.push_back(px * u[0] + py * u[1] + pz * u[2]);
            out.poly_pts.push_back(px * v[0] + py * v[1] + pz * v[2]);
        }
        const int32_t tix = d.polygon_transform ? d.polygon_transform[i] : -1;
        if (tix >= (int32_t)d.n_transforms) {
            err = "polygon transform index out of range";
            return PVX_EINVAL;
        }
        // bounds: the points (then the transform's image of their box), +-SMALL_TOLERANCE
        DBox bb = DBox::empty();
        for (int q = 0; q < number; q++) bb.grow(&P[3 * q]);
        for (int a = 0; a < 3; a++) {
            bb.lo[a] -= kSmallTolerance;
            bb.hi[a] += kSmallTolerance;
        }
        if (tix >= 0) {
            Xform u2;
            std::memcpy(u2.m, d.transforms + 32 * (size_t)tix, 16 * sizeof(double));
            std::memcpy(u2.inv, d.transforms + 32 * (size_t)tix + 16, 16 * sizeof(double));
            xf_compose(x, u2); // only its inverse is kept
            bb = transformed_box(bb.lo, bb.hi, u2.m);
        }
        // S_Normal = MTransNormal(<0,0,1>, Trans): column 2 of the inverse rows
        double nn[3] = {x.inv[2], x.inv[6], x.inv[10]};
        const double nl = std::sqrt(len2(nn));
        for (int a = 0; a < 3; a++) pg.n[a] = nl != 0 ? nn[a] / nl : nn[a];
        pg.xf = (int32_t)out.xforms.size();
        out.xforms.push_back(x);
        poly_boxes.push_back(bb);
    }

    // stand-alone triangles, flat or smooth as their objects say
    {
        std::vector<uint8_t> kind(d.n_triangles, 0);
        for (uint32_t o = 0; o < d.n_objects; o++) {
            const pvx_object &ob = d.objects[o];
            if (ob.type != PVX_TRIANGLE && ob.type != PVX_SMOOTH_TRIANGLE) continue;
            if (ob.index >= d.n_triangles) continue; // reported by check_ref below
            const uint8_t k = ob.type == PVX_SMOOTH_TRIANGLE ? 2 : 1;
            if (kind[ob.index] && kind[ob.index] != k) {
                err = "triangle " + std::to_string(ob.index) + " listed as both triangle and smooth_triangle";
                return PVX_EINVAL;
            }
            if (k == 2 && !d.triangle_normals) {
                err = "smooth triangles need triangle_normals";
                return PVX_EINVAL;
            }
            kind[ob.index] = k;
        }
        out.triangles.resize(d.n_triangles);
        for (uint32_t i = 0; i < d.n_triangles; i++)
            out.triangles[i] = make_triangle(d.triangle_vertices + 9 * (size_t)i,
                                             kind[i] == 2 ? d.triangle_normals + 9 * (size_t)i : nullptr,
                                             kind[i] == 2);
    }

    // lathes and sors: segments, bounding cylinders, placement (always a
    // TRANSFORM, identity when none is given)
    out.revs.resize(d.n_revolutions);
    for (uint32_t i = 0; i < d.n_revolutions; i++) {
        const uint32_t b = d.revolution_first[i], e = d.revolution_first[i + 1];
        if (e < b) {
            err = "lathe / sor point offsets decrease at " + std::to_string(i);
            return PVX_EINVAL;
        }
        RevRec &rec = out.revs[i];
        const std::string msg =
            make_revolution(d.revolution_points + 2 * (size_t)b, (int)(e - b), d.revolution_kind[i],
                            d.revolution_flags ? d.revolution_flags[i] : 0u, rec, out.rev_segs, out.rev_radius,
                            out.rev_height);
        if (!msg.empty()) {
            err = msg + " (lathe / sor " + std::to_string(i) + ")";
            return PVX_EINVAL;
        }
        const int32_t t = d.revolution_transform ? d.revolution_transform[i] : -1;
        if (t < -1 || t >= (int32_t)d.n_transforms) {
            err = "lathe / sor transform index out of range";
            return PVX_EINVAL;
        }
        rec.xf = t >= 0 ? t : get_identity();
    }

    // --- CSG tables ------------------------------------------------------------
    out.csg.resize(d.n_csg);
    out.csg_children.resize(d.n_csg_children);
    std::vector<int32_t> owner_slot(d.n_csg, -1);
    for (uint32_t i = 0; i < d.n_csg; i++) {
        const pvx_csg_node &n = d.csg[i];
        if (n.op < PVX_CSG_UNION || n.op > PVX_CSG_MERGE) {
            err = "bad CSG operator";
            return PVX_EINVAL;
        }
        if ((uint64_t)n.first_child + n.n_children > d.n_csg_children) {
            err = "CSG child range out of bounds";
            return PVX_EINVAL;
        }
        out.csg[i] = CsgNode{n.op, n.first_child, n.n_children, -1};
        for (uint32_t k = 0; k < n.n_children; k++) {
            uint32_t slot = n.first_child + k;
            const pvx_object &c = d.csg_children[slot];
            if (!check_ref(c.type, c.index)) return PVX_EINVAL;
            if (c.type == PVX_MESH || c.type == PVX_TRIANGLE || c.type == PVX_SMOOTH_TRIANGLE ||
                c.type == PVX_POLYGON) {
                err = "meshes, triangles and polygons inside CSG are not supported";
                return PVX_EUNSUPPORTED;
            }
            out.csg_children[slot] = CsgChild{c.type, c.index, c.flags, i};
            if (c.type == PVX_CSG) {
                if (owner_slot[c.index] >= 0) {
                    err = "CSG node shared by two parents";
                    return PVX_EUNSUPPORTED;
                }
                owner_slot[c.index] = (int32_t)slot;
            }
        }
    }
    for (uint32_t i = 0; i < d.n_csg; i++) out.csg[i].parent_slot = owner_slot[i];
    // depth check + leaf lists (tree order) per node
    out.csg_leaf_range.assign(2 * (size_t)d.n_csg, 0);
    std::function<int(uint32_t, int)> depth_of = [&](uint32_t node, int depth) -> int {
        if (depth > 8) return depth;
        int m = depth;
        const CsgNode &n = out.csg[node];
        for (uint32_t k = 0; k < n.count; k++) {
            const CsgChild &c = out.csg_children[n.first + k];
            if (c.type == PVX_CSG) m = std::max(m, depth_of(c.index, depth + 1));
        }
        return m;
    };
    std::function<void(uint32_t)> collect = [&](uint32_t node) {
        const CsgNode &n = out.csg[node];
        for (uint32_t k = 0; k < n.count; k++) {
            const CsgChild &c = out.csg_children[n.first + k];
            if (c.type == PVX_CSG)
                collect(c.index);
            else
                out.csg_leaves.push_back(n.first + k);
        }
    };
    for (uint32_t i = 0; i < d.n_csg; i++) {
        if (owner_slot[i] >= 0) continue; // not a root
        if (depth_of(i, 1) > 8) {
            err = "CSG nesting deeper than 8 levels";
            return PVX_EUNSUPPORTED;
        }
        out.csg_leaf_range[2 * i] = (uint32_t)out.csg_leaves.size();
        collect(i);
        out.csg_leaf_range[2 * i + 1] = (uint32_t)out.csg_leaves.size();
    }

    // --- bounds of shapes (FP64, conservative) ------------------------------
    // Two bounds per shape: S bounds its surfaces (where hits can lie), V
    // bounds the region where Inside can be true.  A CSG hit lies on a child
    // surface and inside every other child of an intersection, so
    //   union/merge:  S = U S(c),            V = U V(c)
    //   intersection: S = (U S(c)) n V,      V = n V(c)
    // Inversion changes V (unbounded) but not S.  Planes are unbounded.
    struct Bnd {
        DBox S, V;
    };
    auto leaf_box = [&](uint32_t type, uint32_t index) -> DBox {
        switch (type) {
        case PVX_SPHERE: {
            DBox b;
            b.bounded = true;
            for (int k = 0; k < 3; k++) {
                b.lo[k] = out.sph[4 * index + k] - out.sph[4 * index + 3];
                b.hi[k] = out.sph[4 * index + k] + out.sph[4 * index + 3];
            }
            return b;
        }
        case PVX_BOX: {
            const BoxRec &bx = out.boxes[index];
            return transformed_box(bx.c1, bx.c2, bx.xf >= 0 ? out.xforms[bx.xf].m : nullptr);
        }
        case PVX_TORUS: {
            const TorusRec &t = out.tori[index];
            double a = t.R + t.r;
            double lo[3] = {-a, -t.r, -a}, hi[3] = {a, t.r, a};
            return transformed_box(lo, hi, out.xforms[t.xf].m);
        }
        case PVX_CONE: { // Cone::Compute_BBox (cone.cpp:870-880)
            const ConeRec &c = out.cones[index];
            double lo[3] = {-1.0, -1.0, c.dist}, hi[3] = {1.0, 1.0, 1.0};
            return transformed_box(lo, hi, out.xforms[c.xf].m);
        }
        case PVX_LATHE:   // Lathe::Compute_BBox (lathe.cpp:455-461)
        case PVX_SOR: {   // Sor::Compute_BBox (sor.cpp:705-711)
            const RevRec &r = out.revs[index];
            double lo[3] = {-r.radius2, r.height1, -r.radius2}, hi[3] = {r.radius2, r.height2, r.radius2};
            return transformed_box(lo, hi, out.xforms[r.xf].m);
        }
        case PVX_DISC: { // Disc::Compute_BBox (disc.cpp): the thin box also bounds the
                         // disc as a CSG intersection operand, as in the reference
            const DiscRec &c = out.discs[index];
            const double rad = std::sqrt(c.or2);
            double lo[3] = {-rad, -rad, -kSmallTolerance}, hi[3] = {rad, rad, kSmallTolerance};
            return transformed_box(lo, hi, out.xforms[c.xf].m);
        }
        default:
            return DBox::unbounded();
        }
    };
    auto join = [](DBox a, const DBox &b) {
        if (!a.bounded || !b.bounded) return DBox::unbounded();
        for (int k = 0; k < 3; k++) {
            a.lo[k] = std::min(a.lo[k], b.lo[k]);
            a.hi[k] = std::max(a.hi[k], b.hi[k]);
        }
        return a;
    };
    auto meet = [](DBox a, const DBox &b) {
        if (!b.bounded) return a;
        if (!a.bounded) return b;
        for (int k = 0; k < 3; k++) {
            a.lo[k] = std::max(a.lo[k], b.lo[k]);
            a.hi[k] = std::min(a.hi[k], b.hi[k]);
        }
        return a;
    };
    std::function<Bnd(uint32_t, uint32_t, uint32_t)> shape_bnd = [&](uint32_t type, uint32_t index,
                                                                       uint32_t flags) -> Bnd {
        if (type != PVX_CSG) {
            DBox b = leaf_box(type, index);
            return Bnd{b, (flags & PVX_FLAG_INVERTED) ? DBox::unbounded() : b};
        }
        const CsgNode &n = out.csg[index];
        DBox S = DBox::empty(), V = (n.op == PVX_CSG_INTERSECTION) ? DBox::unbounded() : DBox::empty();
        for (uint32_t k = 0; k < n.count; k++) {
            const CsgChild &c = out.csg_children[n.first + k];
            Bnd cb = shape_bnd(c.type, c.index, c.flags);
            S = join(S, cb.S);
            V = (n.op == PVX_CSG_INTERSECTION) ? meet(V, cb.V) : join(V, cb.V);
        }
        if (n.op == PVX_CSG_INTERSECTION) S = meet(S, V);
        return Bnd{S, V};
    };

    // --- primitives ------------------------------------------------------------
    std::vector<PrimRec> prims;
    std::vector<float> plo, phi;
    auto add_prim = [&](const PrimRec &p, const float *lo, const float *hi) {
        prims.push_back(p);
        plo.insert(plo.end(), lo, lo + 3);
        phi.insert(phi.end(), hi, hi + 3);
    };
    out.obj_flags.resize(d.n_objects);
    out.obj_type.resize(d.n_objects);
    out.obj_tri_base.assign(d.n_objects, 0);
    out.obj_mesh_xf.assign(d.n_objects, -1);
    for (uint32_t o = 0; o < d.n_objects; o++) {
        const pvx_object &ob = d.objects[o];
        if (!check_ref(ob.type, ob.index)) return PVX_EINVAL;
        if (ob.type == PVX_CSG && owner_slot[ob.index] >= 0) {
            err = "a nested CSG node is listed as a top-level object";
            return PVX_EINVAL;
        }
        out.obj_flags[o] = ob.flags;
        out.obj_type[o] = ob.type;
        if (ob.flags & (PVX_FLAG_NO_SHADOW | PVX_FLAG_NO_IMAGE | PVX_FLAG_NO_REFLECTION |
                        PVX_FLAG_NO_RADIOSITY))
            out.any_flags = 1;
        PrimRec pr;
        std::memset(&pr, 0, sizeof pr);
        pr.obj = (int32_t)o;
        float lo[3], hi[3];
        switch (ob.type) {
        case PVX_SPHERE: {
            pr.kind_axis = kPrimSphere;
            for (int k = 0; k < 3; k++) pr.sph.c[k] = out.sph[4 * ob.index + k];
            pr.sph.r = out.sph[4 * ob.index + 3];
            to_fbox(shape_bnd(ob.type, ob.index, 0).S, lo, hi);
            add_prim(pr, lo, hi);
            break;
        }
        case PVX_BOX:
        case PVX_TORUS:
        case PVX_CONE:
        case PVX_DISC: {
            pr.kind_axis = ob.type == PVX_BOX     ? kPrimBox
                           : ob.type == PVX_TORUS ? kPrimTorus
                           : ob.type == PVX_CONE  ? kPrimCone
                                                  : kPrimDisc;
            pr.ref.index = ob.index;
            to_fbox(shape_bnd(ob.type, ob.index, 0).S, lo, hi);
            add_prim(pr, lo, hi);
            out.has_complex = std::max<uint32_t>(out.has_complex, ob.type == PVX_BOX ? 1u : 2u);
            break;
        }
        case PVX_CSG: {
            DBox b = shape_bnd(ob.type, ob.index, 0).S;
            if (!b.bounded || !(b.lo[0] <= b.hi[0] && b.lo[1] <= b.hi[1] && b.lo[2] <= b.hi[2])) {
                if (!b.bounded) {
                    out.inf.push_back(InfRec{kPrimCsg, ob.index, (int32_t)o, 0});
                    out.has_complex = 2;
                }
                // an empty surface bound: the object can never be hit
                break;
            }
            pr.kind_axis = kPrimCsg;
            pr.ref.index = ob.index;
            to_fbox(b, lo, hi);
            add_prim(pr, lo, hi);
            out.has_complex = 2;
            break;
        }
        case PVX_PLANE:
            out.inf.push_back(InfRec{kPrimPlane, ob.index, (int32_t)o, 0});
            break;
        case PVX_LATHE:
        case PVX_SOR:
            pr.kind_axis = kPrimRevolution;
            pr.ref.index = ob.index;
            to_fbox(shape_bnd(ob.type, ob.index, 0).S, lo, hi);
            add_prim(pr, lo, hi);
            break;
        case PVX_POLYGON: {
            if (out.polygons[ob.index].degenerate) break; // never hit (Polygon::Intersect)
            pr.kind_axis = kPrimPolygon;
            pr.ref.index = ob.index;
            to_fbox(poly_boxes[ob.index], lo, hi);
            add_prim(pr, lo, hi);
            out.has_complex = 2;
            break;
        }
        case PVX_QUADRIC: // unbounded here: Compute_BBox's box only culls (quadric.cpp:764-1000)
            out.inf.push_back(InfRec{kPrimQuadric, ob.index, (int32_t)o, 0});
            out.has_complex = 2;
            break;
        case PVX_TRIANGLE:
        case PVX_SMOOTH_TRIANGLE: {
            const TriangleRec &t = out.triangles[ob.index];
            out.has_complex = std::max<uint32_t>(out.has_complex, 1u);
            if (t.degenerate) break; // never hit
            DBox bb = DBox::empty(); // Triangle::Compute_BBox (triangle.cpp:1296-1311): vertices +- EPSILON
            const double *P = d.triangle_vertices + 9 * (size_t)ob.index;
            for (int k = 0; k < 3; k++) {
                bb.lo[k] = std::min(P[k], std::min(P[3 + k], P[6 + k])) - kEpsilon;
                bb.hi[k] = std::max(P[k], std::max(P[3 + k], P[6 + k])) + kEpsilon;
            }
            pr.kind_axis = kPrimTriangle64;
            pr.ref.index = ob.index;
            to_fbox(bb, lo, hi);
            add_prim(pr, lo, hi);
            break;
        }
        case PVX_MESH: {
            const pvx_mesh_desc &m = d.meshes[ob.index];
            if (!m.vertices || !m.indices) {
                err = "mesh without vertices or indices";
                return PVX_EINVAL;
            }
            const int32_t mxf = d.mesh_transform ? d.mesh_transform[ob.index] : -1;
            if (mxf < -1 || mxf >= (int32_t)d.n_transforms) {
                err = "mesh transform index out of range";
                return PVX_EINVAL;
            }
            out.obj_mesh_xf[o] = mxf;
            if (mxf >= 0) {
                // placed meshes are tested by the level-2 walk (the others keep
                // the lighter triangle walks)
                out.has_complex = std::max<uint32_t>(out.has_complex, 2u);
                if ((uint32_t)mxf + 1u >= (1u << 24)) {
                    err = "mesh transform index above the 24-bit record field";
                    return PVX_EUNSUPPORTED;
                }
            }
            // world bounds of a placed triangle: its vertices through the
            // mesh's TRANSFORM (MTransPoint, matrix.cpp:415-427)
            auto grow_world = [&](DBox &bb, const double *P) {
                if (mxf < 0) {
                    bb.grow(P);
                    return;
                }
                const double *M = out.xforms[mxf].m;
                double q[3];
                for (int k = 0; k < 3; k++) q[k] = P[0] * M[k] + P[1] * M[4 + k] + P[2] * M[8 + k] + M[12 + k];
                bb.grow(q);
            };
            // shading data: vertices, vertex normals then face normals
            // (Normals[] layout of Parse_Mesh2, parser.cpp:4512-4565)
            int32_t vbase = (int32_t)(out.mesh_vtx.size() / 3);
            int32_t nbase = (int32_t)(out.mesh_nrm.size() / 3);
            uint32_t nn = (m.normals && m.n_normals == m.n_vertices) ? m.n_normals : 0;
            out.mesh_vtx.insert(out.mesh_vtx.end(), m.vertices, m.vertices + 3 * (size_t)m.n_vertices);
            if (nn) out.mesh_nrm.insert(out.mesh_nrm.end(), m.normals, m.normals + 3 * (size_t)nn);
            out.mesh_nrm.resize(out.mesh_nrm.size() + 3 * (size_t)m.n_triangles, 0.0f);
            out.obj_tri_base[o] = (uint32_t)out.shade_tris.size();
            out.shade_tris.resize(out.shade_tris.size() + m.n_triangles);
            for (uint32_t t = 0; t < m.n_triangles; t++) {
                int32_t ia = m.indices[3 * t], ib = m.indices[3 * t + 1], ic = m.indices[3 * t + 2];
                if (ia < 0 || ib < 0 || ic < 0 || (uint32_t)ia >= m.n_vertices ||
                    (uint32_t)ib >= m.n_vertices || (uint32_t)ic >= m.n_vertices) {
                    err = "mesh face index out of range";
                    return PVX_EINVAL;
                }
                const float *fa = m.vertices + 3 * (size_t)ia, *fb = m.vertices + 3 * (size_t)ib,
                            *fc = m.vertices + 3 * (size_t)ic;
                double P1[3] = {fa[0], fa[1], fa[2]}, P2[3] = {fb[0], fb[1], fb[2]},
                       P3[3] = {fc[0], fc[1], fc[2]};
                double V1[3] = {P2[0] - P1[0], P2[1] - P1[1], P2[2] - P1[2]};
                double V2[3] = {P3[0] - P1[0], P3[1] - P1[1], P3[2] - P1[2]};
                // S_Normal = cross(V2, V1)
                double S[3] = {(V2[1] * V1[2]) - (V2[2] * V1[1]), (V2[2] * V1[0]) - (V2[0] * V1[2]),
                               (V2[0] * V1[1]) - (V2[1] * V1[0])};
                double L = std::sqrt(S[0] * S[0] + S[1] * S[1] + S[2] * S[2]);
                ShadeTri &st = out.shade_tris[out.obj_tri_base[o] + t];
                std::memset(&st, 0, sizeof st);
                st.fn = nbase + (int32_t)nn + (int32_t)t;
                if (L == 0.0) continue; // zero normal: never hit
                for (int k = 0; k < 3; k++) S[k] = S[k] / L;
                float dist = (float)((S[0] * P1[0]) + (S[1] * P1[1]) + (S[2] * P1[2]));
                dist = (float)(dist * -1.0);
                int axis = max3_axis(std::fabs(S[0]), std::fabs(S[1]), std::fabs(S[2]));
                int a = axis == 0 ? 1 : 0, b = axis == 2 ? 1 : 2;
                bool swap = (P2[a] - P3[a]) * (P2[b] - P1[b]) < (P2[b] - P3[b]) * (P2[a] - P1[a]);
                const float *q1 = swap ? fb : fa, *q2 = swap ? fa : fb;
                for (int k = 0; k < 3; k++) out.mesh_nrm[3 * (size_t)st.fn + k] = (float)S[k];
                st.p1 = vbase + (swap ? ib : ia);
                st.p2 = vbase + (swap ? ia : ib);
                st.p3 = vbase + ic;
                bool smooth = false;
                if (nn) {
                    const float *na = m.normals + 3 * (size_t)ia, *nb = m.normals + 3 * (size_t)ib,
                                *nc = m.normals + 3 * (size_t)ic;
                    double l1 = 0, l2 = 0;
                    for (int k = 0; k < 3; k++) {
                        double e1 = (double)na[k] - (double)nb[k], e2 = (double)na[k] - (double)nc[k];
                        l1 += e1 * e1;
                        l2 += e2 * e2;
                    }
                    smooth = std::fabs(l1) > 1.0e-10 || std::fabs(l2) > 1.0e-10;
                }
                st.n1 = nbase + (smooth && swap ? ib : ia);
                st.n2 = nbase + (smooth && swap ? ia : ib);
                st.n3 = nbase + ic;
                {
                    double Q1[3] = {q1[0], q1[1], q1[2]}, Q2[3] = {q2[0], q2[1], q2[2]};
                    smooth_triangle(Q1, Q2, P3, st);
                    if (smooth) st.vaxis_smooth |= 1u << 4;
                }
                pr.kind_axis = kPrimTriangle | (uint32_t)(axis << 4) | ((uint32_t)(mxf + 1) << 8);
                for (int k = 0; k < 3; k++) {
                    pr.tri.n[k] = (float)S[k];
                    pr.tri.p1[k] = q1[k];
                    pr.tri.p2[k] = q2[k];
                    pr.tri.p3[k] = fc[k];
                }
                pr.tri.dist = dist;
                pr.tri.prim = (int32_t)t;
                DBox bb = DBox::empty();
                grow_world(bb, P1);
                grow_world(bb, P2);
                grow_world(bb, P3);
                to_fbox(bb, lo, hi);
                add_prim(pr, lo, hi);
                out.n_triangles++;
            }
            break;
        }
        }
    }

    if (uses_rev) out.has_complex = 3; // lathes / sors: the level-3 walks
    // --- BVH ---------------------------------------------------------------------
    int leaf = opts && opts->max_leaf_size ? (int)opts->max_leaf_size : 4;
    int bins = opts && opts->sah_bins ? (int)opts->sah_bins : 16;
    if (prims.size() >= (1u << 28)) {
        err = "too many primitives for the 28-bit leaf encoding";
        return PVX_EUNSUPPORTED;
    }
    const uint32_t builder = opts ? opts->builder : (uint32_t)PVX_BUILDER_HOST_SAH;
    if (builder != PVX_BUILDER_HOST_SAH && builder != PVX_BUILDER_GPU) {
        err = "unknown BVH builder";
        return PVX_EINVAL;
    }
    if (builder == PVX_BUILDER_GPU) {
        if (opts->accel != PVX_ACCEL_BVH) {
            err = "the GPU builder builds the BVH only (PVX_ACCEL_BVH)";
            return PVX_EINVAL;
        }
        out.gpu_build = true;
        out.prims = std::move(prims);
        out.prim_lo = std::move(plo);
        out.prim_hi = std::move(phi);
        out.build_ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        return PVX_OK;
    }
    auto tb = std::chrono::steady_clock::now();
    BvhResult bvh = build_bvh(plo.data(), phi.data(), (uint32_t)prims.size(), leaf, bins);
    out.nodes = std::move(bvh.nodes);
    out.root = bvh.root;
    out.depth = bvh.depth;
    out.prims.resize(prims.size());
    for (size_t k = 0; k < prims.size(); k++) out.prims[k] = prims[bvh.order[k]];
    if (out.has_complex == 0) collapse_bvh4(out.nodes, out.root, out.nodes4, out.root4, out.depth4);
    out.bvh_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tb).count();
    if (opts && opts->accel == PVX_ACCEL_BSP) { // kd-tree over the same records, in their final order
        std::vector<float> klo(plo.size()), khi(phi.size());
        for (size_t k = 0; k < prims.size(); k++)
            for (int a = 0; a < 3; a++) {
                klo[3 * k + a] = plo[3 * (size_t)bvh.order[k] + a];
                khi[3 * k + a] = phi[3 * (size_t)bvh.order[k] + a];
            }
        const uint32_t md = opts->bsp_max_depth ? std::min<uint32_t>(opts->bsp_max_depth, kKdMaxDepth) : kKdMaxDepth;
        KdTree kd = build_kd(klo.data(), khi.data(), (uint32_t)prims.size(), md);
        out.kd_nodes = std::move(kd.nodes);
        out.kd_list = std::move(kd.list);
        for (int a = 0; a < 3; a++) {
            out.kd_bmin[a] = kd.bmin[a];
            out.kd_bmax[a] = kd.bmax[a];
        }
        out.kd_depth = kd.depth;
        out.use_kd = true;
    } else if (opts && opts->accel != PVX_ACCEL_BVH) {
        err = "unknown acceleration structure";
        return PVX_EINVAL;
    }
    out.build_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return PVX_OK;
}

} // namespace pvx
