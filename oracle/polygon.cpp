// Polygon (core/shape/polygon.cpp) — FP64 ORACLE, TEST INFRASTRUCTURE ONLY.
#include <algorithm>
#include <cstring>

#include "oracle.hpp"

namespace orc {

// MInvers (matrix.cpp:1154-1202): cofactor inverse of a 4x4 matrix; false
// for a singular one (the reference throws there).
static bool minvers(double r[4][4], const double m[4][4]) {
    const double m00 = m[0][0], m01 = m[0][1], m02 = m[0][2], m03 = m[0][3];
    const double m10 = m[1][0], m11 = m[1][1], m12 = m[1][2], m13 = m[1][3];
    const double m20 = m[2][0], m21 = m[2][1], m22 = m[2][2], m23 = m[2][3];
    const double m30 = m[3][0], m31 = m[3][1], m32 = m[3][2], m33 = m[3][3];
    const double d00 = m11 * m22 * m33 + m12 * m23 * m31 + m13 * m21 * m32 - m31 * m22 * m13 - m32 * m23 * m11 - m33 * m21 * m12;
    const double d01 = m10 * m22 * m33 + m12 * m23 * m30 + m13 * m20 * m32 - m30 * m22 * m13 - m32 * m23 * m10 - m33 * m20 * m12;
    const double d02 = m10 * m21 * m33 + m11 * m23 * m30 + m13 * m20 * m31 - m30 * m21 * m13 - m31 * m23 * m10 - m33 * m20 * m11;
    const double d03 = m10 * m21 * m32 + m11 * m22 * m30 + m12 * m20 * m31 - m30 * m21 * m12 - m31 * m22 * m10 - m32 * m20 * m11;
    const double d10 = m01 * m22 * m33 + m02 * m23 * m31 + m03 * m21 * m32 - m31 * m22 * m03 - m32 * m23 * m01 - m33 * m21 * m02;
    const double d11 = m00 * m22 * m33 + m02 * m23 * m30 + m03 * m20 * m32 - m30 * m22 * m03 - m32 * m23 * m00 - m33 * m20 * m02;
    const double d12 = m00 * m21 * m33 + m01 * m23 * m30 + m03 * m20 * m31 - m30 * m21 * m03 - m31 * m23 * m00 - m33 * m20 * m01;
    const double d13 = m00 * m21 * m32 + m01 * m22 * m30 + m02 * m20 * m31 - m30 * m21 * m02 - m31 * m22 * m00 - m32 * m20 * m01;
    const double d20 = m01 * m12 * m33 + m02 * m13 * m31 + m03 * m11 * m32 - m31 * m12 * m03 - m32 * m13 * m01 - m33 * m11 * m02;
    const double d21 = m00 * m12 * m33 + m02 * m13 * m30 + m03 * m10 * m32 - m30 * m12 * m03 - m32 * m13 * m00 - m33 * m10 * m02;
    const double d22 = m00 * m11 * m33 + m01 * m13 * m30 + m03 * m10 * m31 - m30 * m11 * m03 - m31 * m13 * m00 - m33 * m10 * m01;
    const double d23 = m00 * m11 * m32 + m01 * m12 * m30 + m02 * m10 * m31 - m30 * m11 * m02 - m31 * m12 * m00 - m32 * m10 * m01;
    const double d30 = m01 * m12 * m23 + m02 * m13 * m21 + m03 * m11 * m22 - m21 * m12 * m03 - m22 * m13 * m01 - m23 * m11 * m02;
    const double d31 = m00 * m12 * m23 + m02 * m13 * m20 + m03 * m10 * m22 - m20 * m12 * m03 - m22 * m13 * m00 - m23 * m10 * m02;
    const double d32 = m00 * m11 * m23 + m01 * m13 * m20 + m03 * m10 * m21 - m20 * m11 * m03 - m21 * m13 * m00 - m23 * m10 * m01;
    const double d33 = m00 * m11 * m22 + m01 * m12 * m20 + m02 * m10 * m21 - m20 * m11 * m02 - m21 * m12 * m00 - m22 * m10 * m01;
    const double D = m00 * d00 - m01 * d01 + m02 * d02 - m03 * d03;
    if (D == 0.0) return false;
    r[0][0] = d00 / D; r[0][1] = -d10 / D; r[0][2] = d20 / D; r[0][3] = -d30 / D;
    r[1][0] = -d01 / D; r[1][1] = d11 / D; r[1][2] = -d21 / D; r[1][3] = d31 / D;
    r[2][0] = d02 / D; r[2][1] = -d12 / D; r[2][2] = d22 / D; r[2][3] = -d32 / D;
    r[3][0] = -d03 / D; r[3][1] = d13 / D; r[3][2] = -d23 / D; r[3][3] = d33 / D;
    return true;
}

// Parse_Polygon's closing of the point list (parser.cpp, Parse_Polygon) and
// Compute_Polygon (polygon.cpp:646-805), then Polygon::Transform.
PolygonO polygon_prepare(const double *pts, int number, const Transform *user) {
    std::vector<V3> P;
    for (int k = 0; k < number; k++) P.push_back(V3(pts[3 * k], pts[3 * k + 1], pts[3 * k + 2]));
    bool closed = false;
    V3 Q = P[0];
    for (int i = 1; i < number; i++) {
        closed = false;
        if (std::fabs(Q.x - P[i].x) < kEpsilon && std::fabs(Q.y - P[i].y) < kEpsilon &&
            std::fabs(Q.z - P[i].z) < kEpsilon) {
            P[i] = Q;
            i++;
            if (i < number) Q = P[i];
            closed = true;
        }
    }
    if (!closed) {
        P.push_back(Q);
        number++;
    }
    PolygonO g;
    const V3 o = P[0];
    V3 u, v, w;
    int i = 1;
    for (; i < number; i++) {
        u = P[i] - o;
        if (length_sqr(u) > kEpsilon) break;
    }
    if (i == number) g.degenerate = true;
    for (i++; i < number; i++) {
        v = P[i] - o;
        w = cross(u, v);
        if (length_sqr(v) > kEpsilon && length_sqr(w) > kEpsilon) break;
    }
    if (i == number) g.degenerate = true;
    u = cross(v, w);
    v = cross(w, u);
    for (V3 *x : {&u, &v, &w}) {
        const double l = length(*x);
        if (l != 0) *x = *x / l;
    }
    double A[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {-o.x, -o.y, -o.z, 1}};
    double B[4][4] = {{u.x, v.x, w.x, 0}, {u.y, v.y, w.y, 0}, {u.z, v.z, w.z, 0}, {0, 0, 0, 1}};
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++)
            g.tr.inv[r][c] = A[r][0] * B[0][c] + A[r][1] * B[1][c] + A[r][2] * B[2][c] + A[r][3] * B[3][c];
    if (!minvers(g.tr.m, g.tr.inv)) { // MInvers(Trans->matrix, Trans->inverse)
        g.degenerate = true;          // never hit; keep a finite box
        std::memset(g.tr.m, 0, sizeof g.tr.m);
        for (int k = 0; k < 4; k++) g.tr.m[k][k] = 1.0;
    }
    for (int k = 0; k < number; k++) {
        const double x = P[k].x - o.x, y = P[k].y - o.y, z = P[k].z - o.z;
        if (std::fabs(x * w.x + y * w.y + z * w.z) > 1.0e-10) g.degenerate = true;
        g.pts.push_back(x * u.x + y * u.y + z * u.z);
        g.pts.push_back(x * v.x + y * v.y + z * v.z);
    }
    g.number = number;
    if (user) compose(g.tr, *user); // Polygon::Transform: Compose_Transforms
    // Polygon::Compute_BBox (polygon.cpp:806-842): the points (u, v, 0)
    // through Trans, widened by SMALL_TOLERANCE on the axes thinner than it
    V3 lo(kBoundHuge, kBoundHuge, kBoundHuge), hi(-kBoundHuge, -kBoundHuge, -kBoundHuge);
    for (int k = 0; k < number; k++) {
        const V3 p = mtrans_point(g.tr.m, V3(g.pts[2 * k], g.pts[2 * k + 1], 0.0));
        for (int a = 0; a < 3; a++) {
            lo[a] = std::min(lo[a], p[a]);
            hi[a] = std::max(hi[a], p[a]);
        }
    }
    g.bbox = make_bbox_from_min_max(lo, hi);
    for (int a = 0; a < 3; a++)
        if (std::fabs(g.bbox.size[a]) < kSmallTolerance) {
            g.bbox.ll[a] = (float)(g.bbox.ll[a] - kSmallTolerance);
            g.bbox.size[a] = (float)(g.bbox.size[a] + 2.0 * kSmallTolerance);
        }
    V3 n = minv_trans_normal(g.tr.inv, V3(0.0, 0.0, 1.0));
    g.normal = normalized(n);
    return g;
}

// Polygon::in_polygon (polygon.cpp:905-984)
static bool in_polygon(const std::vector<double> &pts, int number, double tx, double ty) {
    int i = 1;
    const double *vtx0 = &pts[0], *vtx1 = &pts[2], *first = vtx0;
    bool yflag0 = vtx0[1] >= ty, inside = false;
    while (i < number) {
        const bool yflag1 = vtx1[1] >= ty;
        if (yflag0 != yflag1) {
            if ((((vtx1[1] - ty) * (vtx0[0] - vtx1[0])) >= ((vtx1[0] - tx) * (vtx0[1] - vtx1[1]))) == yflag1)
                inside = !inside;
        }
        if (i < number - 2 && vtx1[0] == first[0] && vtx1[1] == first[1]) {
            vtx0 = &pts[2 * (++i)];
            vtx1 = &pts[2 * (++i)];
            yflag0 = vtx0[1] >= ty;
            first = vtx0;
        } else {
            vtx0 = vtx1;
            ++i;
            if (i < number) vtx1 = &pts[2 * i];
            yflag0 = yflag1;
        }
    }
    return inside;
}

// Polygon::Intersect (polygon.cpp:187-230)
bool polygon_intersect(const PolygonO &g, const Ray &r, double *depth) {
    if (g.degenerate) return false;
    const V3 p = mtrans_point(g.tr.inv, r.o);
    V3 d = mtrans_direction(g.tr.inv, r.d);
    const double len = length(d);
    d = d / len;
    if (std::fabs(d.z) < 1.0e-10) return false;
    double t = -p.z / d.z;
    if (t < 1.0e-8 || t > kMaxDistance) return false;
    if (!in_polygon(g.pts, g.number, p.x + t * d.x, p.y + t * d.y)) return false;
    *depth = t / len;
    return true;
}

} // namespace orc
