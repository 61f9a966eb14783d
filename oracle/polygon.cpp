// Polygon (core/shape/polygon.cpp) — FP64 ORACLE, TEST INFRASTRUCTURE ONLY.
#include <cstring>

#include "oracle.hpp"

namespace orc {

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
    std::memset(g.tr.m, 0, sizeof g.tr.m); // not needed for hits
    for (int k = 0; k < number; k++) {
        const double x = P[k].x - o.x, y = P[k].y - o.y, z = P[k].z - o.z;
        if (std::fabs(x * w.x + y * w.y + z * w.z) > 1.0e-10) g.degenerate = true;
        g.pts.push_back(x * u.x + y * u.y + z * u.z);
        g.pts.push_back(x * v.x + y * v.y + z * v.z);
    }
    g.number = number;
    V3 lo(kBoundHuge, kBoundHuge, kBoundHuge), hi(-kBoundHuge, -kBoundHuge, -kBoundHuge);
    for (const V3 &p : P)
        for (int a = 0; a < 3; a++) {
            lo[a] = std::min(lo[a], p[a]);
            hi[a] = std::max(hi[a], p[a]);
        }
    if (user) {
        // Compose_Transforms on the inverse only: inverse = user.inverse * inverse
        double r[4][4];
        for (int a = 0; a < 4; a++)
            for (int c = 0; c < 4; c++)
                r[a][c] = user->inv[a][0] * g.tr.inv[0][c] + user->inv[a][1] * g.tr.inv[1][c] +
                          user->inv[a][2] * g.tr.inv[2][c] + user->inv[a][3] * g.tr.inv[3][c];
        std::memcpy(g.tr.inv, r, sizeof r);
        V3 mn(kBoundHuge, kBoundHuge, kBoundHuge), mx(-kBoundHuge, -kBoundHuge, -kBoundHuge);
        for (int k = 0; k < 8; k++) {
            V3 c((k & 1) ? hi.x : lo.x, (k & 2) ? hi.y : lo.y, (k & 4) ? hi.z : lo.z);
            c = mtrans_point(user->m, c);
            for (int a = 0; a < 3; a++) {
                mn[a] = std::min(mn[a], c[a]);
                mx[a] = std::max(mx[a], c[a]);
            }
        }
        lo = mn;
        hi = mx;
    }
    // the bounding box of Polygon::Compute_BBox, widened by SMALL_TOLERANCE on every axis
    g.bbox = make_bbox_from_min_max(lo - V3(kSmallTolerance, kSmallTolerance, kSmallTolerance),
                                    hi + V3(kSmallTolerance, kSmallTolerance, kSmallTolerance));
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
