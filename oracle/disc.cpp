// Disc (core/shape/disc.cpp) — FP64 ORACLE, TEST INFRASTRUCTURE ONLY.
#include "oracle.hpp"

namespace orc {

// Parse_Disc (parser.cpp, Parse_Disc) + Compute_Disc (disc.cpp:Compute_Disc)
// + Disc::Transform: the normal is normalised, the placement is
// Compute_Coordinate_Transform(center, normal, 1, 1) (which may snap the
// stored normal to +-z), a later transform carries the normal with
// MTransNormal and is composed after.
Disc disc_prepare(const double *center, const double *normal, const double *radius, const Transform *user) {
    Disc c;
    V3 n(normal[0], normal[1], normal[2]);
    const double l = length(n);
    if (l != 0) n = n / l;
    c.oradius2 = radius[0] * radius[0];
    c.iradius2 = radius[1] * radius[1];
    coordinate_transform(c.tr, V3(center[0], center[1], center[2]), n, 1.0, 1.0);
    if (user) {
        n = minv_trans_normal(user->inv, n);
        const double m = length(n);
        if (m != 0) n = n / m;
        compose(c.tr, *user);
    }
    c.normal = n;
    return c;
}

// Disc::Intersect (disc.cpp): the plane z = 0 in disc space, the hit's
// squared radius within [iradius2, oradius2], depth in (1e-6, MAX_DISTANCE).
bool disc_intersect(const Disc &c, const Ray &r, double *depth) {
    const V3 P = mtrans_point(c.tr.inv, r.o);
    V3 D = mtrans_direction(c.tr.inv, r.d);
    const double len = length(D);
    D = D / len;
    if (std::fabs(D.z) > kEpsilon) {
        const double t = -P.z / D.z;
        if (t >= 0.0) {
            const double u = P.x + t * D.x, v = P.y + t * D.y;
            const double r2 = u * u + v * v;
            if (r2 >= c.iradius2 && r2 <= c.oradius2) {
                *depth = t / len;
                if (*depth > 1.0e-6 && *depth < kMaxDistance) return true;
            }
        }
    }
    return false;
}

// Disc::Inside (disc.cpp): the half-space z < 0 of disc space.
bool disc_inside(const Disc &c, V3 p, bool inverted) {
    const V3 q = mtrans_point(c.tr.inv, p);
    return q.z >= 0.0 ? inverted : !inverted;
}

// Disc::Compute_BBox (disc.cpp)
FBox disc_bbox(const Disc &c) {
    const double rad = std::sqrt(c.oradius2);
    return box_bbox(V3(-rad, -rad, -kSmallTolerance), V3(rad, rad, kSmallTolerance), &c.tr);
}

} // namespace orc
