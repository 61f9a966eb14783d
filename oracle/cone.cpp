// Cone / cylinder (core/shape/cone.cpp) — FP64 ORACLE, TEST INFRASTRUCTURE ONLY.
//
// The shape's own placement follows Compute_Cone_Data / Compute_Cylinder_Data
// (cone.cpp:715-840) through Compute_Coordinate_Transform
// (core/math/matrix.cpp:880-906) and Compose_Transforms (matrix.cpp:788-793,
// MTimesA / MTimesB summation order), then the object's own transform.
#include <cstring>
#include <stdexcept>

#include "oracle.hpp"

namespace orc {

static void identity(double (*m)[4]) {
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) m[i][j] = i == j ? 1.0 : 0.0;
}

// Compute_Scaling_Transform (matrix.cpp:535-548)
static void scaling(Transform &t, V3 v) {
    identity(t.m);
    t.m[0][0] = v.x;
    t.m[1][1] = v.y;
    t.m[2][2] = v.z;
    identity(t.inv);
    t.inv[0][0] = 1.0 / v.x;
    t.inv[1][1] = 1.0 / v.y;
    t.inv[2][2] = 1.0 / v.z;
}

// Compute_Translation_Transform (matrix.cpp:644-657)
static void translation(Transform &t, V3 v) {
    identity(t.m);
    t.m[3][0] = v.x;
    t.m[3][1] = v.y;
    t.m[3][2] = v.z;
    identity(t.inv);
    t.inv[3][0] = -v.x;
    t.inv[3][1] = -v.y;
    t.inv[3][2] = -v.z;
}

// Compute_Axis_Rotation_Transform (matrix.cpp:825-850)
static void axis_rotation(Transform &t, V3 axis, double angle) {
    const double l = length(axis);
    const V3 v = l != 0 ? axis / l : axis;
    identity(t.m);
    const double c = std::cos(angle), s = std::sin(angle);
    t.m[0][0] = v.x * v.x + c * (1.0 - v.x * v.x);
    t.m[0][1] = v.x * v.y * (1.0 - c) + v.z * s;
    t.m[0][2] = v.x * v.z * (1.0 - c) - v.y * s;
    t.m[1][0] = v.x * v.y * (1.0 - c) - v.z * s;
    t.m[1][1] = v.y * v.y + c * (1.0 - v.y * v.y);
    t.m[1][2] = v.y * v.z * (1.0 - c) + v.x * s;
    t.m[2][0] = v.x * v.z * (1.0 - c) + v.y * s;
    t.m[2][1] = v.y * v.z * (1.0 - c) - v.x * s;
    t.m[2][2] = v.z * v.z + c * (1.0 - v.z * v.z);
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) t.inv[i][j] = t.m[j][i];
}

// Compose_Transforms: matrix = matrix * add.matrix (MTimesA),
// inverse = add.inverse * inverse (MTimesB).
void compose(Transform &t, const Transform &a) {
    double r[4][4];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            r[i][j] = t.m[i][0] * a.m[0][j] + t.m[i][1] * a.m[1][j] + t.m[i][2] * a.m[2][j] + t.m[i][3] * a.m[3][j];
    std::memcpy(t.m, r, sizeof r);
    for (int j = 0; j < 4; j++)
        for (int i = 0; i < 4; i++)
            r[i][j] = a.inv[i][0] * t.inv[0][j] + a.inv[i][1] * t.inv[1][j] + a.inv[i][2] * t.inv[2][j] +
                      a.inv[i][3] * t.inv[3][j];
    std::memcpy(t.inv, r, sizeof r);
}

// Compute_Coordinate_Transform (matrix.cpp:880-906); like the reference it
// snaps the caller's `up` to +-z when it is within EPSILON of it.
void coordinate_transform(Transform &t, V3 origin, V3 &up, double radius, double len) {
    scaling(t, V3(radius, radius, len));
    V3 tmpv;
    if (std::fabs(up.z) > 1.0 - kEpsilon) {
        tmpv = V3(1.0, 0.0, 0.0);
        up.z = up.z < 0.0 ? -1.0 : 1.0;
    } else {
        tmpv = V3(-up.y, up.x, 0.0);
    }
    Transform t2;
    axis_rotation(t2, tmpv, std::acos(up.z));
    compose(t, t2);
    translation(t2, origin);
    compose(t, t2);
}

// Parse_Cone / Parse_Cylinder + Compute_Cone_Data / Compute_Cylinder_Data
// (parser.cpp:2255-2315, cone.cpp:715-840) + Cone::Transform (cone.cpp:660).
Cone cone_prepare(const double *apex_in, const double *base_in, const double *radius, uint8_t flags,
                  const Transform *user) {
    Cone c;
    V3 apex(apex_in[0], apex_in[1], apex_in[2]), base(base_in[0], base_in[1], base_in[2]);
    double ar = radius[0], br = (flags & PVX_CONE_CYLINDER) ? radius[0] : radius[1];
    c.closed = (flags & PVX_CONE_OPEN) == 0;
    c.cylinder = (flags & PVX_CONE_CYLINDER) != 0;
    auto cylinder_data = [&]() { // Compute_Cylinder_Data (cone.cpp:806-840)
        V3 axis = apex - base;
        const double l = length(axis);
        if (l < kEpsilon) throw std::runtime_error("degenerate cylinder");
        axis = axis / l;
        coordinate_transform(c.tr, base, axis, ar, l);
        c.dist = 0.0;
    };
    if (c.cylinder) {
        cylinder_data();
    } else {
        V3 axis = apex - base;
        const double len = length(axis);
        if (len < kEpsilon) throw std::runtime_error("degenerate cone");
        axis = axis / len;
        if (std::fabs(ar - br) < kEpsilon) {
            c.cylinder = true;
            cylinder_data();
        } else {
            if (ar < br) { // the bigger end at the top
                std::swap(apex, base);
                std::swap(ar, br);
                axis = V3(-axis.x, -axis.y, -axis.z);
            }
            const double tmpf = br * len / (ar - br);
            const V3 origin = base - axis * tmpf;
            const double tlen = tmpf + len;
            if (((ar - br) * len / tlen) < kEpsilon) {
                c.cylinder = true;
                cylinder_data();
            } else {
                c.dist = tmpf / tlen;
                coordinate_transform(c.tr, origin, axis, ar, tlen);
            }
        }
    }
    if (user) compose(c.tr, *user);
    return c;
}

// Cone::Intersect (cone.cpp:140-300); push order of Cone::All_Intersections.
int cone_intersect(const Cone &c, const Ray &r, double *depth, int *part) {
    const double kTol = 1.0e-9; // Cone_Tolerance (cone.cpp:64)
    V3 P = mtrans_point(c.tr.inv, r.o);
    V3 D = mtrans_direction(c.tr.inv, r.d);
    const double len = length(D);
    D = D / len;
    int i = 0;
    if (c.cylinder) {
        const double a = D.x * D.x + D.y * D.y;
        if (a > kEpsilon) {
            const double b = P.x * D.x + P.y * D.y;
            const double cc = P.x * P.x + P.y * P.y - 1.0;
            double d = b * b - a * cc;
            if (d >= 0.0) {
                d = std::sqrt(d);
                const double t1 = (-b + d) / a, t2 = (-b - d) / a;
                double z = P.z + t1 * D.z;
                if (t1 > kTol && t1 < kMaxDistance && z >= 0.0 && z <= 1.0) {
                    depth[i] = t1 / len;
                    part[i++] = 3;
                }
                z = P.z + t2 * D.z;
                if (t2 > kTol && t2 < kMaxDistance && z >= 0.0 && z <= 1.0) {
                    depth[i] = t2 / len;
                    part[i++] = 3;
                }
            }
        }
    } else {
        const double a = D.x * D.x + D.y * D.y - D.z * D.z;
        const double b = D.x * P.x + D.y * P.y - D.z * P.z;
        const double cc = P.x * P.x + P.y * P.y - P.z * P.z;
        if (std::fabs(a) < kEpsilon) {
            if (std::fabs(b) > kEpsilon) {
                const double t1 = -0.5 * cc / b;
                const double z = P.z + t1 * D.z;
                if (t1 > kTol && t1 < kMaxDistance && z >= c.dist && z <= 1.0) {
                    depth[i] = t1 / len;
                    part[i++] = 3;
                }
            }
        } else {
            double d = b * b - a * cc;
            if (d >= 0.0) {
                d = std::sqrt(d);
                const double t1 = (-b - d) / a, t2 = (-b + d) / a;
                double z = P.z + t1 * D.z;
                if (t1 > kTol && t1 < kMaxDistance && z >= c.dist && z <= 1.0) {
                    depth[i] = t1 / len;
                    part[i++] = 3;
                }
                z = P.z + t2 * D.z;
                if (t2 > kTol && t2 < kMaxDistance && z >= c.dist && z <= 1.0) {
                    depth[i] = t2 / len;
                    part[i++] = 3;
                }
            }
        }
    }
    if (c.closed && std::fabs(D.z) > kEpsilon) {
        double d = (1.0 - P.z) / D.z;
        double a = P.x + d * D.x, b = P.y + d * D.y;
        if ((a * a + b * b) <= 1.0 && d > kTol && d < kMaxDistance) {
            depth[i] = d / len;
            part[i++] = 2; // CAP_HIT
        }
        d = (c.dist - P.z) / D.z;
        a = P.x + d * D.x;
        b = P.y + d * D.y;
        if ((a * a + b * b) <= (c.cylinder ? 1.0 : c.dist * c.dist) && d > kTol && d < kMaxDistance) {
            depth[i] = d / len;
            part[i++] = 1; // BASE_HIT
        }
    }
    return i;
}

// Cone::Inside (cone.cpp:360-410)
bool cone_inside(const Cone &c, V3 p, bool inverted) {
    const double offset = c.closed ? -kEpsilon : kEpsilon;
    const V3 q = mtrans_point(c.tr.inv, p);
    const double w2 = q.x * q.x + q.y * q.y;
    bool out;
    if (c.cylinder)
        out = w2 > 1.0 + offset || q.z < 0.0 - offset || q.z > 1.0 + offset;
    else
        out = w2 > q.z * q.z + offset || q.z < c.dist - offset || q.z > 1.0 + offset;
    return out ? inverted : !inverted;
}

// Cone::Normal (cone.cpp:440-480)
V3 cone_normal(const Cone &c, int part, V3 ip) {
    V3 n = mtrans_point(c.tr.inv, ip);
    if (part == 3) {
        if (c.cylinder) n.z = 0.0;
        else n.z = -n.z;
    } else if (part == 1) {
        n = V3(0.0, 0.0, -1.0);
    } else if (part == 2) {
        n = V3(0.0, 0.0, 1.0);
    }
    return normalized(minv_trans_normal(c.tr.inv, n));
}

// Cone::Compute_BBox (cone.cpp:870-880)
FBox cone_bbox(const Cone &c) {
    return box_bbox(V3(-1.0, -1.0, c.dist), V3(1.0, 1.0, 1.0), &c.tr);
}

} // namespace orc
