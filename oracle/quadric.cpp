// Quadric (core/shape/quadric.cpp) — FP64 ORACLE, TEST INFRASTRUCTURE ONLY.
//
// Bounds: Parse_Quadric calls Quadric::Compute_BBox(-BOUND_HUGE, BOUND_HUGE)
// (parser.cpp:5526-5530) before the object's transforms, which then move the
// box with Recompute_BBox (quadric.cpp:368-381).  The box is the object's
// FP32 pretest (ObjectBase::Intersect_BBox, object.cpp:917-942) and decides
// whether the quadric is an infinite object.
#include <algorithm>

#include "oracle.hpp"

namespace orc {

// Quadric::Compute_BBox(ClipMin, ClipMax) (quadric.cpp:764-1240) without
// clipping objects, starting from the default box of ObjectBase
// (object.h:201); returns whether a smaller box was found (else `lo`/`hi`
// keep the default).
static bool quadric_box(const double *coef, V3 clip_min, V3 clip_max, V3 &out_lo, V3 &out_hi) {
    // Square_Terms = (c0, c1, c2), Mixed_Terms = (c3, c4, c5), Terms = (c6, c7, c8), Constant = c9
    double A = coef[0], E = coef[1], H = coef[2];
    double B = coef[3] / 2.0, C = coef[4] / 2.0, F = coef[5] / 2.0;
    double D = coef[6] / 2.0, G = coef[7] / 2.0, I = coef[8] / 2.0, J = coef[9];
    for (double *v : {&A, &B, &C, &D, &E, &F, &G, &H, &I, &J})
        if (std::fabs(*v) < kEpsilon) *v = 0.0;
    const double old_volume = kBoundHuge * kBoundHuge * kBoundHuge; // BOUNDS_VOLUME of the default box
    if (B != 0.0 || C != 0.0 || F != 0.0) {
        const double nv = (clip_max.x - clip_min.x) * (clip_max.y - clip_min.y) * (clip_max.z - clip_min.z);
        if (nv < old_volume) {
            out_lo = clip_min;
            out_hi = clip_max;
            return true;
        }
        return false;
    }
    V3 T1(0.0, 0.0, 0.0);
    if (D != 0.0 || G != 0.0 || I != 0.0) {
        T1.x = A != 0.0 ? -D / A : (D != 0.0 ? -J / (2.0 * D) : 0.0);
        T1.y = E != 0.0 ? -G / E : (G != 0.0 ? -J / (2.0 * G) : 0.0);
        T1.z = H != 0.0 ? -I / H : (I != 0.0 ? -J / (2.0 * I) : 0.0);
        D += A * T1.x;
        G += E * T1.y;
        I += H * T1.z;
        J -= T1.x * (A * T1.x - 2.0 * D) + T1.y * (E * T1.y - 2.0 * G) + T1.z * (H * T1.z - 2.0 * I);
    }
    V3 nmin(-kBoundHuge / 2, -kBoundHuge / 2, -kBoundHuge / 2), nmax(kBoundHuge / 2, kBoundHuge / 2, kBoundHuge / 2);
    clip_min = clip_min - T1;
    clip_max = clip_max - T1;
    if (A < 0.0) {
        A = -A;
        D = -D;
        E = -E;
        G = -G;
        H = -H;
        I = -I;
        J = -J;
    }
    // ellipsoid
    if (A > 0.0 && E > 0.0 && H > 0.0 && J < 0.0) {
        const double a = std::sqrt(-J / A), b = std::sqrt(-J / E), c = std::sqrt(-J / H);
        nmin = V3(-a, -b, -c);
        nmax = V3(a, b, c);
    }
    // cylinders along x, y, z
    if (A == 0.0 && E > 0.0 && H > 0.0 && J < 0.0) {
        const double b = std::sqrt(-J / E), c = std::sqrt(-J / H);
        nmin.y = -b; nmin.z = -c; nmax.y = b; nmax.z = c;
    }
    if (A > 0.0 && E == 0.0 && H > 0.0 && J < 0.0) {
        const double a = std::sqrt(-J / A), c = std::sqrt(-J / H);
        nmin.x = -a; nmin.z = -c; nmax.x = a; nmax.z = c;
    }
    if (A > 0.0 && E > 0.0 && H == 0.0 && J < 0.0) {
        const double a = std::sqrt(-J / A), b = std::sqrt(-J / E);
        nmin.x = -a; nmin.y = -b; nmax.x = a; nmax.y = b;
    }
    // cones along x, y, z: radii at the clip planes of the axis
    if (A > 0.0 && E < 0.0 && H < 0.0 && J == 0.0) {
        const double a = std::sqrt(1.0 / A), b = std::sqrt(-1.0 / E), c = std::sqrt(-1.0 / H);
        const double ry = std::max(std::fabs(clip_min.x * b / a), std::fabs(clip_max.x * b / a));
        const double rz = std::max(std::fabs(clip_min.x * c / a), std::fabs(clip_max.x * c / a));
        nmin.y = -ry; nmin.z = -rz; nmax.y = ry; nmax.z = rz;
    }
    if (A > 0.0 && E < 0.0 && H > 0.0 && J == 0.0) {
        const double a = std::sqrt(1.0 / A), b = std::sqrt(-1.0 / E), c = std::sqrt(1.0 / H);
        const double rx = std::max(std::fabs(clip_min.y * a / b), std::fabs(clip_max.y * a / b));
        const double rz = std::max(std::fabs(clip_min.y * c / b), std::fabs(clip_max.y * c / b));
        nmin.x = -rx; nmin.z = -rz; nmax.x = rx; nmax.z = rz;
    }
    if (A > 0.0 && E > 0.0 && H < 0.0 && J == 0.0) {
        const double a = std::sqrt(1.0 / A), b = std::sqrt(1.0 / E), c = std::sqrt(-1.0 / H);
        const double rx = std::max(std::fabs(clip_min.z * a / c), std::fabs(clip_max.z * a / c));
        const double ry = std::max(std::fabs(clip_min.z * b / c), std::fabs(clip_max.z * b / c));
        nmin.x = -rx; nmin.y = -ry; nmax.x = rx; nmax.y = ry;
    }
    // hyperboloids along x, y, z
    if (A > 0.0 && E < 0.0 && H < 0.0 && J > 0.0) {
        const double d1 = 1.0 + A * (clip_min.x * clip_min.x), d2 = 1.0 + A * (clip_max.x * clip_max.x);
        const double ry = std::max(std::sqrt(-d1 / E), std::sqrt(-d2 / E));
        const double rz = std::max(std::sqrt(-d1 / H), std::sqrt(-d2 / H));
        nmin.y = -ry; nmin.z = -rz; nmax.y = ry; nmax.z = rz;
    }
    if (A > 0.0 && E < 0.0 && H > 0.0 && J < 0.0) {
        const double d1 = 1.0 - E * (clip_min.y * clip_min.y), d2 = 1.0 - E * (clip_max.y * clip_max.y);
        const double rx = std::max(std::sqrt(d1 / A), std::sqrt(d2 / A));
        const double rz = std::max(std::sqrt(d1 / H), std::sqrt(d2 / H));
        nmin.x = -rx; nmin.z = -rz; nmax.x = rx; nmax.z = rz;
    }
    if (A > 0.0 && E > 0.0 && H < 0.0 && J < 0.0) {
        const double d1 = 1.0 - H * (clip_min.z * clip_min.z), d2 = 1.0 - H * (clip_max.z * clip_max.z);
        const double rx = std::max(std::sqrt(d1 / A), std::sqrt(d2 / A));
        const double ry = std::max(std::sqrt(d1 / E), std::sqrt(d2 / E));
        nmin.x = -rx; nmin.y = -ry; nmax.x = rx; nmax.y = ry;
    }
    // paraboloids along x, y, z
    if (A == 0.0 && D != 0.0 && E * H > 0.0 && J == 0.0) {
        double x = D * E < 0 ? std::max(0., clip_min.x) : clip_min.x;
        const double ry1 = std::sqrt(std::fabs(2.0 * D * x / E)), rz1 = std::sqrt(std::fabs(2.0 * D * x / H));
        x = D * E > 0 ? std::min(0., clip_max.x) : clip_max.x;
        const double ry2 = std::sqrt(std::fabs(2.0 * D * x / E)), rz2 = std::sqrt(std::fabs(2.0 * D * x / H));
        const double ry = std::max(ry1, ry2), rz = std::max(rz1, rz2);
        if (D * E < 0) nmin.x = std::max(0., clip_min.x);
        nmin.y = -ry; nmin.z = -rz;
        if (D * E > 0) nmax.x = std::min(0., clip_max.x);
        nmax.y = ry; nmax.z = rz;
    }
    if (E == 0.0 && G != 0.0 && A * H > 0.0 && J == 0.0) {
        double y = G > 0 ? clip_min.y : std::max(0., clip_min.y);
        const double rx1 = std::sqrt(std::fabs(2.0 * G * y / A)), rz1 = std::sqrt(std::fabs(2.0 * G * y / H));
        y = G < 0 ? clip_max.y : std::min(0., clip_max.y);
        const double rx2 = std::sqrt(std::fabs(2.0 * G * y / A)), rz2 = std::sqrt(std::fabs(2.0 * G * y / H));
        const double rx = std::max(rx1, rx2), rz = std::max(rz1, rz2);
        nmin.x = -rx;
        if (G < 0) nmin.y = std::max(0., clip_min.y);
        nmin.z = -rz;
        nmax.x = rx;
        if (G > 0) nmax.y = std::min(0., clip_max.y);
        nmax.z = rz;
    }
    if (H == 0.0 && I != 0.0 && A * E > 0.0 && J == 0.0) {
        double z = I < 0 ? std::max(clip_min.z, 0.) : clip_min.z;
        const double rx1 = std::sqrt(std::fabs(2.0 * I * z / A)), ry1 = std::sqrt(std::fabs(2.0 * I * z / E));
        z = I > 0 ? std::min(0., clip_max.z) : clip_max.z;
        const double rx2 = std::sqrt(std::fabs(2.0 * I * z / A)), ry2 = std::sqrt(std::fabs(2.0 * I * z / E));
        const double rx = std::max(rx1, rx2), ry = std::max(ry1, ry2);
        nmin.x = -rx;
        nmin.y = -ry;
        if (I < 0) nmin.z = std::max(clip_min.z, 0.);
        nmax.x = rx;
        nmax.y = ry;
        if (I > 0) nmax.z = std::min(clip_max.z, 0.);
    }
    for (int k = 0; k < 3; k++) {
        nmin[k] = std::max(nmin[k], clip_min[k]);
        nmax[k] = std::min(nmax[k], clip_max[k]);
    }
    const double nv = (nmax.x - nmin.x) * (nmax.y - nmin.y) * (nmax.z - nmin.z);
    if (!(nv < old_volume)) return false;
    out_lo = nmin + T1;
    out_hi = nmax + T1;
    return true;
}

// Recompute_BBox (boundingbox.cpp): the box's eight corners through the
// transform, their FP64 min/max, Make_BBox_from_min_max.
static FBox recompute_bbox(const FBox &b, const Transform &t) {
    V3 lo(kBoundHuge, kBoundHuge, kBoundHuge), hi(-kBoundHuge, -kBoundHuge, -kBoundHuge);
    for (int k = 0; k < 8; k++) {
        V3 c((k & 1) ? (double)b.ll[0] + (double)b.size[0] : (double)b.ll[0],
             (k & 2) ? (double)b.ll[1] + (double)b.size[1] : (double)b.ll[1],
             (k & 4) ? (double)b.ll[2] + (double)b.size[2] : (double)b.ll[2]);
        c = mtrans_point(t.m, c);
        for (int a = 0; a < 3; a++) {
            lo[a] = std::min(lo[a], c[a]);
            hi[a] = std::max(hi[a], c[a]);
        }
    }
    return make_bbox_from_min_max(lo, hi);
}

// The quadric's BBox after Parse_Quadric's Compute_BBox and its transform.
FBox quadric_bbox(const double *coef, const Transform *user) {
    FBox b; // ObjectBase's default box (object.h:201)
    for (int k = 0; k < 3; k++) {
        b.ll[k] = (float)(-kBoundHuge / 2.0);
        b.size[k] = (float)kBoundHuge;
    }
    V3 lo, hi;
    const V3 cmin(-kBoundHuge, -kBoundHuge, -kBoundHuge), cmax(kBoundHuge, kBoundHuge, kBoundHuge);
    if (quadric_box(coef, cmin, cmax, lo, hi)) {
        b = make_bbox_from_min_max(lo, hi);
        const double kCriticalLength = 1.0e6; // CRITICAL_LENGTH (boundingbox.h)
        if (b.size[0] > kCriticalLength || b.size[1] > kCriticalLength || b.size[2] > kCriticalLength)
            for (int k = 0; k < 3; k++) {
                b.ll[k] = (float)(-kBoundHuge / 2.0);
                b.size[k] = (float)kBoundHuge;
            }
    }
    if (user) b = recompute_bbox(b, *user);
    return b;
}

// Parse_Quadric (parser.cpp) + Quadric::Transform (quadric.cpp): the 4x4
// coefficient matrix (Quadric_To_Matrix) becomes inverse * M (MTimesB),
// then M * inverse^T (MTimesA), read back by Matrix_To_Quadric.
Quadric quadric_prepare(const double *c, const Transform *user) {
    Quadric q;
    for (int k = 0; k < 10; k++) q.c[k] = c[k]; // A B C (square), D E F (mixed), G H I (linear), J
    q.bbox = quadric_bbox(c, user);
    if (!user) return q;
    double M[4][4] = {{c[0], c[3], c[4], c[6]}, {0, c[1], c[5], c[7]}, {0, 0, c[2], c[8]}, {0, 0, 0, c[9]}};
    double R[4][4];
    for (int j = 0; j < 4; j++)
        for (int i = 0; i < 4; i++)
            R[i][j] = user->inv[i][0] * M[0][j] + user->inv[i][1] * M[1][j] + user->inv[i][2] * M[2][j] +
                      user->inv[i][3] * M[3][j];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            M[i][j] = R[i][0] * user->inv[j][0] + R[i][1] * user->inv[j][1] + R[i][2] * user->inv[j][2] +
                      R[i][3] * user->inv[j][3];
    q.c[0] = M[0][0];
    q.c[1] = M[1][1];
    q.c[2] = M[2][2];
    q.c[3] = M[0][1] + M[1][0];
    q.c[4] = M[0][2] + M[2][0];
    q.c[5] = M[1][2] + M[2][1];
    q.c[6] = M[0][3] + M[3][0];
    q.c[7] = M[1][3] + M[3][1];
    q.c[8] = M[2][3] + M[3][2];
    q.c[9] = M[3][3];
    return q;
}

// Quadric::Intersect + All_Intersections (quadric.cpp:100-200)
int quadric_intersect(const Quadric &q, const Ray &r, double *depth) {
    const double QA = q.c[0], QE = q.c[1], QH = q.c[2], QB = q.c[3], QC = q.c[4], QF = q.c[5], QD = q.c[6],
                 QG = q.c[7], QI = q.c[8], QJ = q.c[9];
    const double Xo = r.o.x, Yo = r.o.y, Zo = r.o.z, Xd = r.d.x, Yd = r.d.y, Zd = r.d.z;
    const double a = Xd * (QA * Xd + QB * Yd + QC * Zd) + Yd * (QE * Yd + QF * Zd) + Zd * QH * Zd;
    const double b = Xd * (QA * Xo + 0.5 * (QB * Yo + QC * Zo + QD)) + Yd * (QE * Yo + 0.5 * (QB * Xo + QF * Zo + QG)) +
                     Zd * (QH * Zo + 0.5 * (QC * Xo + QF * Yo + QI));
    const double c = Xo * (QA * Xo + QB * Yo + QC * Zo + QD) + Yo * (QE * Yo + QF * Zo + QG) + Zo * (QH * Zo + QI) + QJ;
    double d1, d2;
    if (a != 0.0) {
        double d = b * b - a * c;
        if (d <= 0.0) return 0;
        d = std::sqrt(d);
        d1 = (-b + d) / a;
        d2 = (-b - d) / a;
    } else {
        if (b == 0.0) return 0;
        d1 = -0.5 * c / b;
        d2 = kMaxDistance;
    }
    int n = 0;
    if (d1 > 1.0e-6 && d1 < kMaxDistance) depth[n++] = d1;
    if (d2 > 1.0e-6 && d2 < kMaxDistance) depth[n++] = d2;
    return n;
}

// Quadric::Inside; Quadric::Invert negates every coefficient.
bool quadric_inside(const Quadric &q, V3 p, bool inverted) {
    const double s = inverted ? -1.0 : 1.0;
    const double QA = s * q.c[0], QE = s * q.c[1], QH = s * q.c[2], QB = s * q.c[3], QC = s * q.c[4],
                 QF = s * q.c[5], QD = s * q.c[6], QG = s * q.c[7], QI = s * q.c[8], QJ = s * q.c[9];
    return (p.x * (QA * p.x + QB * p.y + QD) + p.y * (QE * p.y + QF * p.z + QG) + p.z * (QH * p.z + QC * p.x + QI) +
            QJ) <= 0.0;
}

// Quadric::Normal: the normalised gradient, (1,0,0) where it vanishes.
V3 quadric_normal(const Quadric &q, V3 p) {
    const double QA = q.c[0], QE = q.c[1], QH = q.c[2], QB = q.c[3], QC = q.c[4], QF = q.c[5], QD = q.c[6],
                 QG = q.c[7], QI = q.c[8];
    V3 n(2.0 * QA * p.x + QB * p.y + QC * p.z + QD, QB * p.x + 2.0 * QE * p.y + QF * p.z + QG,
         QC * p.x + QF * p.y + 2.0 * QH * p.z + QI);
    const double len = length(n);
    return len == 0.0 ? V3(1.0, 0.0, 0.0) : n / len;
}

} // namespace orc
