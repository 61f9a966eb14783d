// Quadric (core/shape/quadric.cpp) — FP64 ORACLE, TEST INFRASTRUCTURE ONLY.
//
// Bounds: the oracle treats quadrics as unbounded (tested outside the
// bounding tree, like planes).  Quadric::Compute_BBox (quadric.cpp:764-1000)
// only narrows the box around the surface, which culls no hit of a
// top-level quadric.
#include "oracle.hpp"

namespace orc {

// Parse_Quadric (parser.cpp) + Quadric::Transform (quadric.cpp): the 4x4
// coefficient matrix (Quadric_To_Matrix) becomes inverse * M (MTimesB),
// then M * inverse^T (MTimesA), read back by Matrix_To_Quadric.
Quadric quadric_prepare(const double *c, const Transform *user) {
    Quadric q;
    for (int k = 0; k < 10; k++) q.c[k] = c[k]; // A B C (square), D E F (mixed), G H I (linear), J
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
