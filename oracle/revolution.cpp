// Lathe (core/shape/lathe.cpp) and surface of revolution (core/shape/sor.cpp)
// with the bounding cylinders of core/bounding/boundingcylinder.cpp — FP64
// ORACLE, TEST INFRASTRUCTURE ONLY.
//
// Both shapes rotate a 2-D spline (x = radius, y = height) about the y axis.
// A ray is moved into the shape's space and its direction normalised; the
// spline segments whose bounding cylinder shell (Intersect_BCyl) the ray
// crosses are visited in the order of the shell entry distance, and a
// polynomial in the ray parameter is solved per segment.
#include <algorithm>
#include <cstring>
#include <stdexcept>

#include "oracle.hpp"

namespace orc {

namespace {
const double kDepthTol = 1.0e-4; // lathe.cpp:84, sor.cpp:133 DEPTH_TOLERANCE
const int kBasePlane = 1, kCapPlane = 2, kCurve = 3; // sor.cpp:137-139

void identity(Transform &t) {
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) t.m[i][j] = t.inv[i][j] = (i == j) ? 1.0 : 0.0;
}

// Inverse of a 4x4 matrix as the transposed cofactor matrix over the
// determinant (MInvers, matrix.cpp:1154-1202, which throws on a singular
// one): each cofactor is the 3x3 determinant of its minor by the rule of
// Sarrus, in the reference's term order, so the coefficients round alike.
void invert4(const double (*m)[4], double (*r)[4]) {
    auto minor3 = [&](int row, int col) {
        double a[3][3];
        for (int i = 0, ii = 0; i < 4; i++) {
            if (i == row) continue;
            for (int j = 0, jj = 0; j < 4; j++) {
                if (j == col) continue;
                a[ii][jj++] = m[i][j];
            }
            ii++;
        }
        return a[0][0] * a[1][1] * a[2][2] + a[0][1] * a[1][2] * a[2][0] + a[0][2] * a[1][0] * a[2][1] -
               a[2][0] * a[1][1] * a[0][2] - a[2][1] * a[1][2] * a[0][0] - a[2][2] * a[1][0] * a[0][1];
    };
    double d[4][4];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) d[i][j] = minor3(i, j);
    const double det = m[0][0] * d[0][0] - m[0][1] * d[0][1] + m[0][2] * d[0][2] - m[0][3] * d[0][3];
    if (det == 0.0) throw std::runtime_error("singular matrix in a surface of revolution");
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) r[i][j] = (((i + j) & 1) ? -d[j][i] : d[j][i]) / det;
}

// Create_BCyl (boundingcylinder.cpp:408-560): the distinct radii and heights
// of the segments' cylinders (exact comparison, first appearance order,
// radii stored squared) and each segment's indices into them.
void make_bcyl(Revolution &s, const std::vector<double> &r1, const std::vector<double> &r2,
               const std::vector<double> &h1, const std::vector<double> &h2) {
    std::vector<double> rad, hgt;
    auto find_or_add = [](std::vector<double> &list, double v) {
        for (size_t j = 0; j < list.size(); j++)
            if (list[j] == v) return (int32_t)j;
        list.push_back(v);
        return (int32_t)(list.size() - 1);
    };
    s.entry.clear();
    for (size_t i = 0; i < r1.size(); i++) {
        const int32_t a = find_or_add(rad, r1[i]);
        const int32_t b = find_or_add(rad, r2[i]);
        const int32_t c = find_or_add(hgt, h1[i]);
        const int32_t d = find_or_add(hgt, h2[i]);
        s.entry.insert(s.entry.end(), {a, b, c, d});
    }
    s.radius.resize(rad.size());
    for (size_t i = 0; i < rad.size(); i++) s.radius[i] = rad[i] * rad[i];
    s.height = hgt;
}

// Compute_BBox (lathe.cpp:455-461, sor.cpp:705-711): the overall cylinder's
// box through the shape's transform (Recompute_BBox).
void finish(Revolution &s, const Transform *user) {
    identity(s.tr);
    if (user) s.tr = *user;
    FBox b;
    b.ll[0] = (float)(-s.radius2);
    b.ll[1] = (float)s.height1;
    b.ll[2] = (float)(-s.radius2);
    b.size[0] = (float)(2.0 * s.radius2);
    b.size[1] = (float)(s.height2 - s.height1);
    b.size[2] = (float)(2.0 * s.radius2);
    s.bbox = recompute_bbox(b, user);
}

struct Interval {
    double d0;
    int32_t n;
};

// Intersect_BCyl (boundingcylinder.cpp:295-404) with intersect_bound_elements
// (:186-255) and intersect_thick_cylinder (:80-170): per segment, the
// crossings of the ray with its cap and base rings and its inner and outer
// cylinders, sorted; the entry distance (0 when the ray starts inside)
// inserted into a list ordered by distance, before any equal one (insert_hit).
std::vector<Interval> bcyl_intervals(const Revolution &s, V3 P, V3 D) {
    const double a = D.x * D.x + D.z * D.z;
    const double b = P.x * D.x + P.z * D.z;
    const double bb = b * b, b2 = 2.0 * b;
    const double c = P.x * P.x + P.z * P.z;
    const size_t nh = s.height.size(), nr = s.radius.size();
    std::vector<int> hn(nh, 0), rn(nr, 0);
    std::vector<double> hd(nh), hw(nh), rd(2 * nr), rw(2 * nr);
    if (D.y < -kEpsilon || D.y > kEpsilon) {
        for (size_t i = 0; i < nh; i++) {
            const double k = (s.height[i] - P.y) / D.y;
            hn[i] = 1;
            hd[i] = k;
            hw[i] = k * (a * k + b2) + c;
        }
    }
    for (size_t i = 0; i < nr; i++) {
        if (s.radius[i] > kEpsilon) {
            double d = bb - a * (c - s.radius[i]);
            if (d > 0.0) {
                d = std::sqrt(d);
                double k = (-b + d) / a;
                rn[i] = 2;
                rd[2 * i] = k;
                rw[2 * i] = P.y + k * D.y;
                k = (-b - d) / a;
                rd[2 * i + 1] = k;
                rw[2 * i + 1] = P.y + k * D.y;
            }
        }
    }
    std::vector<Interval> out;
    for (int i = 0; i < s.segments(); i++) {
        const int32_t r1 = s.entry[4 * i], r2 = s.entry[4 * i + 1], h1 = s.entry[4 * i + 2], h2 = s.entry[4 * i + 3];
        double dist[8];
        int n = 0;
        for (int32_t h : {h2, h1}) // cap ring, then base ring
            if (hn[h] && hw[h] >= s.radius[r1] && hw[h] <= s.radius[r2]) dist[n++] = hd[h];
        for (int32_t r : {r1, r2}) // inner, then outer cylinder
            if (rn[r])
                for (int e = 0; e < 2; e++)
                    if (rw[2 * r + e] >= s.height[h1] && rw[2 * r + e] <= s.height[h2]) dist[n++] = rd[2 * r + e];
        for (int p = 0; p < n; p++) // bubble sort, as the reference
            for (int q = 0; q < n - p - 1; q++)
                if (dist[q] > dist[q + 1]) std::swap(dist[q], dist[q + 1]);
        bool add = false;
        double d0 = 0.0;
        if (n == 2 || n == 4) {
            for (int p = 0; p < n && !add; p += 2) {
                if (dist[p] > kEpsilon) {
                    d0 = dist[p];
                    add = true;
                } else if (dist[p + 1] > kEpsilon) {
                    d0 = 0.0;
                    add = true;
                }
            }
        } else if (n != 0) { // an odd count: taken as starting inside
            d0 = dist[0];
            add = true;
        }
        if (add) {
            size_t pos = 0;
            while (pos < out.size() && d0 > out[pos].d0) pos++;
            out.insert(out.begin() + pos, Interval{d0, i});
        }
    }
    return out;
}

// test_hit (lathe.cpp:1291-1308, sor.cpp:1206-1223) without clipped_by.
bool push_hit(std::vector<Isect> &st, double d, int32_t slot, int32_t side, double d1) {
    if (d > kDepthTol && d < kMaxDistance) {
        st.push_back(Isect{d, slot, side, slot, d1});
        return true;
    }
    return false;
}

bool lathe_intersect(const Revolution &s, const Ray &ray, bool child, std::vector<Isect> &st, int32_t slot) {
    // Lathe::Intersect (lathe.cpp:201-365)
    V3 P = mtrans_point(s.tr.inv, ray.o);
    V3 D = mtrans_direction(s.tr.inv, ray.d);
    const double len = length(D);
    D = D / len;
    const double Dylen = D.y * len;
    if ((D.y >= 0.0 && P.y > s.height2) || (D.y <= 0.0 && P.y < s.height1) || (D.x >= 0.0 && P.x > s.radius2) ||
        (D.x <= 0.0 && P.x < -s.radius2))
        return false;
    double r0 = std::fabs(P.x * D.z - P.z * D.x);
    const double r = D.x * D.x + D.z * D.z;
    if (r > 0.0) r0 /= std::sqrt(r);
    if (r0 > s.radius2) return false;
    const std::vector<Interval> iv = bcyl_intervals(s, P, D);
    if (iv.empty()) return false;
    const double m = D.x * P.x + D.z * P.z;
    const double Dy2 = D.y * D.y;
    bool found = false;
    double best = kBoundHuge;
    for (const Interval &in : iv) {
        const double *e = &s.seg[8 * (size_t)in.n];
        const double Ax = e[0], Ay = e[1], Bx = e[2], By = e[3], Cx = e[4], Cy = e[5], Dx = e[6], Dy = e[7];
        if (!child && in.d0 > best) break;
        double x1[7], y1[6];
        int n1 = 0;
        if (s.spline == PVX_LATHE_LINEAR) {
            x1[0] = Cy * Cy * r - Cx * Cx * Dy2;
            x1[1] = 2.0 * (Cy * ((Dy - P.y) * r + D.y * m) - Cx * Dx * Dy2);
            x1[2] = (Dy - P.y) * ((Dy - P.y) * r + 2.0 * D.y * m) + Dy2 * (P.x * P.x + P.z * P.z - Dx * Dx);
            n1 = solve_polynomial(2, x1, y1, false, 0.0);
        } else if (s.spline == PVX_LATHE_QUADRATIC) {
            x1[0] = By * By * r - Bx * Bx * Dy2;
            x1[1] = 2.0 * (By * Cy * r - Bx * Cx * Dy2);
            x1[2] = r * (2.0 * By * (Dy - P.y) + Cy * Cy) + 2.0 * By * D.y * m - (2.0 * Bx * Dx + Cx * Cx) * Dy2;
            x1[3] = 2.0 * (Cy * ((Dy - P.y) * r + D.y * m) - Cx * Dx * Dy2);
            x1[4] = (Dy - P.y) * ((Dy - P.y) * r + 2.0 * D.y * m) + Dy2 * (P.x * P.x + P.z * P.z - Dx * Dx);
            n1 = solve_polynomial(4, x1, y1, s.sturm, 0.0);
        } else { // cubic and Bezier
            x1[0] = Ay * Ay * r - Ax * Ax * Dy2;
            x1[1] = 2.0 * (Ay * By * r - Ax * Bx * Dy2);
            x1[2] = (2.0 * Ay * Cy + By * By) * r - (2.0 * Ax * Cx + Bx * Bx) * Dy2;
            x1[3] = 2.0 * ((Ay * Dy + By * Cy - Ay * P.y) * r + Ay * D.y * m - (Ax * Dx + Bx * Cx) * Dy2);
            x1[4] = (2.0 * By * (Dy - P.y) + Cy * Cy) * r + 2.0 * By * D.y * m - (2.0 * Bx * Dx + Cx * Cx) * Dy2;
            x1[5] = 2.0 * (Cy * ((Dy - P.y) * r + D.y * m) - Cx * Dx * Dy2);
            x1[6] = (Dy - P.y) * ((Dy - P.y) * r + 2.0 * D.y * m) + Dy2 * (P.x * P.x + P.z * P.z - Dx * Dx);
            n1 = solve_polynomial(6, x1, y1, s.sturm, 0.0);
        }
        while (n1--) {
            const double w = y1[n1];
            if (w < 0.0 || w > 1.0) continue;
            if (std::fabs(D.y) > kEpsilon) {
                double k = w * (w * (w * Ay + By) + Cy) + Dy - P.y;
                if (push_hit(st, k / Dylen, slot, in.n, w)) {
                    found = true;
                    k /= D.y;
                    if (k < best) best = k;
                }
            } else {
                const double k = w * (w * (w * Ax + Bx) + Cx) + Dx;
                double x2[3] = {r, 2.0 * m, P.x * P.x + P.z * P.z - k * k}, y2[2];
                int n2 = solve_polynomial(2, x2, y2, false, 0.0);
                while (n2--) {
                    const double k2 = y2[n2];
                    if (push_hit(st, k2 / len, slot, in.n, w)) {
                        found = true;
                        if (k2 < best) best = k2;
                    }
                }
            }
        }
    }
    return found;
}

bool sor_intersect(const Revolution &s, const Ray &ray, bool child, std::vector<Isect> &st, int32_t slot) {
    // Sor::Intersect (sor.cpp:241-420)
    V3 P = mtrans_point(s.tr.inv, ray.o);
    V3 D = mtrans_direction(s.tr.inv, ray.d);
    const double len = length(D);
    D = D / len;
    if ((D.y >= 0.0 && P.y > s.height2) || (D.y <= 0.0 && P.y < s.height1) || (D.x >= 0.0 && P.x > s.radius2) ||
        (D.x <= 0.0 && P.x < -s.radius2))
        return false;
    double r0 = P.x * D.z - P.z * D.x;
    const double a = D.x * D.x + D.z * D.z;
    if (a > 0.0) r0 /= std::sqrt(a);
    if (r0 > s.radius2) return false;
    bool found = false;
    double best = kBoundHuge;
    if (s.closed && std::fabs(D.y) > kEpsilon) {
        if (s.base_r2 > kDepthTol) {
            const double k = (s.height1 - P.y) / D.y;
            const double u = P.x + k * D.x, v = P.z + k * D.z;
            if (u * u + v * v <= s.base_r2 && push_hit(st, k / len, slot, kBasePlane, k)) {
                found = true;
                if (k < best) best = k;
            }
        }
        if (s.cap_r2 > kDepthTol) {
            const double k = (s.height2 - P.y) / D.y;
            const double u = P.x + k * D.x, v = P.z + k * D.z;
            if (u * u + v * v <= s.cap_r2 && push_hit(st, k / len, slot, kCapPlane, k)) {
                found = true;
                if (k < best) best = k;
            }
        }
    }
    const std::vector<Interval> iv = bcyl_intervals(s, P, D);
    for (const Interval &in : iv) {
        const double *e = &s.seg[8 * (size_t)in.n];
        const double A = e[0], B = e[2], C = e[4], Dd = e[6];
        if (!child && in.d0 > best) break;
        double x[4], y[3];
        x[0] = A * D.y * D.y * D.y;
        x[1] = D.y * D.y * (3.0 * A * P.y + B) - a;
        x[2] = D.y * (P.y * (3.0 * A * P.y + 2.0 * B) + C) - 2.0 * (P.x * D.x + P.z * D.z);
        x[3] = P.y * (P.y * (A * P.y + B) + C) + Dd - P.x * P.x - P.z * P.z;
        int n = solve_polynomial(3, x, y, s.sturm, 0.0);
        const double h1 = s.height[s.entry[4 * in.n + 2]], h2 = s.height[s.entry[4 * in.n + 3]];
        while (n--) {
            const double k = y[n];
            const double h = P.y + k * D.y;
            if (h >= h1 && h <= h2 && push_hit(st, k / len, slot, kCurve | (in.n << 2), k)) {
                found = true;
                if (k < best) best = k;
            }
        }
    }
    return found;
}

} // namespace

// Parse_Lathe (parser.cpp:2812-2995) checks and Compute_Lathe (lathe.cpp:488-760).
Revolution lathe_prepare(const double *pts, int n, int spline, bool sturm, const Transform *user) {
    Revolution s;
    s.sor = false;
    s.spline = spline;
    s.sturm = sturm;
    int nseg = 0;
    switch (spline) {
    case PVX_LATHE_LINEAR:
        if (n < 2) throw std::runtime_error("lathe with linear splines needs at least two points");
        nseg = n - 1;
        break;
    case PVX_LATHE_QUADRATIC:
        if (n < 3) throw std::runtime_error("lathe with quadratic splines needs at least three points");
        nseg = n - 2;
        break;
    case PVX_LATHE_CUBIC:
        if (n < 4) throw std::runtime_error("lathe with cubic splines needs at least four points");
        nseg = n - 3;
        break;
    case PVX_LATHE_BEZIER:
        if (n < 4 || (n & 3) != 0) throw std::runtime_error("lathe with Bezier splines needs four points per segment");
        nseg = n / 4;
        break;
    default:
        throw std::runtime_error("unknown lathe spline type");
    }
    for (int i = 0; i < n; i++) { // the v3.8 checks of negative radii
        const double x = pts[2 * i];
        const bool bad = spline == PVX_LATHE_LINEAR      ? x < 0.0
                         : spline == PVX_LATHE_QUADRATIC ? (i > 0 && x < 0.0)
                         : spline == PVX_LATHE_CUBIC     ? (i > 0 && i < n - 1 && x < 0.0)
                                                         : ((i % 4 == 0 || i % 4 == 3) && x < 0.0);
        if (bad) throw std::runtime_error("lathe point with an x value < 0");
    }
    auto P = [&](int i, int k) { return pts[2 * i + k]; };
    std::vector<double> r1(nseg), r2(nseg), h1(nseg), h2(nseg);
    double xmin = kBoundHuge, xmax = -kBoundHuge, ymin = kBoundHuge, ymax = -kBoundHuge;
    s.seg.assign(8 * (size_t)nseg, 0.0);
    for (int seg = 0, i = 0; seg < nseg; seg++) {
        double A[2], B[2], C[2], Dv[2], x[4], y[4];
        for (int k = 0; k < 2; k++) {
            // the points the spline type reads are within the list; the others are unused
            const double p0 = P(i, k), p1 = P(i + 1, k);
            const double p2 = i + 2 < n ? P(i + 2, k) : 0.0;
            const double p3 = i + 3 < n ? P(i + 3, k) : 0.0;
            switch (spline) {
            case PVX_LATHE_LINEAR:
                A[k] = 0.0;
                B[k] = 0.0;
                C[k] = -1.0 * p0 + 1.0 * p1;
                Dv[k] = 1.0 * p0;
                break;
            case PVX_LATHE_QUADRATIC:
                A[k] = 0.0;
                B[k] = 0.5 * p0 - 1.0 * p1 + 0.5 * p2;
                C[k] = -0.5 * p0 + 0.5 * p2;
                Dv[k] = 1.0 * p1;
                break;
            case PVX_LATHE_CUBIC:
                A[k] = -0.5 * p0 + 1.5 * p1 - 1.5 * p2 + 0.5 * p3;
                B[k] = p0 - 2.5 * p1 + 2.0 * p2 - 0.5 * p3;
                C[k] = -0.5 * p0 + 0.5 * p2;
                Dv[k] = p1;
                break;
            default: // Bezier (Bernstein form)
                A[k] = p3 - 3.0 * p2 + 3.0 * p1 - p0;
                B[k] = 3.0 * p2 - 6.0 * p1 + 3.0 * p0;
                C[k] = 3.0 * p1 - 3.0 * p0;
                Dv[k] = p0;
                break;
            }
        }
        if (spline == PVX_LATHE_LINEAR) {
            x[0] = x[2] = P(i, 0);
            x[1] = x[3] = P(i + 1, 0);
            y[0] = y[2] = P(i, 1);
            y[1] = y[3] = P(i + 1, 1);
        } else if (spline == PVX_LATHE_BEZIER) {
            for (int k = 0; k < 4; k++) {
                x[k] = P(i + k, 0);
                y[k] = P(i + k, 1);
            }
        } else {
            x[0] = x[2] = P(i + 1, 0);
            x[1] = x[3] = P(i + 2, 0);
            y[0] = y[2] = P(i + 1, 1);
            y[1] = y[3] = P(i + 2, 1);
            // extremes of the segment inside [0, 1]
            for (int k = 0; k < 2; k++) {
                double c[3] = {3.0 * A[k], 2.0 * B[k], C[k]}, rr[2];
                int nr = solve_polynomial(2, c, rr, false, 0.0);
                double *dst = k == 0 ? x : y;
                while (nr--)
                    if (rr[nr] >= 0.0 && rr[nr] <= 1.0)
                        dst[nr] = rr[nr] * (rr[nr] * (rr[nr] * A[k] + B[k]) + C[k]) + Dv[k];
            }
        }
        double *e = &s.seg[8 * (size_t)seg];
        e[0] = A[0], e[1] = A[1], e[2] = B[0], e[3] = B[1], e[4] = C[0], e[5] = C[1], e[6] = Dv[0], e[7] = Dv[1];
        r1[seg] = std::min(std::min(x[0], x[1]), std::min(x[2], x[3]));
        r2[seg] = std::max(std::max(x[0], x[1]), std::max(x[2], x[3]));
        h1[seg] = std::min(std::min(y[0], y[1]), std::min(y[2], y[3]));
        h2[seg] = std::max(std::max(y[0], y[1]), std::max(y[2], y[3]));
        xmin = std::min(xmin, r1[seg]);
        xmax = std::max(xmax, r2[seg]);
        ymin = std::min(ymin, h1[seg]);
        ymax = std::max(ymax, h2[seg]);
        i += spline == PVX_LATHE_BEZIER ? 4 : 1;
    }
    s.radius1 = xmin;
    s.radius2 = xmax;
    s.height1 = ymin;
    s.height2 = ymax;
    make_bcyl(s, r1, r2, h1, h2);
    finish(s, user);
    return s;
}

// Parse_Sor (parser.cpp:5628-5714) checks and Compute_Sor (sor.cpp:920-1180):
// a cubic r^2(h) = A h^3 + B h^2 + C h + D per segment through points i+1
// and i+2 with the slopes of the neighbouring chords.
Revolution sor_prepare(const double *pts, int n, bool open, bool sturm, const Transform *user) {
    Revolution s;
    s.sor = true;
    s.spline = PVX_SOR_CUBIC;
    s.closed = !open;
    s.sturm = sturm;
    if (n < 4) throw std::runtime_error("surface of revolution needs at least four points");
    auto P = [&](int i, int k) { return pts[2 * i + k]; };
    for (int i = 0; i < n; i++)
        if (P(i, 0) < 0.0 || (i > 1 && i < n - 1 && P(i, 1) <= P(i - 1, 1)))
            throw std::runtime_error("incorrect point in surface of revolution");
    const int nseg = n - 3;
    std::vector<double> r1(nseg), r2(nseg), h1(nseg), h2(nseg);
    double xmin = kBoundHuge, xmax = -kBoundHuge, ymin = kBoundHuge, ymax = -kBoundHuge;
    s.seg.assign(8 * (size_t)nseg, 0.0);
    for (int i = 0; i < nseg; i++) {
        if (std::fabs(P(i + 2, 1) - P(i, 1)) < kEpsilon || std::fabs(P(i + 3, 1) - P(i + 1, 1)) < kEpsilon)
            throw std::runtime_error("incorrect point in surface of revolution");
        double k[4];
        k[0] = P(i + 1, 0) * P(i + 1, 0);
        k[1] = P(i + 2, 0) * P(i + 2, 0);
        k[2] = (P(i + 2, 0) - P(i, 0)) / (P(i + 2, 1) - P(i, 1));
        k[3] = (P(i + 3, 0) - P(i + 1, 0)) / (P(i + 3, 1) - P(i + 1, 1));
        k[2] *= 2.0 * P(i + 1, 0);
        k[3] *= 2.0 * P(i + 2, 0);
        double M[4][4], Mi[4][4];
        double w = P(i + 1, 1);
        M[0][0] = w * w * w, M[0][1] = w * w, M[0][2] = w, M[0][3] = 1.0;
        M[2][0] = 3.0 * w * w, M[2][1] = 2.0 * w, M[2][2] = 1.0, M[2][3] = 0.0;
        w = P(i + 2, 1);
        M[1][0] = w * w * w, M[1][1] = w * w, M[1][2] = w, M[1][3] = 1.0;
        M[3][0] = 3.0 * w * w, M[3][1] = 2.0 * w, M[3][2] = 1.0, M[3][3] = 0.0;
        invert4(M, Mi);
        double c[4];
        for (int r = 0; r < 4; r++) {
            c[r] = k[0] * Mi[r][0] + k[1] * Mi[r][1] + k[2] * Mi[r][2] + k[3] * Mi[r][3];
            if (std::fabs(c[r]) < kEpsilon) c[r] = 0.0;
        }
        const double A = c[0], B = c[1], C = c[2], D = c[3];
        double *e = &s.seg[8 * (size_t)i];
        e[0] = A, e[2] = B, e[4] = C, e[6] = D;
        const double y0 = P(i + 1, 1), y1 = P(i + 2, 1);
        double x[4] = {P(i + 1, 0), P(i + 2, 0), P(i + 1, 0), P(i + 2, 0)};
        double cc[3] = {3.0 * A, 2.0 * B, C}, rr[2];
        int nr = solve_polynomial(2, cc, rr, false, 0.0);
        while (nr--)
            if (rr[nr] >= y0 && rr[nr] <= y1) x[nr] = std::sqrt(rr[nr] * (rr[nr] * (rr[nr] * A + B) + C) + D);
        r1[i] = std::min(std::min(x[0], x[1]), std::min(x[2], x[3]));
        r2[i] = std::max(std::max(x[0], x[1]), std::max(x[2], x[3]));
        h1[i] = y0;
        h2[i] = y1;
        xmin = std::min(xmin, r1[i]);
        xmax = std::max(xmax, r2[i]);
        ymin = std::min(ymin, h1[i]);
        ymax = std::max(ymax, h2[i]);
    }
    s.radius1 = xmin;
    s.radius2 = xmax;
    s.height1 = ymin;
    s.height2 = ymax;
    auto r2_at = [&](int seg, double w) {
        const double *e = &s.seg[8 * (size_t)seg];
        return w * (w * (e[0] * w + e[2]) + e[4]) + e[6];
    };
    s.cap_r2 = std::max(0.0, r2_at(nseg - 1, h2[nseg - 1])); // squared cap and base radii, clamped at 0
    s.base_r2 = std::max(0.0, r2_at(0, h1[0]));
    make_bcyl(s, r1, r2, h1, h2);
    finish(s, user);
    return s;
}

bool revolution_intersect(const Revolution &s, const Ray &r, bool child, std::vector<Isect> &st, int32_t slot) {
    return s.sor ? sor_intersect(s, r, child, st, slot) : lathe_intersect(s, r, child, st, slot);
}

// Lathe::Inside (lathe.cpp:380-450): crossings of the horizontal ray from
// the point outwards with the segments; Sor::Inside (sor.cpp:455-505): the
// point's squared radius against the first segment spanning its height.
bool revolution_inside(const Revolution &s, V3 ip, bool inverted) {
    const V3 P = mtrans_point(s.tr.inv, ip);
    if (s.sor) {
        if (P.y >= s.height1 && P.y <= s.height2) {
            const double r0 = P.x * P.x + P.z * P.z;
            if (r0 <= s.radius2 * s.radius2) {
                for (int i = 0; i < s.segments(); i++) {
                    if (P.y >= s.height[s.entry[4 * i + 2]] && P.y <= s.height[s.entry[4 * i + 3]]) {
                        const double *e = &s.seg[8 * (size_t)i];
                        const double r = P.y * (P.y * (P.y * e[0] + e[2]) + e[4]) + e[6];
                        if (r0 <= r) return !inverted;
                        break;
                    }
                }
            }
        }
        return inverted;
    }
    int nc = 0;
    if (P.y >= s.height1 && P.y <= s.height2) {
        const double r = std::sqrt(P.x * P.x + P.z * P.z);
        if (r >= s.radius1 && r <= s.radius2) {
            for (int i = 0; i < s.segments(); i++) {
                if (P.y >= s.height[s.entry[4 * i + 2]] - kEpsilon && P.y <= s.height[s.entry[4 * i + 3]] + kEpsilon) {
                    const double *e = &s.seg[8 * (size_t)i];
                    double x[4] = {e[1], e[3], e[5], e[7] - P.y}, y[3];
                    int n = solve_polynomial(3, x, y, s.sturm, 0.0);
                    while (n--) {
                        const double w = y[n];
                        if (w >= 0.0 && w <= 1.0 && w * (w * (w * e[0] + e[2]) + e[4]) + e[6] - r >= 0.0) nc++;
                    }
                }
            }
        }
    }
    return (nc & 1) ? !inverted : inverted;
}

// Lathe::Normal (lathe.cpp:466-505), Sor::Normal (sor.cpp:540-585): the
// curve's normal rotated to the hit point, through MTransNormal, normalised.
V3 revolution_normal(const Revolution &s, const Isect &h, V3 ip) {
    const V3 P = mtrans_point(s.tr.inv, ip);
    V3 N;
    if (s.sor) {
        const int part = h.side & 3, seg = h.side >> 2;
        if (part == kCurve) {
            const double *e = &s.seg[8 * (size_t)seg];
            const double k = 0.5 * (P.y * (3.0 * e[0] * P.y + 2.0 * e[2]) + e[4]);
            N = V3(P.x, -k, P.z);
        } else {
            N = V3(0.0, part == kBasePlane ? -1.0 : 1.0, 0.0);
        }
    } else {
        const double *e = &s.seg[8 * (size_t)h.side];
        double r = P.x * P.x + P.z * P.z;
        if (r > kEpsilon) {
            r = std::sqrt(r);
            const double w = h.d1;
            const double dx = w * (3.0 * e[0] * w + 2.0 * e[2]) + e[4];
            const double dy = w * (3.0 * e[1] * w + 2.0 * e[3]) + e[5];
            N = V3(dy * P.x, -dx * r, dy * P.z);
        } else {
            N = V3(0.0, 1.0, 0.0);
        }
    }
    return normalized(minv_trans_normal(s.tr.inv, N));
}

} // namespace orc
