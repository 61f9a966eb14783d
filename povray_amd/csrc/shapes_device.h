// Device-side shape tests (gfx950).  FP64 arithmetic in the reference's
// operation order; the translation unit is compiled with -ffp-contract=off so
// no FMA contraction changes rounding relative to the reference.
#ifndef PVX_SHAPES_DEVICE_H
#define PVX_SHAPES_DEVICE_H

#include <hip/hip_runtime.h>

#include "pvx_types.h"

namespace pvx {

constexpr int kCsgMaxDepth = 8;     // nesting limit of CSG trees on the device
constexpr int kMaxLeafHits = 4;     // hits one leaf shape can push (torus: 4)

struct D3 {
    double x, y, z;
};
__device__ inline D3 d3(double a, double b, double c) { return D3{a, b, c}; }
__device__ inline double dot3(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline D3 add3(D3 a, D3 b) { return D3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ inline D3 sub3(D3 a, D3 b) { return D3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline D3 mul3(D3 a, double s) { return D3{a.x * s, a.y * s, a.z * s}; }
__device__ inline double comp(D3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

// MTransPoint / MTransDirection with a 4x4 row-vector matrix
// (core/math/matrix.cpp:415-466).
__device__ inline D3 xf_point(const double *m, D3 p) {
    return D3{p.x * m[0] + p.y * m[4] + p.z * m[8] + m[12],
              p.x * m[1] + p.y * m[5] + p.z * m[9] + m[13],
              p.x * m[2] + p.y * m[6] + p.z * m[10] + m[14]};
}
__device__ inline D3 xf_dir(const double *m, D3 p) {
    return D3{p.x * m[0] + p.y * m[4] + p.z * m[8],
              p.x * m[1] + p.y * m[5] + p.z * m[9],
              p.x * m[2] + p.y * m[6] + p.z * m[10]};
}

// ---------------------------------------------------------------------------
// mesh2 triangle: Mesh::intersect_mesh_triangle (core/shape/mesh.cpp:1040-1136)
// ---------------------------------------------------------------------------
// Triangle::Intersect (core/shape/triangle.cpp:479-572) on the FP64
// record: plane depth in [DEPTH_TOLERANCE, MAX_DISTANCE], then the three
// edge tests in the dominant-axis projection (an exact tie counts inside).
__device__ inline bool tri_rec_hit(const TriangleRec &t, D3 o, D3 d, double &depth) {
    if (t.degenerate) return false;
    const double ndd = ((t.n[0] * d.x) + (t.n[1] * d.y)) + (t.n[2] * d.z);
    if (fabs(ndd) < kEpsilon) return false;
    const double ndo = ((t.n[0] * o.x) + (t.n[1] * o.y)) + (t.n[2] * o.z);
    depth = -(t.dist + ndo) / ndd;
    if ((depth < 1.0e-6) || (depth > kMaxDistance)) return false;
    const int a = t.axis == 0 ? 1 : 0, b = t.axis == 2 ? 1 : 2; // projected axes
    const double s = comp(o, a) + depth * comp(d, a);
    const double u = comp(o, b) + depth * comp(d, b);
    if ((t.p2[a] - s) * (t.p2[b] - t.p1[b]) < (t.p2[b] - u) * (t.p2[a] - t.p1[a])) return false;
    if ((t.p3[a] - s) * (t.p3[b] - t.p2[b]) < (t.p3[b] - u) * (t.p3[a] - t.p2[a])) return false;
    if ((t.p1[a] - s) * (t.p1[b] - t.p3[b]) < (t.p1[b] - u) * (t.p1[a] - t.p3[a])) return false;
    return true;
}

__device__ inline bool tri_hit64(const TriRec &t, uint32_t axis, D3 o, D3 d, double &depth) {
    double nx = t.n[0], ny = t.n[1], nz = t.n[2];
    double ndd = nx * d.x + ny * d.y + nz * d.z;
    if (fabs(ndd) < kEpsilon) return false;
    double ndo = nx * o.x + ny * o.y + nz * o.z;
    double dep = -((double)t.dist + ndo) / ndd;
    if (dep < 1.0e-6 || dep > kMaxDistance) return false;
    // projection onto the two axes other than the dominant one
    double s, u, a1, a2, a3, b1, b2, b3;
    if (axis == 0) {
        s = o.y + dep * d.y;
        u = o.z + dep * d.z;
        a1 = t.p1[1]; a2 = t.p2[1]; a3 = t.p3[1];
        b1 = t.p1[2]; b2 = t.p2[2]; b3 = t.p3[2];
    } else if (axis == 1) {
        s = o.x + dep * d.x;
        u = o.z + dep * d.z;
        a1 = t.p1[0]; a2 = t.p2[0]; a3 = t.p3[0];
        b1 = t.p1[2]; b2 = t.p2[2]; b3 = t.p3[2];
    } else {
        s = o.x + dep * d.x;
        u = o.y + dep * d.y;
        a1 = t.p1[0]; a2 = t.p2[0]; a3 = t.p3[0];
        b1 = t.p1[1]; b2 = t.p2[1]; b3 = t.p3[1];
    }
    if ((a2 - s) * (b2 - b1) < (b2 - u) * (a2 - a1)) return false;
    if ((a3 - s) * (b3 - b2) < (b3 - u) * (a3 - a2)) return false;
    if ((a1 - s) * (b1 - b3) < (b1 - u) * (a1 - a3)) return false;
    depth = dep;
    return true;
}

// Sphere::Intersect (core/shape/sphere.cpp:214-241); depths filtered by
// DEPTH_TOLERANCE / MAX_DISTANCE as All_Intersections does (sphere.cpp:159-178).
__device__ inline int sphere_hits(D3 c, double r, D3 o, D3 d, double *out) {
    D3 oc = sub3(c, o);
    double oc2 = dot3(oc, oc);
    double tca = dot3(oc, d);
    double r2 = r * r;
    if (oc2 >= r2 && tca < kEpsilon) return 0;
    double thc2 = r2 - oc2 + tca * tca;
    if (!(thc2 > kEpsilon)) return 0;
    double hc = sqrt(thc2);
    double d1 = tca - hc, d2 = tca + hc;
    int n = 0;
    if (d1 > 1.0e-6 && d1 < kMaxDistance) out[n++] = d1;
    if (d2 > 1.0e-6 && d2 < kMaxDistance) out[n++] = d2;
    return n;
}

__device__ inline bool sphere_inside(D3 c, double r, D3 p, bool inv) { // sphere.cpp:271-303
    D3 oc = sub3(c, p);
    double oc2 = dot3(oc, oc);
    return inv ? (oc2 > r * r) : (oc2 < r * r);
}

// Plane::Intersect (core/shape/plane.cpp:110-170)
__device__ inline bool plane_hit(const PlaneRec &pl, D3 o, D3 d, double &depth) {
    D3 n = d3(pl.n[0], pl.n[1], pl.n[2]);
    double ndd = dot3(n, d);
    if (fabs(ndd) < kEpsilon) return false;
    double ndo = dot3(n, o);
    double dep = -(ndo + pl.d) / ndd;
    if (dep >= 1.0e-6 && dep <= kMaxDistance) {
        depth = dep;
        return true;
    }
    return false;
}

// ---------------------------------------------------------------------------
// Box::Intersect (core/shape/box.cpp:172-505): slab test in box space with
// CLOSE_TOLERANCE side tie-breaking; returns pushed depths (Depth1 only when
// > DEPTH_TOLERANCE, Depth2 always) and their sides.
// ---------------------------------------------------------------------------
__device__ inline int box_hits(const BoxRec &b, const Xform *xf, D3 o, D3 d, double *out, int *side) {
    const double kClose = 1.0e-6;
    D3 P = o, D = d;
    if (b.xf >= 0) {
        P = xf_point(xf[b.xf].inv, o);
        D = xf_dir(xf[b.xf].inv, d);
    }
    int smin = 0, smax = 0;
    double t, tmin = 0.0, tmax = kBoundHuge;
    // X: no ties
    if (D.x < -kEpsilon) {
        t = (b.c1[0] - P.x) / D.x;
        if (t < tmin) return 0;
        if (t <= tmax) { smax = 1; tmax = t; }
        t = (b.c2[0] - P.x) / D.x;
        if (t >= tmin) {
            if (t > tmax) return 0;
            smin = 2;
            tmin = t;
        }
    } else if (D.x > kEpsilon) {
        t = (b.c2[0] - P.x) / D.x;
        if (t < tmin) return 0;
        if (t <= tmax) { smax = 2; tmax = t; }
        t = (b.c1[0] - P.x) / D.x;
        if (t >= tmin) {
            if (t > tmax) return 0;
            smin = 1;
            tmin = t;
        }
    } else if (P.x < b.c1[0] || P.x > b.c2[0]) {
        return 0;
    }
    // Y: ties against |D.x|
    if (D.y < -kEpsilon) {
        t = (b.c1[1] - P.y) / D.y;
        if (t < tmin) return 0;
        if (t <= tmax - kClose) { smax = 3; tmax = t; }
        else if (t <= tmax + kClose) { if (-D.y > fabs(D.x)) smax = 3; }
        t = (b.c2[1] - P.y) / D.y;
        if (t >= tmin + kClose) {
            if (t > tmax) return 0;
            smin = 4;
            tmin = t;
        } else if (t >= tmin - kClose) {
            if (-D.y > fabs(D.x)) smin = 4;
        }
    } else if (D.y > kEpsilon) {
        t = (b.c2[1] - P.y) / D.y;
        if (t < tmin) return 0;
        if (t <= tmax - kClose) { smax = 4; tmax = t; }
        else if (t <= tmax + kClose) { if (D.y > fabs(D.x)) smax = 4; }
        t = (b.c1[1] - P.y) / D.y;
        if (t >= tmin + kClose) {
            if (t > tmax) return 0;
            smin = 3;
            tmin = t;
        } else if (t >= tmin - kClose) {
            if (D.y > fabs(D.x)) smin = 3;
        }
    } else if (P.y < b.c1[1] || P.y > b.c2[1]) {
        return 0;
    }
    // Z: ties against the axis of the side chosen so far
    if (D.z < -kEpsilon) {
        double dz = -D.z;
        t = (b.c1[2] - P.z) / D.z;
        if (t < tmin) return 0;
        if (t <= tmax - kClose) { smax = 5; tmax = t; }
        else if (t <= tmax + kClose) {
            if ((smax == 1 || smax == 2) ? dz > fabs(D.x) : ((smax == 3 || smax == 4) && dz > fabs(D.y))) smax = 5;
        }
        t = (b.c2[2] - P.z) / D.z;
        if (t >= tmin + kClose) {
            if (t > tmax) return 0;
            smin = 6;
            tmin = t;
        } else if (t >= tmin - kClose) {
            if ((smin == 1 || smin == 2) ? dz > fabs(D.x) : ((smin == 3 || smin == 4) && dz > fabs(D.y))) smin = 6;
        }
    } else if (D.z > kEpsilon) {
        double dz = D.z;
        t = (b.c2[2] - P.z) / D.z;
        if (t < tmin) return 0;
        if (t <= tmax - kClose) { smax = 6; tmax = t; }
        else if (t <= tmax + kClose) {
            if ((smax == 1 || smax == 2) ? dz > fabs(D.x) : ((smax == 3 || smax == 4) && dz > fabs(D.y))) smax = 6;
        }
        t = (b.c1[2] - P.z) / D.z;
        if (t >= tmin + kClose) {
            if (t > tmax) return 0;
            smin = 5;
            tmin = t;
        } else if (t >= tmin - kClose) {
            if ((smin == 1 || smin == 2) ? dz > fabs(D.x) : ((smin == 3 || smin == 4) && dz > fabs(D.y))) smin = 5;
        }
    } else if (P.z < b.c1[2] || P.z > b.c2[2]) {
        return 0;
    }
    if (tmax < 1.0e-6) return 0;
    int n = 0;
    if (tmin > 1.0e-6) {
        out[n] = tmin;
        side[n++] = smin;
    }
    out[n] = tmax;
    side[n++] = smax;
    return n;
}

__device__ inline bool box_inside(const BoxRec &b, const Xform *xf, D3 p, bool inv) { // box.cpp:535-570
    D3 q = b.xf >= 0 ? xf_point(xf[b.xf].inv, p) : p;
    if (q.x < b.c1[0] || q.x > b.c2[0]) return inv;
    if (q.y < b.c1[1] || q.y > b.c2[1]) return inv;
    if (q.z < b.c1[2] || q.z > b.c2[2]) return inv;
    return !inv;
}

// ---------------------------------------------------------------------------
// Polynomial roots: core/math/polynomialsolver.cpp (default build branches).
// sbisect's recursion becomes an explicit interval stack visited
// left-to-right, which yields the same roots in the same order.
// ---------------------------------------------------------------------------
namespace poly {
constexpr int kMaxIter = 50;
constexpr double kSmall = 1.0e-10;
constexpr double kRel = 1.0e-12;

// A polynomial of order <= N, ascending powers (the reference's `polynomial`
// holds up to MAX_ORDER 35; the shapes here need 4 (torus) and 6 (lathe)).
template <int N>
struct PN {
    int ord;
    double c[N + 1];
};
using P4 = PN<4>;

__device__ inline double eval(double x, int n, const double *c) {
    double v = c[n];
    for (int i = n - 1; i >= 0; i--) v = v * x + c[i];
    return v;
}

template <int N>
__device__ inline int modp(const PN<N> &u, const PN<N> &v, PN<N> &r) {
    r = u;
    if (v.c[v.ord] < 0.0) {
        for (int k = u.ord - v.ord - 1; k >= 0; k -= 2) r.c[k] = -r.c[k];
        for (int k = u.ord - v.ord; k >= 0; k--)
            for (int j = v.ord + k - 1; j >= k; j--) r.c[j] = -r.c[j] - r.c[v.ord + k] * v.c[j - k];
    } else {
        for (int k = u.ord - v.ord; k >= 0; k--)
            for (int j = v.ord + k - 1; j >= k; j--) r.c[j] -= r.c[v.ord + k] * v.c[j - k];
    }
    int k = v.ord - 1;
    while (k >= 0 && fabs(r.c[k]) < kSmall) {
        r.c[k] = 0.0;
        k--;
    }
    r.ord = (k < 0) ? 0 : k;
    return r.ord;
}

template <int N>
__device__ inline int buildsturm(int ord, PN<N> *s) {
    s[0].ord = ord;
    s[1].ord = ord - 1;
    double f = fabs(s[0].c[ord] * ord);
    for (int i = 1; i <= ord; i++) s[1].c[i - 1] = s[0].c[i] * i / f;
    int k = 2;
    while (modp(s[k - 2], s[k - 1], s[k])) {
        f = -fabs(s[k].c[s[k].ord]);
        for (int j = s[k].ord; j >= 0; j--) s[k].c[j] /= f;
        k++;
    }
    s[k].c[0] = -s[k].c[0];
    return k;
}

template <int N>
__device__ inline int numchanges(int np, const PN<N> *s, double a) {
    int ch = 0;
    double lf = eval(a, s[0].ord, s[0].c);
    for (int k = 1; k <= np; k++) {
        double f = eval(a, s[k].ord, s[k].c);
        if (lf == 0.0 || lf * f < 0) ch++;
        lf = f;
    }
    return ch;
}

template <int N>
__device__ inline int visible_roots(int np, const PN<N> *s) {
    int atpos = 0, atzero = 0;
    double lf = s[0].c[s[0].ord];
    for (int k = 1; k <= np; k++) {
        double f = s[k].c[s[k].ord];
        if (lf == 0.0 || lf * f < 0) atpos++;
        lf = f;
    }
    lf = s[0].c[0];
    for (int k = 1; k <= np; k++) {
        double f = s[k].c[0];
        if (lf == 0.0 || lf * f < 0) atzero++;
        lf = f;
    }
    return atzero - atpos;
}

__device__ inline int regula_falsa(int order, const double *coef, double a, double b, double *val) {
    double fa = eval(a, order, coef), fb = eval(b, order, coef), x, fx, lfx;
    if (fa * fb > 0.0) return 0;
    if (fabs(fa) < kSmall) { *val = a; return 1; }
    if (fabs(fb) < kSmall) { *val = b; return 1; }
    lfx = fa;
    for (int its = 0; its < kMaxIter; its++) {
        x = (fb * a - fa * b) / (fb - fa);
        fx = eval(x, order, coef);
        if (fabs(x) > kRel) {
            if (fabs(fx / x) < kRel) { *val = x; return 1; }
        } else if (fabs(fx) < kRel) {
            *val = x;
            return 1;
        }
        if (fa < 0) {
            if (fx < 0) { a = x; fa = fx; if ((lfx * fx) > 0) fb /= 2; }
            else { b = x; fb = fx; if ((lfx * fx) > 0) fa /= 2; }
        } else {
            if (fx < 0) { b = x; fb = fx; if ((lfx * fx) > 0) fa /= 2; }
            else { a = x; fa = fx; if ((lfx * fx) > 0) fb /= 2; }
        }
        if (fabs(b - a) < kRel) { *val = x; return 1; }
        lfx = fx;
    }
    return 0;
}

// sbisect over an explicit stack.
template <int N>
__device__ inline int sbisect(int np, const PN<N> *s, double lo, double hi, int atlo, int athi,
                              double *roots) {
    struct Iv {
        double lo, hi;
        int atlo, athi;
    } st[N + 2]; // each pending interval holds at least one of the <= N roots
    int sp = 0, nr = 0;
    st[sp++] = Iv{lo, hi, atlo, athi};
    while (sp > 0) {
        Iv iv = st[--sp];
        double mn = iv.lo, mx = iv.hi, mid = 0.0;
        int amin = iv.atlo, amax = iv.athi;
        if (amin - amax == 1) {
            if (regula_falsa(s[0].ord, s[0].c, mn, mx, &roots[nr])) {
                nr++;
                continue;
            }
            bool done = false;
            for (int its = 0; its < kMaxIter; its++) {
                mid = (mn + mx) / 2;
                int am = numchanges(np, s, mid);
                if (am < amax || am > amin) { done = true; break; } // no root
                if (fabs(mid) > kRel) {
                    if (fabs((mx - mn) / mid) < kRel) { roots[nr++] = mid; done = true; break; }
                } else if (fabs(mx - mn) < kRel) {
                    roots[nr++] = mid;
                    done = true;
                    break;
                }
                if (amin - am == 0) mn = mid; else mx = mid;
            }
            if (!done) roots[nr++] = mid;
            continue;
        }
        bool handled = false;
        for (int its = 0; its < kMaxIter; its++) {
            mid = (mn + mx) / 2;
            int am = numchanges(np, s, mid);
            if (am < amax || am > amin) { handled = true; break; }
            if (fabs(mid) > kRel) {
                if (fabs((mx - mn) / mid) < kRel) { roots[nr++] = mid; handled = true; break; }
            } else if (fabs(mx - mn) < kRel) {
                roots[nr++] = mid;
                handled = true;
                break;
            }
            int n1 = amin - am, n2 = am - amax;
            if (n1 != 0 && n2 != 0) {
                if (sp + 2 <= N + 2) {
                    st[sp++] = Iv{mid, mx, am, amax}; // right, visited second
                    st[sp++] = Iv{mn, mid, amin, am}; // left, visited first
                }
                handled = true;
                break;
            }
            if (n1 == 0) mn = mid; else mx = mid;
        }
        if (!handled) roots[nr++] = mid;
    }
    return nr;
}

// (an order-N polynomial has at most N roots pending on the interval stack)
template <int N = 4>
__device__ inline int polysolve(int order, const double *coeffs, double *roots) {
    PN<N> s[N + 2];
    for (int i = 0; i <= order; i++) s[0].c[order - i] = coeffs[i] / coeffs[0];
    int np = buildsturm(order, s);
    if (visible_roots(np, s) == 0) return 0;
    int atmin = numchanges(np, s, 0.0);
    int atmax = numchanges(np, s, kMaxDistance);
    if (atmin - atmax == 0) return 0;
    return sbisect(np, s, 0.0, kMaxDistance, atmin, atmax, roots);
}

__device__ inline int solve_quadratic(const double *x, double *y) {
    double a = x[0], b = -x[1], c = x[2];
    if (a == 0.0) {
        if (b == 0.0) return 0;
        y[0] = c / b;
        return 1;
    }
    b /= a;
    c /= a;
    a = 1.0;
    double d = b * b - 4.0 * a * c;
    if (d > -kSmall && d < kSmall) {
        y[0] = 0.5 * b / a;
        return 1;
    }
    if (d < 0.0) return 0;
    d = sqrt(d);
    double t = 2.0 * a;
    y[0] = (b + d) / t;
    y[1] = (b - d) / t;
    return 2;
}

__device__ inline int solve_cubic(const double *x, double *y) {
    double a0 = x[0], a1, a2, a3;
    if (a0 == 0.0) return solve_quadratic(&x[1], y);
    if (a0 != 1.0) {
        a1 = x[1] / a0; a2 = x[2] / a0; a3 = x[3] / a0;
    } else {
        a1 = x[1]; a2 = x[2]; a3 = x[3];
    }
    double A2 = a1 * a1;
    double Q = (A2 - 3.0 * a2) / 9.0;
    double R = (a1 * (A2 - 4.5 * a2) + 13.5 * a3) / 27.0;
    double Q3 = Q * Q * Q, R2 = R * R, dd = Q3 - R2, an = a1 / 3.0;
    if (dd >= 0.0) {
        dd = R / sqrt(Q3);
        double theta = acos(dd) / 3.0;
        double sQ = -2.0 * sqrt(Q);
        y[0] = sQ * cos(theta) - an;
        y[1] = sQ * cos(theta + 2.0943951023931954923084) - an;
        y[2] = sQ * cos(theta + 4.1887902047863909846168) - an;
        return 3;
    }
    double sQ = pow(sqrt(R2 - Q3) + fabs(R), 1.0 / 3.0);
    y[0] = (R < 0) ? (sQ + Q / sQ) - an : -(sQ + Q / sQ) - an;
    return 1;
}

__device__ inline int difficult_coeffs(const double *x) { // signed `biggest`, default branch
    double biggest = 0.0;
    for (int i = 0; i <= 4; i++)
        if (fabs(x[i]) > biggest) biggest = x[i];
    if (biggest == 0.0) return 0;
    for (int i = 0; i <= 4; i++)
        if (x[i] != 0.0 && fabs(biggest / x[i]) > 1.0e12) return 1;
    return 0;
}

__device__ inline int solve_quartic(const double *x, double *res) {
    double c1, c2, c3, c4, c0 = x[0];
    if (c0 != 1.0) {
        c1 = x[1] / c0; c2 = x[2] / c0; c3 = x[3] / c0; c4 = x[4] / c0;
    } else {
        c1 = x[1]; c2 = x[2]; c3 = x[3]; c4 = x[4];
    }
    double c12 = c1 * c1;
    double p = -0.375 * c12 + c2;
    double q = 0.125 * c12 * c1 - 0.5 * c1 * c2 + c3;
    double r = -0.01171875 * c12 * c12 + 0.0625 * c12 * c2 - 0.25 * c1 * c3 + c4;
    double cubic[4] = {1.0, -0.5 * p, -r, 0.5 * r * p - 0.125 * q * q};
    double roots[3];
    if (solve_cubic(cubic, roots) <= 0) return 0;
    double z = roots[0];
    double d1 = 2.0 * z - p, d2;
    if (d1 < 0.0) {
        if (d1 > -kSmall) d1 = 0.0;
        else return 0;
    }
    if (d1 < kSmall) {
        d2 = z * z - r;
        if (d2 < 0.0) return 0;
        d2 = sqrt(d2);
    } else {
        d1 = sqrt(d1);
        d2 = 0.5 * q / d1;
    }
    double q1 = d1 * d1, q2 = -0.25 * c1;
    int i = 0;
    p = q1 - 4.0 * (z - d2);
    if (p == 0) {
        res[i++] = -0.5 * d1 - q2;
    } else if (p > 0) {
        p = sqrt(p);
        res[i++] = -0.5 * (d1 + p) + q2;
        res[i++] = -0.5 * (d1 - p) + q2;
    }
    p = q1 - 4.0 * (z + d2);
    if (p == 0) {
        res[i++] = 0.5 * d1 - q2;
    } else if (p > 0) {
        p = sqrt(p);
        res[i++] = 0.5 * (d1 + p) + q2;
        res[i++] = 0.5 * (d1 - p) + q2;
    }
    return i;
}

// Solve_Polynomial for order 4 (polynomialsolver.cpp:1585-1732), c descending.
__device__ inline int solve4(const double *c0, double *r, bool sturm, double eps) {
    int i = 0, n = 4;
    while (i < n && fabs(c0[i]) < kSmall) i++;
    n -= i;
    const double *c = &c0[i];
    switch (n) {
    case 0:
        return 0;
    case 1:
        if (c[0] != 0.0) { r[0] = -c[1] / c[0]; return 1; }
        return 0;
    case 2:
        return solve_quadratic(c, r);
    case 3:
        if (eps > 0.0 && c[2] != 0.0 && fabs(c[3] / c[2]) < eps) return solve_quadratic(c, r);
        return sturm ? polysolve(3, c, r) : solve_cubic(c, r);
    default:
        if (eps > 0.0 && c[3] != 0.0 && fabs(c[4] / c[3]) < eps)
            return sturm ? polysolve(3, c, r) : solve_cubic(c, r);
        if (difficult_coeffs(c)) sturm = true;
        return sturm ? polysolve(4, c, r) : solve_quartic(c, r);
    }
}

// Solve_Polynomial with epsilon 0 for orders up to 6 (the lathe's and sor's
// calls, lathe.cpp:255-285 and sor.cpp:395), c descending.
__device__ inline int solve6(int n, const double *c0, double *r, bool sturm) {
    int i = 0;
    while (i < n && fabs(c0[i]) < kSmall) i++;
    n -= i;
    const double *c = &c0[i];
    switch (n) {
    case 0:
        return 0;
    case 1:
        if (c[0] != 0.0) { r[0] = -c[1] / c[0]; return 1; }
        return 0;
    case 2:
        return solve_quadratic(c, r);
    case 3:
        return sturm ? polysolve<3>(3, c, r) : solve_cubic(c, r);
    case 4:
        if (difficult_coeffs(c)) sturm = true;
        return sturm ? polysolve<4>(4, c, r) : solve_quartic(c, r);
    default:
        return polysolve<6>(n, c, r);
    }
}
} // namespace poly

// Torus::Test_Thick_Cylinder (core/shape/torus.cpp:932-1085)
__device__ inline bool thick_cylinder(D3 P, D3 D, double h1, double h2, double r1, double r2) {
    double k, u, v, r, h;
    if (fabs(D.y) < kEpsilon) {
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
        for (int w = 0; w < 2; w++) {
            double rr = w == 0 ? r2 : r1;
            double c = P.x * P.x + P.z * P.z - rr;
            double dd = b * b - a * c;
            if (dd >= 0.0) {
                dd = sqrt(dd);
                k = (-b + dd) / a;
                if (k > kEpsilon && k < kMaxDistance) {
                    h = P.y + k * D.y;
                    if (h >= h1 && h <= h2) return true;
                }
                k = (-b - dd) / a;
                if (k > kEpsilon && k < kMaxDistance) {
                    h = P.y + k * D.y;
                    if (h >= h1 && h <= h2) return true;
                }
            }
        }
    }
    return false;
}

// Torus::Intersect + All_Intersections (torus.cpp:133-160, 241-329).
__device__ inline int torus_hits(const TorusRec &tr, const Xform *xf, D3 o, D3 d, double *out) {
    const Xform &x = xf[tr.xf];
    D3 P = xf_point(x.inv, o);
    D3 D = xf_dir(x.inv, d);
    double len = sqrt(dot3(D, D));
    D = D3{D.x / len, D.y / len, D.z / len};
    double R = tr.R, rr = tr.r;
    double r1 = (R - rr) * (R - rr);
    if (R < rr) r1 = 0;
    double r2 = (R + rr) * (R + rr);
    if (!thick_cylinder(P, D, -rr, rr, r1, r2)) return 0;
    double bsr = R + rr + rr;
    double dist = dot3(P, P), closer = 0.0;
    if (dist > bsr * bsr) {
        dist = sqrt(dist);
        closer = dist - bsr;
        P = add3(P, mul3(D, closer));
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
    int nr = poly::solve4(c, roots, tr.sturm != 0, 1.0e-4);
    int n = 0;
    while (nr--) {
        double dep = (roots[nr] + closer) / len;
        if (dep > 1.0e-4 && dep < kMaxDistance) out[n++] = dep;
    }
    return n;
}

__device__ inline bool torus_inside(const TorusRec &tr, const Xform *xf, D3 p, bool inv) { // torus.cpp:364-385
    D3 P = xf_point(xf[tr.xf].inv, p);
    double r = sqrt(P.x * P.x + P.z * P.z);
    double r2 = P.y * P.y + (r - tr.R) * (r - tr.R);
    return (r2 <= tr.r * tr.r) ? !inv : inv;
}

// ---------------------------------------------------------------------------
// Cone / cylinder (core/shape/cone.cpp)
// ---------------------------------------------------------------------------
// Cone::Intersect (cone.cpp:140-300): the ray in canonical space (direction
// normalised, its length kept), side hits, then the cap (z = 1) and the base
// (z = dist) of a closed shape.  part: 1 base, 2 cap, 3 side (the
// Intersection::i1 Cone::Normal reads).  Returns the hit count (<= 4) in
// All_Intersections' push order.
__device__ inline int cone_hits(const ConeRec &c, const Xform *xf, D3 o, D3 d, double *out, int *part) {
    constexpr double kTol = 1.0e-9; // Cone_Tolerance (cone.cpp:64)
    const Xform &x = xf[c.xf];
    const D3 P = xf_point(x.inv, o);
    D3 D = xf_dir(x.inv, d);
    const double len = sqrt(D.x * D.x + D.y * D.y + D.z * D.z);
    D = D3{D.x / len, D.y / len, D.z / len};
    int n = 0;
    auto side = [&](double t, double zlo) {
        const double z = P.z + t * D.z;
        if (t > kTol && t < kMaxDistance && z >= zlo && z <= 1.0) {
            out[n] = t / len;
            part[n++] = 3;
        }
    };
    if (c.flags & kConeCylinder) {
        const double a = D.x * D.x + D.y * D.y;
        if (a > kEpsilon) {
            const double b = P.x * D.x + P.y * D.y;
            const double cc = P.x * P.x + P.y * P.y - 1.0;
            double dd = b * b - a * cc;
            if (dd >= 0.0) {
                dd = sqrt(dd);
                side((-b + dd) / a, 0.0);
                side((-b - dd) / a, 0.0);
            }
        }
    } else {
        const double a = D.x * D.x + D.y * D.y - D.z * D.z;
        const double b = D.x * P.x + D.y * P.y - D.z * P.z;
        const double cc = P.x * P.x + P.y * P.y - P.z * P.z;
        if (fabs(a) < kEpsilon) {
            if (fabs(b) > kEpsilon) side(-0.5 * cc / b, c.dist);
        } else {
            double dd = b * b - a * cc;
            if (dd >= 0.0) {
                dd = sqrt(dd);
                side((-b - dd) / a, c.dist);
                side((-b + dd) / a, c.dist);
            }
        }
    }
    if ((c.flags & kConeClosed) && fabs(D.z) > kEpsilon) {
        double t = (1.0 - P.z) / D.z;
        double u = P.x + t * D.x, v = P.y + t * D.y;
        if ((u * u + v * v) <= 1.0 && t > kTol && t < kMaxDistance) {
            out[n] = t / len;
            part[n++] = 2;
        }
        t = (c.dist - P.z) / D.z;
        u = P.x + t * D.x;
        v = P.y + t * D.y;
        if ((u * u + v * v) <= ((c.flags & kConeCylinder) ? 1.0 : c.dist * c.dist) && t > kTol && t < kMaxDistance) {
            out[n] = t / len;
            part[n++] = 1;
        }
    }
    return n;
}

// Cone::Inside (cone.cpp:360-410): a closed shape shrinks by EPSILON, an
// open one grows by it.
__device__ inline bool cone_inside(const ConeRec &c, const Xform *xf, D3 p, bool inv) {
    const double off = (c.flags & kConeClosed) ? -kEpsilon : kEpsilon;
    const D3 q = xf_point(xf[c.xf].inv, p);
    const double w2 = q.x * q.x + q.y * q.y;
    const bool out = (c.flags & kConeCylinder) ? (w2 > 1.0 + off || q.z < 0.0 - off || q.z > 1.0 + off)
                                               : (w2 > q.z * q.z + off || q.z < c.dist - off || q.z > 1.0 + off);
    return out ? inv : !inv;
}

// ---------------------------------------------------------------------------
// Disc (core/shape/disc.cpp)
// ---------------------------------------------------------------------------
// Disc::Intersect: the plane z = 0 of disc space, squared radius within
// [iradius2, oradius2], depth in (DEPTH_TOLERANCE = 1e-6, MAX_DISTANCE).
__device__ inline bool disc_hit(const DiscRec &c, const Xform *xf, D3 o, D3 d, double &depth) {
    const Xform &x = xf[c.xf];
    const D3 P = xf_point(x.inv, o);
    D3 D = xf_dir(x.inv, d);
    const double len = sqrt(D.x * D.x + D.y * D.y + D.z * D.z);
    D = D3{D.x / len, D.y / len, D.z / len};
    if (fabs(D.z) > kEpsilon) {
        const double t = -P.z / D.z;
        if (t >= 0.0) {
            const double u = P.x + t * D.x, v = P.y + t * D.y;
            const double r2 = u * u + v * v;
            if (r2 >= c.ir2 && r2 <= c.or2) {
                depth = t / len;
                return depth > 1.0e-6 && depth < kMaxDistance;
            }
        }
    }
    return false;
}

// Disc::Inside: the half-space z < 0 of disc space.
__device__ inline bool disc_inside(const DiscRec &c, const Xform *xf, D3 p, bool inv) {
    const D3 q = xf_point(xf[c.xf].inv, p);
    return q.z >= 0.0 ? inv : !inv;
}

// ---------------------------------------------------------------------------
// Quadric (core/shape/quadric.cpp)
// ---------------------------------------------------------------------------
// Quadric::Intersect + All_Intersections (quadric.cpp:100-200): the world
// ray in the quadratic a t^2 + 2 b t + c; roots kept in (1e-6, MAX_DISTANCE).
__device__ inline int quadric_hits(const QuadricRec &q, D3 o, D3 d, double *out) {
    const double a = d.x * (q.qa * d.x + q.qb * d.y + q.qc * d.z) + d.y * (q.qe * d.y + q.qf * d.z) + d.z * q.qh * d.z;
    const double b = d.x * (q.qa * o.x + 0.5 * (q.qb * o.y + q.qc * o.z + q.qd)) +
                     d.y * (q.qe * o.y + 0.5 * (q.qb * o.x + q.qf * o.z + q.qg)) +
                     d.z * (q.qh * o.z + 0.5 * (q.qc * o.x + q.qf * o.y + q.qi));
    const double c = o.x * (q.qa * o.x + q.qb * o.y + q.qc * o.z + q.qd) + o.y * (q.qe * o.y + q.qf * o.z + q.qg) +
                     o.z * (q.qh * o.z + q.qi) + q.qj;
    double t1, t2;
    if (a != 0.0) {
        double dd = b * b - a * c;
        if (dd <= 0.0) return 0;
        dd = sqrt(dd);
        t1 = (-b + dd) / a;
        t2 = (-b - dd) / a;
    } else {
        if (b == 0.0) return 0;
        t1 = -0.5 * c / b;
        t2 = kMaxDistance;
    }
    int n = 0;
    if (t1 > 1.0e-6 && t1 < kMaxDistance) out[n++] = t1;
    if (t2 > 1.0e-6 && t2 < kMaxDistance) out[n++] = t2;
    return n;
}

// Quadric::Inside (quadric.cpp): q(p) <= 0.  Invert negates every
// coefficient (Quadric::Invert), so an inverted quadric holds q(p) >= 0.
__device__ inline bool quadric_inside(const QuadricRec &q, D3 p, bool inv) {
    const double v = p.x * (q.qa * p.x + q.qb * p.y + q.qd) + p.y * (q.qe * p.y + q.qf * p.z + q.qg) +
                     p.z * (q.qh * p.z + q.qc * p.x + q.qi) + q.qj;
    return inv ? (-v <= 0.0) : (v <= 0.0);
}

// ---------------------------------------------------------------------------
// Polygon (core/shape/polygon.cpp)
// ---------------------------------------------------------------------------
// Polygon::in_polygon (polygon.cpp:905-984): crossings of the +u ray from
// (u, v) with the edges; a point equal to its sub-polygon's first point
// closes that sub-polygon and the next starts after it.
__device__ inline bool in_polygon(const double *pts, uint32_t number, double tx, double ty) {
    const double *vtx0 = pts, *vtx1 = pts + 2, *first = pts;
    bool yflag0 = vtx0[1] >= ty, inside = false;
    for (uint32_t i = 1; i < number;) {
        const bool yflag1 = vtx1[1] >= ty;
        if (yflag0 != yflag1 &&
            (((vtx1[1] - ty) * (vtx0[0] - vtx1[0]) >= (vtx1[0] - tx) * (vtx0[1] - vtx1[1])) == yflag1))
            inside = !inside;
        if (i + 2 < number && vtx1[0] == first[0] && vtx1[1] == first[1]) {
            vtx0 = pts + 2 * (size_t)(++i);
            vtx1 = pts + 2 * (size_t)(++i);
            yflag0 = vtx0[1] >= ty;
            first = vtx0;
        } else {
            vtx0 = vtx1;
            ++i;
            if (i < number) vtx1 = pts + 2 * (size_t)i;
            yflag0 = yflag1;
        }
    }
    return inside;
}

// Polygon::Intersect (polygon.cpp:187-230): the plane z = 0 of polygon
// space, depth in [1e-8, MAX_DISTANCE] before the direction's length is
// divided out, then the crossing test.
__device__ inline bool polygon_hit(const PolygonRec &pg, const double *pts, const Xform *xf, D3 o, D3 d,
                                   double &depth) {
    if (pg.degenerate) return false;
    const Xform &x = xf[pg.xf];
    const D3 P = xf_point(x.inv, o);
    D3 D = xf_dir(x.inv, d);
    const double len = sqrt(D.x * D.x + D.y * D.y + D.z * D.z);
    D = D3{D.x / len, D.y / len, D.z / len};
    if (fabs(D.z) < 1.0e-10) return false;
    double t = -P.z / D.z;
    if (t < 1.0e-8 || t > kMaxDistance) return false;
    if (!in_polygon(pts + 2 * (size_t)pg.first, pg.count, P.x + t * D.x, P.y + t * D.y)) return false;
    depth = t / len;
    return true;
}

// Light ray to grid point (u, v) of light `lt` from ip
// (TraceAreaLightSubsetShadowRay's grid points, trace.cpp:2131-2272): unit
// direction and distance.
__device__ inline void light_sample_ray(const LightRec &lt, uint32_t u, uint32_t v, D3 ip, D3 &dir, double &dist) {
    D3 tgt{lt.pos[0], lt.pos[1], lt.pos[2]};
    if (lt.n1 > 1) {
        double f = (double)u / (double)(lt.n1 - 1) - 0.5;
        tgt = add3(tgt, D3{lt.axis1[0] * f, lt.axis1[1] * f, lt.axis1[2] * f});
    }
    if (lt.n2 > 1) {
        double f = (double)v / (double)(lt.n2 - 1) - 0.5;
        tgt = add3(tgt, D3{lt.axis2[0] * f, lt.axis2[1] * f, lt.axis2[2] * f});
    }
    D3 sd = sub3(tgt, ip);
    dist = sqrt(dot3(sd, sd));
    dir = D3{sd.x / dist, sd.y / dist, sd.z / dist};
}

// ---------------------------------------------------------------------------
// Lathe (core/shape/lathe.cpp) and surface of revolution (core/shape/sor.cpp)
// with the bounding cylinders of core/bounding/boundingcylinder.cpp.  The
// ray moves into the shape's space with its direction normalised; each
// segment whose cylinder shell the ray crosses becomes an interval keyed by
// its entry distance; the intervals are visited in that order and a
// polynomial in the ray parameter is solved per segment.  A shape that is
// not a CSG operand stops once the next interval starts beyond its best hit.
// The ring and cylinder crossings of intersect_bound_elements are computed
// per segment from the same expressions (the values are identical).
// ---------------------------------------------------------------------------
constexpr double kRevDepthTol = 1.0e-4; // DEPTH_TOLERANCE (lathe.cpp:84, sor.cpp:133)
constexpr int kSorBase = 1, kSorCap = 2, kSorCurve = 3;

// Entry distance of the ray into segment g's cylinder shell
// (intersect_thick_cylinder + Intersect_BCyl's case analysis); false: missed.
__device__ inline bool rev_shell_entry(const RevSeg &g, const double *rad, const double *hgt, D3 P, D3 D, double a,
                                       double b, double b2, double bb, double c, bool dy, double &d0) {
    double dist[6];
    int n = 0;
    const double R1 = rad[g.r1], R2 = rad[g.r2], H1 = hgt[g.h1], H2 = hgt[g.h2];
    if (dy) { // cap ring, then base ring
        double k = (H2 - P.y) / D.y, w = k * (a * k + b2) + c;
        if (w >= R1 && w <= R2) dist[n++] = k;
        k = (H1 - P.y) / D.y;
        w = k * (a * k + b2) + c;
        if (w >= R1 && w <= R2) dist[n++] = k;
    }
    for (int s = 0; s < 2; s++) { // inner, then outer cylinder
        const double R = s == 0 ? R1 : R2;
        if (R > kEpsilon) {
            double e = bb - a * (c - R);
            if (e > 0.0) {
                e = sqrt(e);
                double k = (-b + e) / a, h = P.y + k * D.y;
                if (h >= H1 && h <= H2) dist[n++] = k;
                k = (-b - e) / a;
                h = P.y + k * D.y;
                if (h >= H1 && h <= H2) dist[n++] = k;
            }
        }
    }
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n - i - 1; j++)
            if (dist[j] > dist[j + 1]) {
                const double t = dist[j];
                dist[j] = dist[j + 1];
                dist[j + 1] = t;
            }
    if (n == 2 || n == 4) {
        for (int p = 0; p < n; p += 2) {
            if (dist[p] > kEpsilon) { d0 = dist[p]; return true; }
            if (dist[p + 1] > kEpsilon) { d0 = 0.0; return true; }
        }
        return false;
    }
    if (n == 0) return false;
    d0 = dist[0]; // an odd count: taken as starting inside the shell
    return true;
}

// Every hit Lathe::Intersect / Sor::Intersect pushes, in push order:
// emit(depth, side, d1) with side = segment (lathe) or part | segment << 2
// (sor) and d1 = the spline parameter (lathe) or the ray parameter (sor).
template <class Emit>
__device__ inline void rev_visit(const SceneView &sv, uint32_t index, D3 o, D3 dw, bool child, Emit &&emit) {
    const RevRec &s = sv.revs[index];
    const Xform &x = sv.xforms[s.xf];
    const D3 P = xf_point(x.inv, o);
    D3 D = xf_dir(x.inv, dw);
    const double len = sqrt(D.x * D.x + D.y * D.y + D.z * D.z);
    D = D3{D.x / len, D.y / len, D.z / len};
    if ((D.y >= 0.0 && P.y > s.height2) || (D.y <= 0.0 && P.y < s.height1) || (D.x >= 0.0 && P.x > s.radius2) ||
        (D.x <= 0.0 && P.x < -s.radius2))
        return;
    const bool sor = (s.flags & 7u) == PVX_SOR_CUBIC;
    const double a = D.x * D.x + D.z * D.z;
    double r0 = P.x * D.z - P.z * D.x;
    if (!sor) r0 = fabs(r0);
    if (a > 0.0) r0 /= sqrt(a);
    if (r0 > s.radius2) return;
    const bool sturm = (s.flags & kRevSturm) != 0;
    double best = kBoundHuge;
    auto test_hit = [&](double dep, int side, double d1) {
        if (dep > kRevDepthTol && dep < kMaxDistance) {
            emit(dep, side, d1);
            return true;
        }
        return false;
    };
    if (sor && (s.flags & kRevClosed) && fabs(D.y) > kEpsilon) { // base and cap discs
        if (s.base_r2 > kRevDepthTol) {
            const double k = (s.height1 - P.y) / D.y;
            const double u = P.x + k * D.x, v = P.z + k * D.z;
            if (u * u + v * v <= s.base_r2 && test_hit(k / len, kSorBase, k) && k < best) best = k;
        }
        if (s.cap_r2 > kRevDepthTol) {
            const double k = (s.height2 - P.y) / D.y;
            const double u = P.x + k * D.x, v = P.z + k * D.z;
            if (u * u + v * v <= s.cap_r2 && test_hit(k / len, kSorCap, k) && k < best) best = k;
        }
    }
    // Intersect_BCyl: intervals ordered by entry distance, a new one before
    // any equal one (insert_hit)
    const double b = P.x * D.x + P.z * D.z, bb = b * b, b2 = 2.0 * b, c = P.x * P.x + P.z * P.z;
    const bool dy = D.y < -kEpsilon || D.y > kEpsilon;
    double iv_d[kRevMaxSegments];
    uint8_t iv_n[kRevMaxSegments];
    int cnt = 0;
    for (uint32_t i = 0; i < s.nseg; i++) {
        double d0;
        if (!rev_shell_entry(sv.rev_segs[s.seg0 + i], sv.rev_radius, sv.rev_height, P, D, a, b, b2, bb, c, dy, d0))
            continue;
        int pos = cnt;
        while (pos > 0 && !(d0 > iv_d[pos - 1])) {
            iv_d[pos] = iv_d[pos - 1];
            iv_n[pos] = iv_n[pos - 1];
            pos--;
        }
        iv_d[pos] = d0;
        iv_n[pos] = (uint8_t)i;
        cnt++;
    }
    const double m = D.x * P.x + D.z * P.z, Dy2 = D.y * D.y, Dylen = D.y * len;
    const uint32_t kind = s.flags & 7u;
    for (int j = 0; j < cnt; j++) {
        if (!child && iv_d[j] > best) break;
        const int seg = iv_n[j];
        const RevSeg &g = sv.rev_segs[s.seg0 + seg];
        if (sor) { // Sor::Intersect (sor.cpp:380-420): the cubic r^2(h) along the ray
            const double A = g.a[0], B = g.b[0], C = g.c[0], Dd = g.d[0];
            double xc[4], yr[3];
            xc[0] = A * D.y * D.y * D.y;
            xc[1] = D.y * D.y * (3.0 * A * P.y + B) - a;
            xc[2] = D.y * (P.y * (3.0 * A * P.y + 2.0 * B) + C) - 2.0 * (P.x * D.x + P.z * D.z);
            xc[3] = P.y * (P.y * (A * P.y + B) + C) + Dd - P.x * P.x - P.z * P.z;
            int n = poly::solve6(3, xc, yr, sturm);
            const double H1 = sv.rev_height[g.h1], H2 = sv.rev_height[g.h2];
            while (n--) {
                const double k = yr[n], h = P.y + k * D.y;
                if (h >= H1 && h <= H2 && test_hit(k / len, kSorCurve | (seg << 2), k) && k < best) best = k;
            }
            continue;
        }
        // Lathe::Intersect (lathe.cpp:245-350): the spline's radius against
        // the ray's distance from the axis at the spline's height
        const double Ax = g.a[0], Ay = g.a[1], Bx = g.b[0], By = g.b[1], Cx = g.c[0], Cy = g.c[1], Dx = g.d[0],
                     Dy = g.d[1];
        const double r = a;
        double x1[7], y1[6];
        int n1;
        if (kind == PVX_LATHE_LINEAR) {
            x1[0] = Cy * Cy * r - Cx * Cx * Dy2;
            x1[1] = 2.0 * (Cy * ((Dy - P.y) * r + D.y * m) - Cx * Dx * Dy2);
            x1[2] = (Dy - P.y) * ((Dy - P.y) * r + 2.0 * D.y * m) + Dy2 * (P.x * P.x + P.z * P.z - Dx * Dx);
            n1 = poly::solve6(2, x1, y1, false);
        } else if (kind == PVX_LATHE_QUADRATIC) {
            x1[0] = By * By * r - Bx * Bx * Dy2;
            x1[1] = 2.0 * (By * Cy * r - Bx * Cx * Dy2);
            x1[2] = r * (2.0 * By * (Dy - P.y) + Cy * Cy) + 2.0 * By * D.y * m - (2.0 * Bx * Dx + Cx * Cx) * Dy2;
            x1[3] = 2.0 * (Cy * ((Dy - P.y) * r + D.y * m) - Cx * Dx * Dy2);
            x1[4] = (Dy - P.y) * ((Dy - P.y) * r + 2.0 * D.y * m) + Dy2 * (P.x * P.x + P.z * P.z - Dx * Dx);
            n1 = poly::solve6(4, x1, y1, sturm);
        } else {
            x1[0] = Ay * Ay * r - Ax * Ax * Dy2;
            x1[1] = 2.0 * (Ay * By * r - Ax * Bx * Dy2);
            x1[2] = (2.0 * Ay * Cy + By * By) * r - (2.0 * Ax * Cx + Bx * Bx) * Dy2;
            x1[3] = 2.0 * ((Ay * Dy + By * Cy - Ay * P.y) * r + Ay * D.y * m - (Ax * Dx + Bx * Cx) * Dy2);
            x1[4] = (2.0 * By * (Dy - P.y) + Cy * Cy) * r + 2.0 * By * D.y * m - (2.0 * Bx * Dx + Cx * Cx) * Dy2;
            x1[5] = 2.0 * (Cy * ((Dy - P.y) * r + D.y * m) - Cx * Dx * Dy2);
            x1[6] = (Dy - P.y) * ((Dy - P.y) * r + 2.0 * D.y * m) + Dy2 * (P.x * P.x + P.z * P.z - Dx * Dx);
            n1 = poly::solve6(6, x1, y1, sturm);
        }
        while (n1--) {
            const double w = y1[n1];
            if (w < 0.0 || w > 1.0) continue;
            if (fabs(D.y) > kEpsilon) {
                double k = w * (w * (w * Ay + By) + Cy) + Dy - P.y;
                if (test_hit(k / Dylen, seg, w)) {
                    k /= D.y;
                    if (k < best) best = k;
                }
            } else {
                const double k = w * (w * (w * Ax + Bx) + Cx) + Dx;
                double x2[3] = {r, 2.0 * m, P.x * P.x + P.z * P.z - k * k}, y2[2];
                int n2 = poly::solve6(2, x2, y2, false);
                while (n2--)
                    if (test_hit(y2[n2] / len, seg, w) && y2[n2] < best) best = y2[n2];
            }
        }
    }
}

// Lathe::Inside (lathe.cpp:380-450) / Sor::Inside (sor.cpp:455-505).
__device__ inline bool rev_inside(const SceneView &sv, uint32_t index, D3 ip, bool inv) {
    const RevRec &s = sv.revs[index];
    const D3 P = xf_point(sv.xforms[s.xf].inv, ip);
    if ((s.flags & 7u) == PVX_SOR_CUBIC) {
        if (P.y >= s.height1 && P.y <= s.height2) {
            const double r0 = P.x * P.x + P.z * P.z;
            if (r0 <= s.radius2 * s.radius2) {
                for (uint32_t i = 0; i < s.nseg; i++) {
                    const RevSeg &g = sv.rev_segs[s.seg0 + i];
                    if (P.y >= sv.rev_height[g.h1] && P.y <= sv.rev_height[g.h2]) {
                        const double r = P.y * (P.y * (P.y * g.a[0] + g.b[0]) + g.c[0]) + g.d[0];
                        if (r0 <= r) return !inv;
                        break;
                    }
                }
            }
        }
        return inv;
    }
    int nc = 0;
    if (P.y >= s.height1 && P.y <= s.height2) {
        const double r = sqrt(P.x * P.x + P.z * P.z);
        if (r >= s.radius1 && r <= s.radius2) {
            const bool sturm = (s.flags & kRevSturm) != 0;
            for (uint32_t i = 0; i < s.nseg; i++) {
                const RevSeg &g = sv.rev_segs[s.seg0 + i];
                if (P.y >= sv.rev_height[g.h1] - kEpsilon && P.y <= sv.rev_height[g.h2] + kEpsilon) {
                    double xc[4] = {g.a[1], g.b[1], g.c[1], g.d[1] - P.y}, yr[3];
                    int n = poly::solve6(3, xc, yr, sturm);
                    while (n--) {
                        const double w = yr[n];
                        if (w >= 0.0 && w <= 1.0 && w * (w * (w * g.a[0] + g.b[0]) + g.c[0]) + g.d[0] - r >= 0.0) nc++;
                    }
                }
            }
        }
    }
    return (nc & 1) ? !inv : inv;
}

// ---------------------------------------------------------------------------
// CSG (core/shape/csg.cpp): point membership with bounded nesting.
// ---------------------------------------------------------------------------
// kRev: lathes and sors may be operands (the walks of scenes without them
// leave their code out, which keeps those kernels' register budget).
template <int D, bool kRev = false>
__device__ bool csg_inside(const SceneView &sv, uint32_t node, D3 p);

template <int D, bool kRev = false>
__device__ inline bool shape_inside(const SceneView &sv, uint32_t type, uint32_t index,
                                    uint32_t flags, D3 p) {
    bool inv = (flags & PVX_FLAG_INVERTED) != 0;
    switch (type) {
    case 1: { // sphere
        const double *s = sv.sph + 4 * (size_t)index;
        return sphere_inside(d3(s[0], s[1], s[2]), s[3], p, inv);
    }
    case 2:
        return box_inside(sv.boxes[index], sv.xforms, p, inv);
    case 3:
        return torus_inside(sv.tori[index], sv.xforms, p, inv);
    case 9:
        return cone_inside(sv.cones[index], sv.xforms, p, inv);
    case 10:
        return disc_inside(sv.discs[index], sv.xforms, p, inv);
    case 11:
        return quadric_inside(sv.quadrics[index], p, inv);
    case 4: { // plane: Plane::Inside (plane.cpp:192-215)
        const PlaneRec &pl = sv.planes[index];
        double t = p.x * pl.n[0] + p.y * pl.n[1] + p.z * pl.n[2];
        bool in = (t + pl.d) < kEpsilon;
        return inv ? !in : in;
    }
    case 6:
        return csg_inside<D, kRev>(sv, index, p);
    case PVX_LATHE:
    case PVX_SOR:
        if constexpr (kRev) return rev_inside(sv, index, p, inv);
        return inv;
    default:
        return inv; // meshes without inside_vector
    }
}

template <int D, bool kRev>
__device__ bool csg_inside(const SceneView &sv, uint32_t node, D3 p) {
    if constexpr (D <= 0) {
        return false;
    } else {
        const CsgNode &n = sv.csg[node];
        for (uint32_t k = 0; k < n.count; k++) {
            const CsgChild &c = sv.csg_children[n.first + k];
            bool in = shape_inside<D - 1, kRev>(sv, c.type, c.index, c.flags, p);
            if (n.op == 2) { // intersection: CSGIntersection::Inside (csg.cpp:440-454)
                if (!in) return false;
            } else if (in) { // union / merge: CSGUnion::Inside (csg.cpp:405-424)
                return true;
            }
        }
        return n.op == 2;
    }
}

// Hits of one leaf shape (sphere/box/torus/cone/disc/quadric/triangle) in
// push order (an inverted quadric pushes its two roots the other way round;
// the set is the same).
__device__ inline int leaf_hits(const SceneView &sv, uint32_t type, uint32_t index, D3 o, D3 d,
                                double *out) {
    int side[2];
    switch (type) {
    case 1: {
        const double *s = sv.sph + 4 * (size_t)index;
        return sphere_hits(d3(s[0], s[1], s[2]), s[3], o, d, out);
    }
    case 2:
        return box_hits(sv.boxes[index], sv.xforms, o, d, out, side);
    case 3:
        return torus_hits(sv.tori[index], sv.xforms, o, d, out);
    case 9: {
        int part[4];
        return cone_hits(sv.cones[index], sv.xforms, o, d, out, part);
    }
    case 10:
        return disc_hit(sv.discs[index], sv.xforms, o, d, out[0]) ? 1 : 0;
    case 11:
        return quadric_hits(sv.quadrics[index], o, d, out);
    case 12: // polygon (not a CSG operand; All_Intersections)
        return polygon_hit(sv.polygons[index], sv.poly_pts, sv.xforms, o, d, out[0]) ? 1 : 0;
    case 7:
    case 8: { // stand-alone triangle (not a CSG child; All_Intersections)
        double dep;
        if (!tri_rec_hit(sv.triangles[index], o, d, dep)) return 0;
        out[0] = dep;
        return 1;
    }
    default:
        return 0;
    }
}

// Every hit of one leaf shape in push order, as emit(depth, side, d1); `child`:
// the shape is a CSG operand (IS_CHILD_OBJECT).
template <bool kRev, class Emit>
__device__ inline void leaf_visit(const SceneView &sv, uint32_t type, uint32_t index, D3 o, D3 d, bool child,
                                  Emit &&emit) {
    if constexpr (kRev) {
        if (type == PVX_LATHE || type == PVX_SOR) {
            rev_visit(sv, index, o, d, child, emit);
            return;
        }
    }
    double hs[kMaxLeafHits];
    const int nh = leaf_hits(sv, type, index, o, d, hs);
    for (int j = 0; j < nh; j++) emit(hs[j], 0, 0.0);
}

// Does a hit at depth `dep` on the leaf in csg_children slot `slot` survive
// every CSG filter between the leaf and the root (CSGIntersection keeps
// hits inside all siblings, CSGMerge hits inside none, CSGUnion all;
// csg.cpp:128-375)?
template <bool kRev = false>
__device__ inline bool csg_survives(const SceneView &sv, uint32_t slot, D3 ip) {
    uint32_t child = slot;
    for (int level = 0; level < kCsgMaxDepth; level++) {
        const CsgChild &cc = sv.csg_children[child];
        const CsgNode &n = sv.csg[cc.node];
        if (n.op != 1) {
            bool is_merge = (n.op == 3);
            for (uint32_t k = 0; k < n.count; k++) {
                uint32_t sib = n.first + k;
                if (sib == child) continue;
                const CsgChild &s = sv.csg_children[sib];
                bool in = shape_inside<kCsgMaxDepth, kRev>(sv, s.type, s.index, s.flags, ip);
                if (is_merge ? in : !in) return false;
            }
        }
        if (n.parent_slot < 0) return true;
        child = (uint32_t)n.parent_slot;
    }
    return true;
}

} // namespace pvx

#endif
