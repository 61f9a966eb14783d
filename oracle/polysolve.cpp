// FP64 CPU ORACLE — TEST INFRASTRUCTURE ONLY (see oracle.hpp).
// Restatement of core/math/polynomialsolver.cpp: Sturm-sequence root
// isolation (polysolve, buildsturm, modp, visible_roots, numchanges, sbisect,
// regula_falsa, polyeval) and the closed-form quadratic/cubic/quartic
// solvers, with the branches the reference compiles by default
// (TEST_SOLVER and USE_NEW_DIFFICULT_COEFFS undefined).
#include <atomic>
#include <cmath>
#include <cstdint>

#include "oracle.hpp"

namespace orc {

// How often Solve_Polynomial took each order-4 branch (for tests that show
// which branches a scene's rays reach): [0] closed-form quartic,
// [1] difficult_coeffs forcing Sturm, [2] the quartic's d1 clamp
// (-SMALL_ENOUGH < d1 < 0 set to 0), [3] Sturm asked for by the caller.
std::atomic<uint64_t> g_branch[4];

namespace {

const double kFudge1 = 1.0e12;           // polynomialsolver.cpp FUDGE_FACTOR1
const double kTwoPi3 = 2.0943951023931954923084;
const double kFourPi3 = 4.1887902047863909846168;
const int kMaxIter = 50;                 // MAX_ITERATIONS
const double kSmallEnough = 1.0e-10;     // SMALL_ENOUGH
const double kRelError = 1.0e-12;        // RELERROR
const int kMaxOrder = 35;

struct Poly {
    int ord;
    double coef[kMaxOrder + 1]; // ascending powers
};

double polyeval(double x, int n, const double *c) { // polynomialsolver.cpp:600-612
    double v = c[n];
    for (int i = n - 1; i >= 0; i--) v = v * x + c[i];
    return v;
}

int modp(const Poly *u, const Poly *v, Poly *r) { // :171-219
    *r = *u;
    if (v->coef[v->ord] < 0.0) {
        for (int k = u->ord - v->ord - 1; k >= 0; k -= 2) r->coef[k] = -r->coef[k];
        for (int k = u->ord - v->ord; k >= 0; k--)
            for (int j = v->ord + k - 1; j >= k; j--)
                r->coef[j] = -r->coef[j] - r->coef[v->ord + k] * v->coef[j - k];
    } else {
        for (int k = u->ord - v->ord; k >= 0; k--)
            for (int j = v->ord + k - 1; j >= k; j--)
                r->coef[j] -= r->coef[v->ord + k] * v->coef[j - k];
    }
    int k = v->ord - 1;
    while (k >= 0 && std::fabs(r->coef[k]) < kSmallEnough) {
        r->coef[k] = 0.0;
        k--;
    }
    r->ord = (k < 0) ? 0 : k;
    return r->ord;
}

int buildsturm(int ord, Poly *sseq) { // :244-281
    sseq[0].ord = ord;
    sseq[1].ord = ord - 1;
    double f = std::fabs(sseq[0].coef[ord] * ord);
    for (int i = 1; i <= ord; i++) sseq[1].coef[i - 1] = sseq[0].coef[i] * i / f;
    Poly *sp;
    for (sp = sseq + 2; modp(sp - 2, sp - 1, sp); sp++) {
        f = -std::fabs(sp->coef[sp->ord]);
        for (double *fp = &sp->coef[sp->ord]; fp >= sp->coef; fp--) *fp /= f;
    }
    sp->coef[0] = -sp->coef[0];
    return (int)(sp - sseq);
}

int visible_roots(int np, const Poly *sseq, int *atzer, int *atpos) { // :303-342
    int atposinf = 0, atzero = 0;
    double lf = sseq[0].coef[sseq[0].ord], f;
    for (const Poly *s = sseq + 1; s <= sseq + np; s++) {
        f = s->coef[s->ord];
        if (lf == 0.0 || lf * f < 0) atposinf++;
        lf = f;
    }
    lf = sseq[0].coef[0];
    for (const Poly *s = sseq + 1; s <= sseq + np; s++) {
        f = s->coef[0];
        if (lf == 0.0 || lf * f < 0) atzero++;
        lf = f;
    }
    *atzer = atzero;
    *atpos = atposinf;
    return atzero - atposinf;
}

int numchanges(int np, const Poly *sseq, double a) { // :364-387
    int changes = 0;
    double lf = polyeval(a, sseq[0].ord, sseq[0].coef), f;
    for (const Poly *s = sseq + 1; s <= sseq + np; s++) {
        f = polyeval(a, s->ord, s->coef);
        if (lf == 0.0 || lf * f < 0) changes++;
        lf = f;
    }
    return changes;
}

int regula_falsa(int order, const double *coef, double a, double b, double *val) { // :634-760
    double fa = polyeval(a, order, coef), fb = polyeval(b, order, coef), x, fx, lfx;
    if (fa * fb > 0.0) return 0;
    if (std::fabs(fa) < kSmallEnough) {
        *val = a;
        return 1;
    }
    if (std::fabs(fb) < kSmallEnough) {
        *val = b;
        return 1;
    }
    lfx = fa;
    for (int its = 0; its < kMaxIter; its++) {
        x = (fb * a - fa * b) / (fb - fa);
        fx = polyeval(x, order, coef);
        if (std::fabs(x) > kRelError) {
            if (std::fabs(fx / x) < kRelError) {
                *val = x;
                return 1;
            }
        } else if (std::fabs(fx) < kRelError) {
            *val = x;
            return 1;
        }
        if (fa < 0) {
            if (fx < 0) {
                a = x;
                fa = fx;
                if ((lfx * fx) > 0) fb /= 2;
            } else {
                b = x;
                fb = fx;
                if ((lfx * fx) > 0) fa /= 2;
            }
        } else {
            if (fx < 0) {
                b = x;
                fb = fx;
                if ((lfx * fx) > 0) fa /= 2;
            } else {
                a = x;
                fa = fx;
                if ((lfx * fx) > 0) fb /= 2;
            }
        }
        if (std::fabs(b - a) < kRelError) {
            *val = x;
            return 1;
        }
        lfx = fx;
    }
    return 0;
}

int sbisect(int np, const Poly *sseq, double minv, double maxv, int atmin, int atmax,
            double *roots) { // :409-578
    double mid = 0;
    int its, atmid;
    if ((atmin - atmax) == 1) {
        if (regula_falsa(sseq->ord, sseq->coef, minv, maxv, roots)) return 1;
        for (its = 0; its < kMaxIter; its++) {
            mid = (minv + maxv) / 2;
            atmid = numchanges(np, sseq, mid);
            if ((atmid < atmax) || (atmid > atmin)) return 0;
            if (std::fabs(mid) > kRelError) {
                if (std::fabs((maxv - minv) / mid) < kRelError) {
                    roots[0] = mid;
                    return 1;
                }
            } else if (std::fabs(maxv - minv) < kRelError) {
                roots[0] = mid;
                return 1;
            }
            if ((atmin - atmid) == 0)
                minv = mid;
            else
                maxv = mid;
        }
        roots[0] = mid;
        return 1;
    }
    for (its = 0; its < kMaxIter; its++) {
        mid = (minv + maxv) / 2;
        atmid = numchanges(np, sseq, mid);
        if ((atmid < atmax) || (atmid > atmin)) return 0;
        if (std::fabs(mid) > kRelError) {
            if (std::fabs((maxv - minv) / mid) < kRelError) {
                roots[0] = mid;
                return 1;
            }
        } else if (std::fabs(maxv - minv) < kRelError) {
            roots[0] = mid;
            return 1;
        }
        int n1 = atmin - atmid, n2 = atmid - atmax;
        if (n1 != 0 && n2 != 0) {
            n1 = sbisect(np, sseq, minv, mid, atmin, atmid, roots);
            n2 = sbisect(np, sseq, mid, maxv, atmid, atmax, &roots[n1]);
            return n1 + n2;
        }
        if (n1 == 0)
            minv = mid;
        else
            maxv = mid;
    }
    roots[0] = mid;
    return 1;
}

int solve_quadratic(const double *x, double *y) { // :780-833
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
    if ((d > -kSmallEnough) && (d < kSmallEnough)) {
        y[0] = 0.5 * b / a;
        return 1;
    }
    if (d < 0.0) return 0;
    d = std::sqrt(d);
    double t = 2.0 * a;
    y[0] = (b + d) / t;
    y[1] = (b - d) / t;
    return 2;
}

int solve_cubic(const double *x, double *y) { // :855-944
    double a0 = x[0], a1, a2, a3;
    if (a0 == 0.0) return solve_quadratic(&x[1], y);
    if (a0 != 1.0) {
        a1 = x[1] / a0;
        a2 = x[2] / a0;
        a3 = x[3] / a0;
    } else {
        a1 = x[1];
        a2 = x[2];
        a3 = x[3];
    }
    double A2 = a1 * a1;
    double Q = (A2 - 3.0 * a2) / 9.0;
    double R = (a1 * (A2 - 4.5 * a2) + 13.5 * a3) / 27.0;
    double Q3 = Q * Q * Q, R2 = R * R, d = Q3 - R2, an = a1 / 3.0;
    if (d >= 0.0) {
        d = R / std::sqrt(Q3);
        double theta = std::acos(d) / 3.0;
        double sQ = -2.0 * std::sqrt(Q);
        y[0] = sQ * std::cos(theta) - an;
        y[1] = sQ * std::cos(theta + kTwoPi3) - an;
        y[2] = sQ * std::cos(theta + kFourPi3) - an;
        return 3;
    }
    double sQ = std::pow(std::sqrt(R2 - Q3) + std::fabs(R), 1.0 / 3.0);
    if (R < 0)
        y[0] = (sQ + Q / sQ) - an;
    else
        y[0] = -(sQ + Q / sQ) - an;
    return 1;
}

// difficult_coeffs, default (#else) branch (:1047-1085): note `biggest`
// keeps the SIGNED value of the largest-magnitude coefficient.
int difficult_coeffs(int n, const double *x) {
    double biggest = 0.0;
    for (int i = 0; i <= n; i++)
        if (std::fabs(x[i]) > biggest) biggest = x[i];
    if (biggest == 0.0) return 0;
    for (int i = 0; i <= n; i++)
        if (x[i] != 0.0 && std::fabs(biggest / x[i]) > kFudge1) return 1;
    return 0;
}

int solve_quartic(const double *x, double *results) { // #else branch :1325-1450
    double cubic[4], roots[3], c1, c2, c3, c4;
    double c0 = x[0];
    if (c0 != 1.0) {
        c1 = x[1] / c0;
        c2 = x[2] / c0;
        c3 = x[3] / c0;
        c4 = x[4] / c0;
    } else {
        c1 = x[1];
        c2 = x[2];
        c3 = x[3];
        c4 = x[4];
    }
    double c12 = c1 * c1;
    double p = -0.375 * c12 + c2;
    double q = 0.125 * c12 * c1 - 0.5 * c1 * c2 + c3;
    double r = -0.01171875 * c12 * c12 + 0.0625 * c12 * c2 - 0.25 * c1 * c3 + c4;
    cubic[0] = 1.0;
    cubic[1] = -0.5 * p;
    cubic[2] = -r;
    cubic[3] = 0.5 * r * p - 0.125 * q * q;
    int i = solve_cubic(cubic, roots);
    if (i <= 0) return 0;
    double z = roots[0];
    double d1 = 2.0 * z - p, d2;
    if (d1 < 0.0) {
        if (d1 > -kSmallEnough) {
            g_branch[2]++;
            d1 = 0.0;
        } else {
            return 0;
        }
    }
    if (d1 < kSmallEnough) {
        d2 = z * z - r;
        if (d2 < 0.0) return 0;
        d2 = std::sqrt(d2);
    } else {
        d1 = std::sqrt(d1);
        d2 = 0.5 * q / d1;
    }
    double q1 = d1 * d1, q2 = -0.25 * c1;
    i = 0;
    p = q1 - 4.0 * (z - d2);
    if (p == 0) {
        results[i++] = -0.5 * d1 - q2;
    } else if (p > 0) {
        p = std::sqrt(p);
        results[i++] = -0.5 * (d1 + p) + q2;
        results[i++] = -0.5 * (d1 - p) + q2;
    }
    p = q1 - 4.0 * (z + d2);
    if (p == 0) {
        results[i++] = 0.5 * d1 - q2;
    } else if (p > 0) {
        p = std::sqrt(p);
        results[i++] = 0.5 * (d1 + p) + q2;
        results[i++] = 0.5 * (d1 - p) + q2;
    }
    return i;
}

int polysolve(int order, const double *coeffs, double *roots) { // :1476-1521
    Poly sseq[kMaxOrder + 1];
    for (int i = 0; i <= order; i++) sseq[0].coef[order - i] = coeffs[i] / coeffs[0];
    int np = buildsturm(order, sseq);
    int atmin, atmax;
    if (visible_roots(np, sseq, &atmin, &atmax) == 0) return 0;
    double minv = 0.0, maxv = kMaxDistance;
    atmin = numchanges(np, sseq, minv);
    atmax = numchanges(np, sseq, maxv);
    if (atmin - atmax == 0) return 0;
    return sbisect(np, sseq, minv, maxv, atmin, atmax, roots);
}

} // namespace

// Solve_Polynomial (:1585-1732); c0 in descending powers, c0[0] leading.
int solve_polynomial(int n, const double *c0, double *r, bool sturm, double epsilon) {
    int roots = 0, i = 0;
    while (i < n && std::fabs(c0[i]) < kSmallEnough) i++;
    n -= i;
    const double *c = &c0[i];
    switch (n) {
    case 0:
        break;
    case 1:
        if (c[0] != 0.0) r[roots++] = -c[1] / c[0];
        break;
    case 2:
        roots = solve_quadratic(c, r);
        break;
    case 3:
        if (epsilon > 0.0 && c[2] != 0.0 && std::fabs(c[3] / c[2]) < epsilon) {
            roots = solve_quadratic(c, r);
            break;
        }
        roots = sturm ? polysolve(3, c, r) : solve_cubic(c, r);
        break;
    case 4:
        if (epsilon > 0.0 && c[3] != 0.0 && std::fabs(c[4] / c[3]) < epsilon) {
            roots = sturm ? polysolve(3, c, r) : solve_cubic(c, r);
            break;
        }
        if (sturm) {
            g_branch[3]++;
        } else if (difficult_coeffs(4, c)) {
            g_branch[1]++;
            sturm = true;
        } else {
            g_branch[0]++;
        }
        roots = sturm ? polysolve(4, c, r) : solve_quartic(c, r);
        break;
    default:
        if (epsilon > 0.0 && c[n - 1] != 0.0 && std::fabs(c[n] / c[n - 1]) < epsilon) {
            roots = polysolve(n - 1, c, r);
            break;
        }
        roots = polysolve(n, c, r);
        break;
    }
    return roots;
}

} // namespace orc

// Single branches of the solver, for tests that pin which branch
// Solve_Polynomial takes: difficult_coeffs, the Ferrari/Vieta quartic and
// the Sturm isolation (coefficients in descending powers).
extern "C" int orc_difficult_coeffs(int n, const double *c) { return orc::difficult_coeffs(n, c); }
extern "C" int orc_solve_quartic(const double *c, double *roots) { return orc::solve_quartic(c, roots); }
extern "C" int orc_polysolve(int order, const double *c, double *roots) { return orc::polysolve(order, c, roots); }
// Reads and clears the order-4 branch counters (four entries, see g_branch).
extern "C" void orc_solver_branches(uint64_t *out) {
    for (int i = 0; i < 4; i++) out[i] = orc::g_branch[i].exchange(0);
}
