// Oracle (test infrastructure only): exact 2-D geometric predicates for the CPU
// restatement of the reference's Delaunay test-and-repair.
//
// The reference evaluates these predicates inside CGAL 4.9's
// Exact_predicates_inexact_constructions_kernel (reference DelaunayCGAL.h:18-42).
// CGAL is not vendored under the reference, so this file restates the published
// method instead: floating-point filter first, then exact evaluation with
// non-overlapping floating-point expansions (J. R. Shewchuk, "Adaptive Precision
// Floating-Point Arithmetic and Fast Robust Geometric Predicates", Discrete Comput.
// Geom. 18:305-363, 1997: TWO-SUM, TWO-PRODUCT, GROW-EXPANSION, SCALE-EXPANSION).
// TWO-PRODUCT uses a fused multiply-add, which is exact.
//
// All predicates take points relative to the cell at the origin, exactly as the
// reference's local triangulations are built from minimum-image displacements.
// Pinned by tests/test_oracle_predicates.py (brute-force rational arithmetic on
// random and near-degenerate inputs).
#pragma once
#include <cmath>

namespace orc {

static inline void two_sum(double a, double b, double &x, double &y) {
    x = a + b;
    double bv = x - a;
    double av = x - bv;
    double br = b - bv;
    double ar = a - av;
    y = ar + br;
}
static inline void fast_two_sum(double a, double b, double &x, double &y) {
    x = a + b;
    double bv = x - a;
    y = b - bv;
}
static inline void two_prod(double a, double b, double &x, double &y) {
    x = a * b;
    y = std::fma(a, b, -x);
}

// h = e + b ; returns length (zero components eliminated)
static inline int grow_expansion(int elen, const double *e, double b, double *h) {
    double Q = b, Qn, hh;
    int hi = 0;
    for (int i = 0; i < elen; ++i) {
        two_sum(Q, e[i], Qn, hh);
        Q = Qn;
        if (hh != 0.0) h[hi++] = hh;
    }
    if (Q != 0.0 || hi == 0) h[hi++] = Q;
    return hi;
}
// h = e + f (repeated growth; output non-overlapping, increasing magnitude)
static inline int expansion_sum(int elen, const double *e, int flen, const double *f, double *h) {
    double tmp[256];
    int hlen = elen;
    for (int i = 0; i < elen; ++i) tmp[i] = e[i];
    for (int j = 0; j < flen; ++j) {
        double out[256];
        hlen = grow_expansion(hlen, tmp, f[j], out);
        for (int i = 0; i < hlen; ++i) tmp[i] = out[i];
    }
    for (int i = 0; i < hlen; ++i) h[i] = tmp[i];
    return hlen;
}
// h = b * e
static inline int scale_expansion(int elen, const double *e, double b, double *h) {
    double Q, sum, hh, p1, p0;
    int hi = 0;
    two_prod(e[0], b, Q, hh);
    if (hh != 0.0) h[hi++] = hh;
    for (int i = 1; i < elen; ++i) {
        two_prod(e[i], b, p1, p0);
        two_sum(Q, p0, sum, hh);
        if (hh != 0.0) h[hi++] = hh;
        fast_two_sum(p1, sum, Q, hh);
        if (hh != 0.0) h[hi++] = hh;
    }
    if (Q != 0.0 || hi == 0) h[hi++] = Q;
    return hi;
}
static inline int negate(int elen, double *e) {
    for (int i = 0; i < elen; ++i) e[i] = -e[i];
    return elen;
}
// sign of a non-overlapping expansion = sign of its largest-magnitude (last) component
static inline int expansion_sign(int elen, const double *e) {
    for (int i = elen - 1; i >= 0; --i) {
        if (e[i] > 0) return 1;
        if (e[i] < 0) return -1;
    }
    return 0;
}
// h = (ax*by - ay*bx) exactly
static inline int cross_exact(double ax, double ay, double bx, double by, double *h) {
    double p[2], q[2];
    two_prod(ax, by, p[1], p[0]);
    two_prod(ay, bx, q[1], q[0]);
    negate(2, q);
    // order components by magnitude for the sum routine: both pairs are non-overlapping
    double t1[2] = {p[0], p[1]};
    double t2[2] = {q[0], q[1]};
    return expansion_sum(2, t1, 2, t2, h);
}
// h = x^2 + y^2 exactly
static inline int norm2_exact(double x, double y, double *h) {
    double p[2], q[2];
    two_prod(x, x, p[1], p[0]);
    two_prod(y, y, q[1], q[0]);
    return expansion_sum(2, p, 2, q, h);
}
// h = e * f (f an expansion)
static inline int mul_expansion(int elen, const double *e, int flen, const double *f, double *h) {
    double acc[256];
    int alen = 0;
    for (int j = 0; j < flen; ++j) {
        double part[256];
        int plen = scale_expansion(elen, e, f[j], part);
        double out[256];
        if (alen == 0) {
            for (int i = 0; i < plen; ++i) acc[i] = part[i];
            alen = plen;
        } else {
            alen = expansion_sum(alen, acc, plen, part, out);
            for (int i = 0; i < alen; ++i) acc[i] = out[i];
        }
    }
    for (int i = 0; i < alen; ++i) h[i] = acc[i];
    return alen;
}

// sign of cross(a, b): >0 when b lies counter-clockwise of a as seen from the origin
static inline int orient0(double ax, double ay, double bx, double by) {
    double l = ax * by, r = ay * bx;
    double det = l - r;
    double bound = 1e-14 * (std::fabs(l) + std::fabs(r));
    if (det > bound) return 1;
    if (-det > bound) return -1;
    double h[8];
    int n = cross_exact(ax, ay, bx, by, h);
    return expansion_sign(n, h);
}

// >0 when d lies strictly inside the circle through the origin, b and c
// (for counter-clockwise (0, b, c)); <0 outside; 0 cocircular.
static inline int incircle0(double bx, double by, double cx, double cy, double dx, double dy) {
    double b2 = bx * bx + by * by, c2 = cx * cx + cy * cy, d2 = dx * dx + dy * dy;
    double m1 = cy * d2 - c2 * dy;
    double m2 = cx * d2 - c2 * dx;
    double m3 = cx * dy - cy * dx;
    double det = bx * m1 - by * m2 + b2 * m3;
    double perm = std::fabs(bx) * (std::fabs(cy * d2) + std::fabs(c2 * dy)) +
                  std::fabs(by) * (std::fabs(cx * d2) + std::fabs(c2 * dx)) +
                  b2 * (std::fabs(cx * dy) + std::fabs(cy * dx));
    double bound = 1e-13 * perm;
    if (det > bound) return -1;
    if (-det > bound) return 1;
    // exact: det = bx*(cy*d2 - c2*dy) - by*(cx*d2 - c2*dx) + b2*(cx*dy - cy*dx)
    double B2[8], C2[8], D2[8];
    int lb = norm2_exact(bx, by, B2), lc = norm2_exact(cx, cy, C2), ld = norm2_exact(dx, dy, D2);
    double t1[64], t2[64], t3[64], u[128], M1[128], M2[128], M3[16];
    int l1 = scale_expansion(ld, D2, cy, t1);
    int l2 = scale_expansion(lc, C2, dy, t2);
    negate(l2, t2);
    int lm1 = expansion_sum(l1, t1, l2, t2, M1);
    l1 = scale_expansion(ld, D2, cx, t1);
    l2 = scale_expansion(lc, C2, dx, t2);
    negate(l2, t2);
    int lm2 = expansion_sum(l1, t1, l2, t2, M2);
    int lm3 = cross_exact(cx, cy, dx, dy, M3);
    double A[256], Bv[256], Cv[256], S[256], T[256];
    int la = scale_expansion(lm1, M1, bx, A);
    int lbv = scale_expansion(lm2, M2, by, Bv);
    negate(lbv, Bv);
    int lcv = mul_expansion(lm3, M3, lb, B2, Cv);
    int ls = expansion_sum(la, A, lbv, Bv, S);
    int lt = expansion_sum(ls, S, lcv, Cv, T);
    (void)u; (void)t3;
    return -expansion_sign(lt, T);
}

// exact comparison of |a|^2 and |b|^2: returns sign(|a|^2 - |b|^2)
static inline int cmp_norm2(double ax, double ay, double bx, double by) {
    double a2 = ax * ax + ay * ay, b2 = bx * bx + by * by;
    double diff = a2 - b2;
    double bound = 1e-14 * (a2 + b2);
    if (diff > bound) return 1;
    if (-diff > bound) return -1;
    double A[8], B[8], H[16];
    int la = norm2_exact(ax, ay, A), lb = norm2_exact(bx, by, B);
    negate(lb, B);
    int lh = expansion_sum(la, A, lb, B, H);
    return expansion_sign(lh, H);
}

}  // namespace orc
