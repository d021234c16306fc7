// Device-side building blocks shared by the cellgpu_amd HIP kernels (gfx950).
//
// Periodic box, Voronoi geometry primitives, exact orientation / in-circle
// predicates, and the counter-based normal generator. Everything here is
// compiled with -ffp-contract=off so each FP64 expression rounds exactly as
// written; FP32 kernels opt into fused multiply-adds explicitly where wanted.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define CGPU_HD __host__ __device__ inline

namespace cgpu {

template <typename R> struct vec2 { R x, y; };
using dvec2 = vec2<double>;

// ------------------------------------------------------------------ periodic box
// Square periodic domain. Same semantics as the reference gpubox
// (inc/utility/gpubox.h:75-84, :139-152, :154-179): minimum image by repeated
// +-L, wrap into [0,L) through the unit square (x*(1/L) -> wrap -> *L).
struct Box {
    double L, half, inv;
};
// XCD-aware block index. Workgroups are dealt round-robin over the 8 XCDs, each
// with its own L2; the remap gives every XCD one contiguous range of logical
// blocks, so cells that are neighbours in memory share one L2.
__device__ inline int xcd_block() {
    const int nb = (int)gridDim.x, b = (int)blockIdx.x;
    const int per = nb >> 3, rem = nb & 7, x = b & 7, k = b >> 3;
    return x * per + min(x, rem) + k;
}
CGPU_HD Box make_box(double L) { return Box{L, L * 0.5, 1. / L}; }
// The wrap loops are capped (WRAP_MAX iterations): for coordinates inside the box
// they run at most twice, and a non-finite or runaway value can never spin.
constexpr int WRAP_MAX = 64;
CGPU_HD dvec2 min_dist(const Box &b, dvec2 p1, dvec2 p2) {
    dvec2 a{p1.x - p2.x, p1.y - p2.y};
    // Two points of the box [0, L) differ by less than L per axis: each loop
    // below then runs at most once and only one of the two per axis, so a
    // select gives the same bits without the loop control flow.
    if (fabs(a.x) < b.L && fabs(a.y) < b.L) {
        a.x = a.x < -b.half ? a.x + b.L : (a.x > b.half ? a.x - b.L : a.x);
        a.y = a.y < -b.half ? a.y + b.L : (a.y > b.half ? a.y - b.L : a.y);
        return a;
    }
    for (int t = 0; t < WRAP_MAX && a.x < -b.half; ++t) a.x += b.L;
    for (int t = 0; t < WRAP_MAX && a.y < -b.half; ++t) a.y += b.L;
    for (int t = 0; t < WRAP_MAX && a.x > b.half; ++t) a.x -= b.L;
    for (int t = 0; t < WRAP_MAX && a.y > b.half; ++t) a.y -= b.L;
    return a;
}
// Fold a unit-square coordinate into [0,1) with the same rounding as the
// reference loops "while(v<0) v+=1; while(v>=1) v-=1" (gpubox.h:139-152), in
// closed form: every +-1 step but the last one that crosses zero is exact
// (|v| < 2^52), so the loops reduce to one exact integer shift plus that step.
CGPU_HD double fold_unit(double v) {
    if (v < 0) {
        double k = ceil(-v) - 1.0;  // exact steps that keep v negative
        v += k;
        if (v < 0) v += 1.0;  // the rounded step
    }
    if (v >= 1.0) v -= floor(v);
    return v;
}
CGPU_HD void put_in_box(const Box &b, dvec2 &p) {
    double vx = fold_unit(b.inv * p.x), vy = fold_unit(b.inv * p.y);
    p.x = b.L * vx;
    p.y = b.L * vy;
}

// ----------------------------------------------------------- geometry primitives
// Circumcenter of (origin, x1, x2) — reference inc/utility/functions.h:34-43.
template <typename R> CGPU_HD vec2<R> circ0(vec2<R> x1, vec2<R> x2) {
    R n1 = x1.x * x1.x + x1.y * x1.y;
    R n2 = x2.x * x2.x + x2.y * x2.y;
    R den = R(0.5) / (x1.x * x2.y - x1.y * x2.x);
    return vec2<R>{den * (n1 * x2.y - x1.y * n2), den * (x1.x * n2 - n1 * x2.x)};
}
// Circumcenter of (x1, x2, x3) — reference functions.h:46-62.
template <typename R> CGPU_HD vec2<R> circ3(vec2<R> x1, vec2<R> x2, vec2<R> x3) {
    R ax = x1.x - x3.x, ay = x1.y - x3.y, an = ax * ax + ay * ay;
    R bx = x2.x - x3.x, by = x2.y - x3.y, bn = bx * bx + by * by;
    R den = R(0.5) / (ax * by - ay * bx);
    return vec2<R>{x3.x + den * (an * by - ay * bn), x3.y + den * (ax * bn - an * bx)};
}
template <typename R> CGPU_HD R tri_area(vec2<R> a, vec2<R> b) {
    R v = R(0.5) * (a.x * b.y - a.y * b.x);
    return v < R(0) ? -v : v;
}
template <typename R> CGPU_HD R vlen(R x, R y) { return sqrt(x * x + y * y); }

// ------------------------------------------------------------- exact predicates
// Floating-point filter, then exact evaluation with non-overlapping expansions
// (Shewchuk 1997: two-sum, two-product via FMA, grow / scale expansion). Inputs
// are minimum-image displacements relative to the cell at the origin.
namespace exact {
CGPU_HD void two_sum(double a, double b, double &x, double &y) {
    x = a + b;
    double bv = x - a, av = x - bv;
    y = (a - av) + (b - bv);
}
CGPU_HD void two_prod(double a, double b, double &x, double &y) {
    x = a * b;
    y = fma(a, b, -x);
}
// e <- e + b (in place, zero elimination); returns new length
CGPU_HD int grow(int n, double *e, double b) {
    double Q = b, Qn, h;
    int k = 0;
    for (int i = 0; i < n; ++i) {
        two_sum(Q, e[i], Qn, h);
        Q = Qn;
        if (h != 0.0) e[k++] = h;
    }
    if (Q != 0.0 || k == 0) e[k++] = Q;
    return k;
}
// acc <- acc + b * (x, with x given as a hi/lo pair list)
CGPU_HD int add_scaled(int n, double *acc, const double *x, int xn, double b) {
    for (int i = 0; i < xn; ++i) {
        double p, q;
        two_prod(x[i], b, p, q);
        n = grow(n, acc, q);
        n = grow(n, acc, p);
    }
    return n;
}
CGPU_HD int sign_of(int n, const double *e) {
    for (int i = n - 1; i >= 0; --i) {
        if (e[i] > 0) return 1;
        if (e[i] < 0) return -1;
    }
    return 0;
}
// exact |v|^2 as up to 4 components
CGPU_HD int norm2(double x, double y, double *e) {
    double p, q, r, s;
    two_prod(x, x, p, q);
    two_prod(y, y, r, s);
    int n = 0;
    e[n++] = q;
    n = grow(n, e, p);
    n = grow(n, e, s);
    n = grow(n, e, r);
    return n;
}
// exact a.x*b.y - a.y*b.x as up to 4 components
CGPU_HD int cross(double ax, double ay, double bx, double by, double *e) {
    double p, q, r, s;
    two_prod(ax, by, p, q);
    two_prod(ay, bx, r, s);
    int n = 0;
    e[n++] = q;
    n = grow(n, e, p);
    n = grow(n, e, -s);
    n = grow(n, e, -r);
    return n;
}
static __host__ __device__ __noinline__ int orient_exact(double ax, double ay, double bx, double by) {
    double e[8];
    int n = cross(ax, ay, bx, by, e);
    return sign_of(n, e);
}
// sign of  |b|^2 (c x d) + |c|^2 (d x b) + |d|^2 (b x c)
static __host__ __device__ __noinline__ int incircle_exact(double bx, double by, double cx, double cy, double dx,
                                                     double dy) {
    // every grow() adds at most one component: norm2 and cross give at most 4
    // each, add_scaled adds 2 per pair, so the three terms reach at most
    // 3 * 4 * 4 * 2 = 96 (the array is per-lane stack, kept to that bound)
    double acc[96];
    int n = 0;
    double nrm[8], crs[8];
    int nn, nc;
    // |b|^2 * (c x d)
    nn = norm2(bx, by, nrm);
    nc = cross(cx, cy, dx, dy, crs);
    for (int j = 0; j < nc; ++j) n = add_scaled(n, acc, nrm, nn, crs[j]);
    nn = norm2(cx, cy, nrm);
    nc = cross(dx, dy, bx, by, crs);
    for (int j = 0; j < nc; ++j) n = add_scaled(n, acc, nrm, nn, crs[j]);
    nn = norm2(dx, dy, nrm);
    nc = cross(bx, by, cx, cy, crs);
    for (int j = 0; j < nc; ++j) n = add_scaled(n, acc, nrm, nn, crs[j]);
    return sign_of(n, acc);
}
static __host__ __device__ __noinline__ int cmp_norm2_exact(double ax, double ay, double bx, double by) {
    double acc[16];
    int n = norm2(ax, ay, acc);
    double p, q;
    two_prod(bx, bx, p, q);
    n = grow(n, acc, -q);
    n = grow(n, acc, -p);
    two_prod(by, by, p, q);
    n = grow(n, acc, -q);
    n = grow(n, acc, -p);
    return sign_of(n, acc);
}
}  // namespace exact

// >0 if b is counter-clockwise of a (seen from the origin)
CGPU_HD int orient0(double ax, double ay, double bx, double by) {
    double l = ax * by, r = ay * bx, det = l - r;
    double bound = 1e-14 * (fabs(l) + fabs(r));
    if (det > bound) return 1;
    if (-det > bound) return -1;
    return exact::orient_exact(ax, ay, bx, by);
}
// >0 if d lies strictly inside the circle through (origin, b, c), (0,b,c) CCW
CGPU_HD int incircle0(double bx, double by, double cx, double cy, double dx, double dy) {
    double b2 = bx * bx + by * by, c2 = cx * cx + cy * cy, d2 = dx * dx + dy * dy;
    double x1 = cx * dy - cy * dx, x2 = dx * by - dy * bx, x3 = bx * cy - by * cx;
    double det = b2 * x1 + c2 * x2 + d2 * x3;
    double perm = b2 * (fabs(cx * dy) + fabs(cy * dx)) + c2 * (fabs(dx * by) + fabs(dy * bx)) +
                  d2 * (fabs(bx * cy) + fabs(by * cx));
    double bound = 1e-13 * perm;
    // the determinant is negative when d is inside
    if (det > bound) return -1;
    if (-det > bound) return 1;
    return -exact::incircle_exact(bx, by, cx, cy, dx, dy);
}
// sign(|a|^2 - |b|^2), exact
CGPU_HD int cmp_norm2(double ax, double ay, double bx, double by) {
    double a2 = ax * ax + ay * ay, b2 = bx * bx + by * by, d = a2 - b2;
    double bound = 1e-14 * (a2 + b2);
    if (d > bound) return 1;
    if (-d > bound) return -1;
    return exact::cmp_norm2_exact(ax, ay, bx, by);
}

// -------------------------------------------------- counter-based normal deviates
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as
// 1, 2, 3", SC'11) keyed by a 64-bit seed, counter (index, step, stream, 0).
// Two 53-bit uniforms -> Box-Muller in FP64. The same function serves the
// pre-generated noise buffer and in-kernel generation, so both give equal bits.
CGPU_HD void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}
// returns a pair of independent N(0,1) deviates for (index, step, stream)
CGPU_HD void normal_pair(uint64_t seed, uint32_t index, uint32_t step, uint32_t stream, double &z0, double &z1) {
    uint32_t c[4] = {index, step, stream, 0u};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    uint64_t a = ((uint64_t)c[0] << 21) ^ (uint64_t)(c[1] >> 11);  // 53 bits
    uint64_t b = ((uint64_t)c[2] << 21) ^ (uint64_t)(c[3] >> 11);
    a &= ((uint64_t)1 << 53) - 1;
    b &= ((uint64_t)1 << 53) - 1;
    const double s = 1.0 / 9007199254740992.0;  // 2^-53
    double u1 = ((double)a + 1.0) * s;          // (0,1]
    double u2 = (double)b * s;                  // [0,1)
    double r = sqrt(-2.0 * log(u1));
    double t = 6.283185307179586476925286766559 * u2;
    z0 = r * cos(t);
    z1 = r * sin(t);
}

}  // namespace cgpu
