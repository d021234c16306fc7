// HIP kernels of the Voronoi-model time step for gfx950 (MI355X), and the
// kernel-caller half of the C ABI declared in include/cellgpu_amd.h.
//
// Layout: structure-of-arrays over all cells of all systems of a batch; rings
// are slot-major (nbr[k*N + i]) so that a wavefront reading slot k of 64
// consecutive cells issues one coalesced load. Positions stay FP64 in every
// mode; precision 1 runs geometry, forces and the integrator in FP32 while the
// Delaunay test, the star construction and the cell list stay FP64.
//
// One step = force/EOM/move -> bin -> test+geometry -> repair chain. The repair
// chain is a fixed sequence of kernels that exit immediately when no system
// needs them, so a whole step is a fixed launch sequence that can be captured
// in a hipGraph with no host round trip.
#include <hip/hip_runtime.h>
#include <climits>
#include <mutex>
#include <string>
#include <type_traits>

#include "../../include/cellgpu_amd.h"
#include "cgpu_device.h"
#include "fire_device.h"

namespace cgpu_host {
void set_last_error(const char *what);  // message of cgpu_last_error (host layer)
}

using namespace cgpu;

namespace {

// Ring storage: ring_cap(b) slots per cell, slot-major. Register copies of a
// ring (serial star forms) hold CAP_MAX entries; the wavefront forms hold one
// neighbour per lane.
constexpr int CAP_MAX = CGPU_RING_CAP_MAX;
__host__ __device__ inline int ring_cap(const cgpu_voronoi_batch &b) {
    return b.ring_cap <= 0 ? CGPU_RING_CAP : (b.ring_cap < CAP_MAX ? b.ring_cap : CAP_MAX);
}

// ------------------------------------------------------------- launch shapes
// A zero (or unsupported) field of the batch's launch shape takes the default.
int g_block = 256;  // default cell_block (cgpu_set_block_size)
__host__ __device__ inline int shape_queue_lanes(const cgpu_voronoi_batch &b) {
    const int q = b.shape.queue_lanes;
    return (q == 4 || q == 8 || q == 16) ? q : 8;
}
__host__ __device__ inline int log2_lanes(int q) { return q == 4 ? 2 : (q == 8 ? 3 : 4); }
inline int shape_block3(int v) { return (v == 128 || v == 256 || v == 512) ? v : 256; }
inline int shape_force_block(const cgpu_voronoi_batch &b) { return shape_block3(b.shape.force_block); }
inline int shape_test_block(const cgpu_voronoi_batch &b) { return shape_block3(b.shape.test_block); }
inline int shape_cell_block(const cgpu_voronoi_batch &b) {
    const int v = b.shape.cell_block;
    return (v >= 64 && v <= 1024 && v % 64 == 0) ? v : g_block;
}
inline int shape_star_blocks(const cgpu_voronoi_batch &b) {
    const int v = b.shape.star_blocks;
    return (v >= 64 && v <= 65536) ? v : 4096;
}
inline int shape_repair_blocks(const cgpu_voronoi_batch &b) {
    const int v = b.shape.repair_blocks;
    return (v >= 64 && v <= 65536) ? v : 1024;
}

// Optional per-launch hook of a profiled step: after each kernel of the step
// the host records an event (cgpu_vor_spv_steps_profiled). Slot ids follow the
// launch order of one step; a slot may hold several kernels (the cell-list scan).
enum ProbeSlot {
    PROBE_FORCE_EOM_MOVE = 0,  // force, EOM, move, reset, binning pass
    PROBE_BIN_SCAN,            // cell-list prefix sum (one or three kernels)
    PROBE_BIN_SCATTER,
    PROBE_TEST_GEOMETRY,       // geometry and the skip rule's checks
    PROBE_LD_REDUCE,           // the rule's per-system sums and the stamps
    PROBE_SELECT_TRIANGLES,    // the rule of each system, the triangles it leaves, queued
    PROBE_TEST_QUEUE,          // their circumcircle test
    PROBE_MARK_NF,
    PROBE_STAR_LOCAL,
    PROBE_DELSETS_LOCAL,
    PROBE_GLOBAL_FALLBACK,
    PROBE_STAR_GLOBAL,
    PROBE_GLOBAL_STATS,
    PROBE_FINALIZE,
    PROBE_SLOTS
};
static_assert(PROBE_SLOTS == CGPU_PROFILE_SLOTS, "profile slots of the C ABI");
struct Probe {
    hipEvent_t *ev;  // ev[slot + 1] is recorded after slot's launches
    hipStream_t st;
    void mark(int slot) const { (void)hipEventRecord(ev[slot + 1], st); }
};
inline void probe(const Probe *p, int slot) {
    if (p) p->mark(slot);
}

__device__ inline dvec2 ldpos(const double *p, int i) {
    double2 v = reinterpret_cast<const double2 *>(p)[i];
    return dvec2{v.x, v.y};
}
__device__ inline void stpos(double *p, int i, dvec2 v) {
    reinterpret_cast<double2 *>(p)[i] = make_double2(v.x, v.y);
}
template <typename R> __device__ inline vec2<R> ld2(const void *a, int i) {
    const R *p = reinterpret_cast<const R *>(a);
    return vec2<R>{p[2 * i], p[2 * i + 1]};
}
template <> __device__ inline vec2<float> ld2<float>(const void *a, int i) {
    float2 v = reinterpret_cast<const float2 *>(a)[i];
    return vec2<float>{v.x, v.y};
}
template <> __device__ inline vec2<double> ld2<double>(const void *a, int i) {
    double2 v = reinterpret_cast<const double2 *>(a)[i];
    return vec2<double>{v.x, v.y};
}
template <typename R> __device__ inline void st2(void *a, int i, vec2<R> v) {
    R *p = reinterpret_cast<R *>(a);
    p[2 * i] = v.x;
    p[2 * i + 1] = v.y;
}
template <> __device__ inline void st2<float>(void *a, int i, vec2<float> v) {
    reinterpret_cast<float2 *>(a)[i] = make_float2(v.x, v.y);
}
template <> __device__ inline void st2<double>(void *a, int i, vec2<double> v) {
    reinterpret_cast<double2 *>(a)[i] = make_double2(v.x, v.y);
}
template <typename R> __device__ inline vec2<R> cvt(dvec2 d) { return vec2<R>{(R)d.x, (R)d.y}; }

// wave-aggregated reservation: lanes with pred get consecutive slots of *counter
// (one atomic per wave instead of one per lane); returns -1 for the others
__device__ inline int wave_reserve(bool pred, int *counter) {
    unsigned long long m = __ballot(pred);
    if (m == 0ull) return -1;
    int lane = __lane_id();
    int leader = __ffsll((long long)m) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(counter, __popcll(m));
    base = __shfl(base, leader);
    return pred ? base + __popcll(m & ((1ull << lane) - 1ull)) : -1;
}
// ------------------------------------------------- precision-dispatched pieces
// FP64 path: the reference expressions, evaluated exactly as written.
// FP32 path: the same expressions in float, except that a circumcenter (or the
// Voronoi-vertex Jacobian) of an ill-conditioned triangle — cross product below
// 1e-2 of |a|_1 |b|_1, i.e. an angle under about half a degree — is evaluated in
// FP64 from the FP64 displacements and rounded once.
template <typename R> struct Prec;
template <> struct Prec<double> {
    static __device__ inline dvec2 c0(dvec2 a, dvec2 b) { return circ0<double>(a, b); }
    static __device__ inline dvec2 c3(dvec2 a, dvec2 b, dvec2 c) { return circ3<double>(a, b, c); }
};
template <> struct Prec<float> {
    static __device__ inline bool well(float cr, vec2<float> a, vec2<float> b) {
        float sc = (fabsf(a.x) + fabsf(a.y)) * (fabsf(b.x) + fabsf(b.y));
        return fabsf(cr) > 1e-2f * sc;
    }
    static __device__ inline vec2<float> c0(dvec2 A, dvec2 B) {
        vec2<float> a = cvt<float>(A), b = cvt<float>(B);
        if (well(a.x * b.y - a.y * b.x, a, b)) return circ0<float>(a, b);
        return cvt<float>(circ0<double>(A, B));
    }
    static __device__ inline vec2<float> c3(dvec2 A, dvec2 B, dvec2 C) {
        vec2<float> a = cvt<float>(A), b = cvt<float>(B), c = cvt<float>(C);
        vec2<float> u{a.x - c.x, a.y - c.y}, v{b.x - c.x, b.y - c.y};
        if (well(u.x * v.y - u.y * v.x, u, v)) return circ3<float>(a, b, c);
        return cvt<float>(circ3<double>(A, B, C));
    }
};

// a * b + c: fused (one rounding) in the FP32 path; in FP64 the product and the
// sum round separately, as in the reference expression written out (c + a b and
// a b + c are the same sum)
template <typename R> __device__ inline R madd(R a, R b, R c);
template <> __device__ inline float madd<float>(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
template <> __device__ inline double madd<double>(double a, double b, double c) { return a * b + c; }

// reference getdhdr (functions.h:234-242): d(Voronoi vertex)/d(r_i) for the
// vertex of (i, j, k) with rij, rik relative to i; m = (x11, x12, x21, x22)
template <typename R> __device__ inline void dhdr_eval(vec2<R> rij, vec2<R> rik, R m[4]) {
    R c = rij.y * rik.x - rij.x * rik.y;
    R den = R(1.0) / (c * c);
    R rik2 = rik.x * rik.x + rik.y * rik.y;
    R rij2 = rij.x * rij.x + rij.y * rij.y;
    m[0] = R(0.5) * den * (rik.y - rij.y) * (rik.y * rij2 - rij.y * rik2);
    m[1] = R(0.5) * den * (rij.y - rik.y) * (rik.x * rij2 - rij.x * rik2);
    m[2] = R(0.5) * den * (rij.x - rik.x) * (rik.y * rij2 - rij.y * rik2);
    m[3] = R(0.5) * den * (rik.x - rij.x) * (rik.x * rij2 - rij.x * rik2);
}
template <typename R> __device__ inline void dhdr(dvec2 rij, dvec2 rik, R m[4]);
template <> __device__ inline void dhdr<double>(dvec2 rij, dvec2 rik, double m[4]) { dhdr_eval<double>(rij, rik, m); }
template <> __device__ inline void dhdr<float>(dvec2 rij, dvec2 rik, float m[4]) {
    vec2<float> a = cvt<float>(rij), b = cvt<float>(rik);
    if (Prec<float>::well(a.y * b.x - a.x * b.y, a, b)) {
        dhdr_eval<float>(a, b, m);
    } else {
        double md[4];
        dhdr_eval<double>(rij, rik, md);
        for (int q = 0; q < 4; ++q) m[q] = (float)md[q];
    }
}
// reference float builds clamp edge lengths at THRESHOLD = 1e-18
// (voronoiQuadraticEnergy.cu:172-177, 201-204); double builds do not
template <typename R> __device__ inline R clamp_len(R x) { return x; }
template <> __device__ inline float clamp_len<float>(float x) { return x < 1e-18f ? 1e-18f : x; }

// ------------------------------------------------------- position accessors
// Gathers of neighbour data go through an accessor: GlobalData reads HBM (via
// the caches); BlockTile first looks in an LDS copy of the block's own cells,
// which with Hilbert-ordered cells holds most Delaunay neighbours.
template <typename R> struct GlobalData {
    const double *pos;
    const void *dAP;
    __device__ dvec2 p(int id) const { return ldpos(pos, id); }
    __device__ vec2<R> q(int id) const { return ld2<R>(dAP, id); }
};
template <typename R> struct BlockTile {
    const double *pos;
    const void *dAP;
    const dvec2 *sp;     // LDS positions of cells c0 .. c0 + n - 1
    const vec2<R> *sq;   // LDS dAP of the same cells (may be null)
    int c0, n;
    __device__ dvec2 p(int id) const {
        const unsigned t = (unsigned)(id - c0);
        return t < (unsigned)n ? sp[t] : ldpos(pos, id);
    }
    __device__ vec2<R> q(int id) const {
        const unsigned t = (unsigned)(id - c0);
        return t < (unsigned)n ? sq[t] : ld2<R>(dAP, id);
    }
};

// ------------------------------------------------- skip rule of the circumcircle test
// If every edge of the triangulation that meets a triangle's circumdisk is
// locally Delaunay (the far vertex of its other triangle lies outside the
// circle), no point lies inside the circle: the segment from the triangle to
// such a point stays in the disk and crosses an edge that is not locally
// Delaunay (the classical walk argument). The reference test then flags nothing
// for that triangle. The geometry pass (cell_geometry<R, true>) checks every
// spoke of its cell and stamps the bins around each spoke that fails (with a
// margin far above rounding) and around each ring that is not a proper
// counter-clockwise fan; the reference scan then runs only for the triangles
// of cells near a stamp (k_select_triangles, k_test_queue). Per system it also sums a hash of the
// ring triangles over all slots and over the owned slots, and their area: the
// rule holds only where the rings are consistent (every triangle in the rings
// of its three cells, all = 3 x owned) and cover the box once (area L^2).
constexpr double LD_MARGIN = 1e-9;  // power of the far vertex, squared cell spacings
constexpr int MARK_DILATE = 3;      // bins of bin_mark_wide around a stamped region
__device__ inline int quadrant(dvec2 d);

struct LdSums {
    unsigned long long all, own;
    long long area;  // units of 2^-30 (integer sums: exact, native atomics)
};
constexpr double AREA_UNIT = 1073741824.0;  // 2^30
__device__ inline unsigned long long mix64(unsigned long long x) {
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebull;
    x ^= x >> 31;
    return x;
}
// hash of the counter-clockwise triangle (a, b, c) rotated to start at its
// smallest index (the same triangle from each of its cells)
__device__ inline unsigned long long tri_hash(int a, int b, int c) {
    if (b < a && b < c) {
        const int t = a;
        a = b;
        b = c;
        c = t;
    } else if (c < a && c < b) {
        const int t = c;
        c = b;
        b = a;
        a = t;
    }
    // (a, b) packed, c spread by an odd multiplier (a bijection), then mixed
    return mix64(((unsigned long long)(unsigned)a << 32 | (unsigned)b) ^ ((unsigned long long)(unsigned)c * 0x9e3779b97f4a7c15ull));
}
// A region for k_ld_reduce to stamp: the rectangle [x0, x1] x [y0, y1] of system
// sys, or (sys = -1 - i) the bounding box of cell i's star. The geometry pass
// only appends records (no loops, no calls); a record that does not fit turns
// the rule off for its system.
struct LdRegion {
    int sys, pad;
    double x0, x1, y0, y1;
};
static_assert(sizeof(LdRegion) == CGPU_LD_REGION_BYTES, "region record size of the C ABI");
__device__ inline void push_region(const cgpu_voronoi_batch &b, int s, int sys, double x0, double x1, double y0,
                                   double y1) {
    const int slot = atomicAdd(&b.counters[8], 1);
    if (slot >= b.ld_region_cap) {
        b.st[s].skip_off = 1;
        return;
    }
    reinterpret_cast<LdRegion *>(b.ld_regions)[slot] = LdRegion{sys, 0, x0, x1, y0, y1};
}
// spoke (i, q) of the triangle whose circle is centred at c (relative to i):
// a region unless the far vertex f lies outside the circle by LD_MARGIN (a
// circle wider than 4 bins counts too: its power is not resolved there)
__device__ inline void check_spoke(const cgpu_voronoi_batch &b, const cgpu_sys_params &sp, int s, dvec2 pi,
                                   vec2<double> c, dvec2 q, dvec2 f) {
    const double pw = f.x * (f.x - 2.0 * c.x) + f.y * (f.y - 2.0 * c.y);
    const double r2 = c.x * c.x + c.y * c.y;
    if (pw > LD_MARGIN && r2 < 16.0 * sp.boxsize * sp.boxsize) return;
    push_region(b, s, s, pi.x + fmin(0.0, q.x), pi.x + fmax(0.0, q.x), pi.y + fmin(0.0, q.y), pi.y + fmax(0.0, q.y));
}

// the skip rule's per-cell state: fan order and winding, largest squared
// circumradius
struct LdFan {
    bool bad;
    int wraps;
    float rmax2;
};
// one ring pair (a, c) = (ring k-1, ring k) with indices ia, ic and circumcenter v
__device__ __forceinline__ void ld_pair(LdFan &f, LdSums &sums, int i, dvec2 a, dvec2 c, int ia, int ic,
                                        vec2<double> v) {
    // a pair not clearly counter-clockwise (FP64, with a margin) counts as a defect
    if (!(a.x * c.y - a.y * c.x > 1e-12 * (a.x * a.x + a.y * a.y + c.x * c.x + c.y * c.y))) f.bad = true;
    f.wraps += quadrant(c) < quadrant(a);
    f.rmax2 = fmaxf(f.rmax2, (float)(v.x * v.x + v.y * v.y));
    const unsigned long long h = tri_hash(i, ia, ic);
    sums.all += h;
    if (i < ia && i < ic) {
        sums.own += h;
        sums.area += llrint(0.5 * AREA_UNIT * (a.x * c.y - a.y * c.x));
    }
}

// ------------------------------------------------------------------ geometry
// Voronoi geometry of cell i (reference gpu_compute_voronoi_geometry_kernel,
// voronoiModelBase.cu:115-201): vertex k = circumcenter(ring k-1, ring k), the
// circumcenters of the cell and consecutive ring pairs. Area and perimeter are
// evaluated in FP64 in both precision modes and rounded once at the store: the
// force uses KA (A - A0) and KP (P - P0), which in relaxed tissue are a few
// percent of A and P, so FP32 vertices would leave them with about 1e-5
// relative error (the dominant FP32 force error). With LD also the skip rule's
// checks of the cell's spokes and fan, its largest circumradius (cell_rmax)
// and its triangles' hash and area sums.
template <typename R, bool LD = false>
__device__ __forceinline__ void cell_geometry(const cgpu_voronoi_batch &b, const double *pos, int i,
                                              const cgpu_sys_params &sp, LdSums *sums = nullptr) {
    using G = double;
    const Box box = make_box(sp.L);
    const int N = b.num_cells;
    const int z = b.nn[i];
    if (z < 3) {
        if (LD) b.st[b.cell_sys[i]].skip_off = 1;
        return;
    }
    const int s = LD ? b.cell_sys[i] : 0;
    dvec2 pi = ldpos(pos, i);
    const int jlast = b.nbr[(z - 1) * N + i], jfirst = b.nbr[i];
    dvec2 rij = min_dist(box, ldpos(pos, jlast), pi);
    dvec2 rik = min_dist(box, ldpos(pos, jfirst), pi);
    const dvec2 r0 = rik;
    vec2<G> vfirst = Prec<G>::c0(rij, rik), vlast = vfirst;
    G A = G(0), P = G(0);
    // skip-rule state: fan order, winding, circumradius, star bounding box
    LdFan fan{};
    int idprev = jfirst;
    if (LD) ld_pair(fan, *sums, i, rij, rik, jlast, jfirst, vfirst);
    // ring positions are fetched one slot ahead of their use, ring indices two
    int idnext = b.nbr[(1 % z) * N + i];
    dvec2 pnext = ldpos(pos, idnext);
    int jnext = b.nbr[(2 % z) * N + i];
    for (int k = 1; k < z; ++k) {
        const dvec2 pk = pnext;
        const int idk = idnext;
        if (k + 1 < z) {
            idnext = jnext;
            pnext = ldpos(pos, jnext);
        }
        if (k + 2 < z) jnext = b.nbr[(k + 2) * N + i];
        rij = rik;
        rik = min_dist(box, pk, pi);
        // spoke (i, ring k-1): far vertex ring k, circle of (i, ring k-2, ring k-1)
        if (LD) check_spoke(b, sp, s, pi, vlast, rij, rik);
        vec2<G> vnext = Prec<G>::c0(rij, rik);
        if (LD) {
            ld_pair(fan, *sums, i, rij, rik, idprev, idk, vnext);
            idprev = idk;
        }
        A += tri_area<G>(vlast, vnext);
        G dx = vlast.x - vnext.x, dy = vlast.y - vnext.y;
        P += sqrt(dx * dx + dy * dy);
        vlast = vnext;
    }
    A += tri_area<G>(vlast, vfirst);
    G dx = vlast.x - vfirst.x, dy = vlast.y - vfirst.y;
    P += sqrt(dx * dx + dy * dy);
    st2<R>(b.AP, i, vec2<R>{(R)A, (R)P});
    const vec2<R> pref = ld2<R>(b.APpref, i);
    st2<R>(b.dAP, i, vec2<R>{(R)(sp.KA * (A - (G)pref.x)), (R)(sp.KP * (P - (G)pref.y))});
    if (LD) {
        // spoke (i, ring z-1): far vertex ring 0
        check_spoke(b, sp, s, pi, vlast, rik, r0);
        if (fan.bad || fan.wraps != 1) push_region(b, s, -1 - i, 0.0, 0.0, 0.0, 0.0);
        b.cell_rmax[i] = sqrtf(fan.rmax2) * 1.0001f;
    }
}

// -------------------------------------------------------------------- forces
// Line-tension terms of one force set (reference voronoiQuadraticEnergyWithTension.cu:
// 27-199 for the matrix, the one-value kernel after it): which pairs of the
// vertex's three cells (I = the cell, J = ring k-2, K = ring k-1) differ in type,
// their tensions, and the reference's way of adding 0.5 (gamma) dT to a term.
template <typename R> struct TensionSlot {
    bool ik, ij, jk;
    R gik, gij, gjk;  // 1 in the one-value mode (gamma multiplies the sum)
    R half;           // 0.5 (matrix) or 0.5 gamma (one value)
    __device__ R add(R e, R dT) const { return e + half * dT; }
};
template <typename R>
__device__ inline TensionSlot<R> tension_slot(const cgpu_voronoi_batch &b, const cgpu_sys_params &sp, int tI, int tJ,
                                              int tK) {
    TensionSlot<R> t;
    t.ik = tI != tK;
    t.ij = tI != tJ;
    t.jk = tJ != tK;
    if (sp.tension == 1) {
        t.gik = t.gij = t.gjk = R(1);
        t.half = R(0.5) * (R)sp.gamma;
    } else {
        const int n = b.num_types;  // reference cellTypeIndexer(b, a) = a*n + b
        t.gik = t.ik ? (R)b.tension[tI * n + tK] : R(0);
        t.gij = t.ij ? (R)b.tension[tI * n + tJ] : R(0);
        t.gjk = t.jk ? (R)b.tension[tJ * n + tK] : R(0);
        t.half = R(0.5);
    }
    return t;
}
// the tension kernels clamp edge lengths at THRESHOLD = 1e-18 in both precisions
template <typename R> __device__ inline R clamp_threshold(R x) { return x < R(1e-18) ? R(1e-18) : x; }

// One reference force set (voronoiQuadraticEnergy.cu:99-234) for the vertex
// vcur of (I = the cell, J = ring k-2, K = ring k-1), from its neighbours vlast,
// vnext (in I's polygon) and vother (across the J-K edge), the unit vectors ul,
// un, uo of vlast - vcur, vcur - vnext, vcur - vother, the Jacobian m of the
// vertex with respect to I, and q = (KA (A - A0), KP (P - P0)) of I, J, K;
// added to sum.
template <typename R>
__device__ __forceinline__ void slot_force_set(const cgpu_voronoi_batch &b, const cgpu_sys_params &sp, bool tens,
                                               int tI, int tB, int tC, vec2<R> qI, vec2<R> qB, vec2<R> qC,
                                               vec2<R> vlast, vec2<R> vcur, vec2<R> vnext, vec2<R> vother,
                                               vec2<R> ul, vec2<R> un, vec2<R> uo, const R m[4], vec2<R> &sum) {
    vec2<R> dAdv{R(0.5) * (vlast.y - vnext.y), R(0.5) * (vnext.x - vlast.x)};
    vec2<R> dPdv{ul.x - un.x, ul.y - un.y};
    vec2<R> dEdv{madd(qI.y, dPdv.x, qI.x * dAdv.x), madd(qI.y, dPdv.y, qI.x * dAdv.y)};
    TensionSlot<R> T{};
    if (tens) {
        // self terms: edges I-K (vcur, vnext) and I-J (vcur, vlast)
        T = tension_slot<R>(b, sp, tI, tB, tC);
        vec2<R> dT{R(0), R(0)};
        if (T.ik) {
            dT.x -= T.gik * un.x;
            dT.y -= T.gik * un.y;
        }
        if (T.ij) {
            dT.x += T.gij * ul.x;
            dT.y += T.gij * ul.y;
        }
        dEdv.x = T.add(dEdv.x, dT.x);
        dEdv.y = T.add(dEdv.y, dT.y);
    }
    // cell ring k-1
    dAdv = vec2<R>{R(0.5) * (vnext.y - vother.y), R(0.5) * (vother.x - vnext.x)};
    dPdv = vec2<R>{-un.x - uo.x, -un.y - uo.y};
    {
        vec2<R> e{madd(qC.y, dPdv.x, qC.x * dAdv.x), madd(qC.y, dPdv.y, qC.x * dAdv.y)};
        if (tens) {  // edges K-I (vcur, vnext) and K-J (vcur, vother)
            vec2<R> dT{R(0), R(0)};
            if (T.ik) {
                dT.x += T.gik * -un.x;
                dT.y += T.gik * -un.y;
            }
            if (T.jk) {
                dT.x -= T.gjk * uo.x;
                dT.y -= T.gjk * uo.y;
            }
            e.x = T.add(e.x, dT.x);
            e.y = T.add(e.y, dT.y);
        }
        dEdv.x += e.x;
        dEdv.y += e.y;
    }
    // cell ring k-2
    dAdv = vec2<R>{R(0.5) * (vother.y - vlast.y), R(0.5) * (vlast.x - vother.x)};
    dPdv = vec2<R>{-uo.x + ul.x, -uo.y + ul.y};
    {
        vec2<R> e{madd(qB.y, dPdv.x, qB.x * dAdv.x), madd(qB.y, dPdv.y, qB.x * dAdv.y)};
        if (tens) {  // edges J-I (vcur, vlast) and J-K (vcur, vother)
            vec2<R> dT{R(0), R(0)};
            if (T.ij) {
                dT.x -= T.gij * -ul.x;
                dT.y -= T.gij * -ul.y;
            }
            if (T.jk) {
                dT.x += T.gjk * -uo.x;
                dT.y += T.gjk * -uo.y;
            }
            e.x = T.add(e.x, dT.x);
            e.y = T.add(e.y, dT.y);
        }
        dEdv.x += e.x;
        dEdv.y += e.y;
    }
    dEdv.x *= R(2.0);
    dEdv.y *= R(2.0);
    sum.x += madd(dEdv.y, m[2], dEdv.x * m[0]);
    sum.y += madd(dEdv.y, m[3], dEdv.x * m[1]);
}

// Unit vectors of the Voronoi edges in the perimeter gradient. Exact identity:
// the edge between the vertices circ(0, X, Y) and circ(0, Y, Z) of cell i (i
// at the origin) lies on the perpendicular bisector of i and Y, and the edge
// between circ(0, X, Y) and circ(X, Y, W) on the bisector of X and Y; each
// edge's direction is +-perp(s) / |s| for the segment s joining the two cells
// it separates. The FP64 path evaluates the reference expression d / |d| of the
// vertex difference d. The FP32 path takes the direction from the identity and
// only the sign from d: a short edge's d is the difference of two nearly equal
// vertices, whose direction float cannot resolve, while the force depends on it
// at O(KP |P - P0|) however short the edge. When the signed length along the
// direction is within AMBIGUOUS_EDGE x (cell-list bin side, about the cell
// spacing) of zero, well above the FP32 vertex error, its sign is evaluated in
// FP64 (edge_sign64).
constexpr float AMBIGUOUS_EDGE = 2e-5f;
// Sign (-1, 0, 1) of the signed length of the Voronoi edge from circ(0, X, Y)
// to circ(0, Y, Z) along perp(Y) / |Y|: both vertices lie on the bisector of 0
// and Y, at c = Y / 2 + t perp(Y) / |Y| with t(P) = |Y| P.(P - Y) / (2 cross(Y, P)),
// so the sign is that of t(X) - t(Z), evaluated without division in FP64 from
// the FP64 displacements.
__device__ inline int edge_sign64(dvec2 X, dvec2 Y, dvec2 Z) {
    const double nx = X.x * (X.x - Y.x) + X.y * (X.y - Y.y), dx = Y.x * X.y - Y.y * X.x;
    const double nz = Z.x * (Z.x - Y.x) + Z.y * (Z.y - Y.y), dz = Y.x * Z.y - Y.y * Z.x;
    const double v = nx * dz - nz * dx;
    const int sv = (v > 0.0) - (v < 0.0);
    return (dx * dz < 0.0) ? -sv : sv;
}
// The FP32 force's rare FP64 fallbacks, out of line: FP64 displacements of the
// cells ix, iy, ... from cell position pi (minimum image in a box of side L),
// read from pos.
__device__ inline dvec2 rare_rel(const double *pos, double L, dvec2 pi, int id) {
    return min_dist(make_box(L), ldpos(pos, id), pi);
}
__device__ __attribute__((noinline)) float2 rare_circ0(const double *pos, double L, dvec2 pi, int ix, int iy) {
    const dvec2 c = circ0<double>(rare_rel(pos, L, pi, ix), rare_rel(pos, L, pi, iy));
    return make_float2((float)c.x, (float)c.y);
}
__device__ __attribute__((noinline)) float2 rare_circ3(const double *pos, double L, dvec2 pi, int ix, int iy,
                                                         int iw) {
    const dvec2 c = circ3<double>(rare_rel(pos, L, pi, ix), rare_rel(pos, L, pi, iy), rare_rel(pos, L, pi, iw));
    return make_float2((float)c.x, (float)c.y);
}
__device__ __attribute__((noinline)) float4 rare_dhdr(const double *pos, double L, dvec2 pi, int ix, int iy);
// sign of the edge from circ(o, X, Y) to circ(o, Y, Z) along perp(Y - o), o the
// cell io (-1: the cell at pi)
__device__ __attribute__((noinline)) int rare_edge_sign(const double *pos, double L, dvec2 pi, int io, int ix, int iy,
                                                        int iz) {
    const dvec2 o = io < 0 ? dvec2{0.0, 0.0} : rare_rel(pos, L, pi, io);
    const dvec2 X = rare_rel(pos, L, pi, ix), Y = rare_rel(pos, L, pi, iy), Z = rare_rel(pos, L, pi, iz);
    return edge_sign64(dvec2{X.x - o.x, X.y - o.y}, dvec2{Y.x - o.x, Y.y - o.y}, dvec2{Z.x - o.x, Z.y - o.y});
}

__device__ __attribute__((noinline)) float4 rare_dhdr(const double *pos, double L, dvec2 pi, int ix, int iy) {
    double md[4];
    dhdr_eval<double>(rare_rel(pos, L, pi, ix), rare_rel(pos, L, pi, iy), md);
    return make_float4((float)md[0], (float)md[1], (float)md[2], (float)md[3]);
}

// FP32 pieces of the batched force: cross product and squared norm with one
// rounding of the product pair, the circumcenter of (0, x, y) from them and
// h = 0.5 / cross(x, y), the Jacobian of that vertex with respect to the origin
// cell (reference getdhdr, functions.h:234-242, with rij = x, rik = y:
// 0.5 / cross^2 = 2 h^2), and the hardware reciprocal and reciprocal square
// root (1 ulp). The force's FP32 error is set by the vertex positions, far
// above these roundings.
__device__ inline float cross_f(vec2<float> x, vec2<float> y) { return __builtin_fmaf(x.x, y.y, -(x.y * y.x)); }
__device__ inline float norm2_f(vec2<float> x) { return __builtin_fmaf(x.x, x.x, x.y * x.y); }
__device__ inline float rcp_f(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ inline float rsq_f(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ inline vec2<float> circ0_f(vec2<float> x, vec2<float> y, float nx, float ny, float h) {
    return vec2<float>{h * __builtin_fmaf(nx, y.y, -(x.y * ny)), h * __builtin_fmaf(x.x, ny, -(nx * y.x))};
}
__device__ inline void dhdr_f(vec2<float> rij, vec2<float> rik, float nij, float nik, float h, float m[4]) {
    const float g = 2.0f * h * h;
    const float P = __builtin_fmaf(rik.y, nij, -(rij.y * nik)), Q = __builtin_fmaf(rik.x, nij, -(rij.x * nik));
    const float gy = g * (rik.y - rij.y), gx = g * (rik.x - rij.x);
    m[0] = gy * P;
    m[1] = -gy * Q;
    m[2] = -gx * P;
    m[3] = gx * Q;
}


// Force on cell i: sum over ring slots k of the reference force set
// (voronoiQuadraticEnergy.cu:99-234): slot k is the Voronoi vertex of
// (i, ring k-2, ring k-1); its neighbours in the three polygons are vertex k-1,
// vertex k+1 and the circumcenter of (ring k-2, ring k-1, delOther k).
// TENS: the batch has cell types; systems with tension mode != 0 add the
// line-tension terms. FP64 path: the reference expressions.
template <typename R, bool TENS, typename Acc>
__device__ vec2<R> cell_force_ref(const cgpu_voronoi_batch &b, const Acc &acc, int i, const cgpu_sys_params &sp) {
    const Box box = make_box(sp.L);
    const int N = b.num_cells;
    const int z = b.nn[i];
    if (z < 3) return vec2<R>{R(0), R(0)};
    const bool tens = TENS && sp.tension != 0;
    auto clampn = [&](R x) { return tens ? clamp_threshold<R>(x) : clamp_len<R>(x); };
    const dvec2 pi = acc.p(i);
    auto rk = [&](int k) { return b.nbr[((k % z + z) % z) * N + i]; };
    auto dd = [&](int id) { return min_dist(box, acc.p(id), pi); };
    auto unit = [&](vec2<R> d) {
        const R n = clampn(sqrt(d.x * d.x + d.y * d.y));
        return vec2<R>{d.x / n, d.y / n};
    };
    int jB = rk(-2), jC = rk(-1);
    dvec2 dA = dd(rk(-3)), dB = dd(jB), dC = dd(jC);
    vec2<R> qB = acc.q(jB), qC = acc.q(jC);
    vec2<R> vL = Prec<R>::c0(dA, dB), vM = Prec<R>::c0(dB, dC);
    // unit vector of vlast - vcur; slot k + 1's is slot k's vcur - vnext (the
    // same operands), carried over
    vec2<R> ul = unit(vec2<R>{vL.x - vM.x, vL.y - vM.y});
    const vec2<R> qI = acc.q(i);
    int tI = 0, tB = 0, tC = 0;
    if (tens) {
        tI = b.cell_type[i];
        tB = b.cell_type[jB];
        tC = b.cell_type[jC];
    }
    vec2<R> sum{R(0), R(0)};
    // slot k's neighbour, delOther and their positions are fetched one slot
    // ahead, the indices two slots ahead
    int jDn = b.nbr[i];
    dvec2 pDn = acc.p(jDn), pOn = acc.p(b.oth[i]);
    int jD2 = b.nbr[(1 % z) * N + i], oD2 = b.oth[(1 % z) * N + i];
    for (int k = 0; k < z; ++k) {
        const int jD = jDn;
        const dvec2 pD = pDn, pO = pOn;
        if (k + 1 < z) {
            jDn = jD2;
            pDn = acc.p(jD2);
            pOn = acc.p(oD2);
        }
        if (k + 2 < z) {
            jD2 = b.nbr[(k + 2) * N + i];
            oD2 = b.oth[(k + 2) * N + i];
        }
        dvec2 dD = min_dist(box, pD, pi);
        vec2<R> vN = Prec<R>::c0(dC, dD);
        dvec2 dO = min_dist(box, pO, pi);
        // --- reference force set for slot k: (nm1, n1) = (ring k-2, ring k-1)
        R m[4];
        dhdr<R>(dB, dC, m);
        const vec2<R> vcur = vM, vlast = vL, vnext = vN;
        vec2<R> vother = Prec<R>::c3(dB, dC, dO);
        const vec2<R> un = unit(vec2<R>{vcur.x - vnext.x, vcur.y - vnext.y});
        const vec2<R> uo = unit(vec2<R>{vcur.x - vother.x, vcur.y - vother.y});
        slot_force_set<R>(b, sp, tens, tI, tB, tC, qI, qB, qC, vlast, vcur, vnext, vother, ul, un, uo, m, sum);
        // --- advance the window
        ul = un;
        dB = dC;
        dC = dD;
        qB = qC;
        qC = acc.q(jD);
        if (tens) {
            tB = tC;
            tC = b.cell_type[jD];
        }
        vL = vM;
        vM = vN;
    }
    return sum;
}

// The same in FP32 with the exact edge directions. Relative positions are held
// in float (the FP64 displacement rounded once, as the FP64-to-float conversion
// of the FP64 path would); the rare FP64 fallbacks (an ill-conditioned
// circumcenter or Voronoi-vertex Jacobian, an ambiguous edge sign) recompute the
// FP64 displacements from the cells' positions, so the loop keeps few FP64
// registers.
template <bool TENS, typename Acc>
__device__ vec2<float> cell_force_f32(const cgpu_voronoi_batch &b, const Acc &acc, int i, const cgpu_sys_params &sp) {
    using R = float;
    const Box box = make_box(sp.L);
    const int N = b.num_cells;
    const int z = b.nn[i];
    if (z < 3) return vec2<R>{R(0), R(0)};
    const bool tens = TENS && sp.tension != 0;
    const dvec2 pi = acc.p(i);
    auto rk = [&](int k) { return b.nbr[((k % z + z) % z) * N + i]; };
    auto d64 = [&](int id) { return min_dist(box, acc.p(id), pi); };
    const R amb = (R)(AMBIGUOUS_EDGE * sp.boxsize);
    // unit vector of the vertex difference d of the edge from circ(0, X, Y) to
    // circ(0, Y, Z) (cells ix, iy, iz; origin io, -1 = cell i), along perp(Y)
    // (|Y|^2 = nY)
    auto unit = [&](vec2<R> d, vec2<R> Y, R nY, int io, int ix, int iy, int iz) -> vec2<R> {
        const R inv = rsq_f(nY);
        const vec2<R> e{-Y.y * inv, Y.x * inv};
        R lam = __builtin_fmaf(d.x, e.x, d.y * e.y);
        if (fabsf(lam) < amb) {
            const int sg = rare_edge_sign(acc.pos, sp.L, pi, io, ix, iy, iz);
            if (sg == 0) return vec2<R>{R(0), R(0)};  // zero-length edge (float build: d / 1e-18 = 0)
            lam = (R)sg;
        }
        return lam >= R(0) ? e : vec2<R>{-e.x, -e.y};
    };
    // The Voronoi vertex circ(0, X, Y) of the ring pair (X, Y) = (ring k-1, ring
    // k) with its cross product's h = 0.5 / cross(X, Y) and well-conditioning
    // flag; the next slot's Jacobian of the same vertex reuses all three.
    int jA = rk(-3), jB = rk(-2), jC = rk(-1);
    vec2<R> fB = cvt<R>(d64(jB)), fC = cvt<R>(d64(jC));
    R nB = norm2_f(fB), nC = norm2_f(fC);
    vec2<R> qB = acc.q(jB), qC = acc.q(jC);
    auto vertex = [&](vec2<R> X, vec2<R> Y, R nX, R nY, int iX, int iY, R &h, bool &ok) -> vec2<R> {
        const R cr = cross_f(X, Y);
        ok = Prec<R>::well(cr, X, Y);
        if (ok) {
            h = R(0.5) * rcp_f(cr);
            return circ0_f(X, Y, nX, nY, h);
        }
        h = R(0);
        const float2 c = rare_circ0(acc.pos, sp.L, pi, iX, iY);
        return vec2<R>{c.x, c.y};
    };
    R hM;
    bool okM;
    vec2<R> vM = vertex(fB, fC, nB, nC, jB, jC, hM, okM);
    vec2<R> vL;
    {
        const vec2<R> fA = cvt<R>(d64(jA));
        R hL;
        bool okL;
        vL = vertex(fA, fB, norm2_f(fA), nB, jA, jB, hL, okL);
    }
    vec2<R> ul = unit(vec2<R>{vL.x - vM.x, vL.y - vM.y}, fB, nB, -1, jA, jB, jC);
    const vec2<R> qI = acc.q(i);
    int tI = 0, tB = 0, tC = 0;
    if (tens) {
        tI = b.cell_type[i];
        tB = b.cell_type[jB];
        tC = b.cell_type[jC];
    }
    vec2<R> sum{R(0), R(0)};
    // slot k's neighbour, delOther and their positions are fetched one slot
    // ahead, the indices two slots ahead
    int jDn = b.nbr[i], oDn = b.oth[i];
    dvec2 pDn = acc.p(jDn), pOn = acc.p(oDn);
    int jD2 = b.nbr[(1 % z) * N + i], oD2 = b.oth[(1 % z) * N + i];
    for (int k = 0; k < z; ++k) {
        const int jD = jDn, jO = oDn;
        const dvec2 pD = pDn, pO = pOn;
        if (k + 1 < z) {
            jDn = jD2;
            oDn = oD2;
            pDn = acc.p(jD2);
            pOn = acc.p(oD2);
        }
        if (k + 2 < z) {
            jD2 = b.nbr[(k + 2) * N + i];
            oD2 = b.oth[(k + 2) * N + i];
        }
        const vec2<R> fD = cvt<R>(min_dist(box, pD, pi)), fO = cvt<R>(min_dist(box, pO, pi));
        const R nD = norm2_f(fD);
        R hN;
        bool okN;
        const vec2<R> vN = vertex(fC, fD, nC, nD, jC, jD, hN, okN);
        // --- reference force set for slot k: (nm1, n1) = (ring k-2, ring k-1);
        // the Jacobian of vertex M = circ(0, B, C)
        R m[4];
        if (okM) {
            dhdr_f(fB, fC, nB, nC, hM, m);
        } else {
            const float4 md = rare_dhdr(acc.pos, sp.L, pi, jB, jC);
            m[0] = md.x;
            m[1] = md.y;
            m[2] = md.z;
            m[3] = md.w;
        }
        const vec2<R> vcur = vM, vlast = vL, vnext = vN;
        // vother = circ(B, C, O)
        vec2<R> vother;
        {
            const vec2<R> u{fB.x - fO.x, fB.y - fO.y}, v{fC.x - fO.x, fC.y - fO.y};
            const R cr = cross_f(u, v);
            if (Prec<R>::well(cr, u, v)) {
                const R h = R(0.5) * rcp_f(cr), nu = norm2_f(u), nv = norm2_f(v);
                vother = vec2<R>{fO.x + h * __builtin_fmaf(nu, v.y, -(u.y * nv)),
                                 fO.y + h * __builtin_fmaf(u.x, nv, -(nu * v.x))};
            } else {
                const float2 c = rare_circ3(acc.pos, sp.L, pi, jB, jC, jO);
                vother = vec2<R>{c.x, c.y};
            }
        }
        // un: vcur - vnext on the bisector of i and ring k-1; uo: vcur - vother
        // on the bisector of ring k-2 and ring k-1 (frame of ring k-2:
        // vcur = circ(B, 0, C), vother = circ(B, O, C))
        const vec2<R> un = unit(vec2<R>{vcur.x - vnext.x, vcur.y - vnext.y}, fC, nC, -1, jB, jC, jD);
        const vec2<R> sBC{fC.x - fB.x, fC.y - fB.y};
        const vec2<R> uo = unit(vec2<R>{vcur.x - vother.x, vcur.y - vother.y}, sBC, norm2_f(sBC), jB, i, jC, jO);
        slot_force_set<R>(b, sp, tens, tI, tB, tC, qI, qB, qC, vlast, vcur, vnext, vother, ul, un, uo, m, sum);
        // --- advance the window
        ul = un;
        fB = fC;
        fC = fD;
        nB = nC;
        nC = nD;
        jB = jC;
        jC = jD;
        qB = qC;
        qC = acc.q(jD);
        if (tens) {
            tB = tC;
            tC = b.cell_type[jD];
        }
        vL = vM;
        vM = vN;
        hM = hN;
        okM = okN;
    }
    return sum;
}

template <typename R, bool TENS, typename Acc>
__device__ inline vec2<R> cell_force(const cgpu_voronoi_batch &b, const Acc &acc, int i, const cgpu_sys_params &sp) {
    if constexpr (std::is_same<R, float>::value)
        return cell_force_f32<TENS>(b, acc, i, sp);
    else
        return cell_force_ref<R, TENS>(b, acc, i, sp);
}

// ------------------------------------------------------- Delaunay star (repair)
// Counter-clockwise Delaunay 1-ring of cell i by gift wrapping over cell-list
// candidates with exact predicates. Each triangle is certified empty by
// requiring its circumcircle to lie inside the scanned bin block; otherwise the
// block grows. Rotation is canonical (smallest neighbour first). Returns the
// degree, -1 when no certified star exists, -2 when the degree exceeds the ring
// capacity.
struct Cand {
    int id;
    dvec2 d;
};
template <typename F>
__device__ inline void scan_block(const cgpu_voronoi_batch &b, const cgpu_sys_params &sp, int ib, int jb, int w,
                                  F &&f) {
    const int xs = sp.xsize;
    const int nb = min(2 * w + 1, xs);
    for (int jj = 0; jj < nb; ++jj) {
        int cy = jb - w + jj;
        cy = ((cy % xs) + xs) % xs;
        for (int ii = 0; ii < nb; ++ii) {
            int cx = ib - w + ii;
            cx = ((cx % xs) + xs) % xs;
            const int bin = sp.bin_off + cy * xs + cx;
            const int k1 = b.bin_start[bin + 1];
            for (int k = b.bin_start[bin]; k < k1; ++k) f(b.sorted_id[k], k);
        }
    }
}

// (the search starts at half-width w0: smaller blocks are known to give no
// certified star)
__device__ int delaunay_star(const cgpu_voronoi_batch &b, const double *pos, int i, const cgpu_sys_params &sp,
                             int *ring, int w0 = 1) {
    const Box box = make_box(sp.L);
    const dvec2 pi = ldpos(pos, i);
    const double bs = sp.boxsize;
    const int xs = sp.xsize;
    int ib = max(0, min(xs - 1, (int)floor(pi.x / bs)));
    int jb = max(0, min(xs - 1, (int)floor(pi.y / bs)));
    const int wmax = max(1, xs / 2);
    for (int w = w0; w <= wmax; ++w) {
        const bool full = (2 * w + 1 >= xs);
        // scanned block relative to pi, shrunk by 1e-9 L for rounding at bin edges
        const double mg = 1e-9 * sp.L;
        const double xlo = (ib - w) * bs - pi.x + mg, xhi = (ib + w + 1) * bs - pi.x - mg;
        const double ylo = (jb - w) * bs - pi.y + mg, yhi = (jb + w + 1) * bs - pi.y - mg;
        // nearest candidate
        Cand first{-1, {0, 0}};
        int count = 0;
        scan_block(b, sp, ib, jb, w, [&](int m, int k) {
            if (m == i) return;
            ++count;
            dvec2 d = min_dist(box, ldpos(b.sorted_pos, k), pi);
            if (first.id < 0) { first = Cand{m, d}; return; }
            int c = cmp_norm2(d.x, d.y, first.d.x, first.d.y);
            if (c < 0 || (c == 0 && m < first.id)) first = Cand{m, d};
        });
        if (count < 3) continue;
        Cand cur = first;
        int z = 0;
        bool ok = true;
        const int cap = ring_cap(b);
        for (int guard = 0; guard <= cap; ++guard) {
            if (z >= cap) return -2;
            ring[z++] = cur.id;
            Cand best{-1, {0, 0}};
            scan_block(b, sp, ib, jb, w, [&](int m, int k) {
                if (m == i || m == cur.id) return;
                dvec2 d = min_dist(box, ldpos(b.sorted_pos, k), pi);
                if (orient0(cur.d.x, cur.d.y, d.x, d.y) <= 0) return;
                if (best.id < 0) { best = Cand{m, d}; return; }
                int ic = incircle0(cur.d.x, cur.d.y, best.d.x, best.d.y, d.x, d.y);
                if (ic > 0 || (ic == 0 && m < best.id)) best = Cand{m, d};
            });
            if (best.id < 0) { ok = false; break; }
            if (!full) {
                // certified when the closed circumdisk lies inside the scanned
                // block: no unscanned point can be inside it
                dvec2 c = circ0<double>(cur.d, best.d);
                double rr = sqrt(c.x * c.x + c.y * c.y);
                if (!(c.x - rr >= xlo && c.x + rr <= xhi && c.y - rr >= ylo && c.y + rr <= yhi)) {
                    ok = false;
                    break;
                }
            }
            cur = best;
            if (cur.id == first.id) break;
        }
        if (!ok || cur.id != first.id) continue;
        // canonical rotation
        int k0 = 0;
        for (int k = 1; k < z; ++k)
            if (ring[k] < ring[k0]) k0 = k;
        if (k0 != 0) {
            int tmp[CAP_MAX];
            for (int k = 0; k < z; ++k) tmp[k] = ring[(k + k0) % z];
            for (int k = 0; k < z; ++k) ring[k] = tmp[k];
        }
        return z;
    }
    return -1;
}

// polar-angle quadrant of a nonzero vector: [0, pi/2) -> 0, [pi/2, pi) -> 1,
// [pi, 3pi/2) -> 2, [3pi/2, 2pi) -> 3 (exact sign tests)
__device__ inline int quadrant(dvec2 d) {
    if (d.y >= 0.0 && d.x > 0.0) return 0;
    if (d.x <= 0.0 && d.y > 0.0) return 1;
    if (d.y <= 0.0 && d.x < 0.0) return 2;
    return 3;
}

// True when the current ring of cell i already is its exact Delaunay star:
// consecutive neighbours strictly counter-clockwise, winding once, and no point
// strictly inside any of its triangles' circumcircles (exact in-circle test on
// the cell-list bins covering each circle's bounding box, widened by 1e-9 L for
// rounding at bin edges). The Delaunay triangulation is unique for points in
// general position, so keeping such a ring gives the same result as rebuilding it
// for a fraction of the work. Circles wider than an eighth of the box are left to
// the rebuild.
__device__ bool verify_ring(const cgpu_voronoi_batch &b, const double *pos, int i, const cgpu_sys_params &sp) {
    const int N = b.num_cells;
    const int z = b.nn[i];
    if (z < 3 || z > ring_cap(b)) return false;
    const Box box = make_box(sp.L);
    const dvec2 pi = ldpos(pos, i);
    const double bs = sp.boxsize;
    const int xs = sp.xsize;
    int wraps = 0;
    int ja = b.nbr[(z - 1) * N + i];
    dvec2 da = min_dist(box, ldpos(pos, ja), pi);
    for (int k = 0; k < z; ++k) {
        const int jb2 = b.nbr[k * N + i];
        const dvec2 db = min_dist(box, ldpos(pos, jb2), pi);
        if (orient0(da.x, da.y, db.x, db.y) <= 0) return false;
        // every step turns counter-clockwise by less than pi, so the polar angle
        // decreases (quadrant index drops) exactly once per full turn
        wraps += (quadrant(db) < quadrant(da)) ? 1 : 0;
        const dvec2 c = circ0<double>(da, db);
        const double r2 = c.x * c.x + c.y * c.y;
        const double r = sqrt(r2);
        if (!(r < 0.125 * sp.L)) return false;
        const double m = r + 1e-9 * sp.L;
        const int x0 = (int)floor((pi.x + c.x - m) / bs), x1 = (int)floor((pi.x + c.x + m) / bs);
        const int y0 = (int)floor((pi.y + c.y - m) / bs), y1 = (int)floor((pi.y + c.y + m) / bs);
        for (int gy = y0; gy <= y1; ++gy) {
            const int cy = ((gy % xs) + xs) % xs;
            for (int gx = x0; gx <= x1; ++gx) {
                const int cx = ((gx % xs) + xs) % xs;
                const int bin = sp.bin_off + cy * xs + cx;
                const int k1 = b.bin_start[bin + 1];
                for (int k = b.bin_start[bin]; k < k1; ++k) {
                    const int mm = b.sorted_id[k];
                    if (mm == i || mm == ja || mm == jb2) continue;
                    const dvec2 dm = min_dist(box, ldpos(b.sorted_pos, k), pi);
                    const double ex = dm.x - c.x, ey = dm.y - c.y;
                    if (ex * ex + ey * ey > r2 * (1.0 + 1e-8)) continue;  // clearly outside
                    if (incircle0(da.x, da.y, db.x, db.y, dm.x, dm.y) > 0) return false;
                }
            }
        }
        ja = jb2;
        da = db;
    }
    return wraps == 1;
}

// ------------------------------------------ wave-cooperative star (repair)
// One wavefront per queued cell. The candidates of a (2w+1)^2 bin block sit one
// per lane, the current ring one member per lane, so ring verification and the
// gift-wrapping selections are wave-wide reductions under exact total orders
// instead of serial scans with dependent loads. The result is the same ring as
// verify_ring / delaunay_star give: a verified ring is the unique Delaunay star,
// and each wrap step selects the minimum of a strict total order, which no
// evaluation order changes.
struct WCand {
    int id;
    double x, y;
};
__device__ inline WCand shfl_xor_cand(WCand c, int o) {
    return WCand{__shfl_xor(c.id, o), __shfl_xor(c.x, o), __shfl_xor(c.y, o)};
}
// nearest-neighbour order: squared norm, then index (exact)
__device__ inline bool nearer(const WCand &a, const WCand &b) {
    if (a.id < 0) return false;
    if (b.id < 0) return true;
    const int c = cmp_norm2(a.x, a.y, b.x, b.y);
    return c < 0 || (c == 0 && a.id < b.id);
}
// wrap order seen from cur (all points strictly left of 0 -> cur): a precedes b
// when a lies strictly inside the circle through 0, cur and b, then by index
__device__ inline bool wraps_before(double cx, double cy, const WCand &a, const WCand &b) {
    if (a.id < 0) return false;
    if (b.id < 0) return true;
    const int ic = incircle0(cx, cy, b.x, b.y, a.x, a.y);
    return ic > 0 || (ic == 0 && a.id < b.id);
}

__device__ inline double uniform_d(double v) {
    const long long u = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readfirstlane((int)(u & 0xffffffffll));
    const int hi = __builtin_amdgcn_readfirstlane((int)(u >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ inline double readlane_d(double v, int l) {
    const long long u = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(u & 0xffffffffll), l);
    const int hi = __builtin_amdgcn_readlane((int)(u >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
// minimum over the 64 lanes (all active) by DPP row moves, in every lane
__device__ inline float wave_min_f32(float v) {
    int x = __float_as_int(v);
    v = fminf(v, __int_as_float(__builtin_amdgcn_update_dpp(x, x, 0xb1, 0xf, 0xf, false)));  // quad_perm 1,0,3,2
    x = __float_as_int(v);
    v = fminf(v, __int_as_float(__builtin_amdgcn_update_dpp(x, x, 0x4e, 0xf, 0xf, false)));  // quad_perm 2,3,0,1
    x = __float_as_int(v);
    v = fminf(v, __int_as_float(__builtin_amdgcn_update_dpp(x, x, 0x141, 0xf, 0xf, false)));  // row_half_mirror
    x = __float_as_int(v);
    v = fminf(v, __int_as_float(__builtin_amdgcn_update_dpp(x, x, 0x140, 0xf, 0xf, false)));  // row_mirror
    x = __float_as_int(v);
    v = fminf(v, __int_as_float(__builtin_amdgcn_update_dpp(x, x, 0x142, 0xa, 0xf, false)));  // row_bcast15
    x = __float_as_int(v);
    v = fminf(v, __int_as_float(__builtin_amdgcn_update_dpp(x, x, 0x143, 0xc, 0xf, false)));  // row_bcast31
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

// One gift-wrapping step seen from cur, decided in floating point: the next
// ring member is the candidate strictly left of 0 -> cur whose circle through
// 0, cur and itself has the smallest centre parameter t along the bisector of
// 0 and cur (t = c.(c - cur) / (2 cross(cur, c)); a lies inside the circle of b
// exactly when t_a < t_b, so this is the wraps_before order). Each t carries a
// rigorous bound e on its rounding error; the step is decided only when one
// candidate's interval [t - e, t + e] lies wholly below every other's. Returns
// false, selecting nothing, otherwise (a tie, a near-tie or an ill-conditioned
// circle): the exact comparison then decides. best.id = -1 when no candidate
// lies left of 0 -> cur.
template <int K>
__device__ inline bool wrap_select_fast(const WCand (&c)[K], const WCand &cur, WCand &best) {
    constexpr double U4 = 4.0 * 0x1p-53;  // four units of roundoff
    double lo[K];
    float key = INFINITY;
    bool amb = false, any = false;
#pragma unroll
    for (int r = 0; r < K; ++r) {
        lo[r] = INFINITY;
        if (c[r].id >= 0 && c[r].id != cur.id && orient0(cur.x, cur.y, c[r].x, c[r].y) > 0) {
            any = true;
            const double px = c[r].x * (c[r].x - cur.x), py = c[r].y * (c[r].y - cur.y);
            const double qx = cur.x * c[r].y, qy = cur.y * c[r].x;
            const double num = px + py, den = 2.0 * (qx - qy);
            // |num - num*| <= 3.02 u (|px| + |py|), |den - den*| <= 6.02 u (|qx| + |qy|)
            const double en = U4 * (fabs(px) + fabs(py)), ed = 2.0 * U4 * (fabs(qx) + fabs(qy));
            if (!(2.0 * ed <= den)) {
                amb = true;  // den* may be far below den (or the sign rests on the exact test alone)
                continue;
            }
            const double inv = 1.0 / den, t = num * inv;
            // |t - t*| <= 2 (|t| ed + en) / den + 2.01 u |t| when den* >= den / 2
            const double e = 4.0 * (fabs(t) * ed + en) * inv + U4 * fabs(t);
            const double hi = t + e;
            if (!(fabs(t) < 1e300 && e < 1e300)) {
                amb = true;
                continue;
            }
            lo[r] = t - e;
            // float threshold at or above hi
            float f = (float)hi;
            f = f + fabsf(f) * 2.4e-7f + 1e-37f;
            key = fminf(key, f);
        }
    }
    if (__ballot(amb) != 0ull) return false;
    const float kmin = wave_min_f32(key);
    if (kmin == INFINITY) {
        if (__ballot(any) != 0ull) return false;
        best = WCand{-1, 0.0, 0.0};
        return true;
    }
    const double th = (double)kmin;
    int cnt = 0;
    WCand mine = c[0];
#pragma unroll
    for (int r = 0; r < K; ++r)
        if (lo[r] <= th) {
            ++cnt;
            mine = c[r];
        }
    const unsigned long long m = __ballot(cnt > 0);
    if (__popcll(m) != 1 || __ballot(cnt > 1) != 0ull) return false;
    const int wl = __ffsll((long long)m) - 1;
    best = WCand{__builtin_amdgcn_readlane(mine.id, wl), readlane_d(mine.x, wl), readlane_d(mine.y, wl)};
    return true;
}

// The candidate nearest to the origin, decided in floating point as above
// (|c|^2 carries at most 2.01 u |c|^2 of rounding error); false when two
// candidates are too close to call (nearer then decides exactly).
template <int K>
__device__ inline bool nearest_fast(const WCand (&c)[K], WCand &first) {
    constexpr double U4 = 4.0 * 0x1p-53;
    double lo[K];
    float key = INFINITY;
#pragma unroll
    for (int r = 0; r < K; ++r) {
        lo[r] = INFINITY;
        if (c[r].id >= 0) {
            const double n2 = c[r].x * c[r].x + c[r].y * c[r].y, e = U4 * n2;
            lo[r] = n2 - e;
            float f = (float)(n2 + e);
            f = f + f * 2.4e-7f + 1e-37f;
            key = fminf(key, f);
        }
    }
    const float kmin = wave_min_f32(key);
    if (kmin == INFINITY) return false;
    const double th = (double)kmin;
    int cnt = 0;
    WCand mine = c[0];
#pragma unroll
    for (int r = 0; r < K; ++r)
        if (lo[r] <= th) {
            ++cnt;
            mine = c[r];
        }
    const unsigned long long m = __ballot(cnt > 0);
    if (__popcll(m) != 1 || __ballot(cnt > 1) != 0ull) return false;
    const int wl = __ffsll((long long)m) - 1;
    first = WCand{__builtin_amdgcn_readlane(mine.id, wl), readlane_d(mine.x, wl), readlane_d(mine.y, wl)};
    return true;
}

// Candidates of the block of half-width w around bin (ib, jb), K per lane
// (slot lane + 64 r in c[r]; cell i itself and empty slots have id -1).
// Returns the number of points in the block (i included), or -1 when they do
// not fit in K slots per lane. Requires nb^2 <= 64.
template <int K>
__device__ inline int wave_load_block(const cgpu_voronoi_batch &b, const double *pos, const Box &box, dvec2 pi,
                                      int i, const cgpu_sys_params &sp, int ib, int jb, int w, WCand (&c)[K]) {
    const int lane = __lane_id();
    const int xs = sp.xsize;
    const int nb = min(2 * w + 1, xs);
    const int B = nb * nb;
    int cnt = 0, bin = 0;
    if (lane < B) {
        int cx = ib - w + lane % nb;
        int cy = jb - w + lane / nb;
        cx = ((cx % xs) + xs) % xs;
        cy = ((cy % xs) + xs) % xs;
        bin = sp.bin_off + cy * xs + cx;
        bin = b.bin_start[bin];  // from here on: the bin's first bin-ordered slot
        cnt = b.bin_start[sp.bin_off + cy * xs + cx + 1] - bin;
    }
    int incl = cnt;
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(incl, o);
        if (lane >= o) incl += t;
    }
    const int total = __builtin_amdgcn_readlane(incl, 63);
#pragma unroll
    for (int r = 0; r < K; ++r) c[r] = WCand{-1, 0.0, 0.0};
    if (total > 64 * K) return -1;
    const int excl = incl - cnt;
    // bin of slot j: the last bin whose exclusive offset is <= j (an empty bin
    // shares its offset with the next one)
    int t[K];
#pragma unroll
    for (int r = 0; r < K; ++r) t[r] = 0;
    for (int u = 1; u < B; ++u) {
        const int e = __builtin_amdgcn_readlane(excl, u);
#pragma unroll
        for (int r = 0; r < K; ++r)
            if (e <= lane + 64 * r) t[r] = u;
    }
#pragma unroll
    for (int r = 0; r < K; ++r) {
        const int j = lane + 64 * r;
        const int tstart = __shfl(bin, t[r]), toff = __shfl(excl, t[r]);
        if (j < total) {
            const int k = tstart + (j - toff);
            const int m = b.sorted_id[k];
            if (m != i) {
                const dvec2 d = min_dist(box, ldpos(b.sorted_pos, k), pi);
                c[r] = WCand{m, d.x, d.y};
            }
        }
    }
    return total;
}

// Star of cell i by one wave. Returns the new degree (member k of the ring in
// lane k's ring_out, canonical rotation), 0 when the current ring is verified
// as the Delaunay star (nothing to write), -1 when no certified star exists
// within the largest block, -2 when the degree exceeds the ring capacity, -3 when a block's
// candidates do not fit in K slots per lane (the caller takes another path).
// The search starts at half-width w0 (blocks below it are known to give no
// certified star); on -3, w_at is the half-width whose block did not fit.
template <int K>
__device__ int wave_star(const cgpu_voronoi_batch &b, const double *pos, int i, const cgpu_sys_params &sp,
                         int &ring_out, int w0, int &w_at) {
    const int lane = __lane_id();
    const int N = b.num_cells;
    const Box box = make_box(sp.L);
    const dvec2 pi{uniform_d(pos[2 * i]), uniform_d(pos[2 * i + 1])};
    const double bs = sp.boxsize;
    const int xs = sp.xsize;
    const int ib = max(0, min(xs - 1, (int)floor(pi.x / bs)));
    const int jb = max(0, min(xs - 1, (int)floor(pi.y / bs)));
    const int wmax = max(1, xs / 2);
    const double mg = 1e-9 * sp.L;

    // current ring: member k in lane k, with the triangle (k-1, k) and its circle
    const int zold = __builtin_amdgcn_readfirstlane(b.nn[i]);
    int verify = (zold >= 3 && zold <= ring_cap(b)) ? 1 : 0;  // 1 undecided, 0 rebuild, 2 verified
    int rid = -1, pid = -1;
    double rx = 0, ry = 0, px = 0, py = 0, ccx = 0, ccy = 0, cr2 = 0, cr = 0;
    if (verify) {
        if (lane < zold) {
            rid = b.nbr[lane * N + i];
            const dvec2 d = min_dist(box, ldpos(pos, rid), pi);
            rx = d.x;
            ry = d.y;
        }
        const int src = (lane + zold - 1) % zold;
        pid = __shfl(rid, src);
        px = __shfl(rx, src);
        py = __shfl(ry, src);
        bool bad = false, wrapped = false;
        if (lane < zold) {
            bad = orient0(px, py, rx, ry) <= 0;
            wrapped = quadrant(dvec2{rx, ry}) < quadrant(dvec2{px, py});
            const dvec2 cc = circ0<double>(dvec2{px, py}, dvec2{rx, ry});
            ccx = cc.x;
            ccy = cc.y;
            cr2 = cc.x * cc.x + cc.y * cc.y;
            cr = sqrt(cr2);
        }
        if (__ballot(bad) != 0ull || __popcll(__ballot(wrapped)) != 1) verify = 0;
    }
    // the first block that holds every current-ring circle decides the
    // verification: start there (no star built from a smaller block is needed,
    // any certified block gives the same ring)
    int wfirst = w0;
    if (verify == 1 && w0 == 1) {
        int need = 1;
        if (lane < zold) {
            while (need < 2) {
                const double xlo = (ib - need) * bs - pi.x + mg, xhi = (ib + need + 1) * bs - pi.x - mg;
                const double ylo = (jb - need) * bs - pi.y + mg, yhi = (jb + need + 1) * bs - pi.y - mg;
                if (ccx - cr >= xlo && ccx + cr <= xhi && ccy - cr >= ylo && ccy + cr <= yhi) break;
                ++need;
            }
        }
        for (int o = 32; o > 0; o >>= 1) need = max(need, __shfl_xor(need, o));
        wfirst = min(need, wmax);
    }

    for (int w = wfirst; w <= wmax; ++w) {
        w_at = w;
        const int nb = min(2 * w + 1, xs);
        if (nb * nb > 64) return -3;
        WCand c[K];
        const int total = wave_load_block<K>(b, pos, box, pi, i, sp, ib, jb, w, c);
        if (total < 0) return -3;
        const bool full = (2 * w + 1 >= xs);
        const double xlo = (ib - w) * bs - pi.x + mg, xhi = (ib + w + 1) * bs - pi.x - mg;
        const double ylo = (jb - w) * bs - pi.y + mg, yhi = (jb + w + 1) * bs - pi.y - mg;

        // a current-ring circle that leaves the block leaves the verification
        // undecided at this width; from the second width on, the star is then
        // built from this block
        bool decide = false;
        if (verify == 1) {
            if (full) {
                verify = 0;
            } else {
                const bool fits = lane >= zold ||
                                  (ccx - cr >= xlo && ccx + cr <= xhi && ccy - cr >= ylo && ccy + cr <= yhi);
                decide = __ballot(!fits) == 0ull;
            }
        }
        if (decide) {
            {
                bool viol = false;
                for (int k = 0; k < zold; ++k) {
                    const double cx = readlane_d(ccx, k), cy = readlane_d(ccy, k), r2 = readlane_d(cr2, k);
                    const double ax = readlane_d(px, k), ay = readlane_d(py, k);
                    const double bx = readlane_d(rx, k), by = readlane_d(ry, k);
                    const int ia = __builtin_amdgcn_readlane(pid, k), ibb = __builtin_amdgcn_readlane(rid, k);
#pragma unroll
                    for (int r = 0; r < K; ++r) {
                        if (c[r].id >= 0 && c[r].id != ia && c[r].id != ibb) {
                            const double ex = c[r].x - cx, ey = c[r].y - cy;
                            if (!(ex * ex + ey * ey > r2 * (1.0 + 1e-8)) &&
                                incircle0(ax, ay, bx, by, c[r].x, c[r].y) > 0)
                                viol = true;
                        }
                    }
                }
                verify = (__ballot(viol) == 0ull) ? 2 : 0;
            }
            if (verify == 2) return 0;
        } else if (verify == 1 && w == 1) {
            continue;  // the next block usually decides; building here rarely certifies
        }

        // gift wrapping over the block's candidates
        int nvalid = 0;
#pragma unroll
        for (int r = 0; r < K; ++r) nvalid += __popcll(__ballot(c[r].id >= 0));
        if (nvalid < 3) continue;
        WCand first = c[0];
        if (!nearest_fast<K>(c, first)) {
            first = c[0];
#pragma unroll
            for (int r = 1; r < K; ++r)
                if (nearer(c[r], first)) first = c[r];
            for (int o = 32; o > 0; o >>= 1) {
                const WCand oth = shfl_xor_cand(first, o);
                if (nearer(oth, first)) first = oth;
            }
        }
        WCand cur = first;
        int z = 0, myring = -1;
        bool ok = true;
        const int cap = ring_cap(b);
        for (int guard = 0; guard <= cap; ++guard) {
            if (z >= cap) return -2;
            if (lane == z) myring = cur.id;
            ++z;
            WCand best{-1, 0.0, 0.0};
            if (!wrap_select_fast<K>(c, cur, best)) {
                best = WCand{-1, 0.0, 0.0};
#pragma unroll
                for (int r = 0; r < K; ++r) {
                    if (c[r].id >= 0 && c[r].id != cur.id && orient0(cur.x, cur.y, c[r].x, c[r].y) > 0 &&
                        wraps_before(cur.x, cur.y, c[r], best))
                        best = c[r];
                }
                for (int o = 32; o > 0; o >>= 1) {
                    const WCand oth = shfl_xor_cand(best, o);
                    if (wraps_before(cur.x, cur.y, oth, best)) best = oth;
                }
            }
            if (best.id < 0) {
                ok = false;
                break;
            }
            if (!full) {
                const dvec2 cc = circ0<double>(dvec2{cur.x, cur.y}, dvec2{best.x, best.y});
                const double rr = sqrt(cc.x * cc.x + cc.y * cc.y);
                if (!(cc.x - rr >= xlo && cc.x + rr <= xhi && cc.y - rr >= ylo && cc.y + rr <= yhi)) {
                    ok = false;
                    break;
                }
            }
            cur = best;
            if (cur.id == first.id) break;
        }
        if (!ok || cur.id != first.id) continue;
        // canonical rotation: smallest index first
        int mn = (lane < z) ? myring : INT_MAX;
        for (int o = 32; o > 0; o >>= 1) mn = min(mn, __shfl_xor(mn, o));
        const int k0 = __ffsll((long long)__ballot(lane < z && myring == mn)) - 1;
        ring_out = __shfl(myring, (lane + k0) % z);
        return z;
    }
    return -1;
}

// number of triangles of the ring held in lanes 0..z-1 owned by cell i
__device__ inline int wave_owned(int i, int ring_lane, int z) {
    const int lane = __lane_id();
    const int nxt = __shfl(ring_lane, (lane + 1) % max(z, 1));
    return __popcll(__ballot(lane < z && i < ring_lane && i < nxt));
}


// slot k of the reference getDelSets (voronoiModelBase.cpp:1236-1275) for cell
// i, slot-major: delSets[k] = (ring k-2, ring k-1) and delOther[k] the opposite
// vertex across that edge, found in the ring of ring k-1; when that ring does
// not contain ring k-2 the slot's delOther stays unchanged. False when the
// slot's delOther is one of the triangle's own cells.
__device__ bool del_set_slot(const cgpu_voronoi_batch &b, int i, int z, int k) {
    const int N = b.num_cells;
    const int nm1 = b.nbr[((k + z - 2) % z) * N + i], n1 = b.nbr[((k + z - 1) % z) * N + i];
    const int zn = min(b.nn[n1], ring_cap(b));
    for (int q = 0; q < zn; ++q) {
        if (b.nbr[q * N + n1] == nm1) {
            b.oth[k * N + i] = b.nbr[((q + 1) % zn) * N + n1];
            break;
        }
    }
    const int o = b.oth[k * N + i];
    return !(nm1 == o || n1 == o || i == o);
}

__device__ inline int owned_triangles(const cgpu_voronoi_batch &b, int i) {
    const int N = b.num_cells;
    const int z = b.nn[i];
    int c = 0;
    for (int k = 0; k < z; ++k) {
        int a = b.nbr[k * N + i], q = b.nbr[((k + 1) % z) * N + i];
        if (i < a && i < q) ++c;
    }
    return c;
}

// ===================================================================== kernels

// the per-step topology state of system s (and, for s = 0, the queue counters)
__device__ inline void reset_system(const cgpu_voronoi_batch &b, int s) {
    if (s == 0) {
        b.counters[0] = 0;
        b.counters[1] = 0;
        b.counters[2] = 0;
        b.counters[7] += 1;  // the circumcircle test's epoch (bin stamps)
        b.counters[8] = 0;   // its region records
        b.counters[9] = 0;   // its triangle queue
    }
    if (s < b.num_systems) {
        cgpu_sys_state &st = b.st[s];
        st.any_fail = 0;
        st.nf_count = 0;
        st.mode = 0;
        st.need_global = 0;
        st.nn_delta = 0;
        st.tri_delta = 0;
        st.max_deg = 0;
        st.nflag = 0;
        st.tri_hash_all = 0ull;
        st.tri_hash_own = 0ull;
        st.tri_area = 0ll;
        st.skip_off = 0;
    }
}
__global__ void k_reset(cgpu_voronoi_batch b) { reset_system(b, blockIdx.x * blockDim.x + threadIdx.x); }

// end of a time step for one system: the step counter, and the serving
// countdown that parks the system when it reaches zero
__device__ inline void advance_system(cgpu_sys_state &st) {
    if (!st.active) return;
    st.step += 1;
    if (st.steps_left > 0) {
        st.steps_left -= 1;
        if (st.steps_left == 0) st.active = 0;
    }
}

// pass 1 for one cell at position p (also run by the step's force kernel)
__device__ inline void bin_cell(const cgpu_voronoi_batch &b, const cgpu_sys_params &sp, int i, dvec2 p) {
    const int ib = max(0, min(sp.xsize - 1, (int)floor(p.x / sp.boxsize)));
    const int jb = max(0, min(sp.xsize - 1, (int)floor(p.y / sp.boxsize)));
    const int bin = sp.bin_off + jb * sp.xsize + ib;
    b.cell_bin[i] = bin;
    b.cell_rank[i] = (int)atomicAdd(&b.bin_count[bin], 1u);
}

// Counting sort of the cells into the cell-list bins (reference
// cellListGPU.cu:20-58: bin = clamp(floor(p / boxsize)); its fixed-capacity bins
// become contiguous runs in bin order, so no bin can overflow). Pass 1: bin and
// rank of each cell.
__global__ void k_bin(cgpu_voronoi_batch b, const double *pos) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= b.num_cells) return;
    bin_cell(b, b.sys[b.cell_sys[i]], i, ldpos(pos, i));
}

// Pass 2, one block per system: exclusive prefix sum of the bin counts. A
// system's cells fill its bins, so the run of system s starts at its first
// cell. Each thread sums 8 consecutive bins, the block scans the thread sums.
constexpr int SCAN_THREADS = 1024, SCAN_PER = 8;
// Exclusive prefix sum of the counts cnt[0 .. n) into start[0 .. n), from
// carry0, by the whole block; the counts are cleared for the next cell list.
// Element e = r * SCAN_THREADS + t of a pass belongs to thread t, so every load
// and store is coalesced. The exclusive prefix of (r, t) is the sum of the
// (row, wavefront) partial sums before (r, t's wavefront) in row-major order,
// plus the lanes before t in its wavefront.
__device__ __forceinline__ void scan_bins(unsigned int *cnt, int *start, int n, int carry0) {
    constexpr int NW = SCAN_THREADS / 64;
    constexpr int NP = SCAN_PER * NW;  // partial sums per pass: two wavefronts
    static_assert(NP == 128, "the partial sums are scanned by two wavefronts");
    __shared__ int part[NP], pex[NP], ptot[2];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = carry0;
    const int lane = __lane_id(), wv = threadIdx.x >> 6;
    for (int base = 0; base < n; base += SCAN_THREADS * SCAN_PER) {
        int c[SCAN_PER], incl[SCAN_PER];
#pragma unroll
        for (int r = 0; r < SCAN_PER; ++r) {
            const int u = base + r * SCAN_THREADS + (int)threadIdx.x;
            c[r] = (u < n) ? (int)cnt[u] : 0;
        }
#pragma unroll
        for (int r = 0; r < SCAN_PER; ++r) {
            const int u = base + r * SCAN_THREADS + (int)threadIdx.x;
            if (u < n) cnt[u] = 0u;
        }
#pragma unroll
        for (int r = 0; r < SCAN_PER; ++r) {
            int x = c[r];
            for (int o = 1; o < 64; o <<= 1) {
                const int t = __shfl_up(x, o);
                if (lane >= o) x += t;
            }
            incl[r] = x;
            if (lane == 63) part[r * NW + wv] = x;
        }
        __syncthreads();
        if ((int)threadIdx.x < NP) {  // wavefronts 0 and 1, whole
            const int v = part[threadIdx.x];
            int x = v;
            for (int o = 1; o < 64; o <<= 1) {
                const int t = __shfl_up(x, o);
                if (lane >= o) x += t;
            }
            pex[threadIdx.x] = x - v;
            if (lane == 63) ptot[wv] = x;
        }
        __syncthreads();
        const int c0 = carry;
#pragma unroll
        for (int r = 0; r < SCAN_PER; ++r) {
            const int q = r * NW + wv;
            const int u = base + r * SCAN_THREADS + (int)threadIdx.x;
            if (u < n) start[u] = c0 + pex[q] + (q >= 64 ? ptot[0] : 0) + incl[r] - c[r];
        }
        __syncthreads();
        if (threadIdx.x == 0) carry = c0 + ptot[0] + ptot[1];
        __syncthreads();
    }
}

// one block per system (every system one bin chunk)
__global__ __launch_bounds__(SCAN_THREADS) void k_bin_scan(cgpu_voronoi_batch b) {
    const int s = blockIdx.x;
    const cgpu_sys_params &sp = b.sys[s];
    scan_bins(b.bin_count + sp.bin_off, b.bin_start + sp.bin_off, sp.xsize * sp.xsize, sp.off);
    if (s == b.num_systems - 1 && threadIdx.x == 0) b.bin_start[b.total_bins] = b.num_cells;
}

// The same prefix sum for systems with more than one bin chunk, in three
// passes over bin chunks: chunk sums, per-system offsets of the chunks, then the
// scan of each chunk from its offset.
__global__ __launch_bounds__(SCAN_THREADS) void k_bin_chunk_sum(cgpu_voronoi_batch b) {
    const int c = blockIdx.x;
    const int u0 = b.bchunk_begin[c], u1 = b.bchunk_end[c];
    int tot = 0;
    for (int u = u0 + (int)threadIdx.x; u < u1; u += blockDim.x) tot += (int)b.bin_count[u];
    for (int o = 32; o > 0; o >>= 1) tot += __shfl_xor(tot, o);
    __shared__ int wsum[SCAN_THREADS / 64];
    if (__lane_id() == 0) wsum[threadIdx.x >> 6] = tot;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) t += wsum[k];
        b.bchunk_sum[c] = t;
    }
}
__global__ void k_bin_chunk_offsets(cgpu_voronoi_batch b) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= b.num_systems) return;
    int off = b.sys[s].off;
    for (int c = b.sys_bchunk[s]; c < b.sys_bchunk[s + 1]; ++c) {
        const int t = b.bchunk_sum[c];
        b.bchunk_sum[c] = off;
        off += t;
    }
}
__global__ __launch_bounds__(SCAN_THREADS) void k_bin_scan_chunks(cgpu_voronoi_batch b) {
    const int cidx = blockIdx.x;
    const int u0 = b.bchunk_begin[cidx], u1 = b.bchunk_end[cidx];
    scan_bins(b.bin_count + u0, b.bin_start + u0, u1 - u0, b.bchunk_sum[cidx]);
    if (cidx == b.num_bin_chunks - 1 && threadIdx.x == 0) b.bin_start[b.total_bins] = b.num_cells;
}

// prefix sum of the bin counts: one pass when every system is one bin chunk
void launch_bin_scan(const cgpu_voronoi_batch &b, hipStream_t st) {
    if (b.num_bin_chunks == b.num_systems) {
        hipLaunchKernelGGL(k_bin_scan, dim3(b.num_systems), dim3(SCAN_THREADS), 0, st, b);
        return;
    }
    hipLaunchKernelGGL(k_bin_chunk_sum, dim3(b.num_bin_chunks), dim3(SCAN_THREADS), 0, st, b);
    hipLaunchKernelGGL(k_bin_chunk_offsets, dim3((b.num_systems + 63) / 64), dim3(64), 0, st, b);
    hipLaunchKernelGGL(k_bin_scan_chunks, dim3(b.num_bin_chunks), dim3(SCAN_THREADS), 0, st, b);
}

// Pass 3: every cell to its bin-ordered slot, with its position. The rank is
// consumed here and replaced by the slot (the circumcircle test skips a
// triangle's own corners by slot).
__global__ void k_bin_scatter(cgpu_voronoi_batch b, const double *pos) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= b.num_cells) return;
    const int k = b.bin_start[b.cell_bin[i]] + b.cell_rank[i];
    b.sorted_id[k] = i;
    reinterpret_cast<double2 *>(b.sorted_pos)[k] = reinterpret_cast<const double2 *>(pos)[i];
    b.cell_rank[i] = k;
}

// flags cell m of system s for repair (once per step) and counts it per system,
// one atomic per wave when the freshly flagged lanes share their system
__device__ inline void flag_cell(const cgpu_voronoi_batch &b, int s, int m) {
    const bool fresh = atomicExch(&b.repair[m], 1) == 0;
    const int q = wave_reserve(fresh, &b.counters[0]);
    if (fresh) b.flagged[q] = m;
    const unsigned long long fm = __ballot(fresh);
    if (fm == 0ull) return;
    const int leader = __ffsll((long long)fm) - 1;
    const int s0 = __shfl(s, leader);
    if (__ballot(fresh && s != s0) == 0ull) {
        if ((int)__lane_id() == leader) atomicAdd(&b.st[s0].nflag, __popcll(fm));
    } else if (fresh) {
        atomicAdd(&b.st[s].nflag, 1);
    }
}
// More than N/6 flagged cells already send the system to a global
// retriangulation (|NeedsFixing| >= |flagged|, voronoiModelBase.cpp:650-653),
// whatever else the test would flag, so its remaining scans can stop.
// The count only grows during the test and only lets work be skipped, so a
// cached (workgroup-scope) read that may lag behind other work-groups' flags is
// enough: no cache bypass on every triangle.
__device__ inline bool system_goes_global(const cgpu_voronoi_batch &b, int s, const cgpu_sys_params &sp) {
    return __hip_atomic_load(&b.st[s].nflag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) > sp.n / 6;
}

// Reference window of one triangle (voronoiModelBase.cu:24-109): circumcircle
// of (0, pt1, pt2) relative to cell i at v, wcheck = ceil(rad / boxsize) capped
// at xsize / 2 (a non-finite radius takes the cap), bins ib +- wcheck around
// i's (unclamped) bin with one periodic wrap per index.
struct TriWindow {
    dvec2 Q;
    double rad2;      // squared radius
    double r1;        // radius + 1e-9 L, for bin rejection
    int wcheck;
    bool all_inside;  // wider than any wrapped distance: every candidate is inside
    bool can_skip;    // bin rejection valid (window and circle well inside half a box)
};
__device__ inline TriWindow tri_window(const cgpu_sys_params &sp, dvec2 pt1, dvec2 pt2) {
    TriWindow t;
    t.Q = circ0<double>(pt1, pt2);
    const double dx = pt1.x - t.Q.x, dy = pt1.y - t.Q.y;
    const double rad = sqrt(dx * dx + dy * dy);
    const int xs = sp.xsize;
    const double wc = ceil(rad / sp.boxsize);
    t.wcheck = (wc < (double)(xs / 2)) ? (int)wc : xs / 2;
    t.rad2 = rad * rad;
    // wrapped distances are at most L/sqrt(2): a wider circle contains every
    // candidate, exactly as the reference's wrap loops decide
    t.all_inside = t.rad2 > 0.5 * sp.L * sp.L * (1.0 + 1e-9);
    // A bin whose rectangle (seen from v, unwrapped) stays farther than the
    // radius from Q (plus 1e-9 L for rounding and bin-edge assignment) holds no
    // point the reference test can find inside. Valid while the window and the
    // circle stay well inside half a box, so every minimum image is the
    // unwrapped one.
    t.r1 = sqrt(t.rad2) + 1e-9 * sp.L;
    t.can_skip = !t.all_inside && (t.wcheck + 2) * sp.boxsize + 2.0 * t.r1 < 0.5 * sp.L;
    return t;
}
// true when bin column ii / row jj (unwrapped indices) may hold a point inside
__device__ inline bool bin_may_hit(const cgpu_sys_params &sp, const TriWindow &t, dvec2 v, int ii, int jj) {
    if (!t.can_skip) return true;
    const double x0 = ii * sp.boxsize - v.x, y0 = jj * sp.boxsize - v.y;
    const double ddx = fmax(fmax(x0 - t.Q.x, t.Q.x - (x0 + sp.boxsize)), 0.0);
    const double ddy = fmax(fmax(y0 - t.Q.y, t.Q.y - (y0 + sp.boxsize)), 0.0);
    return !(ddx * ddx + ddy * ddy > t.r1 * t.r1);
}
// reference inside test of candidate m: minDist(minDist(p_m, v), Q) < rad^2
__device__ inline bool cand_inside(const Box &box, const TriWindow &t, dvec2 a) {
    if (t.all_inside) return true;
    const dvec2 c = min_dist(box, a, t.Q);
    return c.x * c.x + c.y * c.y < t.rad2;
}
__device__ inline void flag_triangle(const cgpu_voronoi_batch &b, int s, int i, int j, int l) {
    b.st[s].any_fail = 1;
    flag_cell(b, s, i);
    flag_cell(b, s, j);
    flag_cell(b, s, l);
}

// serial scan of one triangle's whole window (any wcheck)
__device__ __forceinline__ void test_triangle(const cgpu_voronoi_batch &b, const double *pos, const cgpu_sys_params &sp,
                              const Box &box, int s, int i, dvec2 v, int j, int l, const TriWindow &t) {
    const int xs = sp.xsize;
    const int ib = (int)floor(v.x / sp.boxsize), jb = (int)floor(v.y / sp.boxsize);
    bool bad = false;
    for (int ii = ib - t.wcheck; ii <= ib + t.wcheck; ++ii) {
        if (system_goes_global(b, s, sp)) break;
        int cx = ii;
        if (cx < 0) cx += xs;
        if (cx >= xs) cx -= xs;
        for (int jj = jb - t.wcheck; jj <= jb + t.wcheck; ++jj) {
            int cy = jj;
            if (cy < 0) cy += xs;
            if (cy >= xs) cy -= xs;
            if (!bin_may_hit(sp, t, v, ii, jj)) continue;
            const int bin = sp.bin_off + cy * xs + cx;
            const int k1 = b.bin_start[bin + 1];
            for (int k = b.bin_start[bin]; k < k1; ++k) {
                const int m = b.sorted_id[k];
                if (m == i || m == j || m == l) continue;
                if (t.all_inside || cand_inside(box, t, min_dist(box, ldpos(b.sorted_pos, k), v))) {
                    bad = true;
                    flag_cell(b, s, m);
                }
            }
        }
    }
    if (bad) flag_triangle(b, s, i, j, l);
}

// serial scan of the window of owned triangle (i, j, l), any wcheck
__device__ __forceinline__ void serial_test_wide(const cgpu_voronoi_batch &b, const double *pos, int i, int j, int l) {
    const int s = b.cell_sys[i];
    const cgpu_sys_params &sp = b.sys[s];
    const Box box = make_box(sp.L);
    const dvec2 v = ldpos(pos, i);
    test_triangle(b, pos, sp, box, s, i, v, j, l,
                  tri_window(sp, min_dist(box, ldpos(pos, j), v), min_dist(box, ldpos(pos, l), v)));
}

// Empty-circumcircle test of one owned triangle (i, j, l) (reference
// gpu_test_circumcenters_kernel, voronoiModelBase.cu:24-109, with its search
// window) when its window is the 3 x 3 bins around cell i's bin (wcheck = 1):
// independent loads of the nine bins. Returns false, testing nothing, when the
// window is wider. TEAM > 1: the TEAM lanes of an aligned group (t = 0 .. TEAM-1)
// test the same triangle, lane t taking the candidates p = t mod TEAM of the
// concatenated runs; the team's ballot decides the triangle's flag.
template <int TEAM = 1>
__device__ __forceinline__ bool test_narrow(const cgpu_voronoi_batch &b, const double *pos, int i, int j, int l,
                                            int t = 0) {
    const int s = b.cell_sys[i];
    const cgpu_sys_params &sp = b.sys[s];
    if (system_goes_global(b, s, sp)) return true;
    const Box box = make_box(sp.L);
    const dvec2 v = ldpos(pos, i);
    // the corners' bin-ordered slots (cell_rank after the scatter pass): they are
    // excluded from the scan without loading them
    const int ki = b.cell_rank[i], kj = b.cell_rank[j], kl = b.cell_rank[l];
    const TriWindow w = tri_window(sp, min_dist(box, ldpos(pos, j), v), min_dist(box, ldpos(pos, l), v));
    if (w.wcheck != 1 || w.all_inside) return false;
    const int xs = sp.xsize;
    const int ib = (int)floor(v.x / sp.boxsize), jb = (int)floor(v.y / sp.boxsize);
    // window columns ib-1 .. ib+1 (one periodic wrap each); rows jb-1 .. jb+1.
    // Within a row the three bins are one run of bin-ordered slots unless the
    // row wraps, then the run is cut where the columns wrap.
    auto wrapc = [&](int c) { return c < 0 ? c + xs : (c >= xs ? c - xs : c); };
    const int cc0 = wrapc(ib - 1), cc1 = wrapc(ib), cc2 = wrapc(ib + 1);
    auto ccol = [&](int c) { return c == 0 ? cc0 : (c == 1 ? cc1 : cc2); };
    // columns / rows of the window that the circle's extent (plus 1e-9 L) can
    // reach; all three when bin rejection is not valid (see TriWindow)
    int ca = 0, cb = 2, ra = 0, rb = 2;
    if (w.can_skip) {
        const double bs = sp.boxsize;
        const double xl = w.Q.x - w.r1, xh = w.Q.x + w.r1, yl = w.Q.y - w.r1, yh = w.Q.y + w.r1;
        ca = ((ib) * bs - v.x >= xl) ? 0 : (((ib + 1) * bs - v.x >= xl) ? 1 : 2);
        cb = ((ib + 1) * bs - v.x <= xh) ? 2 : ((ib * bs - v.x <= xh) ? 1 : 0);
        ra = ((jb) * bs - v.y >= yl) ? 0 : (((jb + 1) * bs - v.y >= yl) ? 1 : 2);
        rb = ((jb + 1) * bs - v.y <= yh) ? 2 : ((jb * bs - v.y <= yh) ? 1 : 0);
    }
    // a row's reachable columns form one run of bin-ordered slots unless the
    // columns wrap between ca and cb
    const bool run = !((ca <= 0 && cb >= 1 && cc1 != cc0 + 1) || (ca <= 1 && cb >= 2 && cc2 != cc1 + 1));
    const int ca_bin = ccol(ca), cb_bin = ccol(cb);
    int lo[9], hi[9];  // up to three runs per row
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        int cy = jb - 1 + r;
        if (cy < 0) cy += xs;
        if (cy >= xs) cy -= xs;
        const int rowb = sp.bin_off + cy * xs;
        const bool rowon = r >= ra && r <= rb && ca <= cb;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const bool on = rowon && (run ? (c == 0) : (c >= ca && c <= cb));
            const int first = run ? ca_bin : ccol(c);
            const int last = run ? cb_bin : ccol(c);
            lo[3 * r + c] = on ? b.bin_start[rowb + first] : 0;
            hi[3 * r + c] = on ? b.bin_start[rowb + last + 1] : 0;
        }
    }
    bool bad = false;
    auto corner = [&](int k) { return k == ki || k == kj || k == kl; };
    auto scan_run = [&](int k0, int k1) {
        // two candidates per iteration: their loads are in flight together; a
        // candidate's id is read only when it lies inside
        for (int k = k0; k < k1; k += 2) {
            const bool c0 = !corner(k);
            const bool c1 = k + 1 < k1 && !corner(k + 1);
            const dvec2 q0 = c0 ? ldpos(b.sorted_pos, k) : v;
            const dvec2 q1 = c1 ? ldpos(b.sorted_pos, k + 1) : v;
            const bool in0 = c0 && cand_inside(box, w, min_dist(box, q0, v));
            const bool in1 = c1 && cand_inside(box, w, min_dist(box, q1, v));
            if (in0) flag_cell(b, s, b.sorted_id[k]);
            if (in1) flag_cell(b, s, b.sorted_id[k + 1]);
            bad = bad || in0 || in1;
        }
    };
    if constexpr (TEAM == 1) {
        scan_run(lo[0], hi[0]);
        scan_run(lo[1], hi[1]);
        scan_run(lo[2], hi[2]);
        scan_run(lo[3], hi[3]);
        scan_run(lo[4], hi[4]);
        scan_run(lo[5], hi[5]);
        scan_run(lo[6], hi[6]);
        scan_run(lo[7], hi[7]);
        scan_run(lo[8], hi[8]);
        if (bad) flag_triangle(b, s, i, j, l);
    } else {
        // lane t's candidates are p = t, t + TEAM, ... of the concatenated runs:
        // their slots are found from the run lengths first, then up to CAND of
        // them are loaded together (one round trip instead of one per run)
        constexpr int CAND = 4;
        int len[9], tot = 0;
#pragma unroll
        for (int r = 0; r < 9; ++r) {
            len[r] = hi[r] - lo[r];
            tot += len[r];
        }
        for (int base = t; base < tot; base += CAND * TEAM) {
            int kk[CAND];
            bool ok[CAND];
#pragma unroll
            for (int m = 0; m < CAND; ++m) {
                const int p = base + m * TEAM;
                int acc = 0, k = 0;
#pragma unroll
                for (int r = 0; r < 9; ++r) {
                    if (p >= acc && p < acc + len[r]) k = lo[r] + (p - acc);
                    acc += len[r];
                }
                kk[m] = k;
                ok[m] = p < tot && !corner(k);
            }
            dvec2 q[CAND];
#pragma unroll
            for (int m = 0; m < CAND; ++m) q[m] = ok[m] ? ldpos(b.sorted_pos, kk[m]) : v;
#pragma unroll
            for (int m = 0; m < CAND; ++m)
                if (ok[m] && cand_inside(box, w, min_dist(box, q[m], v))) {
                    flag_cell(b, s, b.sorted_id[kk[m]]);
                    bad = true;
                }
        }
        const unsigned long long m = __ballot(bad);
        const int first = (int)__lane_id() & ~(TEAM - 1);
        if (t == 0 && ((m >> first) & ((1ull << TEAM) - 1ull)) != 0ull) flag_triangle(b, s, i, j, l);
    }
    return true;
}

// The same for a wider window, one wavefront per triangle: the lanes deal out
// the (2 wcheck + 1)^2 bins of the reference window (including its repeated
// columns when the window is wider than the box).
__device__ __forceinline__ void wave_test_wide(const cgpu_voronoi_batch &b, const double *pos, int i, int j, int l) {
    const int s = b.cell_sys[i];
    const cgpu_sys_params &sp = b.sys[s];
    const Box box = make_box(sp.L);
    const dvec2 v = ldpos(pos, i);
    const TriWindow w = tri_window(sp, min_dist(box, ldpos(pos, j), v), min_dist(box, ldpos(pos, l), v));
    const int xs = sp.xsize;
    const int ib = (int)floor(v.x / sp.boxsize), jb = (int)floor(v.y / sp.boxsize);
    const int nb = 2 * w.wcheck + 1;
    const int B = nb * nb;
    bool bad = false;
    for (int u = __lane_id(); u < B; u += 64) {
        if ((u & 511) < 64 && system_goes_global(b, s, sp)) break;
        const int ii = ib - w.wcheck + u / nb, jj = jb - w.wcheck + u % nb;
        if (!bin_may_hit(sp, w, v, ii, jj)) continue;
        int cx = ii, cy = jj;
        if (cx < 0) cx += xs;
        if (cx >= xs) cx -= xs;
        if (cy < 0) cy += xs;
        if (cy >= xs) cy -= xs;
        const int bin = sp.bin_off + cy * xs + cx;
        const int k1 = b.bin_start[bin + 1];
        for (int k = b.bin_start[bin]; k < k1; ++k) {
            const int m = b.sorted_id[k];
            if (m == i || m == j || m == l) continue;
            if (w.all_inside || cand_inside(box, w, min_dist(box, ldpos(b.sorted_pos, k), v))) {
                bad = true;
                flag_cell(b, s, m);
            }
        }
    }
    if (bad) flag_triangle(b, s, i, j, l);
}

// sum of three 64-bit values over the wavefront (all lanes call)
__device__ inline void wave_sum3(unsigned long long &a, unsigned long long &o, unsigned long long &r) {
    for (int off = 32; off > 0; off >>= 1) {
        a += ((unsigned long long)__shfl_xor((unsigned int)(a >> 32), off) << 32) | __shfl_xor((unsigned int)a, off);
        o += ((unsigned long long)__shfl_xor((unsigned int)(o >> 32), off) << 32) | __shfl_xor((unsigned int)o, off);
        r += ((unsigned long long)__shfl_xor((unsigned int)(r >> 32), off) << 32) | __shfl_xor((unsigned int)r, off);
    }
}
// add each system's part of the wavefront's (a, o, r) to its sums with one set
// of atomics per system; lanes hold cells of consecutive systems (all lanes call)
__device__ inline void wave_add_by_system(const cgpu_voronoi_batch &b, bool live, int s, unsigned long long a,
                                          unsigned long long o, unsigned long long r) {
    unsigned long long todo = __ballot(live);
    while (todo != 0ull) {
        const int leader = __ffsll((long long)todo) - 1;
        const int sx = __shfl(s, leader);
        const bool mine = live && s == sx;
        unsigned long long x = mine ? a : 0ull, y = mine ? o : 0ull, z = mine ? r : 0ull;
        wave_sum3(x, y, z);
        if ((int)__lane_id() == leader) {
            atomicAdd(&b.st[sx].tri_hash_all, x);
            atomicAdd(&b.st[sx].tri_hash_own, y);
            atomicAdd((unsigned long long *)&b.st[sx].tri_area, z);
        }
        todo &= ~__ballot(mine);
    }
}
// the skip rule's per-cell sums of a wavefront. The wavefront covers the 64
// cells from a multiple of 64 (i of lane 0). When they are 64 live cells of
// one system its sums go to that group's record in ld_part, which k_ld_reduce
// adds up: no contention on the system's sums. A wavefront at a system
// boundary adds each system's part with atomics. All lanes call.
__device__ inline void add_ld_sums(const cgpu_voronoi_batch &b, int i, bool live, int s, const LdSums &v) {
    const unsigned long long m = __ballot(live);
    if (m == 0ull) return;
    const int s0 = __shfl(s, __ffsll((long long)m) - 1);
    unsigned long long a = live ? v.all : 0ull, o = live ? v.own : 0ull;
    unsigned long long r = live ? (unsigned long long)v.area : 0ull;
    if (m == ~0ull && __ballot(s != s0) == 0ull) {
        wave_sum3(a, o, r);
        if (__lane_id() == 0) {
            unsigned long long *g = b.ld_part + 3 * (size_t)(i >> 6);
            g[0] = a;
            g[1] = o;
            g[2] = r;
        }
        return;
    }
    wave_add_by_system(b, live, s, a, o, r);
}

// Geometry of every cell. With test set also the skip rule's checks (regions
// to stamp, circumradii, per-system sums) for k_select_triangles.
template <typename R>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(5, 8))) void k_test_geometry(
    cgpu_voronoi_batch b, const double *pos, int test) {
    const int i = xcd_block() * blockDim.x + threadIdx.x;
    const bool live = i < b.num_cells && b.st[b.cell_sys[i]].active;
    if (!test) {  // uniform
        if (live) cell_geometry<R>(b, pos, i, b.sys[b.cell_sys[i]]);
        return;
    }
    LdSums sums{0ull, 0ull, 0ll};
    int s = 0;
    if (live) {
        s = b.cell_sys[i];
        cell_geometry<R, true>(b, pos, i, b.sys[s], &sums);
    }
    add_ld_sums(b, i, live, s, sums);
}

// whether the skip rule holds for system s in this test: no region too wide to
// stamp, a grid wide enough for the dilation, consistent rings, one covering
__device__ inline bool skip_rule_on(const cgpu_voronoi_batch &b, int s, const cgpu_sys_params &sp) {
    const cgpu_sys_state &st = b.st[s];
    const double L2 = sp.L * sp.L;
    return !st.skip_off && sp.xsize >= 2 * MARK_DILATE + 2 && st.tri_hash_all == 3ull * st.tri_hash_own &&
           fabs((double)st.tri_area / AREA_UNIT - L2) <= 1e-6 * L2;
}

// After the geometry pass: (1) the records of the 64-cell groups that lie
// inside one active system are added to that system's sums (one thread per
// group, one set of atomics per system and wavefront); (2) the recorded regions
// are stamped, one wavefront per region with a lane per bin (a star's bounding
// box from its ring first).
__global__ void k_ld_reduce(cgpu_voronoi_batch b, const double *pos) {
    const int N = b.num_cells, G = N >> 6;
    for (int base = blockIdx.x * blockDim.x; base < G; base += gridDim.x * blockDim.x) {  // uniform per block
        const int g = base + threadIdx.x;
        bool full = false;
        int s = 0;
        if (g < G) {
            const int c0 = g << 6;
            s = b.cell_sys[c0];
            const cgpu_sys_params &sp = b.sys[s];
            full = c0 >= sp.off && c0 + 64 <= sp.off + sp.n && b.st[s].active;
        }
        const unsigned long long *r = b.ld_part + 3 * (size_t)g;
        wave_add_by_system(b, full, s, full ? r[0] : 0ull, full ? r[1] : 0ull, full ? r[2] : 0ull);
    }
    const int n = min(b.counters[8], b.ld_region_cap);
    const int epoch = b.counters[7];
    const LdRegion *reg = reinterpret_cast<const LdRegion *>(b.ld_regions);
    const int lane = __lane_id();
    const int nw = (gridDim.x * blockDim.x) >> 6;
    for (int q = (int)(blockIdx.x * blockDim.x + threadIdx.x) >> 6; q < n; q += nw) {  // uniform per wavefront
        const LdRegion rg = reg[q];
        int sys = rg.sys;
        double x0 = rg.x0, x1 = rg.x1, y0 = rg.y0, y1 = rg.y1;
        if (sys < 0) {  // star of cell i: bounding box of its ring (relative), then absolute
            const int i = -1 - sys;
            sys = b.cell_sys[i];
            const cgpu_sys_params &sp = b.sys[sys];
            const Box box = make_box(sp.L);
            const int z = min(b.nn[i], ring_cap(b));
            const dvec2 pi = ldpos(pos, i);
            dvec2 d{0.0, 0.0};
            if (lane < z) d = min_dist(box, ldpos(pos, b.nbr[lane * N + i]), pi);
            double a0 = fmin(d.x, 0.0), a1 = fmax(d.x, 0.0), c0 = fmin(d.y, 0.0), c1 = fmax(d.y, 0.0);
            for (int off = 32; off > 0; off >>= 1) {
                a0 = fmin(a0, __shfl_xor(a0, off));
                a1 = fmax(a1, __shfl_xor(a1, off));
                c0 = fmin(c0, __shfl_xor(c0, off));
                c1 = fmax(c1, __shfl_xor(c1, off));
            }
            x0 = pi.x + a0;
            x1 = pi.x + a1;
            y0 = pi.y + c0;
            y1 = pi.y + c1;
        }
        const cgpu_sys_params &sp = b.sys[sys];
        const int xs = sp.xsize;
        if (!(x1 - x0 < 0.25 * sp.L && y1 - y0 < 0.25 * sp.L)) {
            if (lane == 0) b.st[sys].skip_off = 1;
            continue;
        }
        const double inv = 1.0 / sp.boxsize;
        const int ia = (int)floor(x0 * inv), ib = (int)floor(x1 * inv);
        const int ja = (int)floor(y0 * inv), jb = (int)floor(y1 * inv);
        const int nu = ib - ia + 2 * MARK_DILATE + 1, nv = jb - ja + 2 * MARK_DILATE + 1;
        if (!(nu < xs && nv < xs)) {
            if (lane == 0) b.st[sys].skip_off = 1;
            continue;
        }
        for (int t = lane; t < nu * nv; t += 64) {
            const int u = ia - MARK_DILATE + t % nu, w = ja - MARK_DILATE + t / nu;
            const int cu = ((u % xs) + xs) % xs, cw = ((w % xs) + xs) % xs;
            const int bin = sp.bin_off + cw * xs + cu;
            b.bin_mark_wide[bin] = epoch;
        }
    }
}



// The triangles each cell owns (i the smallest index of the three: the
// triangles the reference tests, voronoiModelBase.cpp:380-423) that the skip
// rule leaves to the empty-circumcircle test: all of them where the rule is
// off for the system, else those of cells whose bin is stamped or whose largest
// circumradius exceeds the dilation. They are gathered in LDS and appended to
// the triangle queue with one reservation per block; a triangle beyond the LDS
// list (more than TRI_PER_CELL per cell of the block) or beyond the queue's
// capacity is tested where it is found.
constexpr int TRI_PER_CELL = 5;
// grid of k_test_queue: 8 blocks per CU; the blocks loop over the queue
constexpr int TEST_QUEUE_BLOCKS = 2048;
__global__ __launch_bounds__(512) void k_select_triangles(cgpu_voronoi_batch b, const double *pos) {
    extern __shared__ int3 test_lds[];
    const int TRI_LDS = TRI_PER_CELL * (int)blockDim.x;
    int3 *stri = test_lds;
    __shared__ int scount, sepoch, sbase;
    const int i = xcd_block() * blockDim.x + threadIdx.x;
    if (threadIdx.x == 0) {
        scount = 0;
        sepoch = b.counters[7];
    }
    __syncthreads();
    const bool live = i < b.num_cells && b.st[b.cell_sys[i]].active;
    if (live) {
        const int s = b.cell_sys[i];
        const cgpu_sys_params &sp = b.sys[s];
        // the rule of the cell's system, from the sums the reduce pass completed
        const bool near = !skip_rule_on(b, s, sp) || b.bin_mark_wide[b.cell_bin[i]] == sepoch ||
                          !(2.0 * (double)b.cell_rmax[i] <= (MARK_DILATE - 1) * sp.boxsize);
        if (near) {
            const int N = b.num_cells;
            const int z = b.nn[i];
            int jn = b.nbr[i];
            for (int k = 0; k < z; ++k) {
                const int j = jn;
                jn = b.nbr[((k + 1) % z) * N + i];
                if (!(i < j && i < jn)) continue;
                const int slot = atomicAdd(&scount, 1);
                if (slot < TRI_LDS)
                    stri[slot] = make_int3(i, j, jn);
                else if (!test_narrow(b, pos, i, j, jn))
                    serial_test_wide(b, pos, i, j, jn);
            }
        }
    }
    __syncthreads();
    const int n = min(scount, TRI_LDS);
    if (threadIdx.x == 0) sbase = n > 0 ? atomicAdd(&b.counters[9], n) : 0;
    __syncthreads();
    const int base = sbase;
    for (int q = threadIdx.x; q < n; q += blockDim.x) {
        const int3 t = stri[q];
        const long long at = (long long)base + q;
        if (at < b.ld_tri_cap) {
            int *d = b.ld_tri + 3 * at;
            d[0] = t.x;
            d[1] = t.y;
            d[2] = t.z;
        } else if (!test_narrow(b, pos, t.x, t.y, t.z)) {
            serial_test_wide(b, pos, t.x, t.y, t.z);
        }
    }
}

// The empty-circumcircle test of the queued triangles, spread evenly over the
// grid: a team of TEST_TEAM lanes per triangle for the 3 x 3-bin windows (the
// queue is short, so the lanes shorten each triangle's chain of loads instead
// of idling); triangles with a wider window go to an LDS list that the block's
// wavefronts scan cooperatively, one wavefront per triangle.
constexpr int TEST_TEAM = 8;
__global__ __launch_bounds__(256) void k_test_queue(cgpu_voronoi_batch b, const double *pos) {
    __shared__ int sn, swcount;
    __shared__ int swide[256 / TEST_TEAM];
    if (threadIdx.x == 0) sn = min(b.counters[9], b.ld_tri_cap);
    __syncthreads();
    const int n = sn;
    const int per_block = (int)blockDim.x / TEST_TEAM, t = (int)threadIdx.x & (TEST_TEAM - 1);
    for (int base = blockIdx.x * per_block; base < n; base += gridDim.x * per_block) {  // uniform per block
        if (threadIdx.x == 0) swcount = 0;
        __syncthreads();
        const int q = base + (int)threadIdx.x / TEST_TEAM;
        if (q < n) {
            const int *tr = b.ld_tri + 3 * (size_t)q;
            if (!test_narrow<TEST_TEAM>(b, pos, tr[0], tr[1], tr[2], t) && t == 0) swide[atomicAdd(&swcount, 1)] = q;
        }
        __syncthreads();
        const int nw = swcount;
        for (int w = threadIdx.x >> 6; w < nw; w += blockDim.x >> 6) {
            const int *tw = b.ld_tri + 3 * (size_t)swide[w];
            wave_test_wide(b, pos, tw[0], tw[1], tw[2]);
        }
        __syncthreads();
    }
}

void launch_test_geometry(const cgpu_voronoi_batch &b, const double *pos, int test, hipStream_t st,
                          const Probe *pr = nullptr) {
    const int bs = shape_test_block(b);
    const dim3 g((unsigned)((b.num_cells + bs - 1) / bs)), t(bs);
    if (b.precision)
        hipLaunchKernelGGL(k_test_geometry<float>, g, t, 0, st, b, pos, test);
    else
        hipLaunchKernelGGL(k_test_geometry<double>, g, t, 0, st, b, pos, test);
    if (!test) {
        probe(pr, PROBE_TEST_GEOMETRY);
        return;
    }
    probe(pr, PROBE_TEST_GEOMETRY);
    hipLaunchKernelGGL(k_ld_reduce, dim3(256), dim3(256), 0, st, b, pos);
    probe(pr, PROBE_LD_REDUCE);
    const size_t lds = (size_t)TRI_PER_CELL * bs * sizeof(int3);
    hipLaunchKernelGGL(k_select_triangles, g, t, lds, st, b, pos);
    probe(pr, PROBE_SELECT_TRIANGLES);
    hipLaunchKernelGGL(k_test_queue, dim3(TEST_QUEUE_BLOCKS), dim3(256), 0, st, b, pos);
    probe(pr, PROBE_TEST_QUEUE);
}

// NeedsFixing = flagged cells and their current neighbours (reference
// voronoiModelBase.cpp:617-646), built as a deduplicated device queue.
// (QL = the launch shape's queue_lanes threads per flagged cell: entry e = 0 is
// the cell, e >= 1 its ring slot e - 1; thread k0 takes entries k0, k0 + QL, ...)
__global__ void k_mark_nf(cgpu_voronoi_batch b) {
    const int nflag = b.counters[0];
    const int N = b.num_cells;
    const int QL = shape_queue_lanes(b), lg = log2_lanes(QL);
    const long long total = (long long)nflag * QL;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        const int q = (int)(t >> lg), k0 = (int)(t & (QL - 1));
        const int i = b.flagged[q];
        const int z = b.nn[i];
        for (int e = k0; e <= z; e += QL) {
            const int x = (e == 0) ? i : b.nbr[(e - 1) * N + i];
            const bool fresh = atomicExch(&b.mark[x], 1) == 0;
            const int slot = wave_reserve(fresh, &b.counters[1]);
            const int sx = fresh ? b.cell_sys[x] : 0;
            if (fresh) b.nfq[slot] = x;
            // per-system count, one atomic per wave when the fresh lanes share a system
            const unsigned long long fm = __ballot(fresh);
            if (fm != 0ull) {
                const int leader = __ffsll((long long)fm) - 1;
                const int s0 = __shfl(sx, leader);
                if (__ballot(fresh && sx != s0) == 0ull) {
                    if ((int)__lane_id() == leader) atomicAdd(&b.st[s0].nf_count, __popcll(fm));
                } else if (fresh) {
                    atomicAdd(&b.st[sx].nf_count, 1);
                }
            }
        }
        if (k0 == 0) b.repair[i] = 0;
    }
}

// index of the chunk that holds cell i (binary search over chunk starts)
__device__ inline int chunk_of(const cgpu_voronoi_batch &b, int i) {
    int lo = 0, hi = b.num_chunks - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (b.chunk_begin[mid] <= i)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// enqueue the cells of chunk c into the global queue (one reservation per chunk).
// Called by every thread of a block (uniform per block). The closing barrier
// keeps the next call's reservation from overwriting `base` while lanes of this
// call still read it (callers loop over several chunks per block).
__device__ void enqueue_chunk(const cgpu_voronoi_batch &b, int c) {
    __shared__ int base;
    const int c0 = b.chunk_begin[c], n = b.chunk_end[c] - c0;
    if (threadIdx.x == 0) base = atomicAdd(&b.counters[2], n);
    __syncthreads();
    const int bs = base;
    for (int t = threadIdx.x; t < n; t += blockDim.x) b.gq[bs + t] = c0 + t;
    __syncthreads();
}

// The repair of system s once NeedsFixing is complete: 0 none, 1 local, 2 global
// (|NeedsFixing| > N/6, reference voronoiModelBase.cpp:650-653).
__device__ inline int repair_mode(const cgpu_voronoi_batch &b, int s) {
    const cgpu_sys_state &st = b.st[s];
    return st.any_fail ? (st.nf_count > b.sys[s].n / 6 ? 2 : 1) : 0;
}

// every system to a global retriangulation (reference globalTriangulationCGAL,
// voronoiModelBase.cpp:281-351): one block per cell chunk queues its cells
__global__ void k_global_all(cgpu_voronoi_batch b) {
    const int c = blockIdx.x;
    const int s = b.chunk_sys[c];
    if (threadIdx.x == 0 && b.chunk_begin[c] == b.sys[s].off) b.st[s].mode = 2;
    enqueue_chunk(b, c);
}

// the same for the active systems only
__global__ void k_global_active(cgpu_voronoi_batch b) {
    const int c = blockIdx.x;
    const int s = b.chunk_sys[c];
    if (!b.st[s].active) return;  // uniform per block
    if (threadIdx.x == 0 && b.chunk_begin[c] == b.sys[s].off) b.st[s].mode = 2;
    enqueue_chunk(b, c);
}

// local repair of cell i: new Delaunay star of a NeedsFixing cell (reference
// repairTriangulation, voronoiModelBase.cpp:429-505), with the neighMax rule and
// the changes of the 6N / 2N invariants accumulated per system. Serial form.
__device__ void star_local_serial(const cgpu_voronoi_batch &b, const double *pos, int i, int w0) {
    const int N = b.num_cells;
    const int s = b.cell_sys[i];
    cgpu_sys_state &st = b.st[s];
    const cgpu_sys_params &sp = b.sys[s];
    const int zold = b.nn[i];
    if (verify_ring(b, pos, i, sp)) return;  // unchanged: no invariant change
    const int told = owned_triangles(b, i);
    int ring[CAP_MAX];
    const int z = delaunay_star(b, pos, i, sp, ring, w0);
    if (z < 0 || z > st.neigh_max) {
        st.need_global = 1;
        return;
    }
    b.nn[i] = z;
    b.mark[i] = 2;
    for (int k = 0; k < z; ++k) b.nbr[k * N + i] = ring[k];
    const int tnew = owned_triangles(b, i);
    if (z != zold) atomicAdd(&st.nn_delta, z - zold);
    if (tnew != told) atomicAdd(&st.tri_delta, tnew - told);
}

// Star of queued cell i by one wavefront: one candidate per lane, two where a
// bin block holds more than 64 points. Returns as wave_star; on -3 the serial
// form is left, starting at half-width w_at.
__device__ inline int wave_star_any(const cgpu_voronoi_batch &b, const double *pos, int i,
                                    const cgpu_sys_params &sp, int &ring, int w0, int &w_at) {
    int z = wave_star<1>(b, pos, i, sp, ring, w0, w_at);
    if (z == -3) z = wave_star<2>(b, pos, i, sp, ring, w_at, w_at);
    return z;
}

// The same, one wavefront per cell of the NeedsFixing queue.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_star_local_wave(cgpu_voronoi_batch b, const double *pos) {
    {
        // also records each system's repair decision (k_global_fallback queues
        // the cells of the systems that go global)
        const int nthreads = gridDim.x * blockDim.x;
        for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < b.num_systems; s += nthreads) {
            const int mode = repair_mode(b, s);
            if (mode == 0) continue;
            cgpu_sys_state &st = b.st[s];
            st.mode = mode;
            st.repair_events += 1;
            if (mode == 1) st.local_repairs += st.nf_count;
        }
    }
    const int nq = b.counters[1];
    const int N = b.num_cells;
    const int lane = __lane_id();
    const int nw = (gridDim.x * blockDim.x) >> 6;
    // one wavefront per queue entry: the entry and its cell are wave-uniform
    for (int q = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * blockDim.x + threadIdx.x) >> 6); q < nq;
         q += nw) {
        const int i = __builtin_amdgcn_readfirstlane(b.nfq[q]);
        const int s = b.cell_sys[i];
        cgpu_sys_state &st = b.st[s];
        if (repair_mode(b, s) != 1) continue;
        int ring = -1, w_at = 0;
        const int z = wave_star_any(b, pos, i, b.sys[s], ring, 1, w_at);
        if (z == 0) continue;  // verified: unchanged
        if (z == -3) {
            if (lane == 0) star_local_serial(b, pos, i, w_at);
            continue;
        }
        if (z < 0 || z > st.neigh_max) {
            if (lane == 0) st.need_global = 1;
            continue;
        }
        const int zold = b.nn[i];
        const int cap = ring_cap(b);
        const int rold = (lane < zold && lane < cap) ? b.nbr[lane * N + i] : -1;
        const int told = wave_owned(i, rold, min(zold, cap));
        const int tnew = wave_owned(i, ring, z);
        if (lane < z) b.nbr[lane * N + i] = ring;
        if (lane == 0) {
            b.nn[i] = z;
            b.mark[i] = 2;
            if (z != zold) atomicAdd(&st.nn_delta, z - zold);
            if (tnew != told) atomicAdd(&st.tri_delta, tnew - told);
        }
    }
}

// delOther of repaired cells (reference enforceTopology, voronoiModelBase.cpp:
// 700-718) after the 6N/2N invariant check of getCircumcenterIndices (:411-414)
// (QL threads per cell, each taking slots k, k + QL, ...)
__global__ void k_delsets_local(cgpu_voronoi_batch b) {
    const int nq = b.counters[1];
    const int QL = shape_queue_lanes(b), lg = log2_lanes(QL);
    const long long total = (long long)nq * QL;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        const int q = (int)(t >> lg), k0 = (int)(t & (QL - 1));
        const int i = b.nfq[q];
        cgpu_sys_state &st = b.st[b.cell_sys[i]];
        if (st.mode != 1 || st.need_global) continue;
        if (st.nn_delta != 0 || st.tri_delta != 0) {
            st.need_global = 1;
            continue;
        }
        const int z = b.nn[i];
        if (z < 3) {
            st.need_global = 1;
            continue;
        }
        for (int k = k0; k < z; k += QL)
            if (!del_set_slot(b, i, z, k)) st.need_global = 1;
    }
}

// a failed local repair becomes a global one
__global__ void k_global_fallback(cgpu_voronoi_batch b) {
    // one block per cell chunk; the first chunk of a system switches its mode to
    // 3, which the others still accept. The chunks of systems that went global
    // at once (mode 2, the N/6 rule) are queued here too.
    const int c = blockIdx.x;
    const int s = b.chunk_sys[c];
    cgpu_sys_state &st = b.st[s];
    const int mode = st.mode;
    const bool fallback = (mode == 1 || mode == 3) && st.need_global;
    if (!(fallback || mode == 2)) return;  // uniform per block
    __syncthreads();
    if (fallback && threadIdx.x == 0 && b.chunk_begin[c] == b.sys[s].off) {
        st.mode = 3;
        st.nn_delta = 0;
    }
    enqueue_chunk(b, c);
}

// A global star that could not be built. Beyond the ring capacity (z = -2) the
// system stalls: it stops advancing from here on (its step counter is not
// advanced for this step) until the host grows the capacity, retriangulates it
// and resumes it. Any other failure is a topology error.
__device__ inline void star_failed(const cgpu_voronoi_batch &b, int s, int z) {
    if (z == -2) {
        b.st[s].error = CGPU_ERR_RING_CAPACITY;
        b.st[s].active = 0;
        b.counters[4] = 2;
    } else {
        b.st[s].error = CGPU_ERR_TOPOLOGY;
        if (b.counters[4] == 0) b.counters[4] = 1;
    }
}

// global retriangulation (globalTriangulationCGAL) of cell i, serial form
__device__ void star_global_serial(const cgpu_voronoi_batch &b, const double *pos, int i, int w0) {
    const int N = b.num_cells;
    const int s = b.cell_sys[i];
    if (verify_ring(b, pos, i, b.sys[s])) return;
    int ring[CAP_MAX];
    const int z = delaunay_star(b, pos, i, b.sys[s], ring, w0);
    if (z < 0) {
        star_failed(b, s, z);
        return;
    }
    b.nn[i] = z;
    b.mark[i] = 2;
    for (int k = 0; k < z; ++k) b.nbr[k * N + i] = ring[k];
}

// the same, one wavefront per cell of the global queue
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_star_global_wave(cgpu_voronoi_batch b, const double *pos) {
    const int nq = b.counters[2];
    const int N = b.num_cells;
    const int lane = __lane_id();
    const int nw = (gridDim.x * blockDim.x) >> 6;
    // one wavefront per queue entry: the entry and its cell are wave-uniform
    for (int q = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * blockDim.x + threadIdx.x) >> 6); q < nq;
         q += nw) {
        const int i = __builtin_amdgcn_readfirstlane(b.gq[q]);
        const int s = b.cell_sys[i];
        int ring = -1, w_at = 0;
        const int z = wave_star_any(b, pos, i, b.sys[s], ring, 1, w_at);
        if (z == 0) continue;  // verified: unchanged
        if (z == -3) {
            if (lane == 0) star_global_serial(b, pos, i, w_at);
            continue;
        }
        if (z < 0) {
            if (lane == 0) star_failed(b, s, z);
            continue;
        }
        if (lane < z) b.nbr[lane * N + i] = ring;
        if (lane == 0) {
            b.nn[i] = z;
            b.mark[i] = 2;
        }
    }
}

// degree sum (6N invariant, voronoiModelBase.cpp:263) and maximum of the
// retriangulated systems, one block per cell chunk (k_reset and
// k_global_fallback zero the sums)
__global__ void k_global_stats(cgpu_voronoi_batch b) {
    const int c = blockIdx.x;
    const int s = b.chunk_sys[c];
    cgpu_sys_state &st = b.st[s];
    if (st.mode < 2) return;
    int sum = 0, mx = 0;
    for (int i = b.chunk_begin[c] + (int)threadIdx.x; i < b.chunk_end[c]; i += blockDim.x) {
        const int z = b.nn[i];
        sum += z;
        mx = max(mx, z);
    }
    for (int o = 32; o > 0; o >>= 1) {
        sum += __shfl_xor(sum, o);
        mx = max(mx, __shfl_xor(mx, o));
    }
    __shared__ int ssum[16], smax[16];
    const int wv = threadIdx.x >> 6;
    if (__lane_id() == 0) {
        ssum[wv] = sum;
        smax[wv] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < (int)(blockDim.x >> 6); ++k) {
            sum += ssum[k];
            mx = max(mx, smax[k]);
        }
        atomicAdd(&st.nn_delta, sum);
        atomicMax(&st.max_deg, mx);
    }
}

// delOther of retriangulated systems (QL threads per cell, each taking slots
// k, k + QL, ...); geometry of the cells whose ring a star kernel rewrote
// (mark 2; the test kernel already computed it for unchanged rings);
// clears the marks. NeedsFixing cells of systems that went global are handled
// with their system's global queue.
template <typename R>
__global__ void k_finalize(cgpu_voronoi_batch b, const double *pos, int advance) {
    // per retriangulated system (its degree statistics are complete): neighMax =
    // max degree + 2 (even) or + 1 (odd) (voronoiModelBase.cpp:309-312), the 6N
    // check, the global-fix counter. With advance set (the last kernel of a time
    // step) also the step counter of every active system (k_advance_step).
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < b.num_systems; s += gridDim.x * blockDim.x) {
        cgpu_sys_state &st = b.st[s];
        if (st.mode >= 2) {
            const int mx = st.max_deg;
            st.neigh_max = (mx % 2 == 0) ? mx + 2 : mx + 1;
            st.global_fixes += 1;
            // the first error stands (a capacity overflow also breaks the sum)
            if (st.nn_delta != 6 * b.sys[s].n && st.error == 0) st.error = CGPU_ERR_TOPOLOGY;
        }
        if (advance) advance_system(st);
    }
    const int ng = b.counters[2];
    const int nl = b.counters[1];
    const int QL = shape_queue_lanes(b), lg = log2_lanes(QL);
    const long long total = (long long)(ng + nl) * QL;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        const int q = (int)(t >> lg), k0 = (int)(t & (QL - 1));
        if (q < ng) {
            const int i = b.gq[q];
            const int z = b.nn[i];
            if (z >= 3)
                for (int k = k0; k < z; k += QL) del_set_slot(b, i, z, k);
            if (k0 == 0) {
                if (b.mark[i] == 2) cell_geometry<R>(b, pos, i, b.sys[b.cell_sys[i]]);
                b.mark[i] = 0;
            }
        } else if (k0 == 0) {
            const int i = b.nfq[q - ng];
            if (b.st[b.cell_sys[i]].mode != 1) continue;
            if (b.mark[i] == 2) cell_geometry<R>(b, pos, i, b.sys[b.cell_sys[i]]);
            b.mark[i] = 0;
        }
    }
}

// delOther of every slot of every cell from the current rings (reference
// allDelSets, voronoiModelBase.cpp:1281-1286); QL threads per cell
__global__ void k_delsets_all(cgpu_voronoi_batch b) {
    const int QL = shape_queue_lanes(b), lg = log2_lanes(QL);
    const long long total = (long long)b.num_cells * QL;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(t >> lg), k0 = (int)(t & (QL - 1));
        const int z = b.nn[i];
        if (z < 3 || z > ring_cap(b)) continue;
        for (int k = k0; k < z; k += QL)
            if (!del_set_slot(b, i, z, k)) {
                int &err = b.st[b.cell_sys[i]].error;
                if (err == 0) err = CGPU_ERR_TOPOLOGY;
            }
    }
}

// force of every active cell, optional exclusions (reference
// gpu_sum_forces_with_exclusions_kernel, voronoiQuadraticEnergy.cu:59-93)
template <typename R, bool TENS, typename Acc>
__device__ inline vec2<R> force_with_exclusions(const cgpu_voronoi_batch &b, const Acc &acc, int i,
                                                const cgpu_sys_params &sp) {
    vec2<R> F = cell_force<R, TENS>(b, acc, i, sp);
    if (b.excl) {
        if (b.excl[i] != 0) {
            if (b.ext_force) st2<R>(b.ext_force, i, vec2<R>{-F.x, -F.y});
            F = vec2<R>{R(0), R(0)};
        } else if (b.ext_force) {
            st2<R>(b.ext_force, i, vec2<R>{R(0), R(0)});
        }
    }
    return F;
}

template <typename R, bool TENS>
__global__ void k_forces(cgpu_voronoi_batch b, const double *pos) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= b.num_cells) return;
    const cgpu_sys_params &sp = b.sys[b.cell_sys[i]];
    st2<R>(b.force, i, force_with_exclusions<R, TENS>(b, GlobalData<R>{pos, b.dAP}, i, sp));
}

// reference noise draw for cell i: buffer or in-kernel counter-based deviate
__device__ inline double eta_of(const cgpu_voronoi_batch &b, int i, const cgpu_sys_params &sp, int step, int stream,
                                int which) {
    if (b.noise) return b.noise[(size_t)which * b.num_cells + i];
    double z0, z1;
    normal_pair(sp.seed, (uint32_t)(i - sp.off), (uint32_t)step, (uint32_t)stream, z0, z1);
    return which ? z1 : z0;
}

template <typename R> __device__ inline R cos_r(R x);
template <> __device__ inline float cos_r<float>(float x) { return cosf(x); }
template <> __device__ inline double cos_r<double>(double x) { return cos(x); }
template <typename R> __device__ inline R sin_r(R x);
template <> __device__ inline float sin_r<float>(float x) { return sinf(x); }
template <> __device__ inline double sin_r<double>(double x) { return sin(x); }
template <typename R> __device__ inline R atan2_r(R y, R x);
template <> __device__ inline float atan2_r<float>(float y, float x) { return atan2f(y, x); }
template <> __device__ inline double atan2_r<double>(double y, double x) { return atan2(y, x); }

// self-propelled (KIND 0, reference GPU semantics, selfPropelledParticleDynamics.cu:
// 35-53) or Brownian (KIND 1, brownianParticleDynamics.cu:20-39) update of
// active cell i with force F, then the move (Simple2DCell.cu:21-35) into
// pos_next and, with bin set, the binning pass of the cell list
template <typename R, int KIND>
__device__ inline void eom_move(const cgpu_voronoi_batch &b, int i, const cgpu_sys_params &sp, cgpu_sys_state &st,
                                vec2<R> F, const double *pos, double *pos_next, bool bin) {
    dvec2 p = ldpos(pos, i);
    st2<R>(b.force, i, F);
    const int step = st.step + 1;  // Timestep is incremented before integrating
    const R dt = (R)sp.dt, mu = (R)sp.mu;
    vec2<R> disp;
    if (KIND == 0) {
        vec2<R> v = ld2<R>(b.vel, i);
        disp.x = dt * (v.x + mu * F.x);
        disp.y = dt * (v.y + mu * F.y);
        vec2<R> m = ld2<R>(b.mot, i);  // v0, Dr
        R th = reinterpret_cast<R *>(b.theta)[i];
        R angleDiff = (R)eta_of(b, i, sp, step, 0, 0) * sqrt(R(2.0) * dt * m.y);
        R cur = th;
        v.x = m.x * cos_r<R>(th);
        v.y = m.x * sin_r<R>(th);
        if (v.y != R(0) && v.x != R(0)) cur = atan2_r<R>(v.y, v.x);
        reinterpret_cast<R *>(b.theta)[i] = cur + angleDiff;
        st2<R>(b.vel, i, v);
    } else {
        R sc = sqrt(dt * (R)sp.T * mu * R(2.0));
        disp.x = (R)eta_of(b, i, sp, step, 1, 0) * sc + dt * mu * F.x;
        disp.y = (R)eta_of(b, i, sp, step, 1, 1) * sc + dt * mu * F.y;
    }
    p.x += (double)disp.x;
    p.y += (double)disp.y;
    // a non-finite (or absurdly large) position is an error; any other finite
    // displacement is wrapped back into the box as in the reference
    if (!(fabs(p.x) < 1e12 * sp.L && fabs(p.y) < 1e12 * sp.L)) {
        b.counters[5] = 1;
        st.error = CGPU_ERR_NONFINITE;
        p = ldpos(pos, i);
    }
    put_in_box(make_box(sp.L), p);
    stpos(pos_next, i, p);
    if (bin) bin_cell(b, sp, i, p);
}

// fused force -> EOM -> move of every cell, kind as for eom_move.
// flags: STEP_BIN also runs the binning pass of the cell list on the new
// positions; STEP_RESET also clears the step's topology state (k_reset), which
// this kernel neither reads nor writes.
constexpr int STEP_BIN = 1, STEP_RESET = 2;
// the block size is the launch shape's force_block (at most 512); the LDS tile
// (dynamic shared memory) holds one entry per thread. Two kernels share the
// body: the register allocator's own choice, and one held to 5 waves per SIMD
// (a few bytes of spill); the launch shape's force_waves picks one.
template <typename R, int KIND, bool TENS>
__device__ __forceinline__ void force_eom_move_body(const cgpu_voronoi_batch &b, const double *pos, double *pos_next,
                                                    int flags) {
    if (flags & STEP_RESET)
        for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < b.num_systems; s += gridDim.x * blockDim.x)
            reset_system(b, s);
    const bool bin = (flags & STEP_BIN) != 0;
    // LDS copy of the block's cells (positions, dAP): most ring members and
    // delOther cells of a Hilbert-ordered block are in it
    extern __shared__ double2 force_lds[];
    dvec2 *spos = reinterpret_cast<dvec2 *>(force_lds);
    vec2<R> *sdap = reinterpret_cast<vec2<R> *>(force_lds + blockDim.x);
    const int bs = (int)blockDim.x;
    const int c0 = xcd_block() * bs;
    const int i = c0 + threadIdx.x;
    const int n = min(bs, b.num_cells - c0);
    if ((int)threadIdx.x < n) {
        spos[threadIdx.x] = ldpos(pos, i);
        sdap[threadIdx.x] = ld2<R>(b.dAP, i);
    }
    __syncthreads();
    if (i >= b.num_cells) return;
    const BlockTile<R> tile{pos, b.dAP, spos, sdap, c0, n};
    const int s = b.cell_sys[i];
    const cgpu_sys_params &sp = b.sys[s];
    cgpu_sys_state &st = b.st[s];
    if (!st.active) {
        const dvec2 p = ldpos(pos, i);
        stpos(pos_next, i, p);
        if (bin) bin_cell(b, sp, i, p);
        return;
    }
    eom_move<R, KIND>(b, i, sp, st, force_with_exclusions<R, TENS>(b, tile, i, sp), pos, pos_next, bin);
}
template <typename R, int KIND, bool TENS>
__global__ __launch_bounds__(512) void k_force_eom_move(cgpu_voronoi_batch b, const double *pos, double *pos_next,
                                                        int flags) {
    force_eom_move_body<R, KIND, TENS>(b, pos, pos_next, flags);
}
template <typename R, int KIND, bool TENS>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(5, 8))) void k_force_eom_move_w5(
    cgpu_voronoi_batch b, const double *pos, double *pos_next, int flags) {
    force_eom_move_body<R, KIND, TENS>(b, pos, pos_next, flags);
}

__global__ void k_advance_step(cgpu_voronoi_batch b) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < b.num_systems) advance_system(b.st[s]);
}

__global__ void k_fill_normal(cgpu_voronoi_batch b, int step, int stream, double *out) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= b.num_cells) return;
    const cgpu_sys_params &sp = b.sys[b.cell_sys[i]];
    double z0, z1;
    normal_pair(sp.seed, (uint32_t)(i - sp.off), (uint32_t)step, (uint32_t)stream, z0, z1);
    out[i] = z0;
    out[(size_t)b.num_cells + i] = z1;
}


// ------------------------------------------------------------------------ FIRE
// (shared pieces and the iteration: fire_device.h)
// flags as for k_force_eom_move: STEP_RESET also clears the step's topology
// state, STEP_BIN also runs the binning pass of the cell list on the new
// positions
template <typename R>
__global__ void k_fire_vv_move(cgpu_voronoi_batch b, const cgpu_fire_state *fs, void *vel, const double *pos,
                               double *pos_next, int flags) {
    if (flags & STEP_RESET)
        for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < b.num_systems; s += gridDim.x * blockDim.x)
            reset_system(b, s);
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= b.num_cells) return;
    const int s = b.cell_sys[i];
    const cgpu_sys_params &sp = b.sys[s];
    const cgpu_fire_state &f = fs[s];
    dvec2 p = ldpos(pos, i);
    vec2<R> v = ld2<R>(vel, i), F = ld2<R>(b.force, i);
    fire_mix<R>(v, F, f);
    if (f.converged || !b.st[s].active) {
        st2<R>(vel, i, v);
        stpos(pos_next, i, p);
        if (flags & STEP_BIN) bin_cell(b, sp, i, p);
        return;
    }
    vec2<R> d = fire_vv_first_half<R>(v, F, (R)f.dt);
    st2<R>(vel, i, v);
    p.x += (double)d.x;
    p.y += (double)d.y;
    if (!(fabs(p.x) < 1e12 * sp.L && fabs(p.y) < 1e12 * sp.L)) {
        b.counters[5] = 1;
        b.st[s].error = CGPU_ERR_NONFINITE;
        p = ldpos(pos, i);
    }
    put_in_box(make_box(sp.L), p);
    stpos(pos_next, i, p);
    if (flags & STEP_BIN) bin_cell(b, sp, i, p);
}

__global__ void k_fire_clear_pending(cgpu_fire_state *fs, int S) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < S) fs[s].pending = 0;
}

// clear: also clears every system's pending mix (k_fire_clear_pending), which
// only the move kernel reads
template <typename R, bool TENS>
__global__ void k_fire_force_vv2(cgpu_voronoi_batch b, cgpu_fire_state *fs, void *vel, const double *pos,
                                 double *dots, int clear) {
    if (clear)
        for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < b.num_systems; q += gridDim.x * blockDim.x)
            fs[q].pending = 0;
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= b.num_cells) return;
    const int s = b.cell_sys[i];
    const cgpu_fire_state &f = fs[s];
    if (f.converged || !b.st[s].active) return;
    vec2<R> F = force_with_exclusions<R, TENS>(b, GlobalData<R>{pos, b.dAP}, i, b.sys[s]);
    st2<R>(b.force, i, F);
    vec2<R> v = ld2<R>(vel, i);
    const R dt = (R)f.dt;
    v.x += R(0.5) * dt * F.x;
    v.y += R(0.5) * dt * F.y;
    st2<R>(vel, i, v);
    fire_dots(dots, b.num_cells, i, F.x, F.y, v.x, v.y);
}

// per cell chunk: partial sums (fixed order, deterministic)
__global__ void k_fire_partial(cgpu_voronoi_batch b, const cgpu_fire_state *fs, double *dots) {
    const int c = blockIdx.x;
    const int s = b.chunk_sys[c];
    const int N = b.num_cells;
    fire_chunk_partial(dots, N, b.chunk_begin[c], b.chunk_end[c], fs[s].converged || !b.st[s].active,
                       dots + 3 * (size_t)N + 4 * (size_t)c);
}

// per system: the chunk partials summed in chunk order, then the FIRE update
__global__ void k_fire_reduce(cgpu_voronoi_batch b, cgpu_fire_state *fs, const double *dots) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= b.num_systems) return;
    cgpu_fire_state &f = fs[s];
    if (f.converged || !b.st[s].active) return;
    const cgpu_sys_params &sp = b.sys[s];
    const double *part = dots + 3 * (size_t)b.num_cells;
    // chunks of a system are consecutive; find the first by its cell range
    double forceNorm = 0, power = 0, velNorm = 0, mf = 0;
    for (int c = chunk_of(b, sp.off); c < b.num_chunks && b.chunk_sys[c] == s; ++c) {
        forceNorm += part[4 * c];
        power += part[4 * c + 1];
        velNorm += part[4 * c + 2];
        mf = fmax(mf, part[4 * c + 3]);
    }
    fire_update(f, forceNorm, power, velNorm, mf, sp.n);
}

__global__ void k_fire_park(cgpu_voronoi_batch b, const cgpu_fire_state *fs) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < b.num_systems && fs[s].converged) b.st[s].active = 0;
}

template <typename R>
__global__ void k_fire_apply_mix(cgpu_voronoi_batch b, const cgpu_fire_state *fs, void *vel) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= b.num_cells) return;
    const cgpu_fire_state &f = fs[b.cell_sys[i]];
    vec2<R> v = ld2<R>(vel, i), F = ld2<R>(b.force, i);
    fire_mix<R>(v, F, f);
    st2<R>(vel, i, v);
}

inline int nblocks(int n, int bs) { return (n + bs - 1) / bs; }
inline hipStream_t S(void *s) { return reinterpret_cast<hipStream_t>(s); }
inline int check() { return hipGetLastError() == hipSuccess ? CGPU_OK : CGPU_ERR_HIP; }

// global pass over the queued systems: star kernel, per-system degree
// statistics, delOther and geometry
void global_pass(const cgpu_voronoi_batch &b, const double *pos, hipStream_t st, int advance = 0,
                 const Probe *pr = nullptr) {
    const dim3 rb(shape_repair_blocks(b)), sb(shape_star_blocks(b)), t(256);
    hipLaunchKernelGGL(k_star_global_wave, sb, t, 0, st, b, pos);
    probe(pr, PROBE_STAR_GLOBAL);
    hipLaunchKernelGGL(k_global_stats, dim3(b.num_chunks), t, 0, st, b);
    probe(pr, PROBE_GLOBAL_STATS);
    if (b.precision)
        hipLaunchKernelGGL(k_finalize<float>, rb, t, 0, st, b, pos, advance);
    else
        hipLaunchKernelGGL(k_finalize<double>, rb, t, 0, st, b, pos, advance);
    probe(pr, PROBE_FINALIZE);
}

// advance: the chain ends a time step (k_finalize advances the step counters)
int repair_chain(const cgpu_voronoi_batch &b, const double *pos, hipStream_t st, int advance = 0,
                 const Probe *pr = nullptr) {
    const dim3 rb(shape_repair_blocks(b)), sb(shape_star_blocks(b)), t(256);
    hipLaunchKernelGGL(k_mark_nf, rb, t, 0, st, b);
    probe(pr, PROBE_MARK_NF);
    hipLaunchKernelGGL(k_star_local_wave, sb, t, 0, st, b, pos);
    probe(pr, PROBE_STAR_LOCAL);
    hipLaunchKernelGGL(k_delsets_local, rb, t, 0, st, b);
    probe(pr, PROBE_DELSETS_LOCAL);
    hipLaunchKernelGGL(k_global_fallback, dim3(b.num_chunks), t, 0, st, b);
    probe(pr, PROBE_GLOBAL_FALLBACK);
    global_pass(b, pos, st, advance, pr);
    return check();
}

}  // namespace

// ============================================================== C ABI (kernel layer)
namespace {
// the tension instantiations run only for batches with cell types
template <typename R> void launch_forces(const cgpu_voronoi_batch &b, hipStream_t st) {
    const int bs = shape_cell_block(b);
    dim3 g(nblocks(b.num_cells, bs)), t(bs);
    if (b.cell_type)
        hipLaunchKernelGGL((k_forces<R, true>), g, t, 0, st, b, b.pos);
    else
        hipLaunchKernelGGL((k_forces<R, false>), g, t, 0, st, b, b.pos);
}

template <typename R, int KIND>
void launch_force_eom_move(const cgpu_voronoi_batch &b, hipStream_t st, int flags) {
    // the kernel's LDS tile holds one block of cells: positions and dAP
    const int bs = shape_force_block(b);
    dim3 g(nblocks(b.num_cells, bs)), t(bs);
    const size_t lds = (size_t)bs * (sizeof(double2) + 2 * sizeof(R));
    const bool w5 = b.shape.force_waves == 5;
    if (b.cell_type) {
        if (w5)
            hipLaunchKernelGGL((k_force_eom_move_w5<R, KIND, true>), g, t, lds, st, b, b.pos, b.pos_next, flags);
        else
            hipLaunchKernelGGL((k_force_eom_move<R, KIND, true>), g, t, lds, st, b, b.pos, b.pos_next, flags);
    } else {
        if (w5)
            hipLaunchKernelGGL((k_force_eom_move_w5<R, KIND, false>), g, t, lds, st, b, b.pos, b.pos_next, flags);
        else
            hipLaunchKernelGGL((k_force_eom_move<R, KIND, false>), g, t, lds, st, b, b.pos, b.pos_next, flags);
    }

}

}  // namespace

extern "C" {

int cgpu_set_block_size(int bs) {
    if (bs < 64 || bs > 1024 || bs % 64) return CGPU_ERR_ARGS;
    g_block = bs;
    return CGPU_OK;
}

int cgpu_compute_cell_list(const cgpu_voronoi_batch *b, int use_next, void *stream) {
    if (!b) return CGPU_ERR_ARGS;
    const double *p = use_next ? b->pos_next : b->pos;
    // bin counts are zero here: allocation clears them and k_bin_scan clears
    // them after use
    const int bs = shape_cell_block(*b);
    hipLaunchKernelGGL(k_bin, dim3(nblocks(b->num_cells, bs)), dim3(bs), 0, S(stream), *b, p);
    launch_bin_scan(*b, S(stream));
    hipLaunchKernelGGL(k_bin_scatter, dim3(nblocks(b->num_cells, bs)), dim3(bs), 0, S(stream), *b, p);
    return check();
}

int cgpu_test_and_geometry(const cgpu_voronoi_batch *b, int test, void *stream) {
    if (!b) return CGPU_ERR_ARGS;
    launch_test_geometry(*b, b->pos, test, S(stream));
    return check();
}

int cgpu_repair_topology(const cgpu_voronoi_batch *b, void *stream) {
    if (!b) return CGPU_ERR_ARGS;
    return repair_chain(*b, b->pos, S(stream));
}

int cgpu_global_triangulation(const cgpu_voronoi_batch *b, void *stream) {
    if (!b) return CGPU_ERR_ARGS;
    hipStream_t st = S(stream);
    hipLaunchKernelGGL(k_reset, dim3(nblocks(b->num_systems, 256)), dim3(256), 0, st, *b);
    int r = cgpu_compute_cell_list(b, 0, stream);
    if (r) return r;
    hipLaunchKernelGGL(k_global_all, dim3(b->num_chunks), dim3(256), 0, st, *b);
    global_pass(*b, b->pos, st);
    // the pass recomputes geometry only where a ring changed; positions may have
    // changed everywhere here
    launch_test_geometry(*b, b->pos, 0, st);
    return check();
}

int cgpu_global_triangulation_active(const cgpu_voronoi_batch *b, void *stream) {
    if (!b) return CGPU_ERR_ARGS;
    hipStream_t st = S(stream);
    hipLaunchKernelGGL(k_reset, dim3(nblocks(b->num_systems, 256)), dim3(256), 0, st, *b);
    int r = cgpu_compute_cell_list(b, 0, stream);
    if (r) return r;
    hipLaunchKernelGGL(k_global_active, dim3(b->num_chunks), dim3(256), 0, st, *b);
    global_pass(*b, b->pos, st);
    launch_test_geometry(*b, b->pos, 0, st);
    return check();
}

int cgpu_derive_topology(const cgpu_voronoi_batch *b, void *stream) {
    if (!b) return CGPU_ERR_ARGS;
    hipLaunchKernelGGL(k_delsets_all, dim3(shape_repair_blocks(*b)), dim3(256), 0, S(stream), *b);
    launch_test_geometry(*b, b->pos, 0, S(stream));
    return check();
}

int cgpu_voronoi_forces(const cgpu_voronoi_batch *b, void *stream) {
    if (!b) return CGPU_ERR_ARGS;
    if (b->precision)
        launch_forces<float>(*b, S(stream));
    else
        launch_forces<double>(*b, S(stream));
    return check();
}

static int force_eom_move(const cgpu_voronoi_batch *b, int kind, void *stream, int flags = 0) {
    hipStream_t st = S(stream);
    if (b->precision) {
        if (kind == 0)
            launch_force_eom_move<float, 0>(*b, st, flags);
        else
            launch_force_eom_move<float, 1>(*b, st, flags);
    } else {
        if (kind == 0)
            launch_force_eom_move<double, 0>(*b, st, flags);
        else
            launch_force_eom_move<double, 1>(*b, st, flags);
    }
    return check();
}

int cgpu_spv_force_eom_move(const cgpu_voronoi_batch *b, void *stream) {
    if (!b) return CGPU_ERR_ARGS;
    return force_eom_move(b, 0, stream);
}
int cgpu_brownian_force_eom_move(const cgpu_voronoi_batch *b, void *stream) {
    if (!b) return CGPU_ERR_ARGS;
    return force_eom_move(b, 1, stream);
}

}  // extern "C"

// One step; after it, b->pos_next holds the new positions (the caller swaps).
// pr: optional per-launch event hook (profiled steps).
static int step_common(const cgpu_voronoi_batch *b, int kind, void *stream, const Probe *pr) {
    hipStream_t st = S(stream);
    // the force kernel also clears the step's topology state and runs the
    // binning pass of the cell list on the new positions (bin counts are zero:
    // the previous scan cleared them)
    int r = force_eom_move(b, kind, stream, STEP_BIN | STEP_RESET);
    if (r) return r;
    probe(pr, PROBE_FORCE_EOM_MOVE);
    launch_bin_scan(*b, st);
    probe(pr, PROBE_BIN_SCAN);
    const int bs = shape_cell_block(*b);
    hipLaunchKernelGGL(k_bin_scatter, dim3(nblocks(b->num_cells, bs)), dim3(bs), 0, st, *b, b->pos_next);
    probe(pr, PROBE_BIN_SCATTER);
    r = check();
    if (r) return r;
    launch_test_geometry(*b, b->pos_next, 1, st, pr);
    // its last kernel advances the step counters
    return repair_chain(*b, b->pos_next, st, 1, pr);
}

// One profiled step: ev[0] is recorded before the first launch, ev[k + 1]
// after the launches of slot k (CGPU_PROFILE_SLOTS slots).
int cgpu_profiled_spv_step(const cgpu_voronoi_batch *b, void *stream, hipEvent_t *ev) {
    Probe pr{ev, S(stream)};
    if (hipEventRecord(ev[0], S(stream)) != hipSuccess) return CGPU_ERR_HIP;
    return step_common(b, 0, stream, &pr);
}

extern "C" {

int cgpu_step_begin(const cgpu_voronoi_batch *b, void *stream) {
    if (!b) return CGPU_ERR_ARGS;
    hipLaunchKernelGGL(k_reset, dim3(nblocks(b->num_systems, 256)), dim3(256), 0, S(stream), *b);
    return check();
}
int cgpu_step_end(const cgpu_voronoi_batch *b, void *stream) {
    if (!b) return CGPU_ERR_ARGS;
    hipLaunchKernelGGL(k_advance_step, dim3(nblocks(b->num_systems, 256)), dim3(256), 0, S(stream), *b);
    return check();
}

int cgpu_spv_step(const cgpu_voronoi_batch *b, void *stream) {
    if (!b) return CGPU_ERR_ARGS;
    return step_common(b, 0, stream, nullptr);
}
int cgpu_brownian_step(const cgpu_voronoi_batch *b, void *stream) {
    if (!b) return CGPU_ERR_ARGS;
    return step_common(b, 1, stream, nullptr);
}

int cgpu_fill_normal(const cgpu_voronoi_batch *b, int step, int stream_id, double *out, void *stream) {
    if (!b || !out) return CGPU_ERR_ARGS;
    hipLaunchKernelGGL(k_fill_normal, dim3(nblocks(b->num_cells, 256)), dim3(256), 0, S(stream), *b, step, stream_id,
                       out);
    return check();
}

int cgpu_fire_vv_move(const cgpu_voronoi_batch *b, cgpu_fire_state *fs, void *vel, void *stream) {
    if (!b || !fs || !vel) return CGPU_ERR_ARGS;
    const int bs = shape_cell_block(*b);
    dim3 g(nblocks(b->num_cells, bs)), t(bs);
    if (b->precision)
        hipLaunchKernelGGL(k_fire_vv_move<float>, g, t, 0, S(stream), *b, fs, vel, b->pos, b->pos_next, 0);
    else
        hipLaunchKernelGGL(k_fire_vv_move<double>, g, t, 0, S(stream), *b, fs, vel, b->pos, b->pos_next, 0);
    hipLaunchKernelGGL(k_fire_clear_pending, dim3(nblocks(b->num_systems, 256)), dim3(256), 0, S(stream), fs,
                       b->num_systems);
    return check();
}

}  // extern "C"

namespace {
// forces, second velocity half step and the FIRE reductions on b->pos
void fire_force_reduce(const cgpu_voronoi_batch *b, cgpu_fire_state *fs, void *vel, double *dots, hipStream_t st,
                       int clear) {
    const int bs = shape_cell_block(*b);
    dim3 g(nblocks(b->num_cells, bs)), t(bs);
    const bool tens = b->cell_type != nullptr;
    if (b->precision) {
        if (tens)
            hipLaunchKernelGGL((k_fire_force_vv2<float, true>), g, t, 0, st, *b, fs, vel, b->pos, dots, clear);
        else
            hipLaunchKernelGGL((k_fire_force_vv2<float, false>), g, t, 0, st, *b, fs, vel, b->pos, dots, clear);
    } else {
        if (tens)
            hipLaunchKernelGGL((k_fire_force_vv2<double, true>), g, t, 0, st, *b, fs, vel, b->pos, dots, clear);
        else
            hipLaunchKernelGGL((k_fire_force_vv2<double, false>), g, t, 0, st, *b, fs, vel, b->pos, dots, clear);
    }
    hipLaunchKernelGGL(k_fire_partial, dim3(b->num_chunks), dim3(256), 0, st, *b, fs, dots);
    hipLaunchKernelGGL(k_fire_reduce, dim3((b->num_systems + 63) / 64), dim3(64), 0, st, *b, fs, dots);
}
}  // namespace

extern "C" {

int cgpu_fire_force_reduce(const cgpu_voronoi_batch *b, cgpu_fire_state *fs, void *vel, double *dots, void *stream) {
    if (!b || !fs || !vel || !dots) return CGPU_ERR_ARGS;
    fire_force_reduce(b, fs, vel, dots, S(stream), 0);
    return check();
}

int cgpu_fire_iteration(const cgpu_voronoi_batch *b, cgpu_fire_state *fs, void *vel, double *dots, void *stream) {
    if (!b || !fs || !vel || !dots) return CGPU_ERR_ARGS;
    hipStream_t st = S(stream);
    // the move also clears the topology state and bins the moved cells
    const int bs = shape_cell_block(*b);
    dim3 g(nblocks(b->num_cells, bs)), t(bs);
    if (b->precision)
        hipLaunchKernelGGL(k_fire_vv_move<float>, g, t, 0, st, *b, fs, vel, b->pos, b->pos_next, STEP_RESET | STEP_BIN);
    else
        hipLaunchKernelGGL(k_fire_vv_move<double>, g, t, 0, st, *b, fs, vel, b->pos, b->pos_next, STEP_RESET | STEP_BIN);
    launch_bin_scan(*b, st);
    hipLaunchKernelGGL(k_bin_scatter, dim3(nblocks(b->num_cells, bs)), dim3(bs), 0, st, *b, b->pos_next);
    cgpu_voronoi_batch bn = *b;  // the rest works on the moved positions
    bn.pos = b->pos_next;
    bn.pos_next = b->pos;
    launch_test_geometry(bn, bn.pos, 1, st);
    int r = repair_chain(bn, bn.pos, st);
    if (r) return r;
    // the force pass also clears the pending mix that the move applied
    fire_force_reduce(&bn, fs, vel, dots, st, 1);
    hipLaunchKernelGGL(k_fire_park, dim3(nblocks(b->num_systems, 256)), dim3(256), 0, st, bn, fs);
    return check();
}

}  // extern "C"

// ------------------------------------------------ allocation-free FIRE iteration
// Workspace of cgpu_fire_iteration_noalloc: per-cell FIRE products (3 N doubles)
// and per-chunk partial sums (4 per cell chunk), in a file-scope device array,
// one copy per GPU, held until the process ends. Its device address is looked
// up once per device with hipGetSymbolAddress (the host-side name is only a
// stand-in); calls on one device share it.
constexpr size_t SCRATCH_DOUBLES = size_t(1) << 20;  // 8 MiB
static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];
namespace {
constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[SCRATCH_MAX_DEVICES] = {};
// device address of this device's scratch (looked up on the first call)
int scratch_of_current_device(double **out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= SCRATCH_MAX_DEVICES) return CGPU_ERR_HIP;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return CGPU_ERR_HIP;
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    *out = g_scratch_ptr[dev];
    return CGPU_OK;
}
}  // namespace

extern "C" {

size_t cgpu_fire_noalloc_capacity(void) { return SCRATCH_DOUBLES; }

int cgpu_fire_iteration_noalloc(const cgpu_voronoi_batch *b, cgpu_fire_state *fs, void *vel, void *stream) {
    if (!b || !fs || !vel) return CGPU_ERR_ARGS;
    const size_t need = 3 * (size_t)b->num_cells + 4 * (size_t)b->num_chunks;
    if (need > SCRATCH_DOUBLES) {
        cgpu_host::set_last_error(
            ("cgpu_fire_iteration_noalloc: the workspace of 3 N + 4 chunks = " + std::to_string(need) +
             " doubles exceeds the static scratch of " + std::to_string(SCRATCH_DOUBLES) +
             " doubles (8 MiB); use cgpu_fire_iteration with a caller workspace")
                .c_str());
        return CGPU_ERR_SCRATCH_CAPACITY;
    }
    double *dots = nullptr;
    const int r = scratch_of_current_device(&dots);
    if (r) {
        cgpu_host::set_last_error("cgpu_fire_iteration_noalloc: hipGetSymbolAddress of the scratch array failed");
        return r;
    }
    return cgpu_fire_iteration(b, fs, vel, dots, stream);
}

int cgpu_fire_park_converged(const cgpu_voronoi_batch *b, const cgpu_fire_state *fs, void *stream) {
    if (!b || !fs) return CGPU_ERR_ARGS;
    hipLaunchKernelGGL(k_fire_park, dim3(nblocks(b->num_systems, 256)), dim3(256), 0, S(stream), *b, fs);
    return check();
}

int cgpu_fire_apply_mix(const cgpu_voronoi_batch *b, cgpu_fire_state *fs, void *vel, void *stream) {
    if (!b || !fs || !vel) return CGPU_ERR_ARGS;
    const int bs = shape_cell_block(*b);
    dim3 g(nblocks(b->num_cells, bs)), t(bs);
    if (b->precision)
        hipLaunchKernelGGL(k_fire_apply_mix<float>, g, t, 0, S(stream), *b, fs, vel);
    else
        hipLaunchKernelGGL(k_fire_apply_mix<double>, g, t, 0, S(stream), *b, fs, vel);
    hipLaunchKernelGGL(k_fire_clear_pending, dim3(nblocks(b->num_systems, 256)), dim3(256), 0, S(stream), fs,
                       b->num_systems);
    return check();
}

}  // extern "C"

