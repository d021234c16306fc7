// ORACLE — test infrastructure only. Nothing in the product links, loads or
// calls this file; tests/, the smoke check and bench.py's CPU-baseline leg use it
// as the checker.
//
// CPU FP64 restatement of the reference's Voronoi-model time step
// (abaso/cellGPU), written from the reference's semantics. Every function cites
// the reference file:line it follows. Compiled with -ffp-contract=off so that
// no multiply-add is fused (the reference's FMA contraction under nvcc is a
// compiler choice, not part of the algorithm).
//
// Parity pins (see tests/test_oracle_*.py):
//   * periodic Delaunay rings: scipy Qhull on a 9-sheet copy (golden fixtures),
//     brute-force empty-circle check, Euler invariants (sum z = 6N, 2N triangles);
//   * geometry: regular triangular lattice closed forms, sum of areas = L^2;
//   * forces: central finite differences of the energy, sum F = 0, GPU-style
//     force sets == CPU-style force formula (two independent reference paths);
//   * flagging: hand-built intruder cases; initial conditions: std::mt19937(13377);
//   * SPP / Brownian / FIRE: closed forms at F = 0, quadratic-bowl convergence.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <vector>

#include "fire_oracle.h"
#include "predicates.h"

using std::vector;

namespace orc {

struct d2 { double x, y; };

// ---------------------------------------------------------------- gpubox (square)
// reference inc/utility/gpubox.h:75-84 (setSquare), :139-152 (putInBox),
// :154-179 (minDist), :130-137 (putInBoxReal = invTrans, putInBox, Trans)
struct Box {
    double L, half, inv;
    explicit Box(double l = 1.0) { set(l); }
    void set(double l) { L = l; half = L * 0.5; inv = 1. / L; }
    // wrap loops capped at 64 iterations (at most two run for in-box inputs)
    inline d2 minDist(d2 p1, d2 p2) const {
        d2 a{p1.x - p2.x, p1.y - p2.y};
        for (int t = 0; t < 64 && a.x < -half; ++t) a.x += L;
        for (int t = 0; t < 64 && a.y < -half; ++t) a.y += L;
        for (int t = 0; t < 64 && a.x > half; ++t) a.x -= L;
        for (int t = 0; t < 64 && a.y > half; ++t) a.y -= L;
        return a;
    }
    // reference loops; iteration count bounded by |v| + 2 for finite v
    static inline double wrapUnit(double v) {
        long long guard = 0;
        while (v < 0 && guard++ < (1LL << 40)) v += 1.0;
        while (v >= 1.0 && guard++ < (1LL << 40)) v -= 1.0;
        return v;
    }
    inline void putInBoxReal(d2 &p) const {
        double vx = wrapUnit(inv * p.x), vy = wrapUnit(inv * p.y);
        p.x = L * vx;
        p.y = L * vy;
    }
};

// ------------------------------------------------------------ geometry primitives
// reference inc/utility/functions.h:34-43 (Circumcenter through origin)
static inline d2 circumcenter0(d2 x1, d2 x2) {
    double x1n = x1.x * x1.x + x1.y * x1.y;
    double x2n = x2.x * x2.x + x2.y * x2.y;
    double den = 0.5 / (x1.x * x2.y - x1.y * x2.x);
    return d2{den * (x1n * x2.y - x1.y * x2n), den * (x1.x * x2n - x1n * x2.x)};
}
// reference functions.h:46-62 (Circumcenter through three points)
static inline d2 circumcenter3(d2 x1, d2 x2, d2 x3) {
    double amcx = x1.x - x3.x, amcy = x1.y - x3.y;
    double amcn = amcx * amcx + amcy * amcy;
    double bmcx = x2.x - x3.x, bmcy = x2.y - x3.y;
    double bmcn = bmcx * bmcx + bmcy * bmcy;
    double den = 0.5 / (amcx * bmcy - amcy * bmcx);
    return d2{x3.x + den * (amcn * bmcy - amcy * bmcn), x3.y + den * (amcx * bmcn - amcn * bmcx)};
}
// reference functions.h:180-186 (Circumcircle: center then radius via sqrt)
static inline void circumcircle0(d2 x1, d2 x2, d2 &c, double &r) {
    c = circumcenter0(x1, x2);
    double dx = x1.x - c.x, dy = x1.y - c.y;
    r = std::sqrt(dx * dx + dy * dy);
}
// reference functions.h:224-227
static inline double triangleArea(d2 p1, d2 p2) { return std::fabs(0.5 * (p1.x * p2.y - p1.y * p2.x)); }
// reference functions.h:217-220
static inline double signedPolygonAreaPart(d2 p1, d2 p2) { return 0.5 * (p1.x + p2.x) * (p2.y - p1.y); }

struct M22 { double x11, x12, x21, x22; };
// reference functions.h:234-242
static inline M22 getdhdr(d2 rij, d2 rik) {
    double c = rij.y * rik.x - rij.x * rik.y;
    double den = 1.0 / (c * c);
    double rik2 = rik.x * rik.x + rik.y * rik.y;
    double rij2 = rij.x * rij.x + rij.y * rij.y;
    M22 m;
    m.x11 = 0.5 * den * (rik.y - rij.y) * (rik.y * rij2 - rij.y * rik2);
    m.x12 = 0.5 * den * (rij.y - rik.y) * (rik.x * rij2 - rij.x * rik2);
    m.x21 = 0.5 * den * (rij.x - rik.x) * (rik.y * rij2 - rij.y * rik2);
    m.x22 = 0.5 * den * (rik.x - rij.x) * (rik.x * rij2 - rij.x * rik2);
    return m;
}
// reference inc/utility/Matrix.h:128-134 (row vector times matrix)
static inline d2 vecmat(d2 v, M22 m) { return d2{v.x * m.x11 + v.y * m.x21, v.x * m.x12 + v.y * m.x22}; }

// --------------------------------------------------------------------- cell list
// reference src/utility/cellListGPU.cpp:97-117 (setGridSize), cellListGPU.cu:20-58
struct CellList {
    int xsize = 0;
    double boxsize = 1.0;
    vector<vector<int>> bins;
    void setGrid(double L, double a) {
        xsize = (int)std::floor(L / a);
        if (xsize % 2 == 1) xsize += 1;
        boxsize = L / xsize;
        bins.assign((size_t)xsize * xsize, {});
    }
    inline int binOf(double v) const {
        int b = (int)std::floor(v / boxsize);
        return std::max(0, std::min(xsize - 1, b));
    }
    void compute(const vector<d2> &p) {
        for (auto &b : bins) b.clear();
        for (int i = 0; i < (int)p.size(); ++i) bins[(size_t)binOf(p[i].y) * xsize + binOf(p[i].x)].push_back(i);
    }
};

// ---------------------------------------------------------- Delaunay star (repair)
// The reference obtains the counter-clockwise 1-ring of a cell from a local
// exact-predicate CGAL triangulation of a candidate set (DelaunayLoc.cpp:381-403,
// DelaunayCGAL.cpp:8-48) or from CGAL's periodic triangulation
// (DelaunayCGAL.cpp:164-229). For points in general position both equal the
// unique Delaunay star. Restated here as gift wrapping around the cell with exact
// orient/incircle predicates on minimum-image displacements, each triangle
// certified empty by checking that its circumcircle lies inside the scanned bins.
// Ring rotation is canonical: the smallest neighbour index first (the reference's
// rotation is CGAL-internal and carries no meaning).
// Returns the degree, or -1 on failure.
int delaunayStar(int i, const vector<d2> &P, const Box &box, const CellList &cl, vector<int> &ring) {
    const int xs = cl.xsize;
    const d2 pi = P[i];
    int ib = cl.binOf(pi.x), jb = cl.binOf(pi.y);
    int wmax = std::max(1, xs / 2);
    vector<int> cid;
    vector<d2> cd;
    for (int w = 1; w <= wmax; ++w) {
        cid.clear();
        cd.clear();
        int nbk = std::min(2 * w + 1, xs);
        for (int jj = jb - w; jj < jb - w + nbk; ++jj)
            for (int ii = ib - w; ii < ib - w + nbk; ++ii) {
                int cx = ((ii % xs) + xs) % xs, cy = ((jj % xs) + xs) % xs;
                for (int m : cl.bins[(size_t)cy * xs + cx]) {
                    if (m == i) continue;
                    cid.push_back(m);
                    cd.push_back(box.minDist(P[m], pi));
                }
            }
        // scanned block relative to pi, shrunk by 1e-9 L for rounding at bin edges
        const double mg = 1e-9 * box.L;
        const double xlo = (ib - w) * cl.boxsize - pi.x + mg, xhi = (ib + w + 1) * cl.boxsize - pi.x - mg;
        const double ylo = (jb - w) * cl.boxsize - pi.y + mg, yhi = (jb + w + 1) * cl.boxsize - pi.y - mg;
        bool fullbox = (2 * w + 1 >= xs);
        int nc = (int)cid.size();
        if (nc < 3) continue;
        // nearest candidate (exact tie break on squared norm, then index)
        int first = 0;
        for (int m = 1; m < nc; ++m) {
            int c = cmp_norm2(cd[m].x, cd[m].y, cd[first].x, cd[first].y);
            if (c < 0 || (c == 0 && cid[m] < cid[first])) first = m;
        }
        ring.clear();
        int cur = first;
        bool ok = true;
        for (int guard = 0; guard < 64; ++guard) {
            ring.push_back(cid[cur]);
            int best = -1;
            for (int m = 0; m < nc; ++m) {
                if (m == cur) continue;
                if (orient0(cd[cur].x, cd[cur].y, cd[m].x, cd[m].y) <= 0) continue;
                if (best < 0) { best = m; continue; }
                int ic = incircle0(cd[cur].x, cd[cur].y, cd[best].x, cd[best].y, cd[m].x, cd[m].y);
                if (ic > 0 || (ic == 0 && cid[m] < cid[best])) best = m;
            }
            if (best < 0) { ok = false; break; }
            if (!fullbox) {
                d2 c = circumcenter0(cd[cur], cd[best]);
                double rr = std::sqrt(c.x * c.x + c.y * c.y);
                // certified: the closed circumdisk lies inside the scanned block
                if (!(c.x - rr >= xlo && c.x + rr <= xhi && c.y - rr >= ylo && c.y + rr <= yhi)) {
                    ok = false;
                    break;
                }
            }
            cur = best;
            if (cur == first) break;
        }
        if (!ok || cur != first) continue;
        // canonical rotation: smallest index first
        int k0 = (int)(std::min_element(ring.begin(), ring.end()) - ring.begin());
        std::rotate(ring.begin(), ring.begin() + k0, ring.end());
        return (int)ring.size();
    }
    return -1;
}

// ---------------------------------------------------------------- Voronoi system
struct VSys {
    int N = 0;
    Box box;
    int cap = 32;                  // ring capacity per cell (storage only)
    vector<d2> P;                  // cellPositions
    vector<int> nn;                // cellNeighborNum
    vector<int> nb;                // cellNeighbors, row i*cap + k
    vector<int> dsx, dsy, dother;  // delSets (.x,.y), delOther per slot
    int neighMax = 0;
    vector<d2> AP, APpref, F, mot, vel;  // AreaPeri, prefs, forces, Motility(v0,Dr), cellVelocities
    vector<double> theta;          // cellDirectors
    vector<int> excl;              // exclusions
    vector<d2> extF;
    double KA = 1.0, KP = 1.0, dt = 0.01, mu = 1.0;
    CellList cl;                   // GPU-path cell list (cellsize 1.25)
    // counters (reference voronoiModelBase.h:134-145)
    long long globalFixes = 0, localRepairs = 0, repairEvents = 0, skippedFrames = 0;
    int sppTimestep = 0;           // selfPropelledParticleDynamics::Timestep
    vector<int> lastNeedsFixing;
    int lastRepairMode = 0;        // 0 none, 1 local, 2 global
    // line tension between cell types (VoronoiQuadraticEnergyWithTension):
    // mode 0 off, 1 one value gamma for every unlike pair, 2 matrix g[a*n + b]
    int tensionMode = 0, ntypes = 0;
    double gamma = 0.0;
    vector<double> tmat;
    vector<int> type;              // cellType
    // reference cellTypeIndexer(b, a) = a*n + b (Index2D(n)(i, j) = j*n + i)
    inline double tension(int a, int b) const { return tmat[(size_t)a * ntypes + b]; }

    inline int &ring(int i, int k) { return nb[(size_t)i * cap + k]; }
    inline int ringc(int i, int k) const { return nb[(size_t)i * cap + k]; }
};

// reference voronoiModelBase.cpp:281-351 (globalTriangulationCGAL) incl. neighMax rule :309-312
static bool globalTriangulation(VSys &s) {
    s.globalFixes += 1;
    CellList cl;
    cl.setGrid(s.box.L, 1.25);
    cl.compute(s.P);
    int nmax = 0, tot = 0;
    vector<int> ring;
    vector<vector<int>> all(s.N);
    for (int i = 0; i < s.N; ++i) {
        int z = delaunayStar(i, s.P, s.box, cl, ring);
        if (z < 0) return false;
        all[i] = ring;
        nmax = std::max(nmax, z);
        tot += z;
    }
    s.neighMax = (nmax % 2 == 0) ? nmax + 2 : nmax + 1;
    if (s.neighMax > s.cap) {
        s.cap = s.neighMax + 8;
    }
    s.nn.assign(s.N, 0);
    s.nb.assign((size_t)s.N * s.cap, 0);
    s.dsx.assign((size_t)s.N * s.cap, 0);
    s.dsy.assign((size_t)s.N * s.cap, 0);
    s.dother.assign((size_t)s.N * s.cap, 0);
    for (int i = 0; i < s.N; ++i) {
        s.nn[i] = (int)all[i].size();
        for (int k = 0; k < s.nn[i]; ++k) s.ring(i, k) = all[i][k];
    }
    return tot == 6 * s.N;
}

// reference voronoiModelBase.cpp:1236-1275 (getDelSets). A neighbour ring that
// does not contain nm1 leaves that slot's delOther at its previous value, as in
// the reference.
static bool getDelSets(VSys &s, int i) {
    int z = s.nn[i];
    int nm1 = s.ring(i, z - 2), n1 = s.ring(i, z - 1), n2;
    bool good = true;
    for (int k = 0; k < z; ++k) {
        n2 = s.ring(i, k);
        int zn = s.nn[n1];
        for (int q = 0; q < zn; ++q) {
            if (s.ring(n1, q) == nm1) {
                s.dother[(size_t)i * s.cap + k] = s.ring(n1, (q + 1) % zn);
                break;
            }
        }
        s.dsx[(size_t)i * s.cap + k] = nm1;
        s.dsy[(size_t)i * s.cap + k] = n1;
        int o = s.dother[(size_t)i * s.cap + k];
        if (nm1 == o || n1 == o || i == o) { good = false; return good; }
        nm1 = n1;
        n1 = n2;
    }
    return good;
}
static void allDelSets(VSys &s) {
    for (int i = 0; i < s.N; ++i) getDelSets(s, i);
}

// reference voronoiModelBase.cpp:380-423 (neighbour / triangle-count invariants)
static bool countsConsistent(const VSys &s) {
    long long tot = 0, tri = 0;
    for (int i = 0; i < s.N; ++i) {
        int z = s.nn[i];
        tot += z;
        for (int k = 0; k < z; ++k) {
            int a = s.ringc(i, k), b = s.ringc(i, (k + 1) % z);
            if (i < a && i < b) tri++;
        }
    }
    return tot == 6LL * s.N && tri == 2LL * s.N;
}

// reference voronoiModelBase.cu:24-109 (gpu_test_circumcenters_kernel), N3 window
static bool testTriangulationGPUStyle(VSys &s, vector<int> &repair) {
    s.cl.setGrid(s.box.L, 1.25);
    s.cl.compute(s.P);
    const CellList &cl = s.cl;
    const int xs = cl.xsize;
    repair.assign(s.N, 0);
    bool anyFail = false;
    for (int i = 0; i < s.N; ++i) {
        int z = s.nn[i];
        for (int k = 0; k < z; ++k) {
            int j = s.ring(i, k), l = s.ring(i, (k + 1) % z);
            if (!(i < j && i < l)) continue;
            d2 v = s.P[i];
            int ib = (int)std::floor(v.x / cl.boxsize), jb = (int)std::floor(v.y / cl.boxsize);
            d2 pt1 = s.box.minDist(s.P[j], v), pt2 = s.box.minDist(s.P[l], v);
            d2 Q;
            double rad;
            circumcircle0(pt1, pt2, Q, rad);
            bool bad = false;
            double wc = std::ceil(rad / cl.boxsize);
            int wcheck = (wc < (double)(xs / 2)) ? (int)wc : xs / 2;
            rad = rad * rad;
            // a circle wider than the box: every wrapped distance is inside
            bool all_inside = rad > 0.5 * s.box.L * s.box.L * (1.0 + 1e-9);
            for (int ii = ib - wcheck; ii <= ib + wcheck; ++ii)
                for (int jj = jb - wcheck; jj <= jb + wcheck; ++jj) {
                    int cx = ii;
                    if (cx < 0) cx += xs;
                    if (cx >= xs) cx -= xs;
                    int cy = jj;
                    if (cy < 0) cy += xs;
                    if (cy >= xs) cy -= xs;
                    for (int m : cl.bins[(size_t)cy * xs + cx]) {
                        bool inside = all_inside;
                        if (!inside) {
                            d2 a = s.box.minDist(s.P[m], v);
                            d2 b = s.box.minDist(a, Q);
                            inside = b.x * b.x + b.y * b.y < rad;
                        }
                        if (inside) {
                            if (m != i && m != j && m != l) {
                                bad = true;
                                repair[m] = 1;
                            }
                        }
                    }
                }
            if (bad) {
                anyFail = true;
                repair[i] = repair[j] = repair[l] = 1;
            }
        }
    }
    return anyFail;
}

// reference DelaunayLoc.cpp:517-574 (testPointTriangulation), CPU path; the
// candidate window covers the whole circumcircle.
static bool testTriangulationCPUStyle(VSys &s, vector<int> &repair) {
    s.cl.setGrid(s.box.L, 1.25);
    s.cl.compute(s.P);
    const CellList &cl = s.cl;
    const int xs = cl.xsize;
    repair.assign(s.N, 0);
    bool any = false;
    for (int i = 0; i < s.N; ++i) {
        int z = s.nn[i];
        d2 v = s.P[i];
        bool bad = false;
        int n1 = s.ring(i, z - 1);
        for (int k = 0; k < z && !bad; ++k) {
            int n2 = s.ring(i, k);
            d2 pt1 = s.box.minDist(s.P[n1], v), pt2 = s.box.minDist(s.P[n2], v);
            d2 Q;
            double radius;
            circumcircle0(pt1, pt2, Q, radius);
            double rad2 = radius * radius;
            double qx = v.x + Q.x, qy = v.y + Q.y;
            int cxq = (int)std::floor(qx / cl.boxsize), cyq = (int)std::floor(qy / cl.boxsize);
            int w = (int)std::ceil(radius / cl.boxsize) + 1;
            w = std::min(w, xs / 2);
            for (int jj = -w; jj <= w && !bad; ++jj)
                for (int ii = -w; ii <= w && !bad; ++ii) {
                    int cx = (((cxq + ii) % xs) + xs) % xs, cy = (((cyq + jj) % xs) + xs) % xs;
                    for (int m : cl.bins[(size_t)cy * xs + cx]) {
                        d2 disp = s.box.minDist(s.P[m], v);
                        d2 tc = s.box.minDist(disp, Q);
                        if (tc.x * tc.x + tc.y * tc.y < rad2) {
                            bool inRing = (m == i);
                            for (int q = 0; q < z; ++q)
                                if (s.ring(i, q) == m) inRing = true;
                            if (!inRing) { bad = true; break; }
                        }
                    }
                }
            n1 = n2;
        }
        if (bad) {
            repair[i] = 1;
            any = true;
        }
    }
    return any;
}

// reference voronoiModelBase.cpp:600-663 (testAndRepairTriangulation),
// :429-505 (repairTriangulation), :686-739 (enforceTopology, GPU branch).
// gpuPath=false restates the CPU branch (setCPU(false) local mode).
static bool enforceTopology(VSys &s, bool gpuPath) {
    vector<int> repair;
    bool anyFail = gpuPath ? testTriangulationGPUStyle(s, repair) : testTriangulationCPUStyle(s, repair);
    s.lastRepairMode = 0;
    s.lastNeedsFixing.clear();
    if (!anyFail) { s.skippedFrames++; return true; }
    s.repairEvents++;
    vector<int> nf;
    for (int i = 0; i < s.N; ++i)
        if (repair[i] == 1) {
            nf.push_back(i);
            for (int k = 0; k < s.nn[i]; ++k) nf.push_back(s.ring(i, k));
        }
    std::sort(nf.begin(), nf.end());
    nf.erase(std::unique(nf.begin(), nf.end()), nf.end());
    s.lastNeedsFixing = nf;
    bool complete = false;
    if ((int)nf.size() > s.N / 6) {
        complete = true;
        if (!globalTriangulation(s)) return false;
    } else {
        // local repair
        vector<vector<int>> rings(nf.size());
        bool needGlobal = false;
        for (size_t q = 0; q < nf.size(); ++q) {
            int z = delaunayStar(nf[q], s.P, s.box, s.cl, rings[q]);
            if (z < 0) { needGlobal = true; break; }
            if (z > s.neighMax) needGlobal = true;
        }
        if (needGlobal) {
            complete = true;
            if (!globalTriangulation(s)) return false;
        } else {
            for (size_t q = 0; q < nf.size(); ++q) {
                int i = nf[q];
                s.nn[i] = (int)rings[q].size();
                for (int k = 0; k < s.nn[i]; ++k) s.ring(i, k) = rings[q][k];
            }
            s.localRepairs += (long long)nf.size();
            if (!countsConsistent(s)) {
                complete = true;
                if (!globalTriangulation(s)) return false;
            }
        }
    }
    if (gpuPath) {
        if (complete) {
            allDelSets(s);
        } else {
            bool localFail = false;
            for (int i : nf)
                if (!getDelSets(s, i)) localFail = true;
            if (localFail) {
                complete = true;
                if (!globalTriangulation(s)) return false;
                allDelSets(s);
            }
        }
    }
    s.lastRepairMode = complete ? 2 : 1;
    return true;
}

// reference voronoiModelBase.cu:115-201 (gpu_compute_voronoi_geometry_kernel)
static void geometryGPUStyle(VSys &s) {
    for (int i = 0; i < s.N; ++i) {
        int z = s.nn[i];
        d2 pi = s.P[i];
        d2 rij = s.box.minDist(s.P[s.ring(i, z - 1)], pi);
        d2 rik = s.box.minDist(s.P[s.ring(i, 0)], pi);
        d2 vfirst = circumcenter0(rij, rik), vlast = vfirst, vnext;
        double A = 0.0, Pm = 0.0;
        for (int k = 1; k < z; ++k) {
            rij = rik;
            rik = s.box.minDist(s.P[s.ring(i, k)], pi);
            vnext = circumcenter0(rij, rik);
            A += triangleArea(vlast, vnext);
            double dx = vlast.x - vnext.x, dy = vlast.y - vnext.y;
            Pm += std::sqrt(dx * dx + dy * dy);
            vlast = vnext;
        }
        A += triangleArea(vlast, vfirst);
        double dx = vlast.x - vfirst.x, dy = vlast.y - vfirst.y;
        Pm += std::sqrt(dx * dx + dy * dy);
        s.AP[i] = d2{A, Pm};
    }
}

// Voronoi vertex k of cell i in the GPU convention: circumcenter of ring k-2, k-1
// (reference voronoiModelBase.cu:146-165)
static inline d2 voroCurGPU(const VSys &s, int i, int k) {
    int z = s.nn[i];
    d2 pi = s.P[i];
    int a = s.ringc(i, ((k - 2) % z + z) % z), b = s.ringc(i, ((k - 1) % z + z) % z);
    return circumcenter0(s.box.minDist(s.P[a], pi), s.box.minDist(s.P[b], pi));
}

// reference voronoiQuadraticEnergy.cu:99-234 (gpu_force_sets_kernel) then
// :30-93 (sum with / without exclusions)
// reference voronoiQuadraticEnergyWithTension.cu:27-199 (matrix) and the
// one-value variant after it: the line-tension terms of one force set. Lengths
// are clamped at THRESHOLD = 1e-18 in both precisions there.
struct TensionTerms {
    bool on, simple;
    double g;
    int tI, tJ, tK;
    const VSys *s;
    double gam(int a, int b) const { return simple ? 1.0 : s->tension(a, b); }
};

static void forcesGPUStyle(VSys &s, bool exclusions) {
    const bool tens = s.tensionMode != 0;
    for (int i = 0; i < s.N; ++i) {
        int z = s.nn[i];
        d2 sum{0.0, 0.0};
        for (int k = 0; k < z; ++k) {
            size_t sl = (size_t)i * s.cap + k;
            d2 pi = s.P[i];
            int nx = s.dsx[sl], ny = s.dsy[sl];
            d2 dnext = s.box.minDist(s.P[nx], pi);
            d2 dcl = s.box.minDist(s.P[ny], pi);
            d2 dnc = s.box.minDist(s.P[s.dother[sl]], pi);
            M22 dhdr = getdhdr(dnext, dcl);
            d2 vcur = voroCurGPU(s, i, k);
            d2 vlast = voroCurGPU(s, i, k - 1);
            d2 vnext = voroCurGPU(s, i, k + 1);
            d2 vother = circumcenter3(dnext, dcl, dnc);
            auto clamp = [&](double x) { return (tens && x < 1e-18) ? 1e-18 : x; };
            TensionTerms T{tens, s.tensionMode == 1, s.gamma, 0, 0, 0, &s};
            if (tens) {
                T.tI = s.type[i];
                T.tJ = s.type[nx];
                T.tK = s.type[ny];
            }
            const bool Tik = tens && T.tI != T.tK, Tij = tens && T.tI != T.tJ, Tjk = tens && T.tJ != T.tK;
            // adds 0.5 (gamma dT) to e: the reference's ordering for either variant
            auto add_tension = [&](double &e, double a, double dT) {
                e = T.simple ? a + 0.5 * T.g * dT : a + 0.5 * dT;
            };
            d2 dAdv{0.5 * (vlast.y - vnext.y), 0.5 * (vnext.x - vlast.x)};
            d2 dlast{vlast.x - vcur.x, vlast.y - vcur.y};
            double dlnorm = std::sqrt(dlast.x * dlast.x + dlast.y * dlast.y);
            d2 dn{vcur.x - vnext.x, vcur.y - vnext.y};
            double dnnorm = std::sqrt(dn.x * dn.x + dn.y * dn.y);
            dnnorm = clamp(dnnorm);
            dlnorm = clamp(dlnorm);
            d2 dclv = dlast, dncv = dn;
            double dclnorm = dlnorm, dncnorm = dnnorm;
            d2 dPdv{dlast.x / dlnorm - dn.x / dnnorm, dlast.y / dlnorm - dn.y / dnnorm};
            double Adiff = s.KA * (s.AP[i].x - s.APpref[i].x);
            double Pdiff = s.KP * (s.AP[i].y - s.APpref[i].y);
            d2 dEdv{Adiff * dAdv.x + Pdiff * dPdv.x, Adiff * dAdv.y + Pdiff * dPdv.y};
            if (tens) {
                d2 dT{0.0, 0.0};
                if (Tik) {
                    dT.x -= T.gam(T.tI, T.tK) * dn.x / dnnorm;
                    dT.y -= T.gam(T.tI, T.tK) * dn.y / dnnorm;
                }
                if (Tij) {
                    dT.x += T.gam(T.tI, T.tJ) * dlast.x / dlnorm;
                    dT.y += T.gam(T.tI, T.tJ) * dlast.y / dlnorm;
                }
                add_tension(dEdv.x, Adiff * dAdv.x + Pdiff * dPdv.x, dT.x);
                add_tension(dEdv.y, Adiff * dAdv.y + Pdiff * dPdv.y, dT.y);
            }
            // cell ny ("k")
            dAdv = d2{0.5 * (vnext.y - vother.y), 0.5 * (vother.x - vnext.x)};
            dn = d2{vcur.x - vother.x, vcur.y - vother.y};
            dnnorm = std::sqrt(dn.x * dn.x + dn.y * dn.y);
            dnnorm = clamp(dnnorm);
            dPdv = d2{-dncv.x / dncnorm - dn.x / dnnorm, -dncv.y / dncnorm - dn.y / dnnorm};
            Adiff = s.KA * (s.AP[ny].x - s.APpref[ny].x);
            Pdiff = s.KP * (s.AP[ny].y - s.APpref[ny].y);
            if (tens) {
                d2 dT{0.0, 0.0};
                if (Tik) {  // dnc = -(self dnext)
                    dT.x += T.gam(T.tI, T.tK) * -dncv.x / dncnorm;
                    dT.y += T.gam(T.tI, T.tK) * -dncv.y / dncnorm;
                }
                if (Tjk) {
                    dT.x -= T.gam(T.tJ, T.tK) * dn.x / dnnorm;
                    dT.y -= T.gam(T.tJ, T.tK) * dn.y / dnnorm;
                }
                double ex, ey;
                add_tension(ex, Adiff * dAdv.x + Pdiff * dPdv.x, dT.x);
                add_tension(ey, Adiff * dAdv.y + Pdiff * dPdv.y, dT.y);
                dEdv.x += ex;
                dEdv.y += ey;
            } else {
                dEdv.x += Adiff * dAdv.x + Pdiff * dPdv.x;
                dEdv.y += Adiff * dAdv.y + Pdiff * dPdv.y;
            }
            // cell nx ("j")
            dAdv = d2{0.5 * (vother.y - vlast.y), 0.5 * (vlast.x - vother.x)};
            dPdv = d2{-dn.x / dnnorm + dclv.x / dclnorm, -dn.y / dnnorm + dclv.y / dclnorm};
            Adiff = s.KA * (s.AP[nx].x - s.APpref[nx].x);
            Pdiff = s.KP * (s.AP[nx].y - s.APpref[nx].y);
            if (tens) {
                d2 dT{0.0, 0.0};
                if (Tij) {  // dcl = -(self dlast)
                    dT.x -= T.gam(T.tI, T.tJ) * -dclv.x / dclnorm;
                    dT.y -= T.gam(T.tI, T.tJ) * -dclv.y / dclnorm;
                }
                if (Tjk) {  // this cell's dlast = -dn
                    dT.x += T.gam(T.tJ, T.tK) * -dn.x / dnnorm;
                    dT.y += T.gam(T.tJ, T.tK) * -dn.y / dnnorm;
                }
                double ex, ey;
                add_tension(ex, Adiff * dAdv.x + Pdiff * dPdv.x, dT.x);
                add_tension(ey, Adiff * dAdv.y + Pdiff * dPdv.y, dT.y);
                dEdv.x += ex;
                dEdv.y += ey;
            } else {
                dEdv.x += Adiff * dAdv.x + Pdiff * dPdv.x;
                dEdv.y += Adiff * dAdv.y + Pdiff * dPdv.y;
            }
            dEdv.x *= 2.0;
            dEdv.y *= 2.0;
            d2 fs = vecmat(dEdv, dhdr);
            sum.x += fs.x;
            sum.y += fs.y;
        }
        if (exclusions && s.excl[i] != 0) {
            s.F[i] = d2{0.0, 0.0};
            s.extF[i] = d2{-sum.x, -sum.y};
        } else {
            s.F[i] = sum;
            if (exclusions) s.extF[i] = d2{0.0, 0.0};
        }
    }
}

// reference voronoiModelBase.cpp:783-846 (computeGeometryCPU) and
// voronoiQuadraticEnergy.cpp:172-378 (computeVoronoiForceCPU)
static void geometryAndForcesCPUStyle(VSys &s) {
    vector<vector<d2>> voro(s.N);
    for (int i = 0; i < s.N; ++i) {
        int z = s.nn[i];
        d2 pi = s.P[i];
        voro[i].resize(z);
        d2 rij = s.box.minDist(s.P[s.ring(i, z - 1)], pi);
        for (int k = 0; k < z; ++k) {
            d2 rik = s.box.minDist(s.P[s.ring(i, k)], pi);
            voro[i][k] = circumcenter0(rij, rik);
            rij = rik;
        }
        double A = 0, Pm = 0;
        d2 vlast = voro[i][z - 1];
        for (int k = 0; k < z; ++k) {
            d2 vnext = voro[i][k];
            A += triangleArea(vlast, vnext);
            double dx = vlast.x - vnext.x, dy = vlast.y - vnext.y;
            Pm += std::sqrt(dx * dx + dy * dy);
            vlast = vnext;
        }
        s.AP[i] = d2{A, Pm};
    }
    const double Pth = 1e-18;
    for (int i = 0; i < s.N; ++i) {
        int z = s.nn[i];
        d2 pi = s.P[i];
        vector<M22> dh(z);
        d2 rij = s.box.minDist(s.P[s.ring(i, z - 1)], pi);
        for (int k = 0; k < z; ++k) {
            d2 rik = s.box.minDist(s.P[s.ring(i, k)], pi);
            d2 rjk{rik.x - rij.x, rik.y - rij.y};
            auto dot = [](d2 a, d2 b) { return a.x * b.x + a.y * b.y; };
            double betaD = -dot(rik, rik) * dot(rij, rjk);
            double gammaD = dot(rij, rij) * dot(rik, rjk);
            double cp = rij.x * rjk.y - rij.y * rjk.x;
            double D = 2 * cp * cp;
            d2 zz{betaD * rij.x + gammaD * rik.x, betaD * rij.y + gammaD * rik.y};
            d2 dbD{2 * dot(rij, rjk) * rik.x + dot(rik, rik) * rjk.x, 2 * dot(rij, rjk) * rik.y + dot(rik, rik) * rjk.y};
            d2 dgD{-2 * dot(rik, rjk) * rij.x - dot(rij, rij) * rjk.x, -2 * dot(rik, rjk) * rij.y - dot(rij, rij) * rjk.y};
            d2 dDOD{(-2.0 * rjk.y) / cp, (2.0 * rjk.x) / cp};
            // Id + 1/D * (dyad(rij,dbD) + dyad(rik,dgD) - (betaD+gammaD) Id - dyad(zz,dDOD))
            M22 t;
            t.x11 = rij.x * dbD.x + rik.x * dgD.x;
            t.x12 = rij.x * dbD.y + rik.x * dgD.y;
            t.x21 = rij.y * dbD.x + rik.y * dgD.x;
            t.x22 = rij.y * dbD.y + rik.y * dgD.y;
            t.x11 = t.x11 - (betaD + gammaD);
            t.x22 = t.x22 - (betaD + gammaD);
            t.x11 = t.x11 - zz.x * dDOD.x;
            t.x12 = t.x12 - zz.x * dDOD.y;
            t.x21 = t.x21 - zz.y * dDOD.x;
            t.x22 = t.x22 - zz.y * dDOD.y;
            double invD = 1.0 / D;
            dh[k] = M22{1.0 + invD * t.x11, invD * t.x12, invD * t.x21, 1.0 + invD * t.x22};
            rij = rik;
        }
        d2 fsum{0, 0};
        double Adiff = s.KA * (s.AP[i].x - s.APpref[i].x);
        double Pdiff = s.KP * (s.AP[i].y - s.APpref[i].y);
        d2 vlast = voro[i][z - 1];
        for (int k = 0; k < z; ++k) {
            d2 vcur = voro[i][k], vnext = voro[i][(k + 1) % z];
            int base = s.ring(i, k);
            int other = s.ring(i, (k - 1 + z) % z);
            d2 dAidv{0.5 * (vlast.y - vnext.y), 0.5 * (vnext.x - vlast.x)};
            d2 dl{vlast.x - vcur.x, vlast.y - vcur.y};
            double dln = std::sqrt(dl.x * dl.x + dl.y * dl.y);
            d2 dn{vcur.x - vnext.x, vcur.y - vnext.y};
            double dnn = std::sqrt(dn.x * dn.x + dn.y * dn.y);
            if (dnn < Pth) dnn = Pth;
            if (dln < Pth) dln = Pth;
            d2 dPidv{dl.x / dln - dn.x / dnn, dl.y / dln - dn.y / dnn};
            int z2 = s.nn[base], dto = -1;
            for (int q = 0; q < z2; ++q)
                if (s.ring(base, q) == other) dto = s.ring(base, (q + 1) % z2);
            if (dto == other || dto == base || dto == -1) dto = other;  // reference throws here
            d2 r1 = s.box.minDist(s.P[other], pi), r2 = s.box.minDist(s.P[base], pi), r3 = s.box.minDist(s.P[dto], pi);
            d2 vother = circumcenter3(r1, r2, r3);
            double Akd = s.KA * (s.AP[base].x - s.APpref[base].x), Pkd = s.KP * (s.AP[base].y - s.APpref[base].y);
            double Ajd = s.KA * (s.AP[other].x - s.APpref[other].x), Pjd = s.KP * (s.AP[other].y - s.APpref[other].y);
            d2 dAkdv{0.5 * (vnext.y - vother.y), 0.5 * (vother.x - vnext.x)};
            dl = d2{vnext.x - vcur.x, vnext.y - vcur.y};
            dln = std::sqrt(dl.x * dl.x + dl.y * dl.y);
            dn = d2{vcur.x - vother.x, vcur.y - vother.y};
            dnn = std::sqrt(dn.x * dn.x + dn.y * dn.y);
            if (dnn < Pth) dnn = Pth;
            if (dln < Pth) dln = Pth;
            d2 dPkdv{dl.x / dln - dn.x / dnn, dl.y / dln - dn.y / dnn};
            d2 dAjdv{0.5 * (vother.y - vlast.y), 0.5 * (vlast.x - vother.x)};
            dl = d2{vother.x - vcur.x, vother.y - vcur.y};
            dln = std::sqrt(dl.x * dl.x + dl.y * dl.y);
            dn = d2{vcur.x - vlast.x, vcur.y - vlast.y};
            dnn = std::sqrt(dn.x * dn.x + dn.y * dn.y);
            if (dnn < Pth) dnn = Pth;
            if (dln < Pth) dln = Pth;
            d2 dPjdv{dl.x / dln - dn.x / dnn, dl.y / dln - dn.y / dnn};
            d2 dEdv{2.0 * Adiff * dAidv.x + 2.0 * Pdiff * dPidv.x, 2.0 * Adiff * dAidv.y + 2.0 * Pdiff * dPidv.y};
            dEdv.x += 2.0 * Akd * dAkdv.x + 2.0 * Pkd * dPkdv.x;
            dEdv.y += 2.0 * Akd * dAkdv.y + 2.0 * Pkd * dPkdv.y;
            dEdv.x += 2.0 * Ajd * dAjdv.x + 2.0 * Pjd * dPjdv.x;
            dEdv.y += 2.0 * Ajd * dAjdv.y + 2.0 * Pjd * dPjdv.y;
            d2 t = vecmat(dEdv, dh[k]);
            fsum.x += t.x;
            fsum.y += t.y;
            vlast = vcur;
        }
        s.F[i] = fsum;
    }
}

// reference voronoiQuadraticEnergy.cpp:384-397, plus the line-tension energy
// whose gradient the tension force sets are: every Voronoi edge between cells of
// different types counted from both of its cells, gamma times its length. (The
// reference's computeEnergy, voronoiQuadraticEnergyWithTension.cpp:96-155, counts
// each edge once and, with the GPU vertex convention, pairs the edge with the
// neighbour one ring slot off; its force sets are what this energy matches.)
static double energy(const VSys &s) {
    double E = 0.0;
    for (int i = 0; i < s.N; ++i) {
        E += s.KA * (s.AP[i].x - s.APpref[i].x) * (s.AP[i].x - s.APpref[i].x);
        E += s.KP * (s.AP[i].y - s.APpref[i].y) * (s.AP[i].y - s.APpref[i].y);
    }
    if (s.tensionMode != 0)
        for (int i = 0; i < s.N; ++i)
            for (int k = 0; k < s.nn[i]; ++k) {
                const int j = s.ringc(i, k);
                if (s.type[i] == s.type[j]) continue;
                // edge shared with ring k: vertices (k-1, k) and (k, k+1)
                d2 a = voroCurGPU(s, i, k + 1), b = voroCurGPU(s, i, k + 2);
                double l = std::sqrt((a.x - b.x) * (a.x - b.x) + (a.y - b.y) * (a.y - b.y));
                E += (s.tensionMode == 1 ? s.gamma : s.tension(s.type[i], s.type[j])) * l;
            }
    return E;
}

}  // namespace orc

using namespace orc;

// ======================================================================= C API
extern "C" {

// --- primitives (unit tests) ---------------------------------------------------
int orc_orient0(double ax, double ay, double bx, double by) { return orient0(ax, ay, bx, by); }
int orc_incircle0(double bx, double by, double cx, double cy, double dx, double dy) {
    return incircle0(bx, by, cx, cy, dx, dy);
}
void orc_circumcenter0(double x1, double y1, double x2, double y2, double *out) {
    d2 c = circumcenter0(d2{x1, y1}, d2{x2, y2});
    out[0] = c.x;
    out[1] = c.y;
}
void orc_min_dist(double L, double x1, double y1, double x2, double y2, double *out) {
    Box b(L);
    d2 r = b.minDist(d2{x1, y1}, d2{x2, y2});
    out[0] = r.x;
    out[1] = r.y;
}
void orc_put_in_box(double L, double *x, double *y) {
    Box b(L);
    d2 p{*x, *y};
    b.putInBoxReal(p);
    *x = p.x;
    *y = p.y;
}
// reference Simple2DCell.cpp:85-100 + Simple2DActiveCell.cpp:55-70: mt19937(seed)
// with uniform_real_distribution<double>; x then y per cell, then all directors.
void orc_random_init(int N, unsigned int seed, double L, double *x, double *y, double *theta) {
    std::mt19937 gen(seed);
    for (int i = 0; i < N; ++i) {
        std::uniform_real_distribution<double> u(0.0, L);
        x[i] = u(gen);
        y[i] = u(gen);
    }
    for (int i = 0; i < N; ++i) {
        std::uniform_real_distribution<double> u(0.0, 2.0 * 3.14159265358979323846);
        theta[i] = u(gen);
    }
}
// global periodic Delaunay rings (canonical rotation); returns neighMax or -1
int orc_triangulate(int N, double L, const double *x, const double *y, int cap, int *nn, int *nbr) {
    VSys s;
    s.N = N;
    s.box.set(L);
    s.P.resize(N);
    for (int i = 0; i < N; ++i) s.P[i] = d2{x[i], y[i]};
    if (!globalTriangulation(s)) return -1;
    for (int i = 0; i < N; ++i)
        if (s.nn[i] > cap) return -2;
    for (int i = 0; i < N; ++i) {
        nn[i] = s.nn[i];
        for (int k = 0; k < s.nn[i]; ++k) nbr[(size_t)i * cap + k] = s.ring(i, k);
    }
    return s.neighMax;
}

// --- Voronoi system handle ---------------------------------------------------------
void *orc_vsys_create(int N, double L, const double *x, const double *y, const double *theta) {
    VSys *s = new VSys();
    s->N = N;
    s->box.set(L);
    s->P.resize(N);
    for (int i = 0; i < N; ++i) s->P[i] = d2{x[i], y[i]};
    s->AP.assign(N, d2{0, 0});
    s->APpref.assign(N, d2{1.0, 3.8});
    s->F.assign(N, d2{0, 0});
    s->extF.assign(N, d2{0, 0});
    s->excl.assign(N, 0);
    s->theta.assign(theta, theta + N);
    s->mot.assign(N, d2{0.05, 1.0});
    s->vel.assign(N, d2{0, 0});
    for (int i = 0; i < N; ++i) s->vel[i] = d2{0.05 * std::cos(theta[i]), 0.05 * std::sin(theta[i])};
    if (!globalTriangulation(*s)) {
        delete s;
        return nullptr;
    }
    allDelSets(*s);
    return s;
}
void orc_vsys_destroy(void *h) { delete (VSys *)h; }
void orc_vsys_set_params(void *h, double KA, double KP, double dt, double mu) {
    VSys *s = (VSys *)h;
    s->KA = KA; s->KP = KP; s->dt = dt; s->mu = mu;
}
void orc_vsys_set_prefs(void *h, const double *A0, const double *P0) {
    VSys *s = (VSys *)h;
    for (int i = 0; i < s->N; ++i) s->APpref[i] = d2{A0[i], P0[i]};
}
// reference Simple2DActiveCell.cpp:76-92 (setv0Dr also resets v = v0 e(theta))
void orc_vsys_set_motility(void *h, const double *v0, const double *Dr) {
    VSys *s = (VSys *)h;
    for (int i = 0; i < s->N; ++i) {
        s->mot[i] = d2{v0[i], Dr[i]};
        s->vel[i] = d2{v0[i] * std::cos(s->theta[i]), v0[i] * std::sin(s->theta[i])};
    }
}
void orc_vsys_set_exclusions(void *h, const int *ex) {
    VSys *s = (VSys *)h;
    for (int i = 0; i < s->N; ++i) {
        s->excl[i] = ex[i] ? 1 : 0;
        if (ex[i]) s->mot[i] = d2{0.0, 0.0};
    }
}
void orc_vsys_set_state(void *h, const double *x, const double *y, const double *theta, const double *vx,
                        const double *vy) {
    VSys *s = (VSys *)h;
    for (int i = 0; i < s->N; ++i) {
        s->P[i] = d2{x[i], y[i]};
        if (theta) s->theta[i] = theta[i];
        if (vx) s->vel[i] = d2{vx[i], vy[i]};
    }
}
// replace the topology (rings in canonical rotation, cap-strided) and rebuild delSets
void orc_vsys_set_topology(void *h, const int *nn, const int *nbr, int cap, int neighMax) {
    VSys *s = (VSys *)h;
    if (cap > s->cap) {
        s->cap = cap;
    }
    s->nb.assign((size_t)s->N * s->cap, 0);
    s->dsx.assign((size_t)s->N * s->cap, 0);
    s->dsy.assign((size_t)s->N * s->cap, 0);
    s->dother.assign((size_t)s->N * s->cap, 0);
    for (int i = 0; i < s->N; ++i) {
        s->nn[i] = nn[i];
        for (int k = 0; k < nn[i]; ++k) s->ring(i, k) = nbr[(size_t)i * cap + k];
    }
    s->neighMax = neighMax;
    allDelSets(*s);
}
int orc_vsys_get(void *h, double *x, double *y, double *theta, double *vx, double *vy, double *Fx, double *Fy,
                 double *A, double *P) {
    VSys *s = (VSys *)h;
    for (int i = 0; i < s->N; ++i) {
        if (x) { x[i] = s->P[i].x; y[i] = s->P[i].y; }
        if (theta) theta[i] = s->theta[i];
        if (vx) { vx[i] = s->vel[i].x; vy[i] = s->vel[i].y; }
        if (Fx) { Fx[i] = s->F[i].x; Fy[i] = s->F[i].y; }
        if (A) { A[i] = s->AP[i].x; P[i] = s->AP[i].y; }
    }
    return s->N;
}
int orc_vsys_get_topology(void *h, int *nn, int *nbr, int cap) {
    VSys *s = (VSys *)h;
    for (int i = 0; i < s->N; ++i) {
        nn[i] = s->nn[i];
        for (int k = 0; k < s->nn[i] && k < cap; ++k) nbr[(size_t)i * cap + k] = s->ring(i, k);
    }
    return s->neighMax;
}
void orc_vsys_get_counters(void *h, long long *out) {
    VSys *s = (VSys *)h;
    out[0] = s->globalFixes;
    out[1] = s->localRepairs;
    out[2] = s->repairEvents;
    out[3] = s->skippedFrames;
    out[4] = s->lastRepairMode;
    out[5] = (long long)s->lastNeedsFixing.size();
}
// reference Simple2DCell::cellDeath + Simple2DActiveCell::cellDeath +
// voronoiModelBase::cellDeath (Simple2DCell.cpp:775-810, Simple2DActiveCell.cpp:118-123,
// voronoiModelBase.cpp:1309-1316): the cell's entries are removed, the cells after
// it move down one index, then a global retriangulation
int orc_vsys_cell_death(void *h, int i) {
    VSys *s = (VSys *)h;
    if (i < 0 || i >= s->N || s->N <= 4) return -1;
    s->N -= 1;
    s->P.erase(s->P.begin() + i);
    s->AP.erase(s->AP.begin() + i);
    s->APpref.erase(s->APpref.begin() + i);
    s->F.erase(s->F.begin() + i);
    s->extF.erase(s->extF.begin() + i);
    s->excl.erase(s->excl.begin() + i);
    s->theta.erase(s->theta.begin() + i);
    s->mot.erase(s->mot.begin() + i);
    s->vel.erase(s->vel.begin() + i);
    if (!s->type.empty()) s->type.erase(s->type.begin() + i);
    if (!globalTriangulation(*s)) return -1;
    allDelSets(*s);
    return 0;
}
// reference voronoiModelBase::cellDivision (voronoiModelBase.cpp:1318-1422) with
// Simple2DCell::cellDivision (.cpp:819-850) and Simple2DActiveCell::cellDivision
// (.cpp:132-147): the cell's Voronoi polygon (CPU vertex order, vertex k =
// circumcenter(ring k-1, ring k)), the two crossings of the axis at angle theta,
// the parent moved to +frac * maxSeparation along the axis and a daughter appended
// at -frac * maxSeparation, copying the parent's preferences, motility and type,
// with director dtheta and velocity v0 (cos, sin)(dtheta); then a global
// retriangulation. (One crossing only: the second is taken equal to the first.)
int orc_vsys_cell_division(void *h, int i, double theta, double frac, double dtheta) {
    VSys *s = (VSys *)h;
    if (i < 0 || i >= s->N) return -1;
    const int z = s->nn[i];
    d2 pi = s->P[i];
    vector<d2> voro;
    d2 rij = s->box.minDist(s->P[s->ring(i, z - 1)], pi);
    for (int k = 0; k < z; ++k) {
        d2 rik = s->box.minDist(s->P[s->ring(i, k)], pi);
        voro.push_back(circumcenter0(rij, rik));
        rij = rik;
    }
    d2 v1 = voro[z - 1], int1{0, 0}, int2{0, 0};
    bool first = false, second = false;
    const d2 ray{std::cos(theta), std::sin(theta)}, c{-std::sin(theta), std::cos(theta)};
    for (int k = 0; k < z; ++k) {
        d2 v2 = voro[k];
        d2 a{0.0 - v2.x, 0.0 - v2.y}, b{v1.x - v2.x, v1.y - v2.y};
        double t2 = -1.;
        double bc = b.x * c.x + b.y * c.y;
        if (bc != 0) t2 = (a.x * c.x + a.y * c.y) / bc;
        if (t2 >= 0. && t2 <= 1.0) {
            d2 q{v2.x + t2 * (v1.x - v2.x), v2.y + t2 * (v1.y - v2.y)};
            if (first) {
                int2 = q;
                second = true;
            } else {
                int1 = q;
                first = true;
            }
        }
        v1 = v2;
    }
    if (!first) return -1;
    if (!second) int2 = int1;
    double sep = std::max(std::sqrt(int1.x * int1.x + int1.y * int1.y), std::sqrt(int2.x * int2.x + int2.y * int2.y));
    double f = frac * sep;
    d2 p1{pi.x + f * ray.x, pi.y + f * ray.y}, p2{pi.x - f * ray.x, pi.y - f * ray.y};
    s->box.putInBoxReal(p1);
    s->box.putInBoxReal(p2);
    s->N += 1;
    s->P[i] = p1;
    s->P.push_back(p2);
    s->AP.push_back(d2{0, 0});
    s->APpref.push_back(s->APpref[i]);
    s->F.push_back(d2{0, 0});
    s->extF.push_back(d2{0, 0});
    s->excl.push_back(0);
    s->mot.push_back(s->mot[i]);
    s->theta.push_back(dtheta);
    s->vel.push_back(d2{s->mot[i].x * std::cos(dtheta), s->mot[i].x * std::sin(dtheta)});
    if (!s->type.empty()) s->type.push_back(s->type[i]);
    if (!globalTriangulation(*s)) return -1;
    allDelSets(*s);
    return 0;
}
int orc_vsys_num_cells(void *h) { return ((VSys *)h)->N; }
// line tension (VoronoiQuadraticEnergyWithTension): mode 0 off, 1 one value
// gamma between unlike types, 2 matrix g[a*ntypes + b]; types per cell
void orc_vsys_set_tension(void *h, int mode, double gamma, int ntypes, const double *g, const int *types) {
    VSys *s = (VSys *)h;
    s->tensionMode = mode;
    s->gamma = gamma;
    s->ntypes = ntypes;
    s->tmat.assign(g ? g : nullptr, g ? g + (size_t)ntypes * ntypes : nullptr);
    s->type.assign(types, types + s->N);
}
// mode 0: GPU-path geometry + force sets; mode 1: CPU-path geometry + forces
double orc_vsys_compute_forces(void *h, int mode) {
    VSys *s = (VSys *)h;
    bool ex = false;
    for (int i = 0; i < s->N; ++i) ex = ex || s->excl[i];
    if (mode == 0) {
        geometryGPUStyle(*s);
        forcesGPUStyle(*s, ex);
    } else {
        geometryAndForcesCPUStyle(*s);
    }
    return energy(*s);
}
// flag vector of the reference GPU circumcircle test (N3 window) on the current state
int orc_vsys_test_flags(void *h, int *repair) {
    VSys *s = (VSys *)h;
    vector<int> r;
    bool any = testTriangulationGPUStyle(*s, r);
    for (int i = 0; i < s->N; ++i) repair[i] = r[i];
    return any ? 1 : 0;
}
// the NeedsFixing set of the last enforceTopology (sorted; empty when no
// triangle failed); returns its size, writes it to out when out is not null
int orc_vsys_last_needs_fixing(void *h, int *out) {
    VSys *s = (VSys *)h;
    if (out)
        for (size_t q = 0; q < s->lastNeedsFixing.size(); ++q) out[q] = s->lastNeedsFixing[q];
    return (int)s->lastNeedsFixing.size();
}
// last repair mode: 0 none, 1 local, 2 global (N/6 rule or a failed local repair)
int orc_vsys_last_repair_mode(void *h) { return ((VSys *)h)->lastRepairMode; }
// CPU-path flags (reference DelaunayLoc::testPointTriangulation window around
// the circumcenter, DelaunayLoc.cpp:517-574); returns whether any cell failed
int orc_vsys_test_flags_cpu(void *h, int *repair) {
    VSys *s = (VSys *)h;
    vector<int> r;
    bool any = testTriangulationCPUStyle(*s, r);
    for (int i = 0; i < s->N; ++i) repair[i] = r[i];
    return any ? 1 : 0;
}
// force-free enforceTopology on the current positions (GPU branch semantics)
int orc_vsys_enforce_topology(void *h, int gpuPath) {
    VSys *s = (VSys *)h;
    return enforceTopology(*s, gpuPath != 0) ? 0 : -1;
}
// One self-propelled Voronoi step, reference Simulation.cpp:141-166 ->
// selfPropelledParticleDynamics GPU (.cu:20-57, N1 lag) or CPU (.cpp:67-101) ->
// movePoints -> enforceTopology. eta[i] is the N(0,1) draw of cell i this step.
int orc_vsys_spp_step(void *h, const double *eta, int gpuPath) {
    VSys *s = (VSys *)h;
    s->sppTimestep += 1;
    bool ex = false;
    for (int i = 0; i < s->N; ++i) ex = ex || s->excl[i];
    vector<d2> disp(s->N);
    if (gpuPath) {
        geometryGPUStyle(*s);
        forcesGPUStyle(*s, ex);
        for (int i = 0; i < s->N; ++i) {
            disp[i].x = s->dt * (s->vel[i].x + s->mu * s->F[i].x);
            disp[i].y = s->dt * (s->vel[i].y + s->mu * s->F[i].y);
            double v0 = s->mot[i].x, Dr = s->mot[i].y;
            double angleDiff = eta[i] * std::sqrt(2.0 * s->dt * Dr);
            double th = s->theta[i];
            s->vel[i].x = v0 * std::cos(s->theta[i]);
            s->vel[i].y = v0 * std::sin(s->theta[i]);
            if (s->vel[i].y != 0. && s->vel[i].x != 0.) th = std::atan2(s->vel[i].y, s->vel[i].x);
            s->theta[i] = th + angleDiff;
        }
    } else {
        geometryAndForcesCPUStyle(*s);
        for (int i = 0; i < s->N; ++i) {
            double v0 = s->mot[i].x, Dr = s->mot[i].y;
            s->vel[i].x = v0 * std::cos(s->theta[i]);
            s->vel[i].y = v0 * std::sin(s->theta[i]);
            d2 V = s->vel[i];
            disp[i].x = s->dt * (V.x + s->mu * s->F[i].x);
            disp[i].y = s->dt * (V.y + s->mu * s->F[i].y);
            double th = s->theta[i];
            if (!(V.x == 0. && V.y == 0.)) th = std::atan2(V.y, V.x);
            s->theta[i] = th + eta[i] * std::sqrt(2.0 * s->dt * Dr);
        }
    }
    // reference Simple2DCell.cu:21-35 (move then putInBoxReal)
    for (int i = 0; i < s->N; ++i) {
        s->P[i].x += disp[i].x;
        s->P[i].y += disp[i].y;
        s->box.putInBoxReal(s->P[i]);
    }
    return enforceTopology(*s, gpuPath != 0) ? 0 : -1;
}
// Brownian step on cells (reference brownianParticleDynamics.cu:20-39):
// disp = eta * sqrt(dt*T*mu*2) + dt*mu*F  (GPU operand order), x draw before y
// `steps` self-propelled steps in one call, eta[k * N + i] the noise of cell i at
// step k, timed with std::chrono::steady_clock (the reference driver times its
// loop with clock(), voronoi.cpp:110-126); returns the first failing status
int orc_vsys_spp_run(void *h, const double *eta, int steps, int gpuPath, double *seconds) {
    VSys *s = (VSys *)h;
    const auto t0 = std::chrono::steady_clock::now();
    int r = 0;
    for (int k = 0; k < steps && r == 0; ++k) r = orc_vsys_spp_step(h, eta + (size_t)k * s->N, gpuPath);
    const auto t1 = std::chrono::steady_clock::now();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
    return r;
}

int orc_vsys_brownian_step(void *h, const double *etax, const double *etay, double T) {
    VSys *s = (VSys *)h;
    geometryGPUStyle(*s);
    forcesGPUStyle(*s, false);
    for (int i = 0; i < s->N; ++i) {
        double sc = std::sqrt(s->dt * T * s->mu * 2.0);
        double dx = etax[i] * sc + s->dt * s->mu * s->F[i].x;
        double dy = etay[i] * sc + s->dt * s->mu * s->F[i].y;
        s->P[i].x += dx;
        s->P[i].y += dy;
        s->box.putInBoxReal(s->P[i]);
    }
    return enforceTopology(*s, true) ? 0 : -1;
}

// --- FIRE (reference EnergyMinimizerFIRE2D.cpp:24-41, 93-114, 156-207, 269-284) ---
// the loop is fire_oracle.h; the Voronoi model's topology step is the repair
struct VoronoiFireModel {
    VSys *s;
    int dof() { return s->N; }
    void forces(vector<d2> &F) {
        geometryGPUStyle(*s);
        forcesGPUStyle(*s, false);
        F = s->F;
    }
    void move(const vector<d2> &d) {
        for (int i = 0; i < s->N; ++i) {
            s->P[i].x += d[i].x;
            s->P[i].y += d[i].y;
            s->box.putInBoxReal(s->P[i]);
        }
    }
    bool topology() { return enforceTopology(*s, true); }
};
void *orc_fire_create(void *hsys) {
    (void)hsys;
    return new fireo::State<d2>();
}
void orc_fire_destroy(void *h) { delete (fireo::State<d2> *)h; }
void orc_fire_set(void *h, int maxIt, double fc, double alphaStart, double dt, double dtMax, double dtInc,
                  double dtDec, double alphaDec, int NMin) {
    ((fireo::State<d2> *)h)->set(maxIt, fc, alphaStart, dt, dtMax, dtInc, dtDec, alphaDec, NMin);
}
// GPU-branch FIRE on a Voronoi system; returns iterations done. out: forceMax
// (= sum|F|^2/N), Power, alpha, dt, max|F|^2
// FIRE on an anisotropic quadratic bowl, E = sum_i (kx_i x_i^2 + ky_i y_i^2) / 2
// (F = -k x componentwise, no box, no topology): a model whose trajectory can
// be followed by hand. k, x: 2n values (x then y per degree of freedom).
// log_out: 9 doubles per iteration (power, dt before / after, alpha before /
// after, NSince before / after, velocities zeroed, sum |v|^2 after); x is
// updated in place. Returns the iterations done.
struct BowlModel {
    int n;
    const double *k;
    std::vector<d2> x;
    int dof() const { return n; }
    void forces(std::vector<d2> &F) {
        F.resize(n);
        for (int i = 0; i < n; ++i) F[i] = d2{-k[2 * i] * x[i].x, -k[2 * i + 1] * x[i].y};
    }
    void move(const std::vector<d2> &d) {
        for (int i = 0; i < n; ++i) {
            x[i].x += d[i].x;
            x[i].y += d[i].y;
        }
    }
    bool topology() { return true; }
};
int orc_fire_bowl(int n, const double *k, double *x, int maxIt, double fc, double alphaStart, double dt,
                  double dtMax, double dtInc, double dtDec, double alphaDec, int NMin, double *log_out,
                  double *out) {
    BowlModel m{n, k, std::vector<d2>(n)};
    for (int i = 0; i < n; ++i) m.x[i] = d2{x[2 * i], x[2 * i + 1]};
    fireo::State<d2> st;
    st.set(maxIt, fc, alphaStart, dt, dtMax, dtInc, dtDec, alphaDec, NMin);
    std::vector<fireo::IterLog> lg;
    st.log = &lg;
    const int it = fireo::minimize(st, m, out);
    for (size_t q = 0; q < lg.size(); ++q) {
        const fireo::IterLog &e = lg[q];
        double *o = log_out + 9 * q;
        o[0] = e.power;
        o[1] = e.dt_before;
        o[2] = e.dt_after;
        o[3] = e.alpha_before;
        o[4] = e.alpha_after;
        o[5] = e.nsince_before;
        o[6] = e.nsince_after;
        o[7] = e.zeroed;
        o[8] = e.vv_after;
    }
    for (int i = 0; i < n; ++i) {
        x[2 * i] = m.x[i].x;
        x[2 * i + 1] = m.x[i].y;
    }
    return it;
}

int orc_fire_minimize(void *hf, void *hsys, double *out) {
    VoronoiFireModel m{(VSys *)hsys};
    return fireo::minimize(*(fireo::State<d2> *)hf, m, out);
}

}  // extern "C"
