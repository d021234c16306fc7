// ORACLE — test infrastructure only (see voronoi_oracle.cpp).
//
// CPU FP64 restatement of the reference vertex model GPU path
// (abaso/cellGPU src/models/vertexModelBase.{cpp,cu}, vertexQuadraticEnergy.{cpp,cu},
// updaters/brownianParticleDynamics.cu), with one deliberate change: T1 flips
// are applied in deterministic rounds (a flagged edge slot is flipped when its
// slot index is the smallest among the flagged edges sharing one of its four
// cells) instead of the reference's atomicExch race. For edges
// whose four cells are not shared with another flagged edge both rules flip
// exactly the same edges.
//
// Initial topology: Voronoi tessellation of the periodic Delaunay triangulation;
// vertex ids follow (owner cell, ring slot) of the triangles owned by their
// smallest cell (the reference's CGAL face order is internal to CGAL).
//
// Parity pins: tests/test_oracle_vertex.py (honeycomb closed forms, -grad E by
// finite differences, single-T1 known answer, sum of cellVertexNum invariant).
#include <algorithm>
#include <array>
#include <climits>
#include <cmath>
#include <cstring>
#include <vector>

#include "fire_oracle.h"

using std::vector;

namespace vmo {

struct d2 { double x, y; };

struct Box {
    double L, half, inv;
    void set(double l) { L = l; half = 0.5 * l; inv = 1. / l; }
    d2 minDist(d2 p1, d2 p2) const {  // reference gpubox.h:154-179
        d2 a{p1.x - p2.x, p1.y - p2.y};
        for (int t = 0; t < 64 && a.x < -half; ++t) a.x += L;
        for (int t = 0; t < 64 && a.y < -half; ++t) a.y += L;
        for (int t = 0; t < 64 && a.x > half; ++t) a.x -= L;
        for (int t = 0; t < 64 && a.y > half; ++t) a.y -= L;
        return a;
    }
    static double wrapUnit(double v) {
        long long g = 0;
        while (v < 0 && g++ < (1LL << 40)) v += 1.0;
        while (v >= 1.0 && g++ < (1LL << 40)) v -= 1.0;
        return v;
    }
    void putInBoxReal(d2 &p) const {  // reference gpubox.h:130-152
        double vx = wrapUnit(inv * p.x), vy = wrapUnit(inv * p.y);
        p.x = L * vx;
        p.y = L * vy;
    }
};

static inline d2 circ0(d2 x1, d2 x2) {  // reference functions.h:34-43
    double n1 = x1.x * x1.x + x1.y * x1.y, n2 = x2.x * x2.x + x2.y * x2.y;
    double den = 0.5 / (x1.x * x2.y - x1.y * x2.x);
    return d2{den * (n1 * x2.y - x1.y * n2), den * (x1.x * n2 - n1 * x2.x)};
}

struct VM {
    int Nc = 0, Nv = 0;
    Box box;
    vector<d2> v;                 // vertexPositions
    vector<int> vn, vcn;          // vertexNeighbors, vertexCellNeighbors [3 Nv]
    vector<vector<int>> cv;       // cellVertices (counter-clockwise)
    vector<d2> AP, APpref, F;
    vector<d2> vcur, vlast, vnext;  // per (vertex, cell) slot: voroCur, voroLastNext
    double KA = 1, KP = 1, dt = 0.005, mu = 1, T = 0, thr = 0.04;
    long long t1count = 0;
    int lastFlagged = 0;
    vector<double> theta;  // cellDirectors
    vector<d2> mot;        // Motility (v0, Dr); reference default setv0Dr(0.0, 1.0)
    fireo::State<d2> fire;
};

// reference vertexModelBase.cpp:248-297 / vertexModelBase.cu:20-77
static void geometry(VM &m) {
    for (int c = 0; c < m.Nc; ++c) {
        const vector<int> &cv = m.cv[c];
        int n = (int)cv.size();
        d2 cellPos = m.v[cv[n - 2]];
        d2 vlast{0.0, 0.0}, vcur, vnext;
        double A = 0, P = 0;
        int vidx = cv[n - 1];
        vcur = m.box.minDist(m.v[vidx], cellPos);
        for (int k = 0; k < n; ++k) {
            int fsi = -1;
            for (int ff = 0; ff < 3; ++ff)
                if (fsi == -1 && m.vcn[3 * vidx + ff] == c) fsi = 3 * vidx + ff;
            vidx = cv[k];
            vnext = m.box.minDist(m.v[vidx], cellPos);
            A += 0.5 * (vcur.x + vnext.x) * (vnext.y - vcur.y);
            double dx = vcur.x - vnext.x, dy = vcur.y - vnext.y;
            P += std::sqrt(dx * dx + dy * dy);
            if (fsi >= 0) {
                m.vcur[fsi] = vcur;
                m.vlast[fsi] = vlast;
                m.vnext[fsi] = vnext;
            }
            vlast = vcur;
            vcur = vnext;
        }
        m.AP[c] = d2{A, P};
    }
}

// reference functions.h:330-354 (computeForceSetVertexModel) summed over the
// three slots of each vertex (vertexQuadraticEnergy.cu:21-70)
static void forces(VM &m) {
    for (int vv = 0; vv < m.Nv; ++vv) {
        d2 f{0, 0};
        for (int ff = 0; ff < 3; ++ff) {
            int s = 3 * vv + ff;
            int c = m.vcn[s];
            double Adiff = m.KA * (m.AP[c].x - m.APpref[c].x);
            double Pdiff = m.KP * (m.AP[c].y - m.APpref[c].y);
            d2 vc = m.vcur[s], vl = m.vlast[s], vn = m.vnext[s];
            d2 dAdv{0.5 * (vl.y - vn.y), -0.5 * (vl.x - vn.x)};
            d2 dl{vl.x - vc.x, vl.y - vc.y};
            double dln = std::sqrt(dl.x * dl.x + dl.y * dl.y);
            d2 dn{vc.x - vn.x, vc.y - vn.y};
            double dnn = std::sqrt(dn.x * dn.x + dn.y * dn.y);
            d2 dPdv{dl.x / dln - dn.x / dnn, dl.y / dln - dn.y / dnn};
            d2 dEdv{2.0 * (Adiff * dAdv.x + Pdiff * dPdv.x), 2.0 * (Adiff * dAdv.y + Pdiff * dPdv.y)};
            f.x += dEdv.x;
            f.y += dEdv.y;
        }
        m.F[vv] = f;
    }
}

static double energy(const VM &m) {  // reference vertexQuadraticEnergy.cpp:30-44
    double E = 0;
    for (int c = 0; c < m.Nc; ++c) {
        E += m.KA * (m.AP[c].x - m.APpref[c].x) * (m.AP[c].x - m.APpref[c].x);
        E += m.KP * (m.AP[c].y - m.APpref[c].y) * (m.AP[c].y - m.APpref[c].y);
    }
    return E;
}

// the four cells of the edge (v1, v2): vertexCellNeighbors of v1 plus the cell
// of v2 not among them (reference vertexModelBase.cu:195-205)
static void edgeCells(const VM &m, int v1, int v2, int out[4]) {
    out[0] = m.vcn[3 * v1];
    out[1] = m.vcn[3 * v1 + 1];
    out[2] = m.vcn[3 * v1 + 2];
    out[3] = -1;
    for (int ff = 0; ff < 3; ++ff) {
        int c = m.vcn[3 * v2 + ff];
        if (c != out[0] && c != out[1] && c != out[2]) out[3] = c;
    }
}

// position of v in cell c: (vlast, vcur, vnext) as the reference loop finds them
static void locate(const vector<int> &cv, int vtx, int &vl, int &vnx) {
    int n = (int)cv.size();
    int vlast = cv[n - 2], vcur = cv[n - 1], vnext = cv[0];
    for (int k = 0; k < n; ++k) {
        vnext = cv[k];
        if (vcur == vtx) break;
        vlast = vcur;
        vcur = vnext;
    }
    vl = vlast;
    vnx = vnext;
}

// reference vm_flip_edges_kernel (vertexModelBase.cu:252-544) for one slot
static bool flipEdge(VM &m, int v1, int v2, const int cells[4]) {
    int cs[4] = {-1, -1, -1, cells[3]};  // i, j, k, l
    for (int q = 0; q < 3; ++q) {
        int vl, vnx;
        locate(m.cv[cells[q]], v1, vl, vnx);
        if (vl == v2)
            cs[0] = cells[q];
        else if (vnx == v2)
            cs[2] = cells[q];
        else
            cs[1] = cells[q];
    }
    if (cs[0] < 0 || cs[1] < 0 || cs[2] < 0 || cs[3] < 0) return false;
    if ((int)m.cv[cs[0]].size() == 3 || (int)m.cv[cs[2]].size() == 3) return false;
    int dummy, a, b;
    locate(m.cv[cs[1]], v1, dummy, a);  // vertexSet.x
    locate(m.cv[cs[3]], v2, dummy, b);  // vertexSet.y
    // rotate the edge by 90 degrees about its midpoint, doubling its length
    d2 p1 = m.v[v1], p2 = m.v[v2];
    d2 e = m.box.minDist(p1, p2);
    d2 mid{p2.x + 0.5 * e.x, p2.y + 0.5 * e.y};
    p1 = d2{mid.x - e.y, mid.y + e.x};
    p2 = d2{mid.x + e.y, mid.y - e.x};
    m.box.putInBoxReal(p1);
    m.box.putInBoxReal(p2);
    m.v[v1] = p1;
    m.v[v2] = p2;
    for (int q = 0; q < 3; ++q) {
        if (m.vcn[3 * v1 + q] == cs[2]) m.vcn[3 * v1 + q] = cs[3];
        if (m.vcn[3 * v2 + q] == cs[0]) m.vcn[3 * v2 + q] = cs[1];
        if (m.vn[3 * v1 + q] == a) m.vn[3 * v1 + q] = b;
        if (m.vn[3 * v2 + q] == b) m.vn[3 * v2 + q] = a;
        if (m.vn[3 * a + q] == v1) m.vn[3 * a + q] = v2;
        if (m.vn[3 * b + q] == v2) m.vn[3 * b + q] = v1;
    }
    auto erase = [](vector<int> &l, int x) { l.erase(std::find(l.begin(), l.end(), x)); };
    auto insertBefore = [](vector<int> &l, int before, int x) { l.insert(std::find(l.begin(), l.end(), before), x); };
    erase(m.cv[cs[0]], v2);
    insertBefore(m.cv[cs[1]], a, v2);
    erase(m.cv[cs[2]], v1);
    insertBefore(m.cv[cs[3]], b, v1);
    return true;
}

// reference vm_simple_T1_test_kernel (vertexModelBase.cu:118-164) + flipping in
// deterministic rounds. A flagged edge is the pair (v1 < v2); in each round an
// edge flips when its key (v1, v2) is the smallest among the flagged edges that
// share one of its four cells. An edge that a neighbouring flip has removed
// (v2 no longer a neighbour of v1) is dropped.
static void t1(VM &m) {
    vector<std::pair<int, int>> flagged;
    for (int s = 0; s < 3 * m.Nv; ++s) {
        int v1 = s / 3, v2 = m.vn[s];
        if (v1 < v2) {
            d2 e = m.box.minDist(m.v[v1], m.v[v2]);
            if (std::sqrt(e.x * e.x + e.y * e.y) < m.thr) flagged.push_back({v1, v2});
        }
    }
    m.lastFlagged = (int)flagged.size();
    const unsigned long long NONE = ~0ull;
    vector<unsigned long long> cellMin(m.Nc, NONE);
    auto key = [](int a, int b) { return ((unsigned long long)(unsigned)a << 32) | (unsigned)b; };
    while (!flagged.empty()) {
        vector<std::pair<int, int>> live;
        for (auto &e : flagged) {
            int v1 = e.first, v2 = e.second;
            if (m.vn[3 * v1] == v2 || m.vn[3 * v1 + 1] == v2 || m.vn[3 * v1 + 2] == v2) live.push_back(e);
        }
        flagged.swap(live);
        if (flagged.empty()) break;
        vector<std::array<int, 4>> C(flagged.size());
        for (size_t q = 0; q < flagged.size(); ++q) {
            int c4[4];
            edgeCells(m, flagged[q].first, flagged[q].second, c4);
            unsigned long long k = key(flagged[q].first, flagged[q].second);
            for (int t = 0; t < 4; ++t) {
                C[q][t] = c4[t];
                if (c4[t] >= 0) cellMin[c4[t]] = std::min(cellMin[c4[t]], k);
            }
        }
        vector<std::pair<int, int>> rest;
        vector<int> sel;
        for (size_t q = 0; q < flagged.size(); ++q) {
            unsigned long long k = key(flagged[q].first, flagged[q].second);
            bool win = true;
            for (int t = 0; t < 4; ++t)
                if (C[q][t] >= 0 && cellMin[C[q][t]] != k) win = false;
            if (win)
                sel.push_back((int)q);
            else
                rest.push_back(flagged[q]);
        }
        for (int q : sel) {
            int c4[4] = {C[q][0], C[q][1], C[q][2], C[q][3]};
            if (flipEdge(m, flagged[q].first, flagged[q].second, c4)) m.t1count++;
        }
        for (size_t q = 0; q < flagged.size(); ++q)
            for (int t = 0; t < 4; ++t)
                if (C[q][t] >= 0) cellMin[C[q][t]] = NONE;
        flagged.swap(rest);
    }
}

}  // namespace vmo

using namespace vmo;

extern "C" {

// Vertex model from a periodic Delaunay triangulation: rings nbr (cap-strided,
// counter-clockwise) of Nc cells at (x, y). Vertices = triangles owned by their
// smallest cell, in (cell, slot) order.
void *vmo_create(int Nc, double L, const double *x, const double *y, const int *nn, const int *nbr, int cap) {
    VM *m = new VM();
    m->Nc = Nc;
    m->box.set(L);
    vector<int> tri_id((size_t)Nc * cap, -1);
    vector<std::array<int, 3>> tris;
    for (int i = 0; i < Nc; ++i)
        for (int k = 0; k < nn[i]; ++k) {
            int j = nbr[(size_t)i * cap + k], l = nbr[(size_t)i * cap + (k + 1) % nn[i]];
            if (i < j && i < l) {
                tri_id[(size_t)i * cap + k] = (int)tris.size();
                tris.push_back({i, j, l});
            }
        }
    // id of the triangle (i, n_k, n_k+1) seen from any of its corners
    auto find_tri = [&](int a, int b, int c) {  // counter-clockwise (a, b, c)
        int o = std::min(a, std::min(b, c));
        int p, q;
        if (o == a) { p = b; q = c; } else if (o == b) { p = c; q = a; } else { p = a; q = b; }
        for (int k = 0; k < nn[o]; ++k)
            if (nbr[(size_t)o * cap + k] == p && nbr[(size_t)o * cap + (k + 1) % nn[o]] == q)
                return tri_id[(size_t)o * cap + k];
        return -1;
    };
    m->Nv = (int)tris.size();
    m->v.resize(m->Nv);
    m->vn.assign(3 * m->Nv, -1);
    m->vcn.assign(3 * m->Nv, -1);
    for (int t = 0; t < m->Nv; ++t) {
        int i = tris[t][0], j = tris[t][1], l = tris[t][2];
        d2 pi{x[i], y[i]};
        d2 dj = m->box.minDist(d2{x[j], y[j]}, pi), dl = m->box.minDist(d2{x[l], y[l]}, pi);
        d2 c = circ0(dj, dl);
        d2 p{pi.x + c.x, pi.y + c.y};
        m->box.putInBoxReal(p);
        m->v[t] = p;
        int cells[3] = {i, j, l};
        for (int f = 0; f < 3; ++f) {
            m->vcn[3 * t + f] = cells[f];
            int a = cells[(f + 1) % 3], b = cells[(f + 2) % 3];
            // triangle across edge (a, b): counter-clockwise (b, a, X)
            int X = -1;
            for (int k = 0; k < nn[b]; ++k)
                if (nbr[(size_t)b * cap + (k + 1) % nn[b]] == a) X = nbr[(size_t)b * cap + (k + 2) % nn[b]];
            m->vn[3 * t + f] = find_tri(b, a, X);
        }
    }
    m->cv.assign(Nc, {});
    for (int i = 0; i < Nc; ++i)
        for (int k = 0; k < nn[i]; ++k)
            m->cv[i].push_back(find_tri(i, nbr[(size_t)i * cap + k], nbr[(size_t)i * cap + (k + 1) % nn[i]]));
    m->AP.assign(Nc, d2{0, 0});
    m->APpref.assign(Nc, d2{1.0, 3.9});
    m->F.assign(m->Nv, d2{0, 0});
    m->theta.assign(Nc, 0.0);
    m->mot.assign(Nc, d2{0.0, 1.0});
    m->vcur.assign(3 * m->Nv, d2{0, 0});
    m->vlast = m->vcur;
    m->vnext = m->vcur;
    return m;
}
void vmo_destroy(void *h) { delete (VM *)h; }
int vmo_nv(void *h) { return ((VM *)h)->Nv; }
void vmo_set_params(void *h, double KA, double KP, double dt, double mu, double T, double thr) {
    VM *m = (VM *)h;
    m->KA = KA; m->KP = KP; m->dt = dt; m->mu = mu; m->T = T; m->thr = thr;
}
void vmo_set_prefs(void *h, const double *A0, const double *P0) {
    VM *m = (VM *)h;
    for (int c = 0; c < m->Nc; ++c) m->APpref[c] = d2{A0[c], P0[c]};
}
void vmo_set_positions(void *h, const double *vx, const double *vy) {
    VM *m = (VM *)h;
    for (int i = 0; i < m->Nv; ++i) m->v[i] = d2{vx[i], vy[i]};
}
// topology as arrays: vn, vcn [3Nv]; cvn [Nc]; cv [Nc*cap] (cell lists as stored)
void vmo_get_topology(void *h, int *vn, int *vcn, int *cvn, int *cv, int cap) {
    VM *m = (VM *)h;
    for (int i = 0; i < 3 * m->Nv; ++i) { vn[i] = m->vn[i]; vcn[i] = m->vcn[i]; }
    for (int c = 0; c < m->Nc; ++c) {
        cvn[c] = (int)m->cv[c].size();
        for (int k = 0; k < cvn[c] && k < cap; ++k) cv[(size_t)c * cap + k] = m->cv[c][k];
    }
}
void vmo_set_topology(void *h, const int *vn, const int *vcn, const int *cvn, const int *cv, int cap) {
    VM *m = (VM *)h;
    for (int i = 0; i < 3 * m->Nv; ++i) { m->vn[i] = vn[i]; m->vcn[i] = vcn[i]; }
    for (int c = 0; c < m->Nc; ++c) m->cv[c].assign(cv + (size_t)c * cap, cv + (size_t)c * cap + cvn[c]);
}
double vmo_compute(void *h, double *vx, double *vy, double *Fx, double *Fy, double *A, double *P) {
    VM *m = (VM *)h;
    geometry(*m);
    forces(*m);
    for (int i = 0; i < m->Nv; ++i) {
        if (vx) { vx[i] = m->v[i].x; vy[i] = m->v[i].y; }
        if (Fx) { Fx[i] = m->F[i].x; Fy[i] = m->F[i].y; }
    }
    for (int c = 0; c < m->Nc; ++c)
        if (A) { A[c] = m->AP[c].x; P[c] = m->AP[c].y; }
    return energy(*m);
}
// one Brownian step (reference brownianParticleDynamics GPU branch) with T1s;
// returns the number of T1 flips performed
int vmo_brownian_step(void *h, const double *etax, const double *etay) {
    VM *m = (VM *)h;
    geometry(*m);
    forces(*m);
    long long before = m->t1count;
    for (int i = 0; i < m->Nv; ++i) {
        double sc = std::sqrt(m->dt * m->T * m->mu * 2.0);
        double dx = etax[i] * sc + m->dt * m->mu * m->F[i].x;
        double dy = etay[i] * sc + m->dt * m->mu * m->F[i].y;
        m->v[i].x += dx;
        m->v[i].y += dy;
        m->box.putInBoxReal(m->v[i]);
    }
    t1(*m);
    return (int)(m->t1count - before);
}
// T1 test-and-flip on the current positions only
int vmo_t1(void *h) {
    VM *m = (VM *)h;
    long long before = m->t1count;
    t1(*m);
    return (int)(m->t1count - before);
}
int vmo_last_flagged(void *h) { return ((VM *)h)->lastFlagged; }
// reference vertexModelBase::cellDeath (vertexModelBase.cpp:905-1110): T2
// transition of a triangular cell. The cell's vertices (rotated so that the
// smallest is first) merge into the first, which keeps its position; the three
// surrounding cells replace the other two by it (first occurrences kept); the
// three outer neighbours point to it; vertices and cells above the removed ones
// move down. Returns 0, or -1 when the cell is not a simple triangle.
int vmo_cell_death(void *h, int c) {
    VM *m = (VM *)h;
    if (c < 0 || c >= m->Nc || m->cv[c].size() != 3) return -1;
    vector<int> V = m->cv[c], cells, outer;
    for (int v : V)
        for (int f = 0; f < 3; ++f) {
            int q = m->vcn[3 * v + f];
            if (q != c && std::find(cells.begin(), cells.end(), q) == cells.end()) cells.push_back(q);
        }
    std::rotate(V.begin(), std::min_element(V.begin(), V.end()), V.end());
    for (int v : V)
        for (int f = 0; f < 3; ++f) {
            int u = m->vn[3 * v + f];
            if (u != V[0] && u != V[1] && u != V[2] && std::find(outer.begin(), outer.end(), u) == outer.end())
                outer.push_back(u);
        }
    if (cells.size() != 3 || outer.size() != 3) return -1;
    for (int q : cells) {
        vector<int> l;
        for (int v : m->cv[q]) {
            int w = (v == V[1] || v == V[2]) ? V[0] : v;
            if (std::find(l.begin(), l.end(), w) == l.end()) l.push_back(w);
        }
        m->cv[q] = l;
    }
    for (int k = 0; k < 3; ++k) {
        m->vcn[3 * V[0] + k] = cells[k];
        m->vn[3 * V[0] + k] = outer[k];
        for (int f = 0; f < 3; ++f)
            if (m->vn[3 * outer[k] + f] == V[1] || m->vn[3 * outer[k] + f] == V[2]) m->vn[3 * outer[k] + f] = V[0];
    }
    const int a = std::min(V[1], V[2]), b = std::max(V[1], V[2]);
    auto vshift = [&](int v) { return v - (v > a ? 1 : 0) - (v > b ? 1 : 0); };
    for (auto &l : m->cv)
        for (int &v : l) v = vshift(v);
    for (int &v : m->vn) v = vshift(v);
    for (int &q : m->vcn)
        if (q > c) q -= 1;
    for (int v : {b, a}) {
        m->v.erase(m->v.begin() + v);
        m->vn.erase(m->vn.begin() + 3 * v, m->vn.begin() + 3 * v + 3);
        m->vcn.erase(m->vcn.begin() + 3 * v, m->vcn.begin() + 3 * v + 3);
    }
    m->cv.erase(m->cv.begin() + c);
    m->APpref.erase(m->APpref.begin() + c);
    m->theta.erase(m->theta.begin() + c);
    m->mot.erase(m->mot.begin() + c);
    m->Nv -= 2;
    m->Nc -= 1;
    m->AP.assign(m->Nc, d2{0, 0});
    m->F.assign(m->Nv, d2{0, 0});
    m->vcur.assign(3 * m->Nv, d2{0, 0});
    m->vlast = m->vcur;
    m->vnext = m->vcur;
    return 0;
}
// reference vertexModelBase::cellDivision (vertexModelBase.cpp:1124-1385): new
// vertices at the midpoints of the edges after list positions p1 < p2 of the cell
// (indices Nv, Nv + 1), the vertices from the first new vertex to the second go
// to a new cell (index Nc) with the parent's preferences and motility and
// director dtheta; the cells across the two edges gain the new vertices.
int vmo_cell_division(void *h, int c, int p1, int p2, double dtheta) {
    VM *m = (VM *)h;
    if (c < 0 || c >= m->Nc) return -1;
    const vector<int> l = m->cv[c];
    const int n = (int)l.size();
    const int i1 = std::min(p1, p2), i2 = std::max(p1, p2);
    if (i1 < 0 || i2 >= n || i1 == i2) return -1;
    const int a = l[i1], an = l[(i1 + 1) % n], b = l[i2], bn = l[(i2 + 1) % n];
    auto mid = [&](int u, int w) {
        d2 d = m->box.minDist(m->v[w], m->v[u]);
        d.x = 0.5 * d.x;
        d.y = 0.5 * d.y;
        d2 q{m->v[u].x + d.x, m->v[u].y + d.y};
        m->box.putInBoxReal(q);
        return q;
    };
    auto third = [&](int u, int w) {
        int ans = -1;
        for (int f = 0; f < 3; ++f)
            for (int g = 0; g < 3; ++g)
                if (m->vcn[3 * u + f] == m->vcn[3 * w + g] && m->vcn[3 * u + f] != c) ans = m->vcn[3 * u + f];
        return ans;
    };
    const int na = third(a, an), nb = third(b, bn);
    if (na < 0 || nb < 0) return -1;
    const int A = m->Nv, B = m->Nv + 1, C = m->Nc;
    m->v.push_back(mid(a, an));
    m->v.push_back(mid(b, bn));
    m->vn.insert(m->vn.end(), {a, an, B, A, b, bn});
    for (int f = 0; f < 3; ++f) {
        if (m->vn[3 * a + f] == an) m->vn[3 * a + f] = A;
        if (m->vn[3 * an + f] == a) m->vn[3 * an + f] = A;
        if (m->vn[3 * b + f] == bn) m->vn[3 * b + f] = B;
        if (m->vn[3 * bn + f] == b) m->vn[3 * bn + f] = B;
    }
    m->vcn.insert(m->vcn.end(), {na, C, c, nb, c, C});
    vector<int> child{A}, parent{B};
    for (int k = i1 + 1; k <= i2; ++k) child.push_back(l[k % n]);
    child.push_back(B);
    for (int k = i2 + 1; k <= i1 + n; ++k) parent.push_back(l[k % n]);
    parent.push_back(A);
    for (size_t k = 1; k + 1 < child.size(); ++k)
        for (int f = 0; f < 3; ++f)
            if (m->vcn[3 * child[k] + f] == c) m->vcn[3 * child[k] + f] = C;
    m->cv[c] = parent;
    m->cv.push_back(child);
    auto insert_after = [&](int q, int after, int v) {
        vector<int> r;
        for (int w : m->cv[q]) {
            r.push_back(w);
            if (w == after) r.push_back(v);
        }
        m->cv[q] = r;
    };
    insert_after(na, an, A);
    insert_after(nb, bn, B);
    m->APpref.push_back(m->APpref[c]);
    m->mot.push_back(m->mot[c]);
    m->theta.push_back(dtheta);
    m->Nv += 2;
    m->Nc += 1;
    m->AP.assign(m->Nc, d2{0, 0});
    m->F.assign(m->Nv, d2{0, 0});
    m->vcur.assign(3 * m->Nv, d2{0, 0});
    m->vlast = m->vcur;
    m->vnext = m->vcur;
    return 0;
}
int vmo_nc(void *h) { return ((VM *)h)->Nc; }
void vmo_set_motility(void *h, const double *v0, const double *Dr) {
    VM *m = (VM *)h;
    for (int c = 0; c < m->Nc; ++c) m->mot[c] = d2{v0[c], Dr[c]};
}
void vmo_set_directors(void *h, const double *th) {
    VM *m = (VM *)h;
    for (int c = 0; c < m->Nc; ++c) m->theta[c] = th[c];
}
void vmo_get_directors(void *h, double *th) {
    VM *m = (VM *)h;
    for (int c = 0; c < m->Nc; ++c) th[c] = m->theta[c];
}
// one self-propelled cell-vertex step (reference selfPropelledCellVertexDynamics
// GPU branch, .cpp:51-76, .cu:21-71): forces, vertex displacement
// dt (mean over the three cells of v0 e(theta) + mu F), then the director
// rotation theta += eta sqrt(2 dt Dr), the move and the T1 pass; returns the
// number of T1 flips
int vmo_spp_step(void *h, const double *eta_cell) {
    VM *m = (VM *)h;
    geometry(*m);
    forces(*m);
    vector<d2> disp(m->Nv);
    for (int i = 0; i < m->Nv; ++i) {
        int c1 = m->vcn[3 * i], c2 = m->vcn[3 * i + 1], c3 = m->vcn[3 * i + 2];
        double v1 = m->mot[c1].x, v2 = m->mot[c2].x, v3 = m->mot[c3].x;
        double dirx = (v1 * std::cos(m->theta[c1]) + v2 * std::cos(m->theta[c2]) + v3 * std::cos(m->theta[c3])) / 3.0;
        double diry = (v1 * std::sin(m->theta[c1]) + v2 * std::sin(m->theta[c2]) + v3 * std::sin(m->theta[c3])) / 3.0;
        disp[i].x = m->dt * (dirx + m->mu * m->F[i].x);
        disp[i].y = m->dt * (diry + m->mu * m->F[i].y);
    }
    for (int c = 0; c < m->Nc; ++c) m->theta[c] += eta_cell[c] * std::sqrt(2.0 * m->dt * m->mot[c].y);
    long long before = m->t1count;
    for (int i = 0; i < m->Nv; ++i) {
        m->v[i].x += disp[i].x;
        m->v[i].y += disp[i].y;
        m->box.putInBoxReal(m->v[i]);
    }
    t1(*m);
    return (int)(m->t1count - before);
}
// FIRE (GPU branch, fire_oracle.h) with the vertices as degrees of freedom and
// the T1 pass as the topology step (vertexModelBase.cpp:93-106)
struct VertexFireModel {
    VM *m;
    int dof() { return m->Nv; }
    void forces(vector<d2> &F) {
        geometry(*m);
        ::forces(*m);
        F = m->F;
    }
    void move(const vector<d2> &d) {
        for (int i = 0; i < m->Nv; ++i) {
            m->v[i].x += d[i].x;
            m->v[i].y += d[i].y;
            m->box.putInBoxReal(m->v[i]);
        }
    }
    bool topology() {
        t1(*m);
        return true;
    }
};
void vmo_fire_set(void *h, int maxIt, double fc, double alphaStart, double dt, double dtMax, double dtInc,
                  double dtDec, double alphaDec, int NMin) {
    ((VM *)h)->fire.set(maxIt, fc, alphaStart, dt, dtMax, dtInc, dtDec, alphaDec, NMin);
}
// returns the iterations done; out as fireo::minimize
int vmo_fire_minimize(void *h, double *out) {
    VM *m = (VM *)h;
    VertexFireModel model{m};
    return fireo::minimize(m->fire, model, out);
}

}
