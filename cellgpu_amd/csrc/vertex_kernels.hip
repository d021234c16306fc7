// HIP kernels of the vertex-model time step for gfx950 (MI355X).
//
// Topology convention kept by every kernel here: the three cells of a vertex
// are stored counter-clockwise around it, and vn[3v+f] is the neighbour vertex
// across from cell vcn[3v+f]; then in cell f the polygon predecessor of v is
// vn[3v+(f+2)%3] and its successor vn[3v+(f+1)%3], so the per-vertex force needs
// no search of the cell's vertex list. Cell vertex lists are counter-clockwise,
// slot-major (cv[k*Nc + c]).
//
// Step: geometry (per cell) -> force + equation of motion (Brownian, or
// self-propelled cell-vertex with the director rotation in the next pass) + move (per vertex) ->
// FP64 T1 edge test (per vertex) -> T1 flips in deterministic rounds (one
// work-group per system: an edge (v1, v2) flips when its key is the smallest
// among the flagged edges sharing one of its four cells).
#include <hip/hip_runtime.h>

#include <climits>

#include "../../include/cellgpu_amd.h"
#include "cgpu_device.h"
#include "fire_device.h"

using namespace cgpu;

namespace {

// cell-list slots per cell (slot-major storage), at least the register copy
__host__ __device__ inline int vertex_cap(const cgpu_vertex_batch &b) {
    return b.vertex_cap <= 0 ? CGPU_VERTEX_CAP : b.vertex_cap;
}
// A system that needs more cell-list slots stalls: it stops advancing (this
// step's flips are not made and its counter not advanced) until the host grows
// the storage and finishes the step (reference growCellVerticesList before the
// flips, vertexModelBase.cpp:475-500, 815-824). The T1 test only records the
// need (the director rotation of the same kernel reads the activity flag); the
// flip kernel parks the system.
__device__ inline void vm_need_room(const cgpu_vertex_batch &b, int s) {
    if (b.st[s].error == 0) atomicCAS(&b.st[s].error, 0, (int)CGPU_ERR_RING_CAPACITY);
}

__device__ inline dvec2 ldp(const double *p, int i) {
    double2 v = reinterpret_cast<const double2 *>(p)[i];
    return dvec2{v.x, v.y};
}
__device__ inline void stp(double *p, int i, dvec2 v) { reinterpret_cast<double2 *>(p)[i] = make_double2(v.x, v.y); }
template <typename R> __device__ inline vec2<R> ldr(const void *a, int i) {
    const R *p = reinterpret_cast<const R *>(a);
    return vec2<R>{p[2 * i], p[2 * i + 1]};
}
template <typename R> __device__ inline void str(void *a, int i, vec2<R> v) {
    R *p = reinterpret_cast<R *>(a);
    p[2 * i] = v.x;
    p[2 * i + 1] = v.y;
}
template <typename R> __device__ inline vec2<R> cv2(dvec2 d) { return vec2<R>{(R)d.x, (R)d.y}; }

// reference vm_geometry_kernel (vertexModelBase.cu:20-77): signed polygon area
// and perimeter relative to the cell's vertex n-2
template <typename R>
__global__ void k_vm_geometry(cgpu_vertex_batch b, const double *pos) {
    int c = xcd_block() * blockDim.x + threadIdx.x;
    if (c >= b.num_cells) return;
    const cgpu_vm_params &sp = b.sys[b.cell_sys[c]];
    const Box box = make_box(sp.L);
    const int Nc = b.num_cells;
    const int n = b.cvn[c];
    if (n < 3) return;
    dvec2 cellPos = ldp(pos, b.cv[(n - 2) * Nc + c]);
    vec2<R> vcur = cv2<R>(min_dist(box, ldp(pos, b.cv[(n - 1) * Nc + c]), cellPos));
    R A = R(0), P = R(0);
    for (int k = 0; k < n; ++k) {
        vec2<R> vnext = cv2<R>(min_dist(box, ldp(pos, b.cv[k * Nc + c]), cellPos));
        A += R(0.5) * (vcur.x + vnext.x) * (vnext.y - vcur.y);
        R dx = vcur.x - vnext.x, dy = vcur.y - vnext.y;
        P += sqrt(dx * dx + dy * dy);
        vcur = vnext;
    }
    str<R>(b.AP, c, vec2<R>{A, P});
    vec2<R> pref = ldr<R>(b.APpref, c);
    str<R>(b.dAP, c, vec2<R>{(R)sp.KA * (A - pref.x), (R)sp.KP * (P - pref.y)});
}

// force on vertex v at p: sum over its three cells of the reference force set
// (functions.h:330-354, vertexQuadraticEnergy.cu:21-70)
template <typename R>
__device__ __forceinline__ vec2<R> vm_force(const cgpu_vertex_batch &b, const double *pos, int v, const Box &box,
                                            dvec2 p) {
    vec2<R> nb[3];
    for (int f = 0; f < 3; ++f) nb[f] = cv2<R>(min_dist(box, ldp(pos, b.vn[3 * v + f]), p));
    vec2<R> F{R(0), R(0)};
    for (int f = 0; f < 3; ++f) {
        vec2<R> q = ldr<R>(b.dAP, b.vcn[3 * v + f]);  // KA(A-A0), KP(P-P0)
        const vec2<R> vl = nb[(f + 2) % 3], vn = nb[(f + 1) % 3];
        vec2<R> dAdv{R(0.5) * (vl.y - vn.y), R(-0.5) * (vl.x - vn.x)};
        vec2<R> dl{vl.x, vl.y};
        R dln = sqrt(dl.x * dl.x + dl.y * dl.y);
        vec2<R> dn{-vn.x, -vn.y};
        R dnn = sqrt(dn.x * dn.x + dn.y * dn.y);
        vec2<R> dPdv{dl.x / dln - dn.x / dnn, dl.y / dln - dn.y / dnn};
        F.x += R(2.0) * (q.x * dAdv.x + q.y * dPdv.x);
        F.y += R(2.0) * (q.x * dAdv.y + q.y * dPdv.y);
    }
    return F;
}

// moved position p + d, wrapped; a non-finite or runaway result flags the
// system and keeps the old position
__device__ inline dvec2 vm_moved(cgpu_sys_state &st, const cgpu_vm_params &sp, const Box &box, dvec2 p, double dx,
                                 double dy) {
    dvec2 q{p.x + dx, p.y + dy};
    if (!(fabs(q.x) < 1e12 * sp.L && fabs(q.y) < 1e12 * sp.L)) {
        st.error = CGPU_ERR_NONFINITE;
        q = p;
    }
    put_in_box(box, q);
    return q;
}

// Per vertex: force, then (move != 0) the equation of motion and the move.
// KIND 0: Brownian, disp = sqrt(2 dt T mu) eta + dt mu F (brownianParticleDynamics.cu:20-39).
// KIND 1: self-propelled cell-vertex, disp = dt (director + mu F) with the
// director the mean of v0 e(theta) over the vertex's three cells
// (selfPropelledCellVertexDynamics.cu:21-47).
template <typename R, int KIND>
__global__ void k_vm_force_eom_move(cgpu_vertex_batch b, const double *pos, double *pos_next, int move) {
    int v = xcd_block() * blockDim.x + threadIdx.x;
    if (v >= b.num_vertices) return;
    const int s = b.vert_sys[v];
    const cgpu_vm_params &sp = b.sys[s];
    cgpu_sys_state &st = b.st[s];
    const Box box = make_box(sp.L);
    dvec2 p = ldp(pos, v);
    if (move && !st.active) {
        stp(pos_next, v, p);
        return;
    }
    vec2<R> F = vm_force<R>(b, pos, v, box, p);
    str<R>(b.force, v, F);
    if (!move) return;
    const R dt = (R)sp.dt, mu = (R)sp.mu;
    R dx, dy;
    if (KIND == 0) {
        const int step = st.step + 1;
        double z0, z1;
        if (b.noise) {
            z0 = b.noise[v];
            z1 = b.noise[b.num_vertices + v];
        } else {
            normal_pair(sp.seed, (uint32_t)(v - sp.voff), (uint32_t)step, 1u, z0, z1);
        }
        R sc = sqrt(dt * (R)sp.T * mu * R(2.0));
        dx = (R)z0 * sc + dt * mu * F.x;
        dy = (R)z1 * sc + dt * mu * F.y;
    } else {
        const R *th = reinterpret_cast<const R *>(b.theta);
        const R *mot = reinterpret_cast<const R *>(b.motility);
        R dirx = R(0), diry = R(0);
        for (int f = 0; f < 3; ++f) {
            const int c = b.vcn[3 * v + f];
            const R v0 = mot[2 * c], t = th[c];
            dirx += v0 * cos(t);
            diry += v0 * sin(t);
        }
        dirx /= R(3.0);
        diry /= R(3.0);
        dx = dt * (dirx + mu * F.x);
        dy = dt * (diry + mu * F.y);
    }
    stp(pos_next, v, vm_moved(st, sp, box, p, (double)dx, (double)dy));
}

// director rotation after the displacements (selfPropelledCellVertexDynamics.cu:52-66):
// theta += eta sqrt(2 dt Dr), eta from the cell noise block or the counter-based
// generator keyed by (cell, step, stream 2)
template <typename R> __device__ inline void rotate_director(const cgpu_vertex_batch &b, int c) {
    const int s = b.cell_sys[c];
    const cgpu_sys_state &st = b.st[s];
    if (!st.active) return;
    const cgpu_vm_params &sp = b.sys[s];
    double z0, z1;
    if (b.noise)
        z0 = b.noise[2 * (size_t)b.num_vertices + c];
    else
        normal_pair(sp.seed, (uint32_t)(c - sp.coff), (uint32_t)(st.step + 1), 2u, z0, z1);
    R *th = reinterpret_cast<R *>(b.theta);
    const R Dr = reinterpret_cast<const R *>(b.motility)[2 * c + 1];
    th[c] += (R)z0 * sqrt(R(2.0) * (R)sp.dt * Dr);
}

// reference vm_simple_T1_test_kernel (vertexModelBase.cu:118-164): FP64 edge
// length below the threshold, for v1 < v2; flagged edges go to the system's list
// as (v1, v2) pairs
// rotate != 0: thread c < Nc also rotates cell c's director (the T1 pass of a
// self-propelled cell-vertex step follows its displacement kernel)
__global__ void k_vm_t1_test(cgpu_vertex_batch b, const double *pos, int rotate) {
    int v = xcd_block() * blockDim.x + threadIdx.x;
    if (rotate && v < b.num_cells) {
        if (b.precision)
            rotate_director<float>(b, v);
        else
            rotate_director<double>(b, v);
    }
    if (v >= b.num_vertices) return;
    const int s = b.vert_sys[v];
    if (!b.st[s].active) return;
    const cgpu_vm_params &sp = b.sys[s];
    const Box box = make_box(sp.L);
    dvec2 p = ldp(pos, v);
    for (int f = 0; f < 3; ++f) {
        int v2 = b.vn[3 * v + f];
        if (v < v2) {
            dvec2 e = min_dist(box, p, ldp(pos, v2));
            if (sqrt(e.x * e.x + e.y * e.y) < sp.thr) {
                // a flip adds a vertex to a cell touching the edge: every one of
                // them must have a free slot (reference vm_simple_T1_test_kernel
                // grow flag, vertexModelBase.cu:150-160)
                const int cap = vertex_cap(b);
                for (int g = 0; g < 3; ++g)
                    if (b.cvn[b.vcn[3 * v + g]] >= cap || b.cvn[b.vcn[3 * v2 + g]] >= cap) vm_need_room(b, s);
                int q = atomicAdd(&b.flag_count[s], 1);
                b.flag_list[2 * (3 * sp.voff + q)] = v;
                b.flag_list[2 * (3 * sp.voff + q) + 1] = v2;
            }
        }
    }
}

// A cell's vertex list loaded into registers in one batch of independent loads
// (lists of up to CL vertices; longer lists take the memory loops), so a flip
// is a few round trips to memory instead of one per list entry.
constexpr int CL = 12;
static_assert(CL <= CGPU_VERTEX_CAP, "register copy longer than a cell list");
struct CellRegs {
    int c, n;
    int v[CL];
};
__device__ inline void load_cell(const cgpu_vertex_batch &b, int c, CellRegs &r) {
    const int Nc = b.num_cells;
    r.c = c;
    r.n = b.cvn[c];
    // slots 0 .. CL - 1 exist for every cell (capacity CGPU_VERTEX_CAP), so the
    // loads need not wait for the count
#pragma unroll
    for (int k = 0; k < CL; ++k) r.v[k] = b.cv[k * Nc + c];
#pragma unroll
    for (int k = 0; k < CL; ++k)
        if (k >= r.n) r.v[k] = -1;
}
__device__ inline int reg_at(const CellRegs &r, int idx) {
    int out = -1;
#pragma unroll
    for (int k = 0; k < CL; ++k)
        if (k == idx) out = r.v[k];
    return out;
}
__device__ inline int reg_find(const CellRegs &r, int x) {
    int at = -1;
#pragma unroll
    for (int k = 0; k < CL; ++k)
        if (k < r.n && r.v[k] == x && at < 0) at = k;
    return at;
}

// position of x in cell c's list: predecessor and successor (reference loop)
__device__ inline void locate_mem(const cgpu_vertex_batch &b, int c, int x, int &pred, int &succ) {
    const int Nc = b.num_cells;
    const int n = b.cvn[c];
    int vlast = b.cv[(n - 2) * Nc + c], vcur = b.cv[(n - 1) * Nc + c], vnext = b.cv[c];
    for (int k = 0; k < n; ++k) {
        vnext = b.cv[k * Nc + c];
        if (vcur == x) break;
        vlast = vcur;
        vcur = vnext;
    }
    pred = vlast;
    succ = vnext;
}
// the same on a register copy (identical result: the loop above stops at the
// first k with vcur == x, i.e. the first occurrence scanning from slot n-1)
__device__ inline void locate(const cgpu_vertex_batch &b, const CellRegs &r, int x, int &pred, int &succ) {
    if (r.n > CL) {
        locate_mem(b, r.c, x, pred, succ);
        return;
    }
    const int n = r.n;
    int at = (reg_at(r, n - 1) == x) ? n - 1 : -1;
    if (at < 0) {
        const int f = reg_find(r, x);
        at = (f >= 0 && f < n - 1) ? f : -1;
    }
    if (at < 0) {  // x absent: the reference loop ends with (v[n-2], v[n-1]) -> pred v[n-2], succ v[n-1]
        pred = reg_at(r, n - 2);
        succ = reg_at(r, n - 1);
        return;
    }
    pred = reg_at(r, at == 0 ? n - 1 : at - 1);
    succ = reg_at(r, at == n - 1 ? 0 : at + 1);
}
__device__ inline bool cell_has(const cgpu_vertex_batch &b, const CellRegs &r, int x) {
    if (r.n > CL) {
        const int Nc = b.num_cells;
        for (int k = 0; k < r.n; ++k)
            if (b.cv[k * Nc + r.c] == x) return true;
        return false;
    }
    return reg_find(r, x) >= 0;
}
__device__ inline void edge_cells(const cgpu_vertex_batch &b, int v1, int v2, int C[4]) {
    C[0] = b.vcn[3 * v1];
    C[1] = b.vcn[3 * v1 + 1];
    C[2] = b.vcn[3 * v1 + 2];
    C[3] = -1;
    for (int f = 0; f < 3; ++f) {
        int c = b.vcn[3 * v2 + f];
        if (c != C[0] && c != C[1] && c != C[2]) C[3] = c;
    }
}
// cell list without x (reference compaction), written back
__device__ inline void remove_vertex(const cgpu_vertex_batch &b, const CellRegs &r, int x) {
    const int Nc = b.num_cells;
    const int c = r.c, n = r.n;
    if (n > CL) {
        int w = 0;
        for (int k = 0; k < n; ++k) {
            const int val = b.cv[k * Nc + c];
            if (val != x) b.cv[(w++) * Nc + c] = val;
        }
    } else {
        int w = 0;
#pragma unroll
        for (int k = 0; k < CL; ++k)
            if (k < n && r.v[k] != x) b.cv[(w++) * Nc + c] = r.v[k];
    }
    b.cvn[c] = n - 1;
}
// cell list with x inserted before the last occurrence of `before` (reference
// loop: at = last k with v[k] == before, 0 if absent), written back
__device__ inline void insert_vertex(const cgpu_vertex_batch &b, const CellRegs &r, int x, int before) {
    const int Nc = b.num_cells;
    const int c = r.c, n = r.n;
    if (n > CL) {
        int at = 0;
        for (int k = 0; k < n; ++k)
            if (b.cv[k * Nc + c] == before) at = k;
        for (int k = n; k > at; --k) b.cv[k * Nc + c] = b.cv[(k - 1) * Nc + c];
        b.cv[at * Nc + c] = x;
    } else {
        int at = 0;
#pragma unroll
        for (int k = 0; k < CL; ++k)
            if (k < n && r.v[k] == before) at = k;
#pragma unroll
        for (int k = 0; k <= CL; ++k) {
            if (k > n) break;
            const int val = (k < at) ? reg_at(r, k) : (k == at ? x : reg_at(r, k - 1));
            b.cv[k * Nc + c] = val;
        }
    }
    b.cvn[c] = n + 1;
}

// the T1 of one edge (reference vm_flip_edges_kernel, vertexModelBase.cu:252-544);
// returns the four vertices whose slots need re-ordering, or false if forbidden
__device__ bool flip_edge(const cgpu_vertex_batch &b, double *pos, const cgpu_vm_params &sp, int v1, int v2,
                          const int C[4], int touched[4]) {
    if (C[0] < 0 || C[1] < 0 || C[2] < 0 || C[3] < 0) return false;
    // the edge's positions and slot triples are fetched with the four cell lists
    dvec2 p1 = ldp(pos, v1), p2 = ldp(pos, v2);
    int c1[3], c2[3], n1[3], n2[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        c1[q] = b.vcn[3 * v1 + q];
        c2[q] = b.vcn[3 * v2 + q];
        n1[q] = b.vn[3 * v1 + q];
        n2[q] = b.vn[3 * v2 + q];
    }
    CellRegs R0, R1, R2, R3;
    load_cell(b, C[0], R0);
    load_cell(b, C[1], R1);
    load_cell(b, C[2], R2);
    load_cell(b, C[3], R3);
    // i has v1, v2 with v2 before v1; k has v1 before v2; j has v1 only; l = C[3]
    int si = -1, sj = -1, sk = -1;
    {
        int pr, su;
        locate(b, R0, v1, pr, su);
        if (pr == v2) si = 0; else if (su == v2) sk = 0; else sj = 0;
        locate(b, R1, v1, pr, su);
        if (pr == v2) si = 1; else if (su == v2) sk = 1; else sj = 1;
        locate(b, R2, v1, pr, su);
        if (pr == v2) si = 2; else if (su == v2) sk = 2; else sj = 2;
    }
    if (si < 0 || sj < 0 || sk < 0) return false;
    const CellRegs &Ri = si == 0 ? R0 : (si == 1 ? R1 : R2);
    const CellRegs &Rj = sj == 0 ? R0 : (sj == 1 ? R1 : R2);
    const CellRegs &Rk = sk == 0 ? R0 : (sk == 1 ? R1 : R2);
    const CellRegs &Rl = R3;
    if (Ri.n == 3 || Rk.n == 3) return false;  // would shrink a triangle
    if (Rj.n >= vertex_cap(b) || Rl.n >= vertex_cap(b)) {
        vm_need_room(b, b.vert_sys[v1]);
        b.st[b.vert_sys[v1]].active = 0;
        return false;
    }
    int dummy, a, bb;
    locate(b, Rj, v1, dummy, a);
    locate(b, Rl, v2, dummy, bb);
    int na[3], nbb[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        na[q] = b.vn[3 * a + q];
        nbb[q] = b.vn[3 * bb + q];
    }
    const Box box = make_box(sp.L);
    dvec2 e = min_dist(box, p1, p2);
    dvec2 mid{p2.x + 0.5 * e.x, p2.y + 0.5 * e.y};
    p1 = dvec2{mid.x - e.y, mid.y + e.x};
    p2 = dvec2{mid.x + e.y, mid.y - e.x};
    put_in_box(box, p1);
    put_in_box(box, p2);
    stp(pos, v1, p1);
    stp(pos, v2, p2);
    const int ci = Ri.c, cj = Rj.c, ck = Rk.c, cl = Rl.c;
    // the reference's rewiring, slot by slot, on the register copies. v1, v2, a
    // and bb are distinct except a == bb in a degenerate tiling; then both of
    // their rules act on one triple, in this order.
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        if (c1[q] == ck) c1[q] = cl;
        if (c2[q] == ci) c2[q] = cj;
        if (n1[q] == a) n1[q] = bb;
        if (n2[q] == bb) n2[q] = a;
        if (na[q] == v1) na[q] = v2;
        if (a == bb) {
            if (na[q] == v2) na[q] = v1;
        } else if (nbb[q] == v2) {
            nbb[q] = v1;
        }
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        b.vcn[3 * v1 + q] = c1[q];
        b.vcn[3 * v2 + q] = c2[q];
        b.vn[3 * v1 + q] = n1[q];
        b.vn[3 * v2 + q] = n2[q];
        b.vn[3 * a + q] = na[q];
        if (a != bb) b.vn[3 * bb + q] = nbb[q];
    }
    // cell i loses v2, cell k loses v1; cell j gains v2 before a, cell l gains v1 before b
    remove_vertex(b, Ri, v2);
    remove_vertex(b, Rk, v1);
    insert_vertex(b, Rj, v2, a);
    insert_vertex(b, Rl, v1, bb);
    touched[0] = v1;
    touched[1] = v2;
    touched[2] = a;
    touched[3] = bb;
    return true;
}

// restore the slot convention of vertex w from the cell lists: cells
// counter-clockwise (the next cell shares the edge to w's predecessor in the
// current one) and each neighbour across from the cell that does not contain it
__device__ void normalize_vertex(const cgpu_vertex_batch &b, int w) {
    const int c[3] = {b.vcn[3 * w], b.vcn[3 * w + 1], b.vcn[3 * w + 2]};
    const int n[3] = {b.vn[3 * w], b.vn[3 * w + 1], b.vn[3 * w + 2]};
    CellRegs R0, R1, R2;
    load_cell(b, c[0], R0);
    load_cell(b, c[1], R1);
    load_cell(b, c[2], R2);
    int pr, su;
    locate(b, R0, w, pr, su);
    const bool first = cell_has(b, R1, pr);
    const CellRegs &Rc1 = first ? R1 : R2;
    const CellRegs &Rc2 = first ? R2 : R1;
    const int cc[3] = {c[0], Rc1.c, Rc2.c};
    for (int f = 0; f < 3; ++f) {
        const CellRegs &Rf = f == 0 ? R0 : (f == 1 ? Rc1 : Rc2);
        b.vcn[3 * w + f] = cc[f];
        int opp = n[0];
        for (int q = 0; q < 3; ++q)
            if (!cell_has(b, Rf, n[q])) opp = n[q];
        b.vn[3 * w + f] = opp;
    }
}

// One work-group per system: T1 flips in rounds. An edge flips when its key
// (v1, v2) is the smallest among the flagged edges sharing one of its four cells;
// flips of one round touch disjoint cells and vertices. Edges removed by a
// neighbouring flip are dropped. `list` holds the waiting edges, `cells` four
// and `touched` four ints per edge: LDS while the edges fit, device memory
// otherwise (the same code on either).
template <typename Count>
__device__ __forceinline__ void t1_rounds_loop(const cgpu_vertex_batch &b, double *pos, const cgpu_vm_params &sp,
                                               int *list, int *cells, int *touched, Count &count, int &rounds,
                                               int &flips, int *wave_kept) {
    const unsigned long long NONE = ~0ull;
    __shared__ int waiting;  // some edge lost this round and waits for the next
    while (count > 0 && rounds < 100000) {
        const int nq = count;
        if (threadIdx.x == 0) waiting = 0;
        for (int q = threadIdx.x; q < nq; q += blockDim.x) {
            int v1 = list[2 * q], v2 = list[2 * q + 1];
            // the edge test and the cells of its two vertices are loaded together
            // (every vertex has its three slots)
            bool live = b.vn[3 * v1] == v2 || b.vn[3 * v1 + 1] == v2 || b.vn[3 * v1 + 2] == v2;
            int C[4];
            edge_cells(b, v1, v2, C);
            if (!live) {
                C[0] = C[1] = C[2] = C[3] = -1;
                list[2 * q] = -2;  // dropped
            }
            unsigned long long k = ((unsigned long long)(unsigned)v1 << 32) | (unsigned)v2;
            reinterpret_cast<int4 *>(cells)[q] = make_int4(C[0], C[1], C[2], C[3]);
            for (int t = 0; t < 4; ++t)
                if (C[t] >= 0) atomicMin(&b.cell_min[C[t]], k);
        }
        __syncthreads();
        for (int q = threadIdx.x; q < nq; q += blockDim.x) {
            touched[4 * q] = -1;
            int v1 = list[2 * q], v2 = list[2 * q + 1];
            if (v1 < 0) continue;
            unsigned long long k = ((unsigned long long)(unsigned)v1 << 32) | (unsigned)v2;
            const int4 c4 = reinterpret_cast<const int4 *>(cells)[q];
            int C[4] = {c4.x, c4.y, c4.z, c4.w};
            // the four minima are loaded together (slot 0 stands in for an absent
            // cell), then combined without branches
            unsigned long long m[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) m[t] = b.cell_min[max(C[t], 0)];
            bool win = true;
#pragma unroll
            for (int t = 0; t < 4; ++t) win = win & ((C[t] < 0) | (m[t] == k));
            if (win) {
                if (flip_edge(b, pos, sp, v1, v2, C, &touched[4 * q])) atomicAdd(&flips, 1);
                list[2 * q] = -1;
            } else {
                waiting = 1;
            }
        }
        __syncthreads();
        for (int q = threadIdx.x; q < nq; q += blockDim.x) {
            const int4 c4 = reinterpret_cast<const int4 *>(cells)[q];
            const int C[4] = {c4.x, c4.y, c4.z, c4.w};
#pragma unroll
            for (int t = 0; t < 4; ++t)
                if (C[t] >= 0) b.cell_min[C[t]] = NONE;
        }
        // the four touched vertices of every flip, one thread each
        for (int e = threadIdx.x; e < 4 * nq; e += blockDim.x) {
            const int q = e >> 2;
            if (list[2 * q] == -1 && touched[4 * q] >= 0) normalize_vertex(b, touched[4 * q + (e & 3)]);
        }
        __syncthreads();
        if (!waiting) {  // every edge flipped or was dropped: the last round
            if (threadIdx.x == 0) {
                count = 0;
                rounds += 1;
            }
            __syncthreads();
            continue;
        }
        // the edges still waiting, in list order, staged in `cells` (free until
        // the next round): ballot prefix counts per wavefront, blockDim.x edges
        // at a time
        const int nwaves = blockDim.x >> 6, wv = threadIdx.x >> 6, lane = __lane_id();
        int kept = 0;
        for (int q0 = 0; q0 < nq; q0 += blockDim.x) {
            const int q = q0 + threadIdx.x;
            const int v1 = (q < nq) ? list[2 * q] : -1;
            const int v2 = (q < nq) ? list[2 * q + 1] : -1;
            const bool keep = v1 >= 0;
            const unsigned long long m = __ballot(keep);
            if (lane == 0) wave_kept[wv] = __popcll(m);
            __syncthreads();
            int at = kept + __popcll(m & ((1ull << lane) - 1ull));
            int total = 0;
            for (int k = 0; k < nwaves; ++k) {
                if (k < wv) at += wave_kept[k];
                total += wave_kept[k];
            }
            if (keep) {
                cells[2 * at] = v1;
                cells[2 * at + 1] = v2;
            }
            kept += total;
            __syncthreads();
        }
        for (int q = threadIdx.x; q < 2 * kept; q += blockDim.x) list[q] = cells[q];
        if (threadIdx.x == 0) {
            count = kept;
            rounds += 1;
        }
        __syncthreads();
    }
}

// The T1 flips of every system, one work-group each (t1_rounds_loop). With
// advance set, the work-group then advances its system's step counter (the last
// kernel of a time step).
constexpr int T1_LDS = 512;  // waiting edges kept in LDS
__global__ __launch_bounds__(256) void k_vm_t1_rounds(cgpu_vertex_batch b, double *pos, int advance) {
    const int s = blockIdx.x;
    const cgpu_vm_params &sp = b.sys[s];
    cgpu_sys_state &st = b.st[s];
    int *list = b.flag_list + 2 * 3 * sp.voff;
    __shared__ int count, rounds, flips;
    __shared__ int wave_kept[16];
    __shared__ int s_list[2 * T1_LDS];
    __shared__ int4 s_cells[T1_LDS], s_touched[T1_LDS];
    if (threadIdx.x == 0) {
        // a system that needs longer cell lists stalls here; a stalled (or
        // parked) system makes no flips
        if (st.active && st.error == CGPU_ERR_RING_CAPACITY) st.active = 0;
        count = st.active ? b.flag_count[s] : 0;
        rounds = 0;
        flips = 0;
    }
    __syncthreads();
    const int n0 = count;
    if (n0 <= T1_LDS) {
        // the list only shrinks: it stays in LDS
        for (int q = threadIdx.x; q < 2 * n0; q += blockDim.x) s_list[q] = list[q];
        __syncthreads();
        t1_rounds_loop(b, pos, sp, s_list, reinterpret_cast<int *>(s_cells), reinterpret_cast<int *>(s_touched),
                       count, rounds, flips, wave_kept);
    } else {
        t1_rounds_loop(b, pos, sp, list, b.flag_cells + 12 * sp.voff, b.flag_touched + 12 * sp.voff, count, rounds,
                       flips, wave_kept);
    }
    if (threadIdx.x == 0) {
        st.t1_events += flips;
        b.flag_count[s] = 0;
        if (advance && st.active) {
            st.step += 1;
            if (st.steps_left > 0 && --st.steps_left == 0) st.active = 0;
        }
    }
}

// ------------------------------------------------------------------------ FIRE
// FIRE with the vertices as degrees of freedom (shared pieces: fire_device.h)
template <typename R> __global__ void k_vm_fire_vv_move(cgpu_vertex_batch b, const cgpu_fire_state *fs, void *vel) {
    int v = xcd_block() * blockDim.x + threadIdx.x;
    if (v >= b.num_vertices) return;
    const int s = b.vert_sys[v];
    const cgpu_vm_params &sp = b.sys[s];
    const cgpu_fire_state &f = fs[s];
    dvec2 p = ldp(b.pos, v);
    vec2<R> u = ldr<R>(vel, v), F = ldr<R>(b.force, v);
    fire_mix<R>(u, F, f);
    if (f.converged || !b.st[s].active) {
        str<R>(vel, v, u);
        stp(b.pos_next, v, p);
        return;
    }
    vec2<R> d = fire_vv_first_half<R>(u, F, (R)f.dt);
    str<R>(vel, v, u);
    stp(b.pos_next, v, vm_moved(b.st[s], sp, make_box(sp.L), p, (double)d.x, (double)d.y));
}

__global__ void k_vm_fire_clear_pending(cgpu_fire_state *fs, int S) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < S) fs[s].pending = 0;
}

// forces on the moved (and T1-flipped) vertices, second velocity half step,
// per-vertex products of the reductions; clear: also clears every system's
// pending mix (k_vm_fire_clear_pending), which only the move kernel reads
template <typename R>
__global__ void k_vm_fire_force_vv2(cgpu_vertex_batch b, cgpu_fire_state *fs, void *vel, double *dots, int clear) {
    if (clear)
        for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < b.num_systems; q += gridDim.x * blockDim.x)
            fs[q].pending = 0;
    int v = xcd_block() * blockDim.x + threadIdx.x;
    if (v >= b.num_vertices) return;
    const int s = b.vert_sys[v];
    const cgpu_fire_state &f = fs[s];
    if (f.converged || !b.st[s].active) return;
    const cgpu_vm_params &sp = b.sys[s];
    vec2<R> F = vm_force<R>(b, b.pos, v, make_box(sp.L), ldp(b.pos, v));
    str<R>(b.force, v, F);
    vec2<R> u = ldr<R>(vel, v);
    const R dt = (R)f.dt;
    u.x += R(0.5) * dt * F.x;
    u.y += R(0.5) * dt * F.y;
    str<R>(vel, v, u);
    fire_dots(dots, b.num_vertices, v, F.x, F.y, u.x, u.y);
}

__global__ void k_vm_fire_partial(cgpu_vertex_batch b, const cgpu_fire_state *fs, double *dots) {
    const int c = blockIdx.x;
    const int s = b.chunk_sys[c];
    const int N = b.num_vertices;
    fire_chunk_partial(dots, N, b.chunk_begin[c], b.chunk_end[c], fs[s].converged || !b.st[s].active,
                       dots + 3 * (size_t)N + 4 * (size_t)c);
}

__global__ void k_vm_fire_reduce(cgpu_vertex_batch b, cgpu_fire_state *fs, const double *dots) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= b.num_systems) return;
    cgpu_fire_state &f = fs[s];
    if (f.converged || !b.st[s].active) return;
    const cgpu_vm_params &sp = b.sys[s];
    const double *part = dots + 3 * (size_t)b.num_vertices;
    double forceNorm = 0, power = 0, velNorm = 0, mf = 0;
    for (int c = sp.chunk0; c < b.num_chunks && b.chunk_sys[c] == s; ++c) {
        forceNorm += part[4 * c];
        power += part[4 * c + 1];
        velNorm += part[4 * c + 2];
        mf = fmax(mf, part[4 * c + 3]);
    }
    fire_update(f, forceNorm, power, velNorm, mf, sp.nv);
}

__global__ void k_vm_fire_park(cgpu_vertex_batch b, const cgpu_fire_state *fs) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < b.num_systems && fs[s].converged) b.st[s].active = 0;
}

template <typename R> __global__ void k_vm_fire_apply_mix(cgpu_vertex_batch b, const cgpu_fire_state *fs, void *vel) {
    int v = xcd_block() * blockDim.x + threadIdx.x;
    if (v >= b.num_vertices) return;
    const cgpu_fire_state &f = fs[b.vert_sys[v]];
    vec2<R> u = ldr<R>(vel, v), F = ldr<R>(b.force, v);
    fire_mix<R>(u, F, f);
    str<R>(vel, v, u);
}

// ---------------------------------------------------------------- cell events
// Death (T2) and division of vertex-model cells (reference vertexModelBase::
// cellDeath / cellDivision, vertexModelBase.cpp:905-1385), one record per system.

// Longest cell list the surgery handles in one thread's stack.
constexpr int EV_LIST = 256;

__device__ inline int list_at(const cgpu_vertex_batch &b, int c, int k) { return b.cv[(size_t)k * b.num_cells + c]; }

// the cell other than c that holds both u and w (-1: none)
__device__ inline int third_cell(const cgpu_vertex_batch &b, int u, int w, int c) {
    int ans = -1;
    for (int f = 0; f < 3; ++f)
        for (int g = 0; g < 3; ++g)
            if (b.vcn[3 * u + f] == b.vcn[3 * w + g] && b.vcn[3 * u + f] != c) ans = b.vcn[3 * u + f];
    return ans;
}

__global__ void k_vm_event_prepare(const cgpu_vertex_batch b, cgpu_vm_event *ev) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= b.num_systems) return;
    cgpu_vm_event &e = ev[s];
    if (e.kind < 0) return;
    const cgpu_vm_params &sp = b.sys[s];
    const int c = sp.coff + e.cell;
    const int n = b.cvn[c];
    e.status = 0;
    e.grow = 0;
    if (e.kind == 0) {
        if (n != 3) {
            e.status = 1;
            return;
        }
        if (sp.nc <= 4) {
            e.status = 2;
            return;
        }
        const int V[3] = {list_at(b, c, 0), list_at(b, c, 1), list_at(b, c, 2)};
        // the cells around the triangle, in order of appearance over its list
        int around[3], na = 0;
        for (int k = 0; k < 3; ++k)
            for (int f = 0; f < 3; ++f) {
                const int q = b.vcn[3 * V[k] + f];
                bool seen = q == c;
                for (int t = 0; t < na; ++t) seen = seen || around[t] == q;
                if (!seen) {
                    if (na == 3) {
                        e.status = 3;
                        return;
                    }
                    around[na++] = q;
                }
            }
        // the list rotated to its smallest vertex, then the vertices beyond it
        int m = 0;
        for (int k = 1; k < 3; ++k)
            if (V[k] < V[m]) m = k;
        const int W[3] = {V[m], V[(m + 1) % 3], V[(m + 2) % 3]};
        int outer[3], no = 0;
        for (int k = 0; k < 3; ++k)
            for (int f = 0; f < 3; ++f) {
                const int u = b.vn[3 * W[k] + f];
                bool seen = u == W[0] || u == W[1] || u == W[2];
                for (int t = 0; t < no; ++t) seen = seen || outer[t] == u;
                if (!seen) {
                    if (no == 3) {
                        e.status = 3;
                        return;
                    }
                    outer[no++] = u;
                }
            }
        if (na != 3 || no != 3) {
            e.status = 3;
            return;
        }
        for (int k = 0; k < 3; ++k) {
            e.v[k] = W[k] - sp.voff;
            e.around[k] = around[k] - sp.coff;
            e.outer[k] = outer[k] - sp.voff;
        }
    } else {
        const int i1 = min(e.p1, e.p2), i2 = max(e.p1, e.p2);
        if (i1 < 0 || i2 >= n || i1 == i2) {
            e.status = 4;
            return;
        }
        const int a = list_at(b, c, i1), an = list_at(b, c, (i1 + 1) % n);
        const int bb = list_at(b, c, i2), bn = list_at(b, c, (i2 + 1) % n);
        const int na = third_cell(b, a, an, c), nbc = third_cell(b, bb, bn, c);
        if (na < 0 || nbc < 0) {
            e.status = 5;
            return;
        }
        e.a = a - sp.voff;
        e.an = an - sp.voff;
        e.b = bb - sp.voff;
        e.bn = bn - sp.voff;
        e.na = na - sp.coff;
        e.nb = nbc - sp.coff;
        const int child = (i2 - i1) + 2, parent = n - (i2 - i1) + 2;
        const int grow_n = (na == nbc) ? b.cvn[na] + 2 : max(b.cvn[na], b.cvn[nbc]) + 1;
        e.grow = max(max(child, parent), grow_n);
    }
}

// new local vertex id of old local vertex v (the dying triangle's vertices merge
// into v[0], its smallest; the two others leave)
__device__ inline int ev_vertex(const cgpu_vm_event &e, int v) {
    if (e.kind != 0) return v;
    const int r1 = min(e.v[1], e.v[2]), r2 = max(e.v[1], e.v[2]);
    if (v == e.v[1] || v == e.v[2]) v = e.v[0];
    return v - (v > r1 ? 1 : 0) - (v > r2 ? 1 : 0);
}
__device__ inline int ev_cell(const cgpu_vm_event &e, int q) {
    if (e.kind != 0) return q;
    return q == e.cell ? -1 : q - (q > e.cell ? 1 : 0);
}
// system of new index j: the last s with off(s) <= j
template <typename F> __device__ inline int ev_system(const cgpu_vm_event *ev, int S, int j, F &&off) {
    int lo = 0, hi = S - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off(ev[mid]) <= j) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__global__ void k_vm_events_gather_vertices(const cgpu_vertex_batch b, const cgpu_vm_event *ev, int new_nv,
                                            double *pos, int *vn, int *vcn) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= new_nv) return;
    const int s = ev_system(ev, b.num_systems, j, [](const cgpu_vm_event &e) { return e.new_voff; });
    const cgpu_vm_event &e = ev[s];
    const int local = j - e.new_voff;
    if (e.kind == 1 && local >= e.old_nv) return;  // new vertices: the surgery
    int lo = local;
    if (e.kind == 0) {
        const int r1 = min(e.v[1], e.v[2]), r2 = max(e.v[1], e.v[2]);
        if (lo >= r1) ++lo;
        if (lo >= r2) ++lo;
    }
    const int o = e.old_voff + lo;
    stp(pos, j, ldp(b.pos, o));
    for (int f = 0; f < 3; ++f) {
        vn[3 * j + f] = e.new_voff + ev_vertex(e, b.vn[3 * o + f] - e.old_voff);
        const int q = ev_cell(e, b.vcn[3 * o + f] - e.old_coff);
        vcn[3 * j + f] = q < 0 ? -1 : e.new_coff + q;
    }
}

template <typename R>
__global__ void k_vm_events_gather_cells(const cgpu_vertex_batch b, const cgpu_vm_event *ev, int new_nc, int cap,
                                         int *cvn, int *cv, R *pref, R *theta, R *mot) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= new_nc) return;
    const int s = ev_system(ev, b.num_systems, j, [](const cgpu_vm_event &e) { return e.new_coff; });
    const cgpu_vm_event &e = ev[s];
    const int local = j - e.new_coff;
    const bool daughter = e.kind == 1 && local >= e.old_nc;
    const int lo = daughter ? e.cell : (e.kind == 0 && local >= e.cell ? local + 1 : local);
    const int o = e.old_coff + lo;
    const int n = daughter ? 0 : b.cvn[o];
    cvn[j] = n;
    for (int k = 0; k < cap; ++k)
        cv[(size_t)k * new_nc + j] = k < n ? e.new_voff + ev_vertex(e, list_at(b, o, k) - e.old_voff) : -1;
    const R *p0 = (const R *)b.APpref, *t0 = (const R *)b.theta, *m0 = (const R *)b.motility;
    pref[2 * j] = p0[2 * o];
    pref[2 * j + 1] = p0[2 * o + 1];
    mot[2 * j] = m0[2 * o];
    mot[2 * j + 1] = m0[2 * o + 1];
    theta[j] = daughter ? (R)e.director : t0[o];
}

__device__ inline void set_list(const cgpu_vertex_batch &b, int c, const int *l, int n) {
    const int Nc = b.num_cells;
    const int old = b.cvn[c];
    for (int k = 0; k < n; ++k) b.cv[(size_t)k * Nc + c] = l[k];
    for (int k = n; k < old; ++k) b.cv[(size_t)k * Nc + c] = -1;
    b.cvn[c] = n;
}
__device__ inline void replace_neighbour(const cgpu_vertex_batch &b, int v, int from, int to) {
    for (int f = 0; f < 3; ++f)
        if (b.vn[3 * v + f] == from) b.vn[3 * v + f] = to;
}
// x inserted right after `after` in cell c's list
__device__ inline void insert_after(const cgpu_vertex_batch &b, int c, int after, int x) {
    const int Nc = b.num_cells, n = b.cvn[c];
    int at = n - 1;
    for (int k = 0; k < n; ++k)
        if (b.cv[(size_t)k * Nc + c] == after) at = k;
    for (int k = n; k > at + 1; --k) b.cv[(size_t)k * Nc + c] = b.cv[(size_t)(k - 1) * Nc + c];
    b.cv[(size_t)(at + 1) * Nc + c] = x;
    b.cvn[c] = n + 1;
}

// the local topology change, one thread per system with an event, on the new layout
__global__ void k_vm_events_surgery(const cgpu_vertex_batch b, const cgpu_vm_event *ev) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= b.num_systems) return;
    const cgpu_vm_event &e = ev[s];
    if (e.kind < 0) return;
    const cgpu_vm_params &sp = b.sys[s];
    int l[EV_LIST];
    if (e.kind == 0) {
        // the cells around the triangle lose their now repeated merged vertex;
        // the merged vertex takes them and the vertices beyond
        const int v0 = sp.voff + ev_vertex(e, e.v[0]);
        for (int k = 0; k < 3; ++k) {
            const int q = sp.coff + ev_cell(e, e.around[k]);
            const int n = b.cvn[q];
            int w = 0;
            bool seen = false;
            for (int t = 0; t < n; ++t) {
                const int x = list_at(b, q, t);
                if (x == v0) {
                    if (seen) continue;
                    seen = true;
                }
                l[w++] = x;
            }
            set_list(b, q, l, w);
            b.vcn[3 * v0 + k] = q;
            b.vn[3 * v0 + k] = sp.voff + ev_vertex(e, e.outer[k]);
        }
        normalize_vertex(b, v0);
        return;
    }
    // division: A, B on the two edges' midpoints, the daughter C takes the part
    // of the list from an to b
    const int c = sp.coff + e.cell, C = sp.coff + e.old_nc;
    const int A = sp.voff + e.old_nv, B = A + 1;
    const int a = sp.voff + e.a, an = sp.voff + e.an, bb = sp.voff + e.b, bn = sp.voff + e.bn;
    const int na = sp.coff + e.na, nbc = sp.coff + e.nb;
    const int n = b.cvn[c];
    const int i1 = min(e.p1, e.p2), i2 = max(e.p1, e.p2);
    for (int k = 0; k < n; ++k) l[k] = list_at(b, c, k);
    const Box box = make_box(sp.L);
    auto mid = [&](int u, int w) {
        const dvec2 pu = ldp(b.pos, u);
        const dvec2 d = min_dist(box, ldp(b.pos, w), pu);
        dvec2 q{pu.x + 0.5 * d.x, pu.y + 0.5 * d.y};
        put_in_box(box, q);
        return q;
    };
    stp(b.pos, A, mid(a, an));
    stp(b.pos, B, mid(bb, bn));
    replace_neighbour(b, a, an, A);
    replace_neighbour(b, an, a, A);
    replace_neighbour(b, bb, bn, B);
    replace_neighbour(b, bn, bb, B);
    // the daughter's vertices, an .. b, change cell
    for (int k = i1 + 1; k <= i2; ++k)
        for (int f = 0; f < 3; ++f)
            if (b.vcn[3 * l[k] + f] == c) b.vcn[3 * l[k] + f] = C;
    int child[EV_LIST], m = 0;
    child[m++] = A;
    for (int k = i1 + 1; k <= i2; ++k) child[m++] = l[k];
    child[m++] = B;
    b.cvn[C] = 0;
    set_list(b, C, child, m);
    int par[EV_LIST], q = 0;
    par[q++] = B;
    for (int k = i2 + 1; k <= i1 + n; ++k) par[q++] = l[k % n];
    par[q++] = A;
    set_list(b, c, par, q);
    insert_after(b, na, an, A);
    insert_after(b, nbc, bn, B);
    const int vA[3] = {a, an, B}, cA[3] = {na, C, c}, vB[3] = {A, bb, bn}, cB[3] = {nbc, c, C};
    for (int f = 0; f < 3; ++f) {
        b.vn[3 * A + f] = vA[f];
        b.vcn[3 * A + f] = cA[f];
        b.vn[3 * B + f] = vB[f];
        b.vcn[3 * B + f] = cB[f];
    }
    normalize_vertex(b, A);
    normalize_vertex(b, B);
}

inline int nb(int n, int bs) { return (n + bs - 1) / bs; }
// per-cell and per-vertex kernels: small work-groups, so a single tissue of tens
// of thousands of cells spreads over every CU
constexpr int VM_BLOCK = 64;
inline hipStream_t S(void *s) { return reinterpret_cast<hipStream_t>(s); }
inline int check() { return hipGetLastError() == hipSuccess ? CGPU_OK : CGPU_ERR_HIP; }

}  // namespace

extern "C" {

int cgpu_vm_geometry(const cgpu_vertex_batch *b, void *stream) {
    if (!b) return CGPU_ERR_ARGS;
    if (b->precision)
        hipLaunchKernelGGL(k_vm_geometry<float>, dim3(nb(b->num_cells, VM_BLOCK)), dim3(VM_BLOCK), 0, S(stream), *b, b->pos);
    else
        hipLaunchKernelGGL(k_vm_geometry<double>, dim3(nb(b->num_cells, VM_BLOCK)), dim3(VM_BLOCK), 0, S(stream), *b, b->pos);
    return check();
}

int cgpu_vm_forces(const cgpu_vertex_batch *b, void *stream) {
    if (!b) return CGPU_ERR_ARGS;
    dim3 g(nb(b->num_vertices, VM_BLOCK)), t(VM_BLOCK);
    if (b->precision)
        hipLaunchKernelGGL((k_vm_force_eom_move<float, 0>), g, t, 0, S(stream), *b, b->pos, b->pos_next, 0);
    else
        hipLaunchKernelGGL((k_vm_force_eom_move<double, 0>), g, t, 0, S(stream), *b, b->pos, b->pos_next, 0);
    return check();
}

}  // extern "C"

namespace {
// geometry, force + equation of motion + move, T1 test (+ director rotation),
// T1 flips and the step counter
template <int KIND> int vm_step(const cgpu_vertex_batch *b, void *stream) {
    if (!b) return CGPU_ERR_ARGS;
    if (KIND == 1 && (!b->theta || !b->motility)) return CGPU_ERR_ARGS;
    hipStream_t st = S(stream);
    int r = cgpu_vm_geometry(b, stream);
    if (r) return r;
    dim3 g(nb(b->num_vertices, VM_BLOCK)), t(VM_BLOCK);
    if (b->precision)
        hipLaunchKernelGGL((k_vm_force_eom_move<float, KIND>), g, t, 0, st, *b, b->pos, b->pos_next, 1);
    else
        hipLaunchKernelGGL((k_vm_force_eom_move<double, KIND>), g, t, 0, st, *b, b->pos, b->pos_next, 1);
    hipLaunchKernelGGL(k_vm_t1_test, g, t, 0, st, *b, b->pos_next, KIND == 1 ? 1 : 0);
    hipLaunchKernelGGL(k_vm_t1_rounds, dim3(b->num_systems), dim3(256), 0, st, *b, b->pos_next, 1);
    return check();
}
}  // namespace

extern "C" {

int cgpu_vm_brownian_step(const cgpu_vertex_batch *b, void *stream) { return vm_step<0>(b, stream); }

int cgpu_vm_spp_step(const cgpu_vertex_batch *b, void *stream) { return vm_step<1>(b, stream); }

int cgpu_vm_t1(const cgpu_vertex_batch *b, void *stream) {
    if (!b) return CGPU_ERR_ARGS;
    hipLaunchKernelGGL(k_vm_t1_test, dim3(nb(b->num_vertices, VM_BLOCK)), dim3(VM_BLOCK), 0, S(stream), *b, b->pos, 0);
    hipLaunchKernelGGL(k_vm_t1_rounds, dim3(b->num_systems), dim3(256), 0, S(stream), *b, b->pos, 0);
    return check();
}

int cgpu_vm_fire_vv_move(const cgpu_vertex_batch *b, cgpu_fire_state *fs, void *vel, void *stream) {
    if (!b || !fs || !vel) return CGPU_ERR_ARGS;
    dim3 g(nb(b->num_vertices, VM_BLOCK)), t(VM_BLOCK);
    if (b->precision)
        hipLaunchKernelGGL(k_vm_fire_vv_move<float>, g, t, 0, S(stream), *b, fs, vel);
    else
        hipLaunchKernelGGL(k_vm_fire_vv_move<double>, g, t, 0, S(stream), *b, fs, vel);
    hipLaunchKernelGGL(k_vm_fire_clear_pending, dim3(nb(b->num_systems, 256)), dim3(256), 0, S(stream), fs,
                       b->num_systems);
    return check();
}

}  // extern "C"

namespace {
// geometry, forces, second velocity half step and the FIRE reductions
int vm_fire_force_reduce(const cgpu_vertex_batch *b, cgpu_fire_state *fs, void *vel, double *dots, void *stream,
                         int clear) {
    int r = cgpu_vm_geometry(b, stream);
    if (r) return r;
    dim3 g(nb(b->num_vertices, VM_BLOCK)), t(VM_BLOCK);
    if (b->precision)
        hipLaunchKernelGGL(k_vm_fire_force_vv2<float>, g, t, 0, S(stream), *b, fs, vel, dots, clear);
    else
        hipLaunchKernelGGL(k_vm_fire_force_vv2<double>, g, t, 0, S(stream), *b, fs, vel, dots, clear);
    hipLaunchKernelGGL(k_vm_fire_partial, dim3(b->num_chunks), dim3(256), 0, S(stream), *b, fs, dots);
    hipLaunchKernelGGL(k_vm_fire_reduce, dim3(nb(b->num_systems, 64)), dim3(64), 0, S(stream), *b, fs, dots);
    return check();
}
}  // namespace

extern "C" {

int cgpu_vm_fire_force_reduce(const cgpu_vertex_batch *b, cgpu_fire_state *fs, void *vel, double *dots,
                              void *stream) {
    if (!b || !fs || !vel || !dots || !b->chunk_sys) return CGPU_ERR_ARGS;
    return vm_fire_force_reduce(b, fs, vel, dots, stream, 0);
}

int cgpu_vm_fire_iteration(const cgpu_vertex_batch *b, cgpu_fire_state *fs, void *vel, double *dots, void *stream) {
    if (!b || !fs || !vel || !dots || !b->chunk_sys) return CGPU_ERR_ARGS;
    dim3 g(nb(b->num_vertices, VM_BLOCK)), t(VM_BLOCK);
    if (b->precision)
        hipLaunchKernelGGL(k_vm_fire_vv_move<float>, g, t, 0, S(stream), *b, fs, vel);
    else
        hipLaunchKernelGGL(k_vm_fire_vv_move<double>, g, t, 0, S(stream), *b, fs, vel);
    cgpu_vertex_batch bn = *b;  // the rest works on the moved positions
    bn.pos = b->pos_next;
    bn.pos_next = b->pos;
    hipLaunchKernelGGL(k_vm_t1_test, g, t, 0, S(stream), bn, bn.pos, 0);
    hipLaunchKernelGGL(k_vm_t1_rounds, dim3(bn.num_systems), dim3(256), 0, S(stream), bn, bn.pos, 0);
    // the force pass also clears the pending mix that the move applied
    const int r = vm_fire_force_reduce(&bn, fs, vel, dots, stream, 1);
    if (r) return r;
    hipLaunchKernelGGL(k_vm_fire_park, dim3(nb(bn.num_systems, 256)), dim3(256), 0, S(stream), bn, fs);
    return check();
}

int cgpu_vm_fire_park_converged(const cgpu_vertex_batch *b, const cgpu_fire_state *fs, void *stream) {
    if (!b || !fs) return CGPU_ERR_ARGS;
    hipLaunchKernelGGL(k_vm_fire_park, dim3(nb(b->num_systems, 256)), dim3(256), 0, S(stream), *b, fs);
    return check();
}

int cgpu_vm_fire_apply_mix(const cgpu_vertex_batch *b, cgpu_fire_state *fs, void *vel, void *stream) {
    if (!b || !fs || !vel) return CGPU_ERR_ARGS;
    dim3 g(nb(b->num_vertices, VM_BLOCK)), t(VM_BLOCK);
    if (b->precision)
        hipLaunchKernelGGL(k_vm_fire_apply_mix<float>, g, t, 0, S(stream), *b, fs, vel);
    else
        hipLaunchKernelGGL(k_vm_fire_apply_mix<double>, g, t, 0, S(stream), *b, fs, vel);
    hipLaunchKernelGGL(k_vm_fire_clear_pending, dim3(nb(b->num_systems, 256)), dim3(256), 0, S(stream), fs,
                       b->num_systems);
    return check();
}

}  // extern "C"

extern "C" {

int cgpu_vm_event_prepare(const cgpu_vertex_batch *b, cgpu_vm_event *ev, void *stream) {
    if (!b || !ev) return CGPU_ERR_ARGS;
    hipLaunchKernelGGL(k_vm_event_prepare, dim3(nb(b->num_systems, 64)), dim3(64), 0, S(stream), *b, ev);
    return check();
}

int cgpu_vm_events_gather(const cgpu_vertex_batch *b, const cgpu_vm_event *ev, int new_nv, int new_nc, int cap,
                          double *pos, int *vn, int *vcn, int *cvn, int *cv, void *APpref, void *theta, void *mot,
                          void *stream) {
    if (!b || !ev || new_nv <= 0 || new_nc <= 0 || cap < CGPU_VERTEX_CAP || !pos || !vn || !vcn || !cvn || !cv ||
        !APpref || !theta || !mot)
        return CGPU_ERR_ARGS;
    hipStream_t st = S(stream);
    hipLaunchKernelGGL(k_vm_events_gather_vertices, dim3(nb(new_nv, 256)), dim3(256), 0, st, *b, ev, new_nv, pos, vn,
                       vcn);
    if (b->precision)
        hipLaunchKernelGGL(k_vm_events_gather_cells<float>, dim3(nb(new_nc, 256)), dim3(256), 0, st, *b, ev, new_nc,
                           cap, cvn, cv, (float *)APpref, (float *)theta, (float *)mot);
    else
        hipLaunchKernelGGL(k_vm_events_gather_cells<double>, dim3(nb(new_nc, 256)), dim3(256), 0, st, *b, ev, new_nc,
                           cap, cvn, cv, (double *)APpref, (double *)theta, (double *)mot);
    return check();
}

int cgpu_vm_events_surgery(const cgpu_vertex_batch *nb_, const cgpu_vm_event *ev, void *stream) {
    if (!nb_ || !ev) return CGPU_ERR_ARGS;
    hipLaunchKernelGGL(k_vm_events_surgery, dim3(nb(nb_->num_systems, 64)), dim3(64), 0, S(stream), *nb_, ev);
    return check();
}

}  // extern "C"
