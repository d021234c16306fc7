// HIP kernels of the vertex-model time step for gfx950 (MI355X).
//
// Topology convention kept by every kernel here: the three cells of a vertex
// are stored counter-clockwise around it, and vn[3v+f] is the neighbour vertex
// across from cell vcn[3v+f]; then in cell f the polygon predecessor of v is
// vn[3v+(f+2)%3] and its successor vn[3v+(f+1)%3], so the per-vertex force needs
// no search of the cell's vertex list. Cell vertex lists are counter-clockwise,
// slot-major (cv[k*Nc + c]).
//
// Step: geometry (per cell) -> force + Brownian update + move (per vertex) ->
// FP64 T1 edge test (per vertex) -> T1 flips in deterministic rounds (one
// work-group per system: an edge (v1, v2) flips when its key is the smallest
// among the flagged edges sharing one of its four cells).
#include <hip/hip_runtime.h>

#include <climits>

#include "../../include/cellgpu_amd.h"
#include "cgpu_device.h"

using namespace cgpu;

namespace {

constexpr int VCAP = CGPU_VERTEX_CAP;

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

// force on vertex v: sum over its three cells of the reference force set
// (functions.h:330-354, vertexQuadraticEnergy.cu:21-70), then the Brownian
// update (brownianParticleDynamics.cu:20-39) and the move
template <typename R>
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
    str<R>(b.force, v, F);
    if (!move) return;
    const int step = st.step + 1;
    double z0, z1;
    if (b.noise) {
        z0 = b.noise[v];
        z1 = b.noise[b.num_vertices + v];
    } else {
        normal_pair(sp.seed, (uint32_t)(v - sp.voff), (uint32_t)step, 1u, z0, z1);
    }
    const R dt = (R)sp.dt, mu = (R)sp.mu;
    R sc = sqrt(dt * (R)sp.T * mu * R(2.0));
    R dx = (R)z0 * sc + dt * mu * F.x;
    R dy = (R)z1 * sc + dt * mu * F.y;
    p.x += (double)dx;
    p.y += (double)dy;
    if (!(fabs(p.x) < 1e12 * sp.L && fabs(p.y) < 1e12 * sp.L)) {
        st.error = CGPU_ERR_NONFINITE;
        p = ldp(pos, v);
    }
    put_in_box(box, p);
    stp(pos_next, v, p);
}

// reference vm_simple_T1_test_kernel (vertexModelBase.cu:118-164): FP64 edge
// length below the threshold, for v1 < v2; flagged edges go to the system's list
// as (v1, v2) pairs
__global__ void k_vm_t1_test(cgpu_vertex_batch b, const double *pos) {
    int v = xcd_block() * blockDim.x + threadIdx.x;
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
                int q = atomicAdd(&b.flag_count[s], 1);
                b.flag_list[2 * (3 * sp.voff + q)] = v;
                b.flag_list[2 * (3 * sp.voff + q) + 1] = v2;
            }
        }
    }
}

// position of x in cell c's list: predecessor and successor (reference loop)
__device__ inline void locate(const cgpu_vertex_batch &b, int c, int x, int &pred, int &succ) {
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
__device__ inline bool cell_has(const cgpu_vertex_batch &b, int c, int x) {
    const int Nc = b.num_cells;
    for (int k = 0; k < b.cvn[c]; ++k)
        if (b.cv[k * Nc + c] == x) return true;
    return false;
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

// the T1 of one edge (reference vm_flip_edges_kernel, vertexModelBase.cu:252-544);
// returns the four vertices whose slots need re-ordering, or false if forbidden
__device__ bool flip_edge(const cgpu_vertex_batch &b, double *pos, const cgpu_vm_params &sp, int v1, int v2,
                          const int C[4], int touched[4]) {
    const int Nc = b.num_cells;
    int cs[4] = {-1, -1, -1, C[3]};  // i, j, k, l
    for (int q = 0; q < 3; ++q) {
        int pr, su;
        locate(b, C[q], v1, pr, su);
        if (pr == v2)
            cs[0] = C[q];
        else if (su == v2)
            cs[2] = C[q];
        else
            cs[1] = C[q];
    }
    if (cs[0] < 0 || cs[1] < 0 || cs[2] < 0 || cs[3] < 0) return false;
    if (b.cvn[cs[0]] == 3 || b.cvn[cs[2]] == 3) return false;  // would shrink a triangle
    if (b.cvn[cs[1]] >= VCAP || b.cvn[cs[3]] >= VCAP) {
        b.st[b.vert_sys[v1]].error = CGPU_ERR_RING_CAPACITY;
        return false;
    }
    int dummy, a, bb;
    locate(b, cs[1], v1, dummy, a);
    locate(b, cs[3], v2, dummy, bb);
    const Box box = make_box(sp.L);
    dvec2 p1 = ldp(pos, v1), p2 = ldp(pos, v2);
    dvec2 e = min_dist(box, p1, p2);
    dvec2 mid{p2.x + 0.5 * e.x, p2.y + 0.5 * e.y};
    p1 = dvec2{mid.x - e.y, mid.y + e.x};
    p2 = dvec2{mid.x + e.y, mid.y - e.x};
    put_in_box(box, p1);
    put_in_box(box, p2);
    stp(pos, v1, p1);
    stp(pos, v2, p2);
    for (int q = 0; q < 3; ++q) {
        if (b.vcn[3 * v1 + q] == cs[2]) b.vcn[3 * v1 + q] = cs[3];
        if (b.vcn[3 * v2 + q] == cs[0]) b.vcn[3 * v2 + q] = cs[1];
        if (b.vn[3 * v1 + q] == a) b.vn[3 * v1 + q] = bb;
        if (b.vn[3 * v2 + q] == bb) b.vn[3 * v2 + q] = a;
        if (b.vn[3 * a + q] == v1) b.vn[3 * a + q] = v2;
        if (b.vn[3 * bb + q] == v2) b.vn[3 * bb + q] = v1;
    }
    // cell i loses v2, cell k loses v1
    for (int t = 0; t < 2; ++t) {
        int c = t == 0 ? cs[0] : cs[2], x = t == 0 ? v2 : v1;
        int n = b.cvn[c], w = 0;
        for (int k = 0; k < n; ++k) {
            int val = b.cv[k * Nc + c];
            if (val != x) b.cv[(w++) * Nc + c] = val;
        }
        b.cvn[c] = n - 1;
    }
    // cell j gains v2 before a, cell l gains v1 before b
    for (int t = 0; t < 2; ++t) {
        int c = t == 0 ? cs[1] : cs[3], x = t == 0 ? v2 : v1, before = t == 0 ? a : bb;
        int n = b.cvn[c];
        int at = 0;
        for (int k = 0; k < n; ++k)
            if (b.cv[k * Nc + c] == before) at = k;
        for (int k = n; k > at; --k) b.cv[k * Nc + c] = b.cv[(k - 1) * Nc + c];
        b.cv[at * Nc + c] = x;
        b.cvn[c] = n + 1;
    }
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
    int c[3] = {b.vcn[3 * w], b.vcn[3 * w + 1], b.vcn[3 * w + 2]};
    int n[3] = {b.vn[3 * w], b.vn[3 * w + 1], b.vn[3 * w + 2]};
    int pr, su;
    locate(b, c[0], w, pr, su);
    int c1 = cell_has(b, c[1], pr) ? c[1] : c[2];
    int c2 = (c1 == c[1]) ? c[2] : c[1];
    int cc[3] = {c[0], c1, c2};
    for (int f = 0; f < 3; ++f) {
        b.vcn[3 * w + f] = cc[f];
        int opp = n[0];
        for (int q = 0; q < 3; ++q)
            if (!cell_has(b, cc[f], n[q])) opp = n[q];
        b.vn[3 * w + f] = opp;
    }
}

// one work-group per system: T1 flips in rounds. An edge flips when its key
// (v1, v2) is the smallest among the flagged edges sharing one of its four cells;
// flips of one round touch disjoint cells and vertices. Edges removed by a
// neighbouring flip are dropped.
__global__ void k_vm_t1_rounds(cgpu_vertex_batch b, double *pos) {
    const int s = blockIdx.x;
    const cgpu_vm_params &sp = b.sys[s];
    cgpu_sys_state &st = b.st[s];
    int *list = b.flag_list + 2 * 3 * sp.voff;
    int *cells = b.flag_cells + 12 * sp.voff;
    int *touched = b.flag_touched + 12 * sp.voff;
    const unsigned long long NONE = ~0ull;
    __shared__ int count, rounds, flips;
    if (threadIdx.x == 0) {
        count = b.flag_count[s];
        rounds = 0;
        flips = 0;
    }
    __syncthreads();
    while (count > 0 && rounds < 100000) {
        const int nq = count;
        for (int q = threadIdx.x; q < nq; q += blockDim.x) {
            int v1 = list[2 * q], v2 = list[2 * q + 1];
            bool live = b.vn[3 * v1] == v2 || b.vn[3 * v1 + 1] == v2 || b.vn[3 * v1 + 2] == v2;
            int C[4] = {-1, -1, -1, -1};
            if (live) edge_cells(b, v1, v2, C);
            else list[2 * q] = -2;  // dropped
            unsigned long long k = ((unsigned long long)(unsigned)v1 << 32) | (unsigned)v2;
            for (int t = 0; t < 4; ++t) {
                cells[4 * q + t] = C[t];
                if (C[t] >= 0) atomicMin(&b.cell_min[C[t]], k);
            }
        }
        __syncthreads();
        for (int q = threadIdx.x; q < nq; q += blockDim.x) {
            touched[4 * q] = -1;
            int v1 = list[2 * q], v2 = list[2 * q + 1];
            if (v1 < 0) continue;
            unsigned long long k = ((unsigned long long)(unsigned)v1 << 32) | (unsigned)v2;
            bool win = true;
            for (int t = 0; t < 4; ++t)
                if (cells[4 * q + t] >= 0 && b.cell_min[cells[4 * q + t]] != k) win = false;
            if (win) {
                int C[4] = {cells[4 * q], cells[4 * q + 1], cells[4 * q + 2], cells[4 * q + 3]};
                if (flip_edge(b, pos, sp, v1, v2, C, &touched[4 * q])) atomicAdd(&flips, 1);
                list[2 * q] = -1;
            }
        }
        __syncthreads();
        for (int q = threadIdx.x; q < nq; q += blockDim.x) {
            for (int t = 0; t < 4; ++t)
                if (cells[4 * q + t] >= 0) b.cell_min[cells[4 * q + t]] = NONE;
            if (list[2 * q] == -1 && touched[4 * q] >= 0)
                for (int t = 0; t < 4; ++t) normalize_vertex(b, touched[4 * q + t]);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int w = 0;
            for (int q = 0; q < nq; ++q)
                if (list[2 * q] >= 0) {
                    list[2 * w] = list[2 * q];
                    list[2 * w + 1] = list[2 * q + 1];
                    ++w;
                }
            count = w;
            rounds += 1;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        st.t1_events += flips;
        b.flag_count[s] = 0;
    }
}

__global__ void k_vm_advance(cgpu_vertex_batch b) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= b.num_systems) return;
    cgpu_sys_state &st = b.st[s];
    if (!st.active) return;
    st.step += 1;
    if (st.steps_left > 0 && --st.steps_left == 0) st.active = 0;
}

inline int nb(int n, int bs) { return (n + bs - 1) / bs; }
inline hipStream_t S(void *s) { return reinterpret_cast<hipStream_t>(s); }
inline int check() { return hipGetLastError() == hipSuccess ? CGPU_OK : CGPU_ERR_HIP; }

}  // namespace

extern "C" {

int cgpu_vm_geometry(const cgpu_vertex_batch *b, void *stream) {
    if (!b) return CGPU_ERR_ARGS;
    if (b->precision)
        hipLaunchKernelGGL(k_vm_geometry<float>, dim3(nb(b->num_cells, 256)), dim3(256), 0, S(stream), *b, b->pos);
    else
        hipLaunchKernelGGL(k_vm_geometry<double>, dim3(nb(b->num_cells, 256)), dim3(256), 0, S(stream), *b, b->pos);
    return check();
}

int cgpu_vm_forces(const cgpu_vertex_batch *b, void *stream) {
    if (!b) return CGPU_ERR_ARGS;
    dim3 g(nb(b->num_vertices, 256)), t(256);
    if (b->precision)
        hipLaunchKernelGGL(k_vm_force_eom_move<float>, g, t, 0, S(stream), *b, b->pos, b->pos_next, 0);
    else
        hipLaunchKernelGGL(k_vm_force_eom_move<double>, g, t, 0, S(stream), *b, b->pos, b->pos_next, 0);
    return check();
}

int cgpu_vm_brownian_step(const cgpu_vertex_batch *b, void *stream) {
    if (!b) return CGPU_ERR_ARGS;
    hipStream_t st = S(stream);
    int r = cgpu_vm_geometry(b, stream);
    if (r) return r;
    dim3 g(nb(b->num_vertices, 256)), t(256);
    if (b->precision)
        hipLaunchKernelGGL(k_vm_force_eom_move<float>, g, t, 0, st, *b, b->pos, b->pos_next, 1);
    else
        hipLaunchKernelGGL(k_vm_force_eom_move<double>, g, t, 0, st, *b, b->pos, b->pos_next, 1);
    hipLaunchKernelGGL(k_vm_t1_test, g, t, 0, st, *b, b->pos_next);
    hipLaunchKernelGGL(k_vm_t1_rounds, dim3(b->num_systems), dim3(256), 0, st, *b, b->pos_next);
    hipLaunchKernelGGL(k_vm_advance, dim3(nb(b->num_systems, 256)), dim3(256), 0, st, *b);
    return check();
}

int cgpu_vm_t1(const cgpu_vertex_batch *b, void *stream) {
    if (!b) return CGPU_ERR_ARGS;
    hipLaunchKernelGGL(k_vm_t1_test, dim3(nb(b->num_vertices, 256)), dim3(256), 0, S(stream), *b, b->pos);
    hipLaunchKernelGGL(k_vm_t1_rounds, dim3(b->num_systems), dim3(256), 0, S(stream), *b, b->pos);
    return check();
}

}  // extern "C"
