// VertexBatch: device-resident batch of vertex-model systems (reference
// vertexModelBase + VertexQuadraticEnergy + brownianParticleDynamics, GPU
// branch), built from the Voronoi tessellation of cell positions.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "../cgpu_device.h"
#include "fire.hpp"
#include "voronoi_batch.hpp"

namespace cgpu_host {

// longest cell list a division or death may make (one thread's stack in the
// event surgery kernel)
constexpr int EV_LIST_MAX = 256;

// The convention the kernels rely on, rebuilt from the vertex-cell incidences and
// the vertex-neighbour sets of one system (local indices), using topology only
// (geometry can be locally inverted right after a T1 flip):
// 1. one consistent cyclic order of the three cells at every vertex, by
//    propagation over edges: where v and u share cells P, Q with P -> Q at v,
//    the order at u must be Q -> P (the two ends of an edge see it from opposite
//    sides);
// 2. each neighbour stored across from the cell that does not contain it;
// 3. cell lists by walking polygon successors (the successor of v in its cell f
//    is the neighbour stored at f + 1), from the cell's smallest vertex;
// 4. the global sense from the total signed area (counter-clockwise > 0).
static void canonical_topology(int nv, int nc, const double *lx, const double *ly, const cgpu::Box &box, int cap,
                        std::vector<int> &lvn, std::vector<int> &lvcn, std::vector<std::vector<int>> &cells) {
    auto nxt = [&](int v, int c) {
        for (int f = 0; f < 3; ++f)
            if (lvcn[3 * v + f] == c) return lvcn[3 * v + (f + 1) % 3];
        throw std::invalid_argument("topology inconsistent");
    };
    auto shared = [&](int v, int u, int &P, int &Q) {
        int got = 0, sc[2] = {-1, -1};
        for (int f = 0; f < 3; ++f)
            for (int g = 0; g < 3; ++g)
                if (lvcn[3 * v + f] == lvcn[3 * u + g] && got < 2) sc[got++] = lvcn[3 * v + f];
        if (got != 2) throw std::invalid_argument("topology inconsistent");
        if (nxt(v, sc[0]) == sc[1]) {
            P = sc[0];
            Q = sc[1];
        } else {
            P = sc[1];
            Q = sc[0];
        }
    };
    std::vector<char> seen(nv, 0);
    std::vector<int> queue;
    for (int root = 0; root < nv; ++root) {
        if (seen[root]) continue;
        seen[root] = 1;
        queue.assign(1, root);
        for (size_t h = 0; h < queue.size(); ++h) {
            const int v = queue[h];
            for (int k = 0; k < 3; ++k) {
                const int u = lvn[3 * v + k];
                int P, Q;
                shared(v, u, P, Q);
                const bool ok = nxt(u, Q) == P;
                if (!seen[u]) {
                    if (!ok) std::swap(lvcn[3 * u + 1], lvcn[3 * u + 2]);
                    seen[u] = 1;
                    queue.push_back(u);
                } else if (!ok) {
                    throw std::invalid_argument("topology is not orientable");
                }
            }
        }
    }
    auto assign_neighbours = [&]() {
        for (int v = 0; v < nv; ++v) {
            const int n3[3] = {lvn[3 * v], lvn[3 * v + 1], lvn[3 * v + 2]};
            for (int f = 0; f < 3; ++f) {
                int across = -1;
                for (int u : n3) {
                    bool in_u = false;
                    for (int g = 0; g < 3; ++g) in_u = in_u || lvcn[3 * u + g] == lvcn[3 * v + f];
                    if (!in_u) across = u;
                }
                if (across < 0) throw std::invalid_argument("topology inconsistent");
                lvn[3 * v + f] = across;
            }
        }
    };
    std::vector<int> first(nc, INT_MAX), count(nc, 0);
    for (int v = 0; v < nv; ++v)
        for (int f = 0; f < 3; ++f) {
            const int c = lvcn[3 * v + f];
            first[c] = std::min(first[c], v);
            ++count[c];
        }
    cells.assign(nc, {});
    auto walk = [&]() {
        double area = 0.0;
        for (int c = 0; c < nc; ++c) {
            if (count[c] < 3 || count[c] > cap) throw std::invalid_argument("cell size out of range");
            cells[c].clear();
            int v = first[c];
            const cgpu::dvec2 base{lx[v], ly[v]};
            cgpu::dvec2 prev{0.0, 0.0};
            for (int k = 0; k < count[c]; ++k) {
                cells[c].push_back(v);
                int f = 0;
                while (f < 3 && lvcn[3 * v + f] != c) ++f;
                if (f == 3) throw std::invalid_argument("topology inconsistent");
                v = lvn[3 * v + (f + 1) % 3];
                const cgpu::dvec2 d = cgpu::min_dist(box, cgpu::dvec2{lx[v], ly[v]}, base);
                area += 0.5 * (prev.x * d.y - prev.y * d.x);
                prev = d;
            }
            if (v != first[c]) throw std::invalid_argument("cell polygon does not close");
        }
        return area;
    };
    assign_neighbours();
    if (walk() < 0) {  // the propagated sense is clockwise: reverse every vertex
        for (int v = 0; v < nv; ++v) std::swap(lvcn[3 * v + 1], lvcn[3 * v + 2]);
        assign_neighbours();
        walk();
    }
}


class VertexBatch {
   public:
    VertexBatch(int S, const int *ncells, const double *L, const uint64_t *seed, const double *x, const double *y,
                int precision, int device, bool spv_initialize)
        : S_(S), prec_(precision ? 1 : 0) {
        hip_check(hipSetDevice(device), "hipSetDevice");
        rsz_ = prec_ ? sizeof(float) : sizeof(double);
        // Delaunay rings of the cell positions, computed on the device in FP64
        std::vector<double> Ls(S);
        for (int s = 0; s < S; ++s) Ls[s] = L ? L[s] : std::sqrt((double)ncells[s]);
        VoronoiBatch vor(S, ncells, Ls.data(), seed, x, y, nullptr, 0, device);
        const int Nc = vor.num_cells();
        // positions, then the directors drawn after them (reference
        // initializeSimple2DActiveCell: positions, setCellDirectorsRandomly)
        std::vector<double> px(Nc), py(Nc), pth(Nc);
        vor.get(px.data(), py.data(), pth.data(), nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
        if (spv_initialize) {
            // reference setCellsVoronoiTesselation(true), vertexModelBase.cpp:118-140:
            // 100 self-propelled Voronoi steps smooth the random point set
            std::vector<double> a0(Nc, 1.0), p0(Nc, 3.8), v0(Nc, 0.1), dr(Nc, 1.0);
            vor.set_params(-1, 1.0, 1.0, 0.1, 1.0, 0.0);
            vor.set_prefs(a0.data(), p0.data());
            vor.set_motility(v0.data(), dr.data());
            if (vor.spv_steps(100, true) != CGPU_OK || vor.status() != CGPU_OK)
                throw std::runtime_error("spv initialisation failed");
            vor.get(px.data(), py.data(), nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
        }
        const int cap = vor.ring_capacity();
        std::vector<int> nn(Nc), nbr((size_t)Nc * cap);
        vor.get_topology(nn.data(), nbr.data(), cap, nullptr);

        hp_.resize(S);
        std::vector<int> tri_of((size_t)Nc * cap, -1);
        std::vector<int> tri_cells;  // 3 per vertex
        int voff = 0, coff = 0;
        for (int s = 0; s < S; ++s) {
            cgpu_vm_params &p = hp_[s];
            std::memset(&p, 0, sizeof(p));
            p.coff = coff;
            p.nc = ncells[s];
            p.voff = voff;
            p.L = Ls[s];
            p.KA = p.KP = 1.0;
            p.dt = 0.005;
            p.mu = 1.0;
            p.T = 0.0;
            p.thr = 0.04;
            p.seed = seed ? seed[s] : 13377ull;
            // vertices: triangles owned by their smallest cell, (cell, slot) order
            for (int i = coff; i < coff + p.nc; ++i)
                for (int k = 0; k < nn[i]; ++k) {
                    int j = nbr[(size_t)i * cap + k], l = nbr[(size_t)i * cap + (k + 1) % nn[i]];
                    if (i < j && i < l) {
                        tri_of[(size_t)i * cap + k] = (int)tri_cells.size() / 3;
                        tri_cells.push_back(i);
                        tri_cells.push_back(j);
                        tri_cells.push_back(l);
                    }
                }
            p.nv = (int)tri_cells.size() / 3 - voff;
            voff += p.nv;
            coff += p.nc;
        }
        Nv_ = voff;
        Nc_ = coff;

        auto find_tri = [&](int a, int b, int c) {  // counter-clockwise (a, b, c)
            int o = std::min(a, std::min(b, c)), p, q;
            if (o == a) { p = b; q = c; } else if (o == b) { p = c; q = a; } else { p = a; q = b; }
            for (int k = 0; k < nn[o]; ++k)
                if (nbr[(size_t)o * cap + k] == p && nbr[(size_t)o * cap + (k + 1) % nn[o]] == q)
                    return tri_of[(size_t)o * cap + k];
            return -1;
        };
        std::vector<int> vn(3 * (size_t)Nv_), vcn(3 * (size_t)Nv_), vsys(Nv_), csys(Nc_);
        std::vector<double> vpos(2 * (size_t)Nv_);
        for (int s = 0; s < S; ++s) {
            const cgpu_vm_params &p = hp_[s];
            cgpu::Box box = cgpu::make_box(p.L);
            for (int t = p.voff; t < p.voff + p.nv; ++t) {
                vsys[t] = s;
                int c3[3] = {tri_cells[3 * t], tri_cells[3 * t + 1], tri_cells[3 * t + 2]};
                cgpu::dvec2 pi{px[c3[0]], py[c3[0]]};
                cgpu::dvec2 dj = cgpu::min_dist(box, cgpu::dvec2{px[c3[1]], py[c3[1]]}, pi);
                cgpu::dvec2 dl = cgpu::min_dist(box, cgpu::dvec2{px[c3[2]], py[c3[2]]}, pi);
                cgpu::dvec2 c = cgpu::circ0<double>(dj, dl);
                cgpu::dvec2 q{pi.x + c.x, pi.y + c.y};
                cgpu::put_in_box(box, q);
                vpos[2 * t] = q.x;
                vpos[2 * t + 1] = q.y;
                for (int f = 0; f < 3; ++f) {
                    vcn[3 * t + f] = c3[f];
                    int a = c3[(f + 1) % 3], b = c3[(f + 2) % 3], X = -1;
                    for (int k = 0; k < nn[b]; ++k)
                        if (nbr[(size_t)b * cap + (k + 1) % nn[b]] == a) X = nbr[(size_t)b * cap + (k + 2) % nn[b]];
                    vn[3 * t + f] = find_tri(b, a, X);
                }
            }
            for (int c = p.coff; c < p.coff + p.nc; ++c) csys[c] = s;
        }
        for (int i = 0; i < Nc_; ++i) vertex_cap_ = std::max(vertex_cap_, (nn[i] + 1) & ~1);
        std::vector<int> cvn(Nc_), cv((size_t)vertex_cap_ * Nc_, -1);
        for (int i = 0; i < Nc_; ++i) {
            cvn[i] = nn[i];
            for (int k = 0; k < nn[i]; ++k)
                cv[(size_t)k * Nc_ + i] =
                    find_tri(i, nbr[(size_t)i * cap + k], nbr[(size_t)i * cap + (k + 1) % nn[i]]);
        }
        hip_check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "stream");
        sys_.alloc(S);
        st_.alloc(S);
        flag_count_.alloc(S);
        std::vector<cgpu_sys_state> st(S);
        std::memset(st.data(), 0, sizeof(cgpu_sys_state) * S);
        for (auto &q : st) {
            q.active = 1;
            q.steps_left = -1;
        }
        st_.upload(st.data(), S);
        layout();
        pos0_.upload(vpos.data(), 2 * (size_t)Nv_);
        vn_.upload(vn.data(), vn.size());
        vcn_.upload(vcn.data(), vcn.size());
        cvn_.upload(cvn.data(), Nc_);
        cv_.upload(cv.data(), cv.size());
        std::vector<double> a0(Nc_, 1.0), p0(Nc_, 3.9);
        set_prefs(a0.data(), p0.data());
        real_upload(theta_.p, pth.data(), Nc_);
        std::vector<double> v0(Nc_, 0.0), Dr(Nc_, 1.0);  // reference setv0Dr(0.0, 1.0)
        set_motility(v0.data(), Dr.data());
    }
    // Offsets, vertex chunks and every per-vertex and per-cell device array for
    // the sizes in hp_ (nv, nc of each system); arrays start zeroed.
    void layout() {
        int voff = 0, coff = 0;
        std::vector<int> ch_sys, ch_begin, ch_end, vsys, csys;
        for (int s = 0; s < S_; ++s) {
            cgpu_vm_params &p = hp_[s];
            p.voff = voff;
            p.coff = coff;
            p.chunk0 = (int)ch_sys.size();
            for (int b = voff; b < voff + p.nv; b += CGPU_CHUNK) {
                ch_sys.push_back(s);
                ch_begin.push_back(b);
                ch_end.push_back(std::min(b + CGPU_CHUNK, voff + p.nv));
            }
            vsys.insert(vsys.end(), p.nv, s);
            csys.insert(csys.end(), p.nc, s);
            voff += p.nv;
            coff += p.nc;
        }
        Nv_ = voff;
        Nc_ = coff;
        num_chunks_ = (int)ch_sys.size();
        sys_.upload(hp_.data(), S_);
        vsys_.alloc(Nv_);
        vsys_.upload(vsys.data(), Nv_);
        csys_.alloc(Nc_);
        csys_.upload(csys.data(), Nc_);
        pos0_.alloc(2 * (size_t)Nv_);
        pos1_.alloc(2 * (size_t)Nv_);
        vn_.alloc(3 * (size_t)Nv_);
        vcn_.alloc(3 * (size_t)Nv_);
        cvn_.alloc(Nc_);
        cv_.alloc((size_t)vertex_cap_ * Nc_);
        AP_.alloc(2 * Nc_ * rsz_);
        dAP_.alloc(2 * Nc_ * rsz_);
        APpref_.alloc(2 * Nc_ * rsz_);
        force_.alloc(2 * Nv_ * rsz_);
        flag_list_.alloc(6 * (size_t)Nv_);
        flag_cells_.alloc(12 * (size_t)Nv_);
        flag_touched_.alloc(12 * (size_t)Nv_);
        cell_min_.alloc(Nc_);
        std::vector<unsigned long long> ones(Nc_, ~0ull);
        cell_min_.upload(ones.data(), Nc_);
        theta_.alloc(Nc_ * rsz_);
        mot_.alloc(2 * Nc_ * rsz_);
        chunk_sys_.alloc(num_chunks_);
        chunk_sys_.upload(ch_sys.data(), num_chunks_);
        chunk_begin_.alloc(num_chunks_);
        chunk_begin_.upload(ch_begin.data(), num_chunks_);
        chunk_end_.alloc(num_chunks_);
        chunk_end_.upload(ch_end.data(), num_chunks_);
        if (noise_.p) noise_.alloc(2 * (size_t)Nv_ + Nc_);
        parity_ = 0;
        drop_graph();
    }
    ~VertexBatch() {
        if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
        if (graph_) (void)hipGraphDestroy(graph_);
        if (stream_) (void)hipStreamDestroy(stream_);
    }
    int num_vertices() const { return Nv_; }
    int num_systems() const { return S_; }
    int num_cells() const { return Nc_; }
    const cgpu_vm_params &params(int s) const { return hp_.at(s); }
    size_t real_size() const { return rsz_; }
    hipStream_t stream() const { return stream_; }
    void swap_parity() { parity_ ^= 1; }
    cgpu_vertex_batch view() const {
        cgpu_vertex_batch b;
        std::memset(&b, 0, sizeof(b));
        b.num_systems = S_;
        b.num_vertices = Nv_;
        b.num_cells = Nc_;
        b.precision = prec_;
        b.sys = sys_.p;
        b.st = st_.p;
        b.vert_sys = vsys_.p;
        b.cell_sys = csys_.p;
        b.pos = parity_ ? pos1_.p : pos0_.p;
        b.pos_next = parity_ ? pos0_.p : pos1_.p;
        b.vn = vn_.p;
        b.vcn = vcn_.p;
        b.cvn = cvn_.p;
        b.cv = cv_.p;
        b.AP = AP_.p;
        b.dAP = dAP_.p;
        b.APpref = APpref_.p;
        b.force = force_.p;
        b.flag_count = flag_count_.p;
        b.flag_list = flag_list_.p;
        b.flag_cells = flag_cells_.p;
        b.flag_touched = flag_touched_.p;
        b.cell_min = cell_min_.p;
        b.noise = noise_buffer_ ? noise_.p : nullptr;
        b.theta = theta_.p;
        b.motility = mot_.p;
        b.chunk_sys = chunk_sys_.p;
        b.chunk_begin = chunk_begin_.p;
        b.chunk_end = chunk_end_.p;
        b.num_chunks = num_chunks_;
        b.vertex_cap = vertex_cap_;
        return b;
    }
    void set_params(int s, double KA, double KP, double dt, double mu, double T, double thr) {
        sync();
        int lo = s < 0 ? 0 : s, hi = s < 0 ? S_ : s + 1;
        for (int q = lo; q < hi; ++q) {
            hp_[q].KA = KA;
            hp_[q].KP = KP;
            hp_[q].dt = dt;
            hp_[q].mu = mu;
            hp_[q].T = T;
            hp_[q].thr = thr;
        }
        sys_.upload(hp_.data(), S_);
        drop_graph();
    }
    void set_prefs(const double *A0, const double *P0) {
        sync();
        std::vector<double> v(2 * (size_t)Nc_);
        for (int i = 0; i < Nc_; ++i) {
            v[2 * i] = A0[i];
            v[2 * i + 1] = P0[i];
        }
        real_upload(APpref_.p, v.data(), v.size());
    }
    void set_positions(const double *vx, const double *vy) {
        sync();
        std::vector<double> xy(2 * (size_t)Nv_);
        for (int i = 0; i < Nv_; ++i) {
            if (!std::isfinite(vx[i]) || !std::isfinite(vy[i])) throw std::invalid_argument("non-finite position");
            xy[2 * i] = vx[i];
            xy[2 * i + 1] = vy[i];
        }
        (parity_ ? pos1_ : pos0_).upload(xy.data(), xy.size());
    }
    // noise buffer: x then y per vertex, then one per cell (director rotation)
    void set_noise_mode(bool on) {
        noise_buffer_ = on;
        if (on && !noise_.p) noise_.alloc(2 * (size_t)Nv_ + Nc_);
        drop_graph();
    }
    void upload_noise(const double *eta) {
        sync();
        if (!noise_.p) noise_.alloc(2 * (size_t)Nv_ + Nc_);
        noise_.upload(eta, 2 * (size_t)Nv_);
    }
    void upload_cell_noise(const double *eta) {
        sync();
        if (!noise_.p) noise_.alloc(2 * (size_t)Nv_ + Nc_);
        noise_.upload(eta, Nc_, 2 * (size_t)Nv_);
    }
    void set_motility(const double *v0, const double *Dr) {
        sync();
        std::vector<double> m(2 * (size_t)Nc_);
        for (int i = 0; i < Nc_; ++i) {
            m[2 * i] = v0[i];
            m[2 * i + 1] = Dr[i];
        }
        real_upload(mot_.p, m.data(), m.size());
    }
    void set_directors(const double *theta) {
        sync();
        real_upload(theta_.p, theta, Nc_);
    }
    void get_directors(double *theta) {
        sync();
        real_download(theta_.p, theta, Nc_);
    }
    int compute_forces() {
        cgpu_vertex_batch b = view();
        int r = cgpu_vm_geometry(&b, stream_);
        if (!r) r = cgpu_vm_forces(&b, stream_);
        sync();
        return r;
    }
    // reference vertexModelBase::enforceTopology (vertexModelBase.cpp:93-106):
    // T1 test and flips on the current positions, then the geometry
    int enforce_topology() {
        cgpu_vertex_batch b = view();
        int r = cgpu_vm_t1(&b, stream_);
        if (!r) r = cgpu_vm_geometry(&b, stream_);
        if (r) return r;
        return status();
    }
    // kind 0: Brownian step, 1: self-propelled cell-vertex step
    int step_launch(int kind) {
        cgpu_vertex_batch b = view();
        int r = kind ? cgpu_vm_spp_step(&b, stream_) : cgpu_vm_brownian_step(&b, stream_);
        parity_ ^= 1;
        return r;
    }
    int vertex_capacity() const { return vertex_cap_; }
    // Cell-list storage of cap slots per cell: slot-major rows keep their place
    // (new rows -1); a captured step graph is dropped (it holds the old address).
    void grow_vertex_capacity(int cap) {
        if (cap <= vertex_cap_) return;
        sync();
        DevBuf<int> g;
        g.alloc((size_t)cap * Nc_);
        hip_check(hipMemset(g.p, 0xff, sizeof(int) * (size_t)cap * Nc_), "grow cell lists");
        hip_check(hipMemcpy(g.p, cv_.p, sizeof(int) * (size_t)vertex_cap_ * Nc_, hipMemcpyDeviceToDevice),
                  "grow cell lists");
        cv_.swap(g);
        vertex_cap_ = cap;
        drop_graph();
    }
    // Systems stalled by a T1 flip beyond the cell-list capacity (error
    // CGPU_ERR_RING_CAPACITY, inactive). False when none.
    bool take_stalled(std::vector<int> &stalled) {
        sync();
        std::vector<cgpu_sys_state> st(S_);
        st_.download(st.data(), S_);
        stalled.clear();
        for (int s = 0; s < S_; ++s)
            if (st[s].error == CGPU_ERR_RING_CAPACITY && !st[s].active) stalled.push_back(s);
        return !stalled.empty();
    }
    // Finish stalled systems as the reference does when a flip needs a longer
    // cell list (growCellVerticesList, then the flips, vertexModelBase.cpp:
    // 815-824): grown storage, the T1 pass of the stalled step on its moved
    // positions, with the other systems parked. complete: the rest of the
    // stalled step or iteration, run for the stalled systems alone.
    template <typename F> int finish_stalled(F &&complete) {
        std::vector<int> stalled;
        while (take_stalled(stalled)) {
            grow_vertex_capacity(vertex_cap_ + 8);
            std::vector<cgpu_sys_state> st(S_);
            st_.download(st.data(), S_);
            const std::vector<cgpu_sys_state> saved = st;
            for (int s = 0; s < S_; ++s) st[s].active = 0;
            for (int s : stalled) {
                st[s].error = 0;
                st[s].active = 1;
            }
            st_.upload(st.data(), S_);
            cgpu_vertex_batch b = view();
            if (int r = cgpu_vm_t1(&b, stream_)) return r;
            sync();
            if (int r = complete(stalled, saved)) return r;
            sync();
            st_.download(st.data(), S_);
            for (int s = 0; s < S_; ++s)
                if (std::find(stalled.begin(), stalled.end(), s) == stalled.end()) {
                    st[s].active = saved[s].active;
                    st[s].steps_left = saved[s].steps_left;
                }
            st_.upload(st.data(), S_);
        }
        return CGPU_OK;
    }
    // steps: the stalled step is counted, then the steps still owed run for the
    // stalled systems alone (a countdown system keeps its countdown instead)
    int run_steps(int kind, int steps, bool use_graph) {
        if (steps <= 0) return CGPU_OK;
        std::vector<cgpu_sys_state> before(S_);
        if (st_call_.n != st_.n) st_call_.alloc(st_.n);
        hip_check(hipMemcpyAsync(st_call_.p, st_.p, sizeof(cgpu_sys_state) * S_, hipMemcpyDeviceToDevice, stream_),
                  "state copy");
        if (int r = launch_steps(kind, steps, use_graph)) return r;
        return finish_stalled([&](const std::vector<int> &stalled, const std::vector<cgpu_sys_state> &saved) {
            st_call_.download(before.data(), S_);
            std::vector<cgpu_sys_state> st(S_);
            st_.download(st.data(), S_);
            int owe = 0;
            for (int s : stalled) {
                st[s].step += 1;
                if (saved[s].steps_left >= 0) {
                    st[s].steps_left = saved[s].steps_left > 0 ? saved[s].steps_left - 1 : 0;
                    if (st[s].steps_left == 0 && saved[s].steps_left > 0) st[s].active = 0;
                    continue;
                }
                const int left = steps - (st[s].step - before[s].step);
                if (left > 0) {
                    st[s].steps_left = left;
                    owe = std::max(owe, left);
                }
            }
            st_.upload(st.data(), S_);
            for (int k = 0; k < owe; ++k)
                if (int r = step_launch(kind)) return r;
            sync();
            st_.download(st.data(), S_);
            for (int s : stalled)
                if (saved[s].steps_left < 0 && st[s].error == 0) {
                    st[s].active = 1;
                    st[s].steps_left = -1;
                }
            st_.upload(st.data(), S_);
            return (int)CGPU_OK;
        });
    }
    int launch_steps(int kind, int steps, bool use_graph) {
        if (!use_graph || noise_buffer_) {
            for (int k = 0; k < steps; ++k)
                if (int r = step_launch(kind)) return r;
            return CGPU_OK;
        }
        if (graph_exec_ && graph_kind_ != kind) drop_graph();
        if (graph_exec_ && graph_parity_ != parity_ && steps > 0) {
            if (int r = step_launch(kind)) return r;
            --steps;
        }
        if (!graph_exec_ && steps >= 2) {
            hip_check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "capture");
            graph_parity_ = parity_;
            graph_kind_ = kind;
            step_launch(kind);
            step_launch(kind);
            hip_check(hipStreamEndCapture(stream_, &graph_), "end capture");
            hip_check(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0), "instantiate");
            // only the executable graph is kept: the template would leak at every re-capture
            (void)hipGraphDestroy(graph_);
            graph_ = nullptr;
        }
        while (steps >= 2) {
            hip_check(hipGraphLaunch(graph_exec_, stream_), "graph launch");
            steps -= 2;
        }
        if (steps == 1) return step_launch(kind);
        return CGPU_OK;
    }
    void sync() { hip_check(hipStreamSynchronize(stream_), "sync"); }
    int status() {
        sync();
        std::vector<cgpu_sys_state> st(S_);
        st_.download(st.data(), S_);
        for (auto &q : st)
            if (q.error) return q.error;
        return CGPU_OK;
    }
    void get(double *vx, double *vy, double *Fx, double *Fy, double *A, double *P) {
        sync();
        std::vector<double> t(2 * (size_t)std::max(Nv_, Nc_));
        if (vx || vy) {
            (parity_ ? pos1_ : pos0_).download(t.data(), 2 * (size_t)Nv_);
            for (int i = 0; i < Nv_; ++i) {
                if (vx) vx[i] = t[2 * i];
                if (vy) vy[i] = t[2 * i + 1];
            }
        }
        if (Fx || Fy) {
            real_download(force_.p, t.data(), 2 * (size_t)Nv_);
            for (int i = 0; i < Nv_; ++i) {
                if (Fx) Fx[i] = t[2 * i];
                if (Fy) Fy[i] = t[2 * i + 1];
            }
        }
        if (A || P) {
            real_download(AP_.p, t.data(), 2 * (size_t)Nc_);
            for (int i = 0; i < Nc_; ++i) {
                if (A) A[i] = t[2 * i];
                if (P) P[i] = t[2 * i + 1];
            }
        }
    }
    void get_topology(int *vn, int *vcn, int *cvn, int *cv, int cap) {
        sync();
        vn_.download(vn, 3 * (size_t)Nv_);
        vcn_.download(vcn, 3 * (size_t)Nv_);
        std::vector<int> n(Nc_), c((size_t)vertex_cap_ * Nc_);
        cvn_.download(n.data(), Nc_);
        cv_.download(c.data(), c.size());
        for (int i = 0; i < Nc_; ++i) {
            cvn[i] = n[i];
            for (int k = 0; k < n[i] && k < cap; ++k) cv[(size_t)i * cap + k] = c[(size_t)k * Nc_ + i];
        }
    }
    // replace the topology after checking that it is consistent
    void set_topology(const int *vn, const int *vcn, const int *cvn, const int *cv, int cap) {
        if (!vn || !vcn || !cvn || !cv || cap <= 0) throw std::invalid_argument("topology arrays required");
        std::vector<int> vsys(Nv_), csys(Nc_);
        vsys_.download(vsys.data(), Nv_);
        csys_.download(csys.data(), Nc_);
        long long total = 0;
        for (int c = 0; c < Nc_; ++c) {
            if (cvn[c] < 3 || cvn[c] > cap) throw std::invalid_argument("cell vertex count outside [3, capacity]");
            total += cvn[c];
            for (int k = 0; k < cvn[c]; ++k) {
                const int v = cv[(size_t)c * cap + k];
                if (v < 0 || v >= Nv_ || vsys[v] != csys[c]) throw std::invalid_argument("cell vertex out of range");
            }
        }
        if (total != 3LL * Nv_) throw std::invalid_argument("cell vertex counts do not sum to 3 Nv");
        for (int v = 0; v < Nv_; ++v)
            for (int f = 0; f < 3; ++f) {
                const int u = vn[3 * v + f], c = vcn[3 * v + f];
                if (u < 0 || u >= Nv_ || vsys[u] != vsys[v] || c < 0 || c >= Nc_ || csys[c] != vsys[v])
                    throw std::invalid_argument("vertex neighbour out of range");
                bool in = false;
                for (int k = 0; k < cvn[c]; ++k) in = in || cv[(size_t)c * cap + k] == v;
                if (!in) throw std::invalid_argument("vertex missing from the list of one of its cells");
            }
        sync();
        int longest = 0;
        for (int c = 0; c < Nc_; ++c) longest = std::max(longest, cvn[c]);
        if (longest > vertex_cap_) grow_vertex_capacity((longest + 1) & ~1);
        vn_.upload(vn, 3 * (size_t)Nv_);
        vcn_.upload(vcn, 3 * (size_t)Nv_);
        cvn_.upload(cvn, Nc_);
        std::vector<int> sm((size_t)vertex_cap_ * Nc_, -1);
        for (int c = 0; c < Nc_; ++c)
            for (int k = 0; k < cvn[c]; ++k) sm[(size_t)k * Nc_ + c] = cv[(size_t)c * cap + k];
        cv_.upload(sm.data(), sm.size());
    }
    int cell_events(int count, const int *kind, const int *cell, const int *p1, const int *p2,
                    const double *new_director);
    // reference vertexModelBase::spatialSorting (vertexModelBase.cpp:236-242,
    // Simple2DCell.cpp:484-599): vertices along a Hilbert curve, cells by first
    // appearance around them, every per-vertex and per-cell array re-indexed on
    // the device; vperm[new] = old vertex, cperm[new] = old cell (either may be
    // NULL). The device arrays keep their addresses, so captured graphs stay valid.
    int spatial_sort(int *vperm_out, int *cperm_out) {
        cgpu_vertex_batch b = view();
        const size_t need = std::max(cgpu_vm_sort_workspace(Nv_, Nc_), cgpu_vm_permute_workspace(Nv_, Nc_));
        if (sort_work_.n < need) sort_work_.alloc(need);
        if (sort_perm_.n < (size_t)Nv_ + Nc_) sort_perm_.alloc((size_t)Nv_ + Nc_);
        int *vp = sort_perm_.p, *cp = sort_perm_.p + Nv_;
        int r = cgpu_vm_hilbert_order(&b, vp, cp, sort_work_.p, sort_work_.n, stream_);
        if (!r) r = cgpu_vm_permute(&b, vp, cp, sort_work_.p, sort_work_.n, stream_);
        if (!r) r = cgpu_vm_geometry(&b, stream_);
        if (r) return r;
        if (vperm_out)
            hip_check(hipMemcpyAsync(vperm_out, vp, sizeof(int) * (size_t)Nv_, hipMemcpyDeviceToHost, stream_),
                      "download");
        if (cperm_out)
            hip_check(hipMemcpyAsync(cperm_out, cp, sizeof(int) * (size_t)Nc_, hipMemcpyDeviceToHost, stream_),
                      "download");
        sync();
        return CGPU_OK;
    }
    void set_active(const int *active) {
        sync();
        std::vector<cgpu_sys_state> st(S_);
        st_.download(st.data(), S_);
        for (int s = 0; s < S_; ++s) st[s].active = active[s] ? 1 : 0;
        st_.upload(st.data(), S_);
    }
    void get_active(int *active) {
        sync();
        std::vector<cgpu_sys_state> st(S_);
        st_.download(st.data(), S_);
        for (int s = 0; s < S_; ++s) active[s] = st[s].active;
    }
    void get_counters(int s, long long *out) {
        sync();
        std::vector<cgpu_sys_state> st(S_);
        st_.download(st.data(), S_);
        out[0] = st[s].t1_events;
        out[1] = st[s].step;
        out[2] = st[s].error;
        out[3] = hp_[s].nv;
    }

   private:
    void drop_graph() {
        if (graph_exec_) {
            (void)hipGraphExecDestroy(graph_exec_);
            graph_exec_ = nullptr;
        }
    }
    void real_upload(void *dst, const double *src, size_t n) {
        if (prec_) {
            std::vector<float> f(src, src + n);
            hip_check(hipMemcpy(dst, f.data(), n * 4, hipMemcpyHostToDevice), "upload");
        } else {
            hip_check(hipMemcpy(dst, src, n * 8, hipMemcpyHostToDevice), "upload");
        }
    }
    void real_download(const void *src, double *dst, size_t n) {
        if (prec_) {
            std::vector<float> f(n);
            hip_check(hipMemcpy(f.data(), src, n * 4, hipMemcpyDeviceToHost), "download");
            for (size_t i = 0; i < n; ++i) dst[i] = f[i];
        } else {
            hip_check(hipMemcpy(dst, src, n * 8, hipMemcpyDeviceToHost), "download");
        }
    }
    int S_, Nv_ = 0, Nc_ = 0, prec_;
    size_t rsz_;
    int parity_ = 0, graph_parity_ = -1, graph_kind_ = -1, num_chunks_ = 0;
    bool noise_buffer_ = false;
    hipStream_t stream_ = nullptr;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t graph_exec_ = nullptr;
    std::vector<cgpu_vm_params> hp_;
    DevBuf<cgpu_vm_params> sys_;
    DevBuf<cgpu_sys_state> st_, st_call_;
    int vertex_cap_ = CGPU_VERTEX_CAP;
    DevBuf<int> vsys_, csys_, vn_, vcn_, cvn_, cv_, flag_count_, flag_list_, flag_cells_, flag_touched_;
    DevBuf<int> chunk_sys_, chunk_begin_, chunk_end_;
    DevBuf<unsigned long long> cell_min_;
    DevBuf<double> pos0_, pos1_, noise_;
    DevBuf<unsigned char> AP_, dAP_, APpref_, force_, theta_, mot_, sort_work_;
    DevBuf<int> sort_perm_;
};

// FIRE on a vertex batch: the vertices are the degrees of freedom and the T1
// pass is enforceTopology (vertexModelBase.cpp:93-106)
class VertexFire : public FireBase {
   public:
    explicit VertexFire(VertexBatch *vm)
        : FireBase(vm->num_systems(), vm->num_vertices(), vm->real_size(), vm->view().num_chunks), vm_(vm) {}

   protected:
    hipStream_t stream() override { return vm_->stream(); }
    cgpu_sys_state *device_state() override { return vm_->view().st; }
    int forces() override {
        cgpu_vertex_batch b = vm_->view();
        int r = cgpu_vm_geometry(&b, stream());
        return r ? r : cgpu_vm_forces(&b, stream());
    }
    int iterate() override {
        cgpu_vertex_batch b = vm_->view();
        const int r = cgpu_vm_fire_iteration(&b, fs_.p, vel_.p, dots_.p, stream());
        vm_->swap_parity();
        return r;
    }
    int apply_mix() override {
        cgpu_vertex_batch b = vm_->view();
        return cgpu_vm_fire_apply_mix(&b, fs_.p, vel_.p, stream());
    }
    void sync() override { vm_->sync(); }
    int status() override { return vm_->status(); }
    // a system stalled inside an iteration: its T1 pass on grown cell lists,
    // then the iteration's force half for it alone
    int recover_stalled() override {
        return vm_->finish_stalled([&](const std::vector<int> &, const std::vector<cgpu_sys_state> &) {
            cgpu_vertex_batch b = vm_->view();
            int r = cgpu_vm_fire_force_reduce(&b, fs_.p, vel_.p, dots_.p, vm_->stream());
            return r ? r : cgpu_vm_fire_park_converged(&b, fs_.p, vm_->stream());
        });
    }
    size_t dof() override { return vm_->num_vertices(); }
    size_t chunks() override { return vm_->view().num_chunks; }

   private:
    VertexBatch *vm_;
};


// Cell death (a T2 transition of a triangular cell) and division through two of
// its edges (reference vertexModelBase::cellDeath / cellDivision,
// vertexModelBase.cpp:905-1110, 1124-1385, with Simple2DActiveCell.cpp:118-147).
// At most one event per system. The state is gathered on the host and each
// system with an event is edited as the reference does: death merges the
// triangle's three vertices into its smallest one (which keeps its position),
// removes the other two and the cell; division inserts a vertex at the midpoint
// of the edges after list positions p1 and p2 of the cell and moves the vertices
// between them into a new cell appended to the system, with the parent's
// preferences and motility and director new_director (NaN: a uniform draw). The
// edited topology is brought back to the kernels' convention and the batch is
// laid out again.
// On the device (cgpu_vm_event_prepare, cgpu_vm_events_gather,
// cgpu_vm_events_surgery): the host sees one record per system.
int VertexBatch::cell_events(int count, const int *kind, const int *cell, const int *p1, const int *p2,
                             const double *new_director) {
    if (count <= 0) return CGPU_OK;
    if (!kind || !cell) throw std::invalid_argument("cell events need kinds and cells");
    std::vector<cgpu_vm_event> rec(S_);
    for (int s = 0; s < S_; ++s) {
        cgpu_vm_event &r = rec[s];
        std::memset(&r, 0, sizeof(r));
        r.kind = -1;
        r.old_voff = hp_[s].voff;
        r.old_nv = hp_[s].nv;
        r.old_coff = hp_[s].coff;
        r.old_nc = hp_[s].nc;
    }
    std::vector<int> ev_of(S_, -1);
    bool draw = false;
    for (int e = 0; e < count; ++e) {
        if (cell[e] < 0 || cell[e] >= Nc_) throw std::invalid_argument("cell index out of range");
        if (kind[e] != 0 && kind[e] != 1) throw std::invalid_argument("event kind is 0 (death) or 1 (division)");
        int s = 0;
        while (s + 1 < S_ && hp_[s + 1].coff <= cell[e]) ++s;
        if (ev_of[s] >= 0) throw std::invalid_argument("at most one event per system per call");
        ev_of[s] = e;
        if (kind[e] == 1 && (!p1 || !p2)) throw std::invalid_argument("a division needs two list positions");
        rec[s].kind = kind[e];
        rec[s].cell = cell[e] - hp_[s].coff;
        rec[s].p1 = kind[e] == 1 ? p1[e] : 0;
        rec[s].p2 = kind[e] == 1 ? p2[e] : 0;
        if (kind[e] == 1 && !(new_director && std::isfinite(new_director[e]))) draw = true;
    }
    sync();
    DevBuf<cgpu_vm_event> dev;
    dev.alloc(S_);
    dev.upload(rec.data(), S_);
    {
        cgpu_vertex_batch b = view();
        if (int r = cgpu_vm_event_prepare(&b, dev.p, stream_)) return r;
        sync();
    }
    dev.download(rec.data(), S_);
    static const char *why[] = {"", "cell death (T2) needs a triangular cell", "a system keeps at least 4 cells",
                                "cell death: not a simple triangle", "division list positions out of range",
                                "division: edge without a neighbouring cell"};
    for (const auto &r : rec)
        if (r.kind >= 0 && r.status != 0) throw std::invalid_argument(why[r.status >= 1 && r.status <= 5 ? r.status : 0]);
    std::vector<cgpu_sys_state> st(S_);
    if (draw) st_.download(st.data(), S_);
    // the new layout and the daughters' directors
    int longest = 0, voff = 0, coff = 0;
    std::vector<int> nv(S_), nc(S_);
    for (int s = 0; s < S_; ++s) {
        cgpu_vm_event &r = rec[s];
        const int dv = r.kind == 1 ? 2 : (r.kind == 0 ? -2 : 0), dc = r.kind == 1 ? 1 : (r.kind == 0 ? -1 : 0);
        nv[s] = r.old_nv + dv;
        nc[s] = r.old_nc + dc;
        r.new_voff = voff;
        r.new_coff = coff;
        voff += nv[s];
        coff += nc[s];
        longest = std::max(longest, r.grow);
        if (r.kind == 1) {
            const int e = ev_of[s];
            double d = new_director ? new_director[e] : NAN;
            if (!std::isfinite(d)) {
                const cgpu_vm_params &p = hp_[s];
                uint32_t ctr[4] = {(uint32_t)r.old_nc, (uint32_t)st[s].step, 4u, 0u};
                cgpu::philox4x32_10(ctr, (uint32_t)p.seed, (uint32_t)(p.seed >> 32));
                const uint64_t u = (((uint64_t)ctr[0] << 21) ^ (uint64_t)(ctr[1] >> 11)) & (((uint64_t)1 << 53) - 1);
                d = 6.283185307179586476925286766559 * ((double)u / 9007199254740992.0);
            }
            r.director = d;
        }
    }
    if (longest > EV_LIST_MAX) throw std::invalid_argument("a cell would exceed the longest list an event handles");
    if (longest > vertex_cap_) vertex_cap_ = (longest + 1) & ~1;
    dev.upload(rec.data(), S_);
    const size_t NV = (size_t)voff, NC = (size_t)coff, cap = (size_t)vertex_cap_;
    DevBuf<double> npos;
    DevBuf<int> nvn, nvcn, ncvn, ncv;
    DevBuf<unsigned char> npref, nth, nmot;
    npos.alloc(2 * NV);
    nvn.alloc(3 * NV);
    nvcn.alloc(3 * NV);
    ncvn.alloc(NC);
    ncv.alloc(cap * NC);
    npref.alloc(2 * NC * rsz_);
    nth.alloc(NC * rsz_);
    nmot.alloc(2 * NC * rsz_);
    {
        cgpu_vertex_batch b = view();
        if (int r = cgpu_vm_events_gather(&b, dev.p, (int)NV, (int)NC, (int)cap, npos.p, nvn.p, nvcn.p, ncvn.p, ncv.p,
                                          npref.p, nth.p, nmot.p, stream_))
            return r;
        sync();
    }
    for (int s = 0; s < S_; ++s) {
        hp_[s].nv = nv[s];
        hp_[s].nc = nc[s];
    }
    layout();
    pos0_.swap(npos);
    vn_.swap(nvn);
    vcn_.swap(nvcn);
    cvn_.swap(ncvn);
    cv_.swap(ncv);
    APpref_.swap(npref);
    theta_.swap(nth);
    mot_.swap(nmot);
    {
        cgpu_vertex_batch b = view();
        if (int r = cgpu_vm_events_surgery(&b, dev.p, stream_)) return r;
        sync();
    }
    return compute_forces();
}

}  // namespace cgpu_host

using cgpu_host::VertexBatch;

namespace {
template <typename F> int vguard(F &&f) {
    try {
        return f();
    } catch (const std::invalid_argument &e) {
        cgpu_host::set_last_error(e.what());
        return CGPU_ERR_ARGS;
    } catch (const std::exception &e) {
        cgpu_host::set_last_error(e.what());
        return CGPU_ERR_HIP;
    }
}
}  // namespace

extern "C" {
void *cgpu_vm_create(int S, const int *ncells, const double *L, const uint64_t *seed, const double *x, const double *y,
                     int precision, int device, int spv_initialize) {
    try {
        return new VertexBatch(S, ncells, L, seed, x, y, precision, device, spv_initialize != 0);
    } catch (...) {
        return nullptr;
    }
}
void cgpu_vm_destroy(void *h) { delete (VertexBatch *)h; }
int cgpu_vm_num_vertices(void *h) { return ((VertexBatch *)h)->num_vertices(); }
int cgpu_vm_set_params(void *h, int s, double KA, double KP, double dt, double mu, double T, double thr) {
    return vguard([&] { ((VertexBatch *)h)->set_params(s, KA, KP, dt, mu, T, thr); return 0; });
}
int cgpu_vm_set_prefs(void *h, const double *A0, const double *P0) {
    return vguard([&] { ((VertexBatch *)h)->set_prefs(A0, P0); return 0; });
}
int cgpu_vm_set_positions(void *h, const double *vx, const double *vy) {
    return vguard([&] { ((VertexBatch *)h)->set_positions(vx, vy); return 0; });
}
int cgpu_vm_set_noise_mode(void *h, int on) {
    return vguard([&] { ((VertexBatch *)h)->set_noise_mode(on != 0); return 0; });
}
int cgpu_vm_upload_noise(void *h, const double *eta) {
    return vguard([&] { ((VertexBatch *)h)->upload_noise(eta); return 0; });
}
int cgpu_vm_compute_forces(void *h) { return vguard([&] { return ((VertexBatch *)h)->compute_forces(); }); }
int cgpu_vm_enforce_topology(void *h) { return vguard([&] { return ((VertexBatch *)h)->enforce_topology(); }); }
int cgpu_vm_upload_cell_noise(void *h, const double *eta) {
    return vguard([&] {
        ((VertexBatch *)h)->upload_cell_noise(eta);
        return CGPU_OK;
    });
}
int cgpu_vm_set_motility(void *h, const double *v0, const double *Dr) {
    return vguard([&] {
        ((VertexBatch *)h)->set_motility(v0, Dr);
        return CGPU_OK;
    });
}
int cgpu_vm_set_directors(void *h, const double *theta) {
    return vguard([&] {
        ((VertexBatch *)h)->set_directors(theta);
        return CGPU_OK;
    });
}
int cgpu_vm_get_directors(void *h, double *theta) {
    return vguard([&] {
        ((VertexBatch *)h)->get_directors(theta);
        return CGPU_OK;
    });
}
int cgpu_vm_set_active(void *h, const int *active) {
    return vguard([&] {
        ((VertexBatch *)h)->set_active(active);
        return CGPU_OK;
    });
}
int cgpu_vm_get_active(void *h, int *active) {
    return vguard([&] {
        ((VertexBatch *)h)->get_active(active);
        return CGPU_OK;
    });
}
int cgpu_vm_cell_events(void *h, int count, const int *kind, const int *cell, const int *p1, const int *p2,
                        const double *new_director) {
    return vguard([&] { return ((VertexBatch *)h)->cell_events(count, kind, cell, p1, p2, new_director); });
}
int cgpu_vm_vertex_capacity(void *h) { return ((VertexBatch *)h)->vertex_capacity(); }
int cgpu_vm_spatial_sort(void *h, int *vperm, int *cperm) {
    return vguard([&] { return ((VertexBatch *)h)->spatial_sort(vperm, cperm); });
}
int cgpu_vm_view(void *h, cgpu_vertex_batch *out) {
    *out = ((VertexBatch *)h)->view();
    return CGPU_OK;
}
void *cgpu_vm_stream(void *h) { return (void *)((VertexBatch *)h)->stream(); }
int cgpu_vm_get_sizes(void *h, int *nc, int *nv) {
    VertexBatch &vm = *(VertexBatch *)h;
    for (int s = 0; s < vm.num_systems(); ++s) {
        if (nc) nc[s] = vm.params(s).nc;
        if (nv) nv[s] = vm.params(s).nv;
    }
    return vm.num_vertices();
}
int cgpu_vm_set_topology(void *h, const int *vn, const int *vcn, const int *cvn, const int *cv, int cap) {
    return vguard([&] {
        ((VertexBatch *)h)->set_topology(vn, vcn, cvn, cv, cap);
        return CGPU_OK;
    });
}
int cgpu_vm_spp_steps(void *h, int steps, int use_graph) {
    return vguard([&] { return ((VertexBatch *)h)->run_steps(1, steps, use_graph != 0); });
}
void *cgpu_vm_fire_create(void *vm) {
    try {
        return static_cast<cgpu_host::FireBase *>(new cgpu_host::VertexFire((VertexBatch *)vm));
    } catch (...) {
        return nullptr;
    }
}
int cgpu_vm_brownian_steps(void *h, int steps, int use_graph) {
    return vguard([&] { return ((VertexBatch *)h)->run_steps(0, steps, use_graph != 0); });
}
int cgpu_vm_get(void *h, double *vx, double *vy, double *Fx, double *Fy, double *A, double *P) {
    return vguard([&] { ((VertexBatch *)h)->get(vx, vy, Fx, Fy, A, P); return 0; });
}
int cgpu_vm_get_topology(void *h, int *vn, int *vcn, int *cvn, int *cv, int cap) {
    return vguard([&] { ((VertexBatch *)h)->get_topology(vn, vcn, cvn, cv, cap); return 0; });
}
int cgpu_vm_get_counters(void *h, int s, long long *out) {
    return vguard([&] { ((VertexBatch *)h)->get_counters(s, out); return 0; });
}
int cgpu_vm_status(void *h) { return vguard([&] { return ((VertexBatch *)h)->status(); }); }
}

// --------------------------------------------------------- database glue
// AVMDatabaseNetCDF::WriteState / ReadState (reference DatabaseNetCDFAVM.cpp:
// 192-271, 78-190) for one system of a vertex batch. The file holds vertex
// positions, forces, Vneighs and VertexCellNeighbors; reading rebuilds every
// cell's vertex list in counter-clockwise order around the mean of its
// vertices (as the reference reader does), puts each vertex's three cells in
// counter-clockwise order and takes Vneighs across from them.
namespace {

struct SysRange {
    int voff, nv, coff, nc;
    double L, dt;
};
SysRange sys_range(VertexBatch &vm, int s) {
    long long c[4];
    vm.get_counters(s, c);
    SysRange r{0, (int)c[3], 0, 0, 0.0, 0.0};
    for (int q = 0; q < s; ++q) {
        vm.get_counters(q, c);
        r.voff += (int)c[3];
    }
    r.coff = r.voff / 2;
    r.nc = r.nv / 2;
    const cgpu_vm_params &p = vm.params(s);
    r.L = p.L;
    r.dt = p.dt;
    return r;
}

}  // namespace

extern "C" int cgpu_db_avm_write_batch(void *db, void *h, int s, double time, int rec) {
    return vguard([&] {
        VertexBatch &vm = *(VertexBatch *)h;
        if (s < 0 || s >= vm.num_systems()) throw std::invalid_argument("no such system");
        const SysRange r = sys_range(vm, s);
        const int Nv = vm.num_vertices(), Nc = vm.num_cells(), cap = vm.vertex_capacity();
        std::vector<double> vx(Nv), vy(Nv), fx(Nv), fy(Nv), A(Nc), P(Nc), th(Nc);
        vm.get(vx.data(), vy.data(), fx.data(), fy.data(), A.data(), P.data());
        vm.get_directors(th.data());
        std::vector<int> vn(3 * (size_t)Nv), vcn(3 * (size_t)Nv), cvn(Nc), cv((size_t)Nc * cap);
        vm.get_topology(vn.data(), vcn.data(), cvn.data(), cv.data(), cap);
        const cgpu::Box box = cgpu::make_box(r.L);
        std::vector<int> lvn(3 * (size_t)r.nv), lvcn(3 * (size_t)r.nv);
        for (int k = 0; k < 3 * r.nv; ++k) {
            lvn[k] = vn[3 * (size_t)r.voff + k] - r.voff;
            lvcn[k] = vcn[3 * (size_t)r.voff + k] - r.coff;
        }
        // cell position: mean of its vertices (reference getCellPositionsCPU, vertexModelBase.cpp:401-429)
        std::vector<double> cx(r.nc), cy(r.nc), ct(r.nc, 0.0);
        std::vector<int> types(r.nc, 0);
        double q = 0.0;
        for (int c = 0; c < r.nc; ++c) {
            const int g = r.coff + c, n = cvn[g];
            const cgpu::dvec2 base{vx[cv[(size_t)g * cap]], vy[cv[(size_t)g * cap]]};
            cgpu::dvec2 pos{0.0, 0.0};
            for (int k = 1; k < n; ++k) {
                const int v = cv[(size_t)g * cap + k];
                cgpu::dvec2 d = cgpu::min_dist(box, cgpu::dvec2{vx[v], vy[v]}, base);
                pos.x += d.x;
                pos.y += d.y;
            }
            pos.x = pos.x / n + base.x;
            pos.y = pos.y / n + base.y;
            cgpu::put_in_box(box, pos);
            cx[c] = pos.x;
            cy[c] = pos.y;
            q += P[g] / std::sqrt(A[g]);
        }
        q /= r.nc;
        if (time < 0) {
            long long c[4];
            vm.get_counters(s, c);
            time = (double)c[1] * r.dt;
        }
        return cgpu_db_avm_write(db, rec, time, r.L, q, vx.data() + r.voff, vy.data() + r.voff, fx.data() + r.voff,
                                 fy.data() + r.voff, lvn.data(), lvcn.data(), types.data(), th.data() + r.coff,
                                 cx.data(), cy.data());
    });
}

extern "C" int cgpu_db_avm_read_batch(void *db, void *h, int s, int rec) {
    return vguard([&] {
        VertexBatch &vm = *(VertexBatch *)h;
        if (s < 0 || s >= vm.num_systems()) throw std::invalid_argument("no such system");
        const SysRange r = sys_range(vm, s);
        const int Nv = vm.num_vertices(), Nc = vm.num_cells(), cap = std::max(vm.vertex_capacity(), 64);
        std::vector<double> lx(r.nv), ly(r.nv), box4(4), ld(r.nc);
        std::vector<int> lvn(3 * (size_t)r.nv), lvcn(3 * (size_t)r.nv);
        int rc = cgpu_db_avm_read(db, rec, nullptr, box4.data(), lx.data(), ly.data(), nullptr, nullptr, lvn.data(),
                                  lvcn.data(), ld.data());
        if (rc) return rc;
        if (std::fabs(box4[0] - r.L) > 1e-12 * r.L || box4[1] != 0.0 || box4[2] != 0.0 || box4[3] != box4[0])
            throw std::invalid_argument("the record's box differs from the system's square box");
        const cgpu::Box box = cgpu::make_box(r.L);
        for (int k = 0; k < 3 * r.nv; ++k)
            if (lvn[k] < 0 || lvn[k] >= r.nv || lvcn[k] < 0 || lvcn[k] >= r.nc)
                throw std::invalid_argument("record topology out of range");
        std::vector<std::vector<int>> cells;
        cgpu_host::canonical_topology(r.nv, r.nc, lx.data(), ly.data(), box, cap, lvn, lvcn, cells);
        // into the batch's full arrays
        std::vector<double> vx(Nv), vy(Nv), th(Nc);
        vm.get(vx.data(), vy.data(), nullptr, nullptr, nullptr, nullptr);
        vm.get_directors(th.data());
        std::vector<int> vn(3 * (size_t)Nv), vcn(3 * (size_t)Nv), cvn(Nc), cv((size_t)Nc * cap, -1);
        vm.get_topology(vn.data(), vcn.data(), cvn.data(), cv.data(), cap);
        for (int v = 0; v < r.nv; ++v) {
            vx[r.voff + v] = lx[v];
            vy[r.voff + v] = ly[v];
            for (int f = 0; f < 3; ++f) {
                vn[3 * (size_t)(r.voff + v) + f] = lvn[3 * v + f] + r.voff;
                vcn[3 * (size_t)(r.voff + v) + f] = lvcn[3 * v + f] + r.coff;
            }
        }
        for (int c = 0; c < r.nc; ++c) {
            th[r.coff + c] = ld[c];
            cvn[r.coff + c] = (int)cells[c].size();
            for (size_t k = 0; k < cells[c].size(); ++k) cv[(size_t)(r.coff + c) * cap + k] = cells[c][k] + r.voff;
        }
        vm.set_topology(vn.data(), vcn.data(), cvn.data(), cv.data(), cap);
        vm.set_positions(vx.data(), vy.data());
        vm.set_directors(th.data());
        return vm.compute_forces();
    });
}
