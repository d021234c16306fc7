// VoronoiBatch: allocation, initial triangulation, step loop and host access of
// a device-resident batch of Voronoi systems, plus the object-layer C ABI.
#include "voronoi_batch.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <random>

#include "../cgpu_device.h"


namespace cgpu_host {

void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) throw std::runtime_error(std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
}

void reference_random_init(int n, uint32_t seed, double L, double *x, double *y, double *theta) {
    std::mt19937 gen(seed);
    for (int i = 0; i < n; ++i) {
        std::uniform_real_distribution<double> u(0.0, L);
        x[i] = u(gen);
        y[i] = u(gen);
    }
    for (int i = 0; i < n; ++i) {
        std::uniform_real_distribution<double> u(0.0, 2.0 * 3.14159265358979323846);
        theta[i] = u(gen);
    }
}

cgpu_sys_params default_sys_params(int n, double L, uint64_t seed) {
    if (n < 4) throw std::invalid_argument("each system needs at least 4 cells");
    cgpu_sys_params p;
    std::memset(&p, 0, sizeof(p));
    p.n = n;
    p.ncells = n;
    p.L = L > 0 ? L : std::sqrt((double)n);
    // reference cellListGPU::setGridSize(1.25): even number of bins per side
    int xs = (int)std::floor(p.L / 1.25);
    if (xs % 2 == 1) xs += 1;
    if (xs < 2) xs = 2;
    p.xsize = xs;
    p.boxsize = p.L / xs;
    p.KA = 1.0;
    p.KP = 1.0;
    p.dt = 0.01;
    p.mu = 1.0;
    p.T = 0.0;
    p.seed = seed;
    return p;
}

BatchTables batch_tables(std::vector<cgpu_sys_params> &hp) {
    BatchTables t;
    int off = 0, boff = 0;
    for (auto &p : hp) {
        if (p.n < 4) throw std::invalid_argument("each system needs at least 4 cells");
        if (p.xsize < 2) throw std::invalid_argument("each system needs at least 2 bins per side");
        if ((long long)off + p.n >= (1LL << 26))
            throw std::invalid_argument("a batch holds fewer than 2^26 cells (queue entries pack the cell index)");
        p.off = off;
        p.ncells = p.n;
        p.bin_off = boff;
        off += p.n;
        boff += p.xsize * p.xsize;
    }
    t.N = off;
    t.total_bins = boff;
    t.cell_sys.assign(t.N, 0);
    for (int s = 0; s < (int)hp.size(); ++s) {
        const cgpu_sys_params &p = hp[s];
        for (int i = 0; i < p.n; ++i) t.cell_sys[p.off + i] = s;
        // work chunks inside systems (cells and cell-list bins)
        for (int c = p.off; c < p.off + p.n; c += CGPU_CHUNK) {
            t.cs.push_back(s);
            t.cb.push_back(c);
            t.ce.push_back(std::min(c + CGPU_CHUNK, p.off + p.n));
        }
        t.sb.push_back((int)t.bs.size());
        const int nb = p.xsize * p.xsize;
        for (int u = 0; u < nb; u += CGPU_BIN_CHUNK) {
            t.bs.push_back(s);
            t.bb.push_back(p.bin_off + u);
            t.be.push_back(p.bin_off + std::min(u + CGPU_BIN_CHUNK, nb));
        }
    }
    t.sb.push_back((int)t.bs.size());
    return t;
}

VoronoiBatch::VoronoiBatch(int num_systems, const int *n, const double *L, const uint64_t *seed, const double *x,
                           const double *y, const double *theta, int precision, int device)
    : S_(num_systems), prec_(precision ? 1 : 0), device_(device) {
    if (S_ <= 0) throw std::invalid_argument("num_systems must be positive");
    hip_check(hipSetDevice(device_), "hipSetDevice");
    rsz_ = prec_ ? sizeof(float) : sizeof(double);
    hp_.resize(S_);
    for (int s = 0; s < S_; ++s) hp_[s] = default_sys_params(n[s], L ? L[s] : 0.0, seed ? seed[s] : 13377ull);
    std::vector<int> ns(n, n + S_);
    hip_check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
    sys_.alloc(S_);
    st_.alloc(S_);
    counters_.alloc(CGPU_COUNTERS);
    layout(ns);
    tag_.resize(N_);
    for (const auto &p : hp_)
        for (int i = 0; i < p.n; ++i) tag_[p.off + i] = i;
    std::vector<cgpu_sys_state> st(S_);
    std::memset(st.data(), 0, sizeof(cgpu_sys_state) * S_);
    for (auto &q : st) {
        q.active = 1;
        q.steps_left = -1;
    }
    st_.upload(st.data(), S_);

    // initial condition
    std::vector<double> hx(N_), hy(N_), ht(N_);
    for (int s = 0; s < S_; ++s) {
        const cgpu_sys_params &p = hp_[s];
        if (x && y) {
            for (int i = 0; i < p.n; ++i) {
                hx[p.off + i] = x[p.off + i];
                hy[p.off + i] = y[p.off + i];
                ht[p.off + i] = theta ? theta[p.off + i] : 0.0;
            }
        } else {
            reference_random_init(p.n, (uint32_t)p.seed, p.L, &hx[p.off], &hy[p.off], &ht[p.off]);
        }
    }
    std::vector<double> xy(2 * (size_t)N_);
    for (int i = 0; i < N_; ++i) {
        if (!std::isfinite(hx[i]) || !std::isfinite(hy[i]) || !std::isfinite(ht[i]))
            throw std::invalid_argument("non-finite initial position or director");
        xy[2 * i] = hx[i];
        xy[2 * i + 1] = hy[i];
    }
    pos0_.upload(xy.data(), 2 * (size_t)N_);
    real_upload(theta_.p, ht.data(), N_);
    std::vector<double> a0(N_, 1.0), p0(N_, 3.8), v0(N_, 0.05), dr(N_, 1.0);
    set_prefs(a0.data(), p0.data());
    // reference VoronoiQuadraticEnergy(n, A0, P0, reprod) -> setv0Dr(0.05, 1.0)
    {
        std::vector<double> mv(2 * (size_t)N_), vv(2 * (size_t)N_);
        for (int i = 0; i < N_; ++i) {
            mv[2 * i] = 0.05;
            mv[2 * i + 1] = 1.0;
            vv[2 * i] = 0.05 * std::cos(ht[i]);
            vv[2 * i + 1] = 0.05 * std::sin(ht[i]);
        }
        real_upload(mot_.p, mv.data(), 2 * (size_t)N_);
        real_upload(vel_.p, vv.data(), 2 * (size_t)N_);
    }
    int r = global_triangulation();
    if (r == CGPU_ERR_RING_CAPACITY)
        throw std::runtime_error("initial triangulation failed: a Delaunay degree exceeds the largest ring capacity (" +
                                 std::to_string(CGPU_RING_CAP_MAX) + ")");
    if (r != CGPU_OK) throw std::runtime_error("initial triangulation failed, status " + std::to_string(r));
}

VoronoiBatch::VoronoiBatch(Empty, int num_systems, int precision, int device)
    : S_(num_systems), prec_(precision ? 1 : 0), device_(device) {
    if (S_ <= 0) throw std::invalid_argument("num_systems must be positive");
    hip_check(hipSetDevice(device_), "hipSetDevice");
    rsz_ = prec_ ? sizeof(float) : sizeof(double);
    hp_.resize(S_);
    hip_check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
    sys_.alloc(S_);
    st_.alloc(S_);
    counters_.alloc(CGPU_COUNTERS);
}

VoronoiBatch *VoronoiBatch::merge(int count, VoronoiBatch *const *src, const int *sys) {
    if (count <= 0 || !src || !sys) throw std::invalid_argument("merge needs at least one system");
    for (int k = 0; k < count; ++k) {
        const VoronoiBatch *b = src[k];
        if (!b || sys[k] < 0 || sys[k] >= b->S_) throw std::invalid_argument("merge: no such system");
        if (b->prec_ != src[0]->prec_ || b->device_ != src[0]->device_)
            throw std::invalid_argument("merge: the sources differ in precision or device");
        if (b->noise_buffer_ || b->excl_.p || b->tension_on_)
            throw std::invalid_argument("merge: exclusions, noise buffers and line tension are not carried");
    }
    // host copies of every source, once each
    struct Host {
        std::vector<double> pos;
        std::vector<int> nn, nbr, oth;
        std::vector<unsigned char> AP, dAP, pref, mot, vel, theta, force;
        std::vector<cgpu_sys_state> st;
    };
    std::vector<const VoronoiBatch *> uniq;
    std::vector<Host> host;
    std::vector<int> which(count);
    for (int k = 0; k < count; ++k) {
        auto it = std::find(uniq.begin(), uniq.end(), src[k]);
        if (it == uniq.end()) {
            VoronoiBatch *b = src[k];
            b->sync();
            Host h;
            const size_t N = b->N_, rs = b->rsz_;
            h.pos.resize(2 * N);
            hip_check(hipMemcpy(h.pos.data(), b->parity_ ? b->pos1_.p : b->pos0_.p, 16 * N, hipMemcpyDeviceToHost),
                      "merge download");
            h.nn.resize(N);
            b->nn_.download(h.nn.data(), N);
            h.nbr.resize((size_t)b->ring_cap_ * N);
            b->nbr_.download(h.nbr.data(), h.nbr.size());
            h.oth.resize(h.nbr.size());
            b->oth_.download(h.oth.data(), h.oth.size());
            auto raw = [&](std::vector<unsigned char> &v, const DevBuf<unsigned char> &d, size_t per) {
                v.resize(per * N * rs);
                d.download(v.data(), v.size());
            };
            raw(h.AP, b->AP_, 2);
            raw(h.dAP, b->dAP_, 2);
            raw(h.pref, b->APpref_, 2);
            raw(h.mot, b->mot_, 2);
            raw(h.vel, b->vel_, 2);
            raw(h.theta, b->theta_, 1);
            raw(h.force, b->force_, 2);
            h.st.resize(b->S_);
            b->st_.download(h.st.data(), b->S_);
            uniq.push_back(b);
            host.push_back(std::move(h));
            it = uniq.end() - 1;
        }
        which[k] = (int)(it - uniq.begin());
    }

    std::unique_ptr<VoronoiBatch> m(new VoronoiBatch(Empty{}, count, src[0]->prec_, src[0]->device_));
    for (int k = 0; k < count; ++k) m->ring_cap_ = std::max(m->ring_cap_, src[k]->ring_cap_);
    std::vector<int> ns(count);
    for (int k = 0; k < count; ++k) {
        m->hp_[k] = src[k]->hp_[sys[k]];
        ns[k] = m->hp_[k].n;
    }
    m->layout(ns);  // offsets, tables, zeroed arrays, parameters
    const size_t N = m->N_, rs = m->rsz_, CAP = m->ring_cap_;
    std::vector<double> pos(2 * N);
    std::vector<int> nn(N), nbr(CAP * N, 0), oth(CAP * N, 0), tag(N);
    std::vector<unsigned char> AP(2 * N * rs), dAP(2 * N * rs), pref(2 * N * rs), mot(2 * N * rs), vel(2 * N * rs),
        theta(N * rs), force(2 * N * rs);
    std::vector<cgpu_sys_state> st(count);
    for (int k = 0; k < count; ++k) {
        const VoronoiBatch &b = *src[k];
        const Host &h = host[which[k]];
        const cgpu_sys_params &op = b.hp_[sys[k]];
        const int o = op.off, no = m->hp_[k].off, n = op.n;
        const size_t Nold = b.N_;
        std::copy(h.pos.begin() + 2 * (size_t)o, h.pos.begin() + 2 * (size_t)(o + n), pos.begin() + 2 * (size_t)no);
        for (int i = 0; i < n; ++i) {
            const int z = h.nn[o + i];
            nn[no + i] = z;
            tag[no + i] = b.tag_[o + i];
            for (int q = 0; q < z; ++q) {
                const int j = h.nbr[q * Nold + o + i], w = h.oth[q * Nold + o + i];
                if (j < o || j >= o + n || w < o || w >= o + n)
                    throw std::runtime_error("merge: a ring entry lies outside its system");
                nbr[q * N + no + i] = j - o + no;
                oth[q * N + no + i] = w - o + no;
            }
        }
        auto cp = [&](std::vector<unsigned char> &dst, const std::vector<unsigned char> &srcv, size_t per) {
            std::copy(srcv.begin() + per * rs * o, srcv.begin() + per * rs * (o + n), dst.begin() + per * rs * no);
        };
        cp(AP, h.AP, 2);
        cp(dAP, h.dAP, 2);
        cp(pref, h.pref, 2);
        cp(mot, h.mot, 2);
        cp(vel, h.vel, 2);
        cp(theta, h.theta, 1);
        cp(force, h.force, 2);
        st[k] = h.st[sys[k]];
    }
    hip_check(hipMemcpy(m->pos0_.p, pos.data(), 16 * N, hipMemcpyHostToDevice), "merge upload");
    m->nn_.upload(nn.data(), N);
    m->nbr_.upload(nbr.data(), nbr.size());
    m->oth_.upload(oth.data(), oth.size());
    m->AP_.upload(AP.data(), AP.size());
    m->dAP_.upload(dAP.data(), dAP.size());
    m->APpref_.upload(pref.data(), pref.size());
    m->mot_.upload(mot.data(), mot.size());
    m->vel_.upload(vel.data(), vel.size());
    m->theta_.upload(theta.data(), theta.size());
    m->force_.upload(force.data(), force.size());
    m->st_.upload(st.data(), count);
    m->tag_ = tag;
    m->shape_ = src[0]->shape_;
    return m.release();
}

VoronoiBatch::~VoronoiBatch() {
    // queued work may still read the graph and the buffers freed below
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    if (stream_) (void)hipStreamDestroy(stream_);
}

// Offsets, cell-list bins, work chunks and every per-cell device array for
// system sizes n (the per-system constants in hp_ are kept). Arrays start zeroed.
void VoronoiBatch::layout(const std::vector<int> &n) {
    for (int s = 0; s < S_; ++s) hp_[s].n = n[s];
    const BatchTables t = batch_tables(hp_);
    N_ = t.N;
    total_bins_ = t.total_bins;
    cell_sys_h_ = t.cell_sys;
    num_chunks_ = (int)t.cs.size();
    num_bin_chunks_ = (int)t.bs.size();
    chunk_sys_.alloc(t.cs.size());
    chunk_sys_.upload(t.cs.data(), t.cs.size());
    chunk_begin_.alloc(t.cb.size());
    chunk_begin_.upload(t.cb.data(), t.cb.size());
    chunk_end_.alloc(t.ce.size());
    chunk_end_.upload(t.ce.data(), t.ce.size());
    bchunk_sys_.alloc(t.bs.size());
    bchunk_sys_.upload(t.bs.data(), t.bs.size());
    bchunk_begin_.alloc(t.bb.size());
    bchunk_begin_.upload(t.bb.data(), t.bb.size());
    bchunk_end_.alloc(t.be.size());
    bchunk_end_.upload(t.be.data(), t.be.size());
    sys_bchunk_.alloc(t.sb.size());
    sys_bchunk_.upload(t.sb.data(), t.sb.size());
    bchunk_sum_.alloc(t.bs.size());
    cell_sys_.alloc(N_);
    nn_.alloc(N_);
    nbr_.alloc((size_t)ring_cap_ * N_);
    oth_.alloc((size_t)ring_cap_ * N_);
    repair_.alloc(N_);
    mark_.alloc(N_);
    flagged_.alloc(N_);
    nfq_.alloc(N_);
    gq_.alloc(N_);
    bin_count_.alloc(total_bins_);
    bin_start_.alloc((size_t)total_bins_ + 1);
    ld_tri_cap_ = ld_tri_capacity(N_);
    ld_tri_.alloc(3 * (size_t)ld_tri_cap_);
    bin_mark_wide_.alloc(total_bins_);
    cell_rmax_.alloc(N_);
    ld_region_cap_ = ld_region_capacity(N_);
    ld_regions_.alloc((size_t)ld_region_cap_ * CGPU_LD_REGION_BYTES);
    ld_part_.alloc(3 * ((size_t)N_ / 64 + 1));
    cell_bin_.alloc(N_);
    cell_rank_.alloc(N_);
    sorted_id_.alloc(N_);
    sorted_pos_.alloc(2 * (size_t)N_);
    pos0_.alloc(2 * (size_t)N_);
    pos1_.alloc(2 * (size_t)N_);
    AP_.alloc(2 * N_ * rsz_);
    dAP_.alloc(2 * N_ * rsz_);
    APpref_.alloc(2 * N_ * rsz_);
    mot_.alloc(2 * N_ * rsz_);
    vel_.alloc(2 * N_ * rsz_);
    theta_.alloc(N_ * rsz_);
    force_.alloc(2 * N_ * rsz_);
    device_bytes_ = sys_.n * sizeof(cgpu_sys_params) + st_.n * sizeof(cgpu_sys_state) +
                    sizeof(int) * (cell_sys_.n + nn_.n + nbr_.n + oth_.n + repair_.n + mark_.n + flagged_.n +
                                   nfq_.n + gq_.n + counters_.n + bin_start_.n + cell_bin_.n + cell_rank_.n +
                                   sorted_id_.n + bin_count_.n + ld_tri_.n + bin_mark_wide_.n) +
                    sizeof(float) * cell_rmax_.n + ld_regions_.n + sizeof(unsigned long long) * ld_part_.n +
                    sizeof(double) * (pos0_.n + pos1_.n + sorted_pos_.n) +
                    (AP_.n + dAP_.n + APpref_.n + mot_.n + vel_.n + theta_.n + force_.n);
    cell_sys_.upload(cell_sys_h_.data(), N_);
    upload_params();
    parity_ = 0;
    drop_graph();
}

void VoronoiBatch::upload_params() { sys_.upload(hp_.data(), S_); }

void VoronoiBatch::real_upload(void *dst, const double *src, size_t count) {
    if (prec_) {
        std::vector<float> f(count);
        for (size_t i = 0; i < count; ++i) f[i] = (float)src[i];
        hip_check(hipMemcpy(dst, f.data(), count * sizeof(float), hipMemcpyHostToDevice), "upload");
    } else {
        hip_check(hipMemcpy(dst, src, count * sizeof(double), hipMemcpyHostToDevice), "upload");
    }
}
void VoronoiBatch::real_download(const void *src, double *dst, size_t count) {
    if (prec_) {
        std::vector<float> f(count);
        hip_check(hipMemcpy(f.data(), src, count * sizeof(float), hipMemcpyDeviceToHost), "download");
        for (size_t i = 0; i < count; ++i) dst[i] = f[i];
    } else {
        hip_check(hipMemcpy(dst, src, count * sizeof(double), hipMemcpyDeviceToHost), "download");
    }
}

cgpu_voronoi_batch VoronoiBatch::view() const {
    cgpu_voronoi_batch b;
    std::memset(&b, 0, sizeof(b));
    b.num_systems = S_;
    b.num_cells = N_;
    b.total_bins = total_bins_;
    b.precision = prec_;
    b.sys = sys_.p;
    b.st = st_.p;
    b.cell_sys = cell_sys_.p;
    b.pos = parity_ ? pos1_.p : pos0_.p;
    b.pos_next = parity_ ? pos0_.p : pos1_.p;
    b.nn = nn_.p;
    b.nbr = nbr_.p;
    b.oth = oth_.p;
    b.AP = AP_.p;
    b.dAP = dAP_.p;
    b.APpref = APpref_.p;
    b.mot = mot_.p;
    b.vel = vel_.p;
    b.theta = theta_.p;
    b.force = force_.p;
    b.excl = excl_.p;
    b.ext_force = ext_.p;
    b.bin_count = bin_count_.p;
    b.bin_start = bin_start_.p;
    b.cell_bin = cell_bin_.p;
    b.cell_rank = cell_rank_.p;
    b.sorted_id = sorted_id_.p;
    b.sorted_pos = sorted_pos_.p;
    b.repair = repair_.p;
    b.mark = mark_.p;
    b.flagged = flagged_.p;
    b.nfq = nfq_.p;
    b.gq = gq_.p;
    b.counters = counters_.p;
    b.noise = noise_buffer_ ? noise_.p : nullptr;
    b.chunk_sys = chunk_sys_.p;
    b.chunk_begin = chunk_begin_.p;
    b.chunk_end = chunk_end_.p;
    b.bchunk_sys = bchunk_sys_.p;
    b.bchunk_begin = bchunk_begin_.p;
    b.bchunk_end = bchunk_end_.p;
    b.sys_bchunk = sys_bchunk_.p;
    b.bchunk_sum = bchunk_sum_.p;
    b.num_chunks = num_chunks_;
    b.num_bin_chunks = num_bin_chunks_;
    b.cell_type = tension_on_ ? cell_type_.p : nullptr;
    b.tension = tension_.p;
    b.num_types = ntypes_;
    b.ring_cap = ring_cap_;
    b.ld_tri = ld_tri_.p;
    b.ld_tri_cap = ld_tri_cap_;
    b.bin_mark_wide = bin_mark_wide_.p;
    b.cell_rmax = cell_rmax_.p;
    b.ld_regions = ld_regions_.p;
    b.ld_region_cap = ld_region_cap_;
    b.ld_part = ld_part_.p;
    b.shape = shape_;
    return b;
}

void VoronoiBatch::drop_graph() {
    if (graph_exec_) {
        (void)hipGraphExecDestroy(graph_exec_);
        graph_exec_ = nullptr;
    }
}

void VoronoiBatch::set_shape(const cgpu_launch_shape &sh) {
    sync();
    shape_ = sh;
    drop_graph();  // the captured launches carry the old shapes
}

void VoronoiBatch::set_params(int s, double KA, double KP, double dt, double mu, double T) {
    sync();  // host copies below must not race the batch stream
    int lo = s < 0 ? 0 : s, hi = s < 0 ? S_ : s + 1;
    for (int q = lo; q < hi; ++q) {
        hp_[q].KA = KA;
        hp_[q].KP = KP;
        hp_[q].dt = dt;
        hp_[q].mu = mu;
        hp_[q].T = T;
    }
    upload_params();
    drop_graph();
    // dAP depends on KA, KP
    cgpu_voronoi_batch b = view();
    cgpu_test_and_geometry(&b, 0, stream_);
    sync();
}

void VoronoiBatch::set_prefs(const double *A0, const double *P0) {
    sync();  // host copies below must not race the batch stream
    std::vector<double> v(2 * (size_t)N_);
    for (int i = 0; i < N_; ++i) {
        v[2 * i] = A0[i];
        v[2 * i + 1] = P0[i];
    }
    real_upload(APpref_.p, v.data(), v.size());
    if (nn_.p && hp_.size()) {
        cgpu_voronoi_batch b = view();
        // geometry only valid once a triangulation exists (nn > 0)
        std::vector<int> probe(1, 0);
        nn_.download(probe.data(), 1);
        if (probe[0] > 0) {
            cgpu_test_and_geometry(&b, 0, stream_);
            sync();
        }
    }
}

// reference Simple2DActiveCell::setv0Dr (.cpp:76-92) resets v = v0 (cos, sin)(theta)
void VoronoiBatch::set_motility(const double *v0, const double *Dr) {
    sync();  // host copies below must not race the batch stream
    std::vector<double> th(N_);
    real_download(theta_.p, th.data(), N_);
    std::vector<double> mv(2 * (size_t)N_), vv(2 * (size_t)N_);
    for (int i = 0; i < N_; ++i) {
        mv[2 * i] = v0[i];
        mv[2 * i + 1] = Dr[i];
        vv[2 * i] = v0[i] * std::cos(th[i]);
        vv[2 * i + 1] = v0[i] * std::sin(th[i]);
    }
    real_upload(mot_.p, mv.data(), mv.size());
    real_upload(vel_.p, vv.data(), vv.size());
}

// reference voronoiModelBase::setExclusions (.cpp:89-107): excluded cells get
// v0 = Dr = 0 and their force is reported as an external force
void VoronoiBatch::set_exclusions(const int *ex) {
    sync();  // host copies below must not race the batch stream
    if (!excl_.p) {
        excl_.alloc(N_);
        ext_.alloc(2 * N_ * rsz_);
    }
    excl_.upload(ex, N_);
    std::vector<double> mv(2 * (size_t)N_);
    real_download(mot_.p, mv.data(), mv.size());
    for (int i = 0; i < N_; ++i)
        if (ex[i]) mv[2 * i] = mv[2 * i + 1] = 0.0;
    real_upload(mot_.p, mv.data(), mv.size());
    drop_graph();
}

// a system in matrix mode may only hold types below the matrix size (the
// kernels index the matrix with them)
void VoronoiBatch::update_tension_flag() {
    bool any = false;
    for (const auto &p : hp_) {
        any = any || p.tension != 0;
        if (p.tension == 2 && !h_type_.empty())
            for (int i = p.off; i < p.off + p.n; ++i)
                if (h_type_[i] >= ntypes_) throw std::invalid_argument("cell type outside the tension matrix");
    }
    tension_on_ = any && !h_type_.empty();
    drop_graph();
}

void VoronoiBatch::set_cell_types(const int *types) {
    sync();
    if (!types) {
        cell_type_.release();
        h_type_.clear();
    } else {
        for (int i = 0; i < N_; ++i)
            if (types[i] < 0) throw std::invalid_argument("negative cell type");
        std::vector<int> keep(h_type_);
        h_type_.assign(types, types + N_);
        try {
            update_tension_flag();
        } catch (...) {
            h_type_ = keep;
            throw;
        }
        if (!cell_type_.p) cell_type_.alloc(N_);
        cell_type_.upload(types, N_);
    }
    update_tension_flag();
}

void VoronoiBatch::set_tension(int s, int mode, double gamma) {
    if (mode < 0 || mode > 2) throw std::invalid_argument("tension mode must be 0, 1 or 2");
    if (mode == 2 && !tension_.p) throw std::invalid_argument("tension matrix not set");
    sync();
    std::vector<cgpu_sys_params> keep(hp_);
    int lo = s < 0 ? 0 : s, hi = s < 0 ? S_ : s + 1;
    for (int q = lo; q < hi; ++q) {
        hp_[q].tension = mode;
        hp_[q].gamma = gamma;
    }
    try {
        update_tension_flag();
    } catch (...) {
        hp_ = keep;
        throw;
    }
    sys_.upload(hp_.data(), S_);
}

void VoronoiBatch::set_tension_matrix(int ntypes, const double *g) {
    if (ntypes <= 0 || !g) throw std::invalid_argument("tension matrix needs ntypes > 0");
    sync();
    const int keep = ntypes_;
    ntypes_ = ntypes;
    try {
        update_tension_flag();
    } catch (...) {
        ntypes_ = keep;
        throw;
    }
    tension_.alloc((size_t)ntypes * ntypes);
    tension_.upload(g, (size_t)ntypes * ntypes);
}

void VoronoiBatch::set_positions(const double *x, const double *y, bool retri) {
    sync();  // host copies below must not race the batch stream
    std::vector<double> xy(2 * (size_t)N_);
    for (int i = 0; i < N_; ++i) {
        if (!std::isfinite(x[i]) || !std::isfinite(y[i])) throw std::invalid_argument("non-finite position");
        xy[2 * i] = x[i];
        xy[2 * i + 1] = y[i];
    }
    double *p = parity_ ? pos1_.p : pos0_.p;
    hip_check(hipMemcpy(p, xy.data(), xy.size() * sizeof(double), hipMemcpyHostToDevice), "set_positions");
    if (retri) global_triangulation();
}

void VoronoiBatch::get_del_other(int *oth, int cap) {
    sync();
    std::vector<int> hn(N_), ho((size_t)ring_cap_ * N_);
    nn_.download(hn.data(), N_);
    oth_.download(ho.data(), ho.size());
    for (int i = 0; i < N_; ++i)
        for (int k = 0; k < hn[i] && k < cap; ++k) oth[(size_t)i * cap + k] = ho[(size_t)k * N_ + i];
}

void VoronoiBatch::set_topology(const int *nn, const int *nbr, int cap, const int *oth) {
    if (!nn || !nbr || cap <= 0) throw std::invalid_argument("set_topology needs nn, nbr and cap > 0");
    int zmax = 0;
    for (int i = 0; i < N_; ++i) zmax = std::max(zmax, nn[i]);
    if (zmax > ring_cap_ && zmax <= CGPU_RING_CAP_MAX && zmax <= cap) grow_ring_capacity(zmax);
    std::vector<int> ring((size_t)ring_cap_ * N_, 0);
    std::vector<int> deg(N_);
    std::vector<cgpu_sys_state> st(S_);
    for (int s = 0; s < S_; ++s) {
        const cgpu_sys_params &p = hp_[s];
        long long sum = 0;
        int mx = 0;
        for (int i = p.off; i < p.off + p.n; ++i) {
            const int z = nn[i];
            if (z < 3 || z > ring_cap_ || z > cap)
                throw std::invalid_argument("degree of cell " + std::to_string(i) + " outside [3, capacity]");
            for (int k = 0; k < z; ++k) {
                const int j = nbr[(size_t)i * cap + k];
                if (j < p.off || j >= p.off + p.n || j == i)
                    throw std::invalid_argument("neighbour of cell " + std::to_string(i) + " outside its system");
                for (int q = 0; q < k; ++q)
                    if (nbr[(size_t)i * cap + q] == j)
                        throw std::invalid_argument("repeated neighbour of cell " + std::to_string(i));
                ring[(size_t)k * N_ + i] = j;
                if (oth) {
                    const int o = oth[(size_t)i * cap + k];
                    if (o < p.off || o >= p.off + p.n)
                        throw std::invalid_argument("delOther of cell " + std::to_string(i) + " outside its system");
                }
            }
            deg[i] = z;
            sum += z;
            mx = std::max(mx, z);
        }
        if (sum != 6LL * p.n) throw std::invalid_argument("degree sum of system " + std::to_string(s) + " is not 6N");
        st[s].neigh_max = (mx % 2 == 0) ? mx + 2 : mx + 1;  // voronoiModelBase.cpp:309-312
    }
    sync();
    std::vector<cgpu_sys_state> cur(S_);
    st_.download(cur.data(), S_);
    for (int s = 0; s < S_; ++s) cur[s].neigh_max = st[s].neigh_max;
    st_.upload(cur.data(), S_);
    nn_.upload(deg.data(), N_);
    nbr_.upload(ring.data(), ring.size());
    cgpu_voronoi_batch b = view();
    if (oth) {
        // the caller's delOther (e.g. another batch's, including values a
        // repair kept because a neighbour's ring lacked the edge)
        std::vector<int> o((size_t)ring_cap_ * N_, 0);
        for (int i = 0; i < N_; ++i)
            for (int k = 0; k < deg[i]; ++k) o[(size_t)k * N_ + i] = oth[(size_t)i * cap + k];
        oth_.upload(o.data(), o.size());
        launch_geometry_only(b);
    } else {
        int r = cgpu_derive_topology(&b, stream_);
        if (r) throw std::runtime_error("derive topology failed");
    }
    sync();
}

void VoronoiBatch::launch_geometry_only(const cgpu_voronoi_batch &b) {
    if (cgpu_test_and_geometry(&b, 0, stream_)) throw std::runtime_error("geometry failed");
}

void VoronoiBatch::set_directors(const double *theta) {
    sync();
    real_upload(theta_.p, theta, N_);
}

// stored self-propulsion velocities (the reference keeps them between steps:
// the displacement of a step uses the velocity of the step before)
void VoronoiBatch::set_velocities(const double *vx, const double *vy) {
    sync();
    std::vector<double> v(2 * (size_t)N_);
    for (int i = 0; i < N_; ++i) {
        if (!std::isfinite(vx[i]) || !std::isfinite(vy[i])) throw std::invalid_argument("non-finite velocity");
        v[2 * i] = vx[i];
        v[2 * i + 1] = vy[i];
    }
    real_upload(vel_.p, v.data(), v.size());
}

// integrator step counters (reference Simulation::setCurrentTimestep /
// Timestep): the counter also keys the in-kernel noise of the next step
void VoronoiBatch::set_step_counters(const int *step) {
    sync();
    std::vector<cgpu_sys_state> st(S_);
    st_.download(st.data(), S_);
    for (int s = 0; s < S_; ++s) {
        if (step[s] < 0) throw std::invalid_argument("negative step counter");
        st[s].step = step[s];
    }
    st_.upload(st.data(), S_);
}

void VoronoiBatch::set_noise_mode(bool use_buffer) {
    sync();  // host copies below must not race the batch stream
    noise_buffer_ = use_buffer;
    if (use_buffer && !noise_.p) noise_.alloc(2 * (size_t)N_);
    drop_graph();
}
void VoronoiBatch::upload_noise(const double *eta, int count) {
    sync();  // host copies below must not race the batch stream
    if (!noise_.p) noise_.alloc(2 * (size_t)N_);
    noise_.upload(eta, std::min<size_t>(count, 2 * (size_t)N_));
}

void VoronoiBatch::set_active(const int *active, const int *steps_left) {
    sync();  // host copies below must not race the batch stream
    std::vector<cgpu_sys_state> st(S_);
    st_.download(st.data(), S_);
    for (int s = 0; s < S_; ++s) {
        if (active) st[s].active = active[s];
        if (steps_left) st[s].steps_left = steps_left[s];
    }
    st_.upload(st.data(), S_);
}

// With growth: a star beyond the ring capacity stalls its system; the capacity
// then doubles (up to CGPU_RING_CAP_MAX) and the whole triangulation is redone,
// as the reference's globalTriangulationCGAL grows neighMax and its arrays
// (voronoiModelBase.cpp:309-312).
int VoronoiBatch::global_triangulation() {
    for (;;) {
        cgpu_voronoi_batch b = view();
        int r = cgpu_global_triangulation(&b, stream_);
        if (r) return r;
        std::vector<int> stalled;
        if (!take_stalled(stalled)) return status();
        if (ring_cap_ >= CGPU_RING_CAP_MAX) return CGPU_ERR_RING_CAPACITY;
        grow_ring_capacity(std::min(CGPU_RING_CAP_MAX, 2 * ring_cap_));
        resume(stalled);
    }
}

// Ring storage of cap slots per cell. Slot-major rows keep their place, so the
// old rows are copied as one block; a captured step graph is dropped (it holds
// the old addresses).
void VoronoiBatch::grow_ring_capacity(int cap) {
    if (cap <= ring_cap_) return;
    if (cap > CGPU_RING_CAP_MAX) throw std::invalid_argument("ring capacity above CGPU_RING_CAP_MAX");
    sync();
    for (DevBuf<int> *a : {&nbr_, &oth_}) {
        DevBuf<int> g;
        g.alloc((size_t)cap * N_);
        hip_check(hipMemcpy(g.p, a->p, sizeof(int) * (size_t)ring_cap_ * N_, hipMemcpyDeviceToDevice), "grow ring");
        a->swap(g);
    }
    device_bytes_ += 2 * sizeof(int) * (size_t)(cap - ring_cap_) * N_;
    ring_cap_ = cap;
    drop_graph();
}

// Systems stalled by a star beyond the ring capacity (device flag counters[4]
// = 2; their error CGPU_ERR_RING_CAPACITY and active 0). False when none.
bool VoronoiBatch::take_stalled(std::vector<int> &stalled) {
    sync();
    int c4 = 0;
    hip_check(hipMemcpy(&c4, counters_.p + 4, sizeof(int), hipMemcpyDeviceToHost), "stall flag");
    stalled.clear();
    if (c4 != 2) return false;
    std::vector<cgpu_sys_state> st(S_);
    st_.download(st.data(), S_);
    for (int s = 0; s < S_; ++s)
        if (st[s].error == CGPU_ERR_RING_CAPACITY && !st[s].active) stalled.push_back(s);
    return !stalled.empty();
}

// clears the stall of the given systems (error, active) and the device flag
void VoronoiBatch::resume(const std::vector<int> &stalled) {
    std::vector<cgpu_sys_state> st(S_);
    st_.download(st.data(), S_);
    for (int s : stalled) {
        st[s].error = 0;
        st[s].active = 1;
    }
    st_.upload(st.data(), S_);
    int zero = 0;
    hip_check(hipMemcpy(counters_.p + 4, &zero, sizeof(int), hipMemcpyHostToDevice), "stall flag");
}

// After a call of 'steps' steps of kind (0 SPV, 1 Brownian): systems that
// stalled on the ring capacity are finished as the reference finishes a repair
// that exceeds neighMax, by a global retriangulation with grown storage
// (voronoiModelBase.cpp:456-460, 309-312), at their current positions (the
// stalled step's move is done); their step counter then counts that step and
// they run the steps they still owe alone, with the other systems parked.
int VoronoiBatch::finish_stalled(int steps, int kind, const std::vector<int> &step_before) {
    std::vector<int> stalled;
    while (take_stalled(stalled)) {
        if (ring_cap_ >= CGPU_RING_CAP_MAX) return CGPU_ERR_RING_CAPACITY;
        grow_ring_capacity(std::min(CGPU_RING_CAP_MAX, 2 * ring_cap_));
        std::vector<cgpu_sys_state> st(S_);
        st_.download(st.data(), S_);
        const std::vector<cgpu_sys_state> saved = st;
        // only the stalled systems take part in the retriangulation and the rest
        for (int s = 0; s < S_; ++s) st[s].active = 0;
        for (int s : stalled) {
            st[s].error = 0;
            st[s].active = 1;
            st[s].step += 1;  // the stalled step, completed below
        }
        st_.upload(st.data(), S_);
        int zero = 0;
        hip_check(hipMemcpy(counters_.p + 4, &zero, sizeof(int), hipMemcpyHostToDevice), "stall flag");
        int r = retriangulate_active();
        if (r) return r;
        // steps still owed (a countdown system keeps its countdown instead)
        st_.download(st.data(), S_);
        int owe = 0;
        for (int s : stalled) {
            if (saved[s].steps_left >= 0) {
                st[s].steps_left = saved[s].steps_left > 0 ? saved[s].steps_left - 1 : 0;
                if (st[s].steps_left == 0 && saved[s].steps_left > 0) st[s].active = 0;
                continue;
            }
            const int left = steps - (st[s].step - step_before[s]);
            if (left > 0) {
                st[s].steps_left = left;
                owe = std::max(owe, left);
            } else {
                st[s].active = 1;
            }
        }
        st_.upload(st.data(), S_);
        for (int k = 0; k < owe; ++k) {
            r = step_launch(kind);
            if (r) return r;
        }
        // restore: the other systems as they were, the stalled ones running on
        sync();
        st_.download(st.data(), S_);
        for (int s = 0; s < S_; ++s) {
            const bool was_stalled = std::find(stalled.begin(), stalled.end(), s) != stalled.end();
            if (!was_stalled) {
                st[s].active = saved[s].active;
                st[s].steps_left = saved[s].steps_left;
            } else if (saved[s].steps_left < 0 && st[s].error == 0) {
                st[s].active = 1;
                st[s].steps_left = -1;
            }
        }
        st_.upload(st.data(), S_);
    }
    return CGPU_OK;
}

// global retriangulation of the active systems only (the others keep their
// rings): every cell of an active system is queued for the global star pass
int VoronoiBatch::retriangulate_active() {
    cgpu_voronoi_batch b = view();
    int r = cgpu_global_triangulation_active(&b, stream_);
    if (r) return r;
    sync();
    return CGPU_OK;
}

// reference voronoiModelBase::enforceTopology (voronoiModelBase.cpp:686-739) on
// the current positions: cell list, circumcircle test, repair; the step
// counters do not advance
int VoronoiBatch::enforce_topology() {
    cgpu_voronoi_batch b = view();
    int r = cgpu_step_begin(&b, stream_);
    if (!r) r = cgpu_compute_cell_list(&b, 0, stream_);
    if (!r) r = cgpu_test_and_geometry(&b, 1, stream_);
    if (!r) r = cgpu_repair_topology(&b, stream_);
    if (r) return r;
    // a stalled system is finished by a global retriangulation with grown storage
    std::vector<int> stalled;
    while (take_stalled(stalled)) {
        if (ring_cap_ >= CGPU_RING_CAP_MAX) return CGPU_ERR_RING_CAPACITY;
        grow_ring_capacity(std::min(CGPU_RING_CAP_MAX, 2 * ring_cap_));
        std::vector<cgpu_sys_state> st(S_);
        st_.download(st.data(), S_);
        const std::vector<cgpu_sys_state> saved = st;
        for (int s = 0; s < S_; ++s) st[s].active = 0;
        for (int s : stalled) {
            st[s].error = 0;
            st[s].active = 1;
        }
        st_.upload(st.data(), S_);
        int zero = 0;
        hip_check(hipMemcpy(counters_.p + 4, &zero, sizeof(int), hipMemcpyHostToDevice), "stall flag");
        if ((r = retriangulate_active())) return r;
        st_.download(st.data(), S_);
        for (int s = 0; s < S_; ++s) st[s].active = saved[s].active || (st[s].error == 0 && std::find(stalled.begin(), stalled.end(), s) != stalled.end());
        st_.upload(st.data(), S_);
    }
    return status();
}

int VoronoiBatch::compute_forces() {
    cgpu_voronoi_batch b = view();
    int r = cgpu_voronoi_forces(&b, stream_);
    if (r) return r;
    return sync();
}

int VoronoiBatch::step_launch(int kind) {
    cgpu_voronoi_batch b = view();
    int r = kind == 0 ? cgpu_spv_step(&b, stream_) : cgpu_brownian_step(&b, stream_);
    parity_ ^= 1;
    return r;
}

int VoronoiBatch::spv_steps(int steps, bool use_graph) {
    if (steps <= 0) return CGPU_OK;
    mark_call_start();
    int r = launch_spv_steps(steps, use_graph);
    return r ? r : after_steps(steps, 0);
}

// the step counters at the start of a stepping call, copied on the device
// (no host wait), for the steps a stalled system still owes
void VoronoiBatch::mark_call_start() {
    if (st_call_.n != st_.n) st_call_.alloc(st_.n);
    hip_check(hipMemcpyAsync(st_call_.p, st_.p, sizeof(cgpu_sys_state) * st_.n, hipMemcpyDeviceToDevice, stream_),
              "state copy");
}

int VoronoiBatch::after_steps(int steps, int kind) {
    std::vector<int> stalled;
    if (!take_stalled(stalled)) return CGPU_OK;
    std::vector<cgpu_sys_state> before(S_);
    st_call_.download(before.data(), S_);
    std::vector<int> step_before(S_);
    for (int s = 0; s < S_; ++s) step_before[s] = before[s].step;
    return finish_stalled(steps, kind, step_before);
}

int VoronoiBatch::launch_spv_steps(int steps, bool use_graph) {
    if (!use_graph || noise_buffer_) {
        for (int k = 0; k < steps; ++k) {
            int r = step_launch(0);
            if (r) return r;
        }
        return CGPU_OK;
    }
    // a graph of two steps returns to the same position buffer parity
    if (graph_exec_ && graph_parity_ != parity_) {
        int r = step_launch(0);
        if (r) return r;
        --steps;
    }
    if (!graph_exec_ && steps >= 2) {
        // capture only records the launches; the two parity flips cancel
        hip_check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "capture");
        graph_parity_ = parity_;
        step_launch(0);
        step_launch(0);
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
    if (steps == 1) return step_launch(0);
    return CGPU_OK;
}

int VoronoiBatch::spv_steps_timed(int steps, double *ms_out) {
    hipEvent_t ev[5];
    for (auto &e : ev) hip_check(hipEventCreate(&e), "event");
    double acc[4] = {0, 0, 0, 0};
    int r = CGPU_OK;
    for (int k = 0; k < steps && r == CGPU_OK; ++k) {
        cgpu_voronoi_batch b = view();
        cgpu_voronoi_batch bn = b;  // stages after the move work on the new positions
        bn.pos = b.pos_next;
        bn.pos_next = b.pos;
        r = cgpu_step_begin(&b, stream_);
        hip_check(hipEventRecord(ev[0], stream_), "event record");
        if (!r) r = cgpu_spv_force_eom_move(&b, stream_);
        hip_check(hipEventRecord(ev[1], stream_), "event record");
        if (!r) r = cgpu_compute_cell_list(&b, 1, stream_);
        hip_check(hipEventRecord(ev[2], stream_), "event record");
        if (!r) r = cgpu_test_and_geometry(&bn, 1, stream_);
        hip_check(hipEventRecord(ev[3], stream_), "event record");
        if (!r) r = cgpu_repair_topology(&bn, stream_);
        hip_check(hipEventRecord(ev[4], stream_), "event record");
        if (!r) r = cgpu_step_end(&b, stream_);
        parity_ ^= 1;
        hip_check(hipEventSynchronize(ev[4]), "event sync");
        for (int q = 0; q < 4; ++q) {
            float ms = 0;
            hip_check(hipEventElapsedTime(&ms, ev[q], ev[q + 1]), "event time");
            acc[q] += ms;
        }
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
    for (int q = 0; q < 4; ++q) ms_out[q] = acc[q];
    return r;
}

int VoronoiBatch::brownian_steps(int steps) {
    if (steps <= 0) return CGPU_OK;
    mark_call_start();
    for (int k = 0; k < steps; ++k) {
        int r = step_launch(1);
        if (r) return r;
    }
    return after_steps(steps, 1);
}

// FIRE: systems that stalled inside an iteration (after its move) are
// retriangulated with grown storage and the iteration's force half is run for
// them alone (complete), then they rejoin the loop
int VoronoiBatch::finish_stalled_fire(cgpu_fire_state *fs, void *vel, double *dots) {
    std::vector<int> stalled;
    while (take_stalled(stalled)) {
        if (ring_cap_ >= CGPU_RING_CAP_MAX) return CGPU_ERR_RING_CAPACITY;
        grow_ring_capacity(std::min(CGPU_RING_CAP_MAX, 2 * ring_cap_));
        std::vector<cgpu_sys_state> st(S_);
        st_.download(st.data(), S_);
        const std::vector<cgpu_sys_state> saved = st;
        for (int s = 0; s < S_; ++s) st[s].active = 0;
        for (int s : stalled) {
            st[s].error = 0;
            st[s].active = 1;
        }
        st_.upload(st.data(), S_);
        int zero = 0;
        hip_check(hipMemcpy(counters_.p + 4, &zero, sizeof(int), hipMemcpyHostToDevice), "stall flag");
        int r = retriangulate_active();
        cgpu_voronoi_batch b = view();
        if (!r) r = cgpu_fire_force_reduce(&b, fs, vel, dots, stream_);
        if (!r) r = cgpu_fire_park_converged(&b, fs, stream_);
        if (r) return r;
        sync();
        st_.download(st.data(), S_);
        for (int s = 0; s < S_; ++s)
            if (std::find(stalled.begin(), stalled.end(), s) == stalled.end()) st[s].active = saved[s].active;
        st_.upload(st.data(), S_);
    }
    return CGPU_OK;
}

int VoronoiBatch::sync() {
    hip_check(hipStreamSynchronize(stream_), "sync");
    return CGPU_OK;
}

int VoronoiBatch::status() {
    sync();
    std::vector<int> c(8);
    counters_.download(c.data(), 8);
    if (c[3]) return CGPU_ERR_BIN_CAPACITY;
    if (c[5]) return CGPU_ERR_NONFINITE;
    std::vector<cgpu_sys_state> st(S_);
    st_.download(st.data(), S_);
    for (auto &q : st)
        if (q.error) return q.error;
    if (c[4]) return CGPU_ERR_TOPOLOGY;
    return CGPU_OK;
}

void VoronoiBatch::get(double *x, double *y, double *theta, double *vx, double *vy, double *Fx, double *Fy,
                       double *A, double *P) {
    sync();
    std::vector<double> t(2 * (size_t)N_);
    if (x || y) {
        (parity_ ? pos1_ : pos0_).download(t.data(), t.size());
        for (int i = 0; i < N_; ++i) {
            if (x) x[i] = t[2 * i];
            if (y) y[i] = t[2 * i + 1];
        }
    }
    if (theta) real_download(theta_.p, theta, N_);
    if (vx || vy) {
        real_download(vel_.p, t.data(), t.size());
        for (int i = 0; i < N_; ++i) {
            if (vx) vx[i] = t[2 * i];
            if (vy) vy[i] = t[2 * i + 1];
        }
    }
    if (Fx || Fy) {
        real_download(force_.p, t.data(), t.size());
        for (int i = 0; i < N_; ++i) {
            if (Fx) Fx[i] = t[2 * i];
            if (Fy) Fy[i] = t[2 * i + 1];
        }
    }
    if (A || P) {
        real_download(AP_.p, t.data(), t.size());
        for (int i = 0; i < N_; ++i) {
            if (A) A[i] = t[2 * i];
            if (P) P[i] = t[2 * i + 1];
        }
    }
}

void VoronoiBatch::get_topology(int *nn, int *nbr, int cap, int *neigh_max) {
    sync();
    std::vector<int> hn(N_), hr((size_t)ring_cap_ * N_);
    nn_.download(hn.data(), N_);
    nbr_.download(hr.data(), hr.size());
    for (int i = 0; i < N_; ++i) {
        nn[i] = hn[i];
        for (int k = 0; k < hn[i] && k < cap; ++k) nbr[(size_t)i * cap + k] = hr[(size_t)k * N_ + i];
    }
    if (neigh_max) {
        std::vector<cgpu_sys_state> st(S_);
        st_.download(st.data(), S_);
        for (int s = 0; s < S_; ++s) neigh_max[s] = st[s].neigh_max;
    }
}

void VoronoiBatch::get_counters(int s, long long *out) {
    sync();
    std::vector<cgpu_sys_state> st(S_);
    st_.download(st.data(), S_);
    const cgpu_sys_state &q = st[s];
    out[0] = q.global_fixes;
    out[1] = q.local_repairs;
    out[2] = q.repair_events;
    out[3] = q.step;
    out[4] = q.mode;
    out[5] = q.nf_count;
    out[6] = q.neigh_max;
    out[7] = q.error;
}

// reference VoronoiQuadraticEnergy::computeEnergy (.cpp:384-397), FP64 sum on the host
double VoronoiBatch::energy(int s) {
    const cgpu_sys_params &p = hp_[s];
    std::vector<double> A(N_), P(N_);
    get(nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, A.data(), P.data());
    std::vector<double> pref(2 * (size_t)N_);
    real_download(APpref_.p, pref.data(), pref.size());
    double E = 0.0;
    for (int i = p.off; i < p.off + p.n; ++i) {
        E += p.KA * (A[i] - pref[2 * i]) * (A[i] - pref[2 * i]);
        E += p.KP * (P[i] - pref[2 * i + 1]) * (P[i] - pref[2 * i + 1]);
    }
    if (p.tension != 0 && !h_type_.empty()) {
        // line tension: every Voronoi edge between cells of different types,
        // counted from both of its cells (the energy whose gradient the tension
        // force sets are), gamma times its length
        std::vector<double> x(N_), y(N_);
        get(x.data(), y.data(), nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
        const int cap = ring_cap_;
        std::vector<int> nn(N_), nbr((size_t)N_ * cap);
        get_topology(nn.data(), nbr.data(), cap, nullptr);
        std::vector<double> g;
        if (p.tension == 2) {
            g.resize((size_t)ntypes_ * ntypes_);
            tension_.download(g.data(), g.size());
        }
        const cgpu::Box box = cgpu::make_box(p.L);
        for (int i = p.off; i < p.off + p.n; ++i) {
            const int z = nn[i];
            const cgpu::dvec2 pi{x[i], y[i]};
            auto rel = [&](int k) {
                const int j = nbr[(size_t)i * cap + ((k % z) + z) % z];
                return cgpu::min_dist(box, cgpu::dvec2{x[j], y[j]}, pi);
            };
            for (int k = 0; k < z; ++k) {
                const int j = nbr[(size_t)i * cap + k];
                if (h_type_[i] == h_type_[j]) continue;
                // the edge shared with ring k joins circumcenters (k-1, k) and (k, k+1)
                cgpu::dvec2 a = cgpu::circ0<double>(rel(k - 1), rel(k)), c = cgpu::circ0<double>(rel(k), rel(k + 1));
                const double l = std::sqrt((a.x - c.x) * (a.x - c.x) + (a.y - c.y) * (a.y - c.y));
                E += (p.tension == 1 ? p.gamma : g[(size_t)h_type_[i] * ntypes_ + h_type_[j]]) * l;
            }
        }
    }
    return E;
}

// Keys, sort and re-indexing run on the device (cgpu_vor_hilbert_order,
// cgpu_vor_permute_cells); the cell list's scratch arrays, rebuilt by the
// retriangulation, hold the permutation (sorted_id) and the gather staging
// (sorted_pos, 16 bytes per cell). Only the permutation comes to the host, for
// the tags and the host copy of the cell types.
int VoronoiBatch::spatial_sort(int *perm_out) {
    cgpu_voronoi_batch b = view();
    const size_t wb = cgpu_vor_sort_workspace(N_);
    if (sort_work_.n < wb) sort_work_.alloc(wb);
    int r = cgpu_vor_hilbert_order(&b, sorted_id_.p, sort_work_.p, sort_work_.n, stream_);
    if (!r) r = cgpu_vor_permute_cells(&b, sorted_id_.p, sorted_pos_.p, stream_);
    if (r) return r;
    std::vector<int> perm(N_);
    hip_check(hipMemcpyAsync(perm.data(), sorted_id_.p, sizeof(int) * (size_t)N_, hipMemcpyDeviceToHost, stream_),
              "download");
    hip_check(hipStreamSynchronize(stream_), "sync");
    {
        std::vector<int> t(N_);
        for (int i = 0; i < N_; ++i) t[i] = tag_[perm[i]];
        tag_.swap(t);
    }
    if (!h_type_.empty()) {
        std::vector<int> t(N_);
        for (int i = 0; i < N_; ++i) t[i] = h_type_[perm[i]];
        h_type_.swap(t);
    }
    if (perm_out) std::copy(perm.begin(), perm.end(), perm_out);
    return global_triangulation();
}

// ------------------------------------------------------------------ profiling
int VoronoiBatch::spv_steps_profiled(int steps, double *ms_slots, long long *queue_sums) {
    constexpr int NS = CGPU_PROFILE_SLOTS;
    hipEvent_t ev[NS + 1];
    for (auto &e : ev) hip_check(hipEventCreate(&e), "event");
    for (int q = 0; q < NS; ++q) ms_slots[q] = 0.0;
    for (int q = 0; q < 5; ++q) queue_sums[q] = 0;
    drop_graph();
    int r = CGPU_OK;
    std::vector<int> c(CGPU_COUNTERS);
    for (int k = 0; k < steps && r == CGPU_OK; ++k) {
        cgpu_voronoi_batch b = view();
        r = cgpu_profiled_spv_step(&b, stream_, ev);
        parity_ ^= 1;
        hip_check(hipEventSynchronize(ev[NS]), "event sync");
        for (int q = 0; q < NS; ++q) {
            float ms = 0;
            hip_check(hipEventElapsedTime(&ms, ev[q], ev[q + 1]), "event time");
            ms_slots[q] += ms;
        }
        // queue sizes of this step: flagged cells, NeedsFixing, global queue,
        // triangles left to the circumcircle test, regions stamped
        counters_.download(c.data(), CGPU_COUNTERS);
        queue_sums[0] += c[0];
        queue_sums[1] += c[1];
        queue_sums[2] += c[2];
        queue_sums[3] += c[9];
        queue_sums[4] += c[8];
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
    return r;
}

// ---------------------------------------------------------------- autotuner
// Every array a time step writes, copied on the device.
struct VoronoiBatch::Snapshot {
    DevBuf<unsigned char> data;
    std::vector<std::pair<void *, size_t>> parts;  // (device array, bytes)
    int parity = 0;
};

void VoronoiBatch::take_snapshot(Snapshot &sn) {
    sync();
    sn.parts = {{pos0_.p, pos0_.n * sizeof(double)}, {pos1_.p, pos1_.n * sizeof(double)},
                {theta_.p, theta_.n}, {vel_.p, vel_.n}, {force_.p, force_.n}, {AP_.p, AP_.n},
                {dAP_.p, dAP_.n}, {nn_.p, nn_.n * sizeof(int)}, {nbr_.p, nbr_.n * sizeof(int)},
                {oth_.p, oth_.n * sizeof(int)}, {st_.p, st_.n * sizeof(cgpu_sys_state)},
                {counters_.p, counters_.n * sizeof(int)}};
    if (ext_.p) sn.parts.push_back({ext_.p, ext_.n});
    size_t total = 0;
    for (auto &p : sn.parts) total += (p.second + 255) & ~(size_t)255;
    sn.data.alloc(total);
    size_t off = 0;
    for (auto &p : sn.parts) {
        hip_check(hipMemcpyAsync(sn.data.p + off, p.first, p.second, hipMemcpyDeviceToDevice, stream_), "snapshot");
        off += (p.second + 255) & ~(size_t)255;
    }
    sn.parity = parity_;
    sync();
}

void VoronoiBatch::restore_snapshot(const Snapshot &sn) {
    sync();
    size_t off = 0;
    for (auto &p : sn.parts) {
        hip_check(hipMemcpyAsync(p.first, sn.data.p + off, p.second, hipMemcpyDeviceToDevice, stream_), "restore");
        off += (p.second + 255) & ~(size_t)255;
    }
    parity_ = sn.parity;
    sync();
}

// wall time (ms, HIP events) of `steps` SPV steps run as the production loop
// runs them (graph of two steps)
double VoronoiBatch::time_steps(int steps) {
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "event");
    hip_check(hipEventCreate(&e1), "event");
    hip_check(hipEventRecord(e0, stream_), "event record");
    int r = spv_steps(steps, true);
    hip_check(hipEventRecord(e1, stream_), "event record");
    hip_check(hipEventSynchronize(e1), "event sync");
    float ms = 0;
    hip_check(hipEventElapsedTime(&ms, e0, e1), "event time");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (r != CGPU_OK) throw std::runtime_error("step failed during autotuning, status " + std::to_string(r));
    return ms;
}

namespace {
// tuned shapes per (device, workload shape), for the life of the process
std::vector<std::pair<std::string, cgpu_launch_shape>> g_tuned;
}  // namespace

int VoronoiBatch::autotune(int trial_steps, cgpu_launch_shape *out, int *from_cache) {
    if (trial_steps < 2) trial_steps = 2;
    if (trial_steps % 2) ++trial_steps;  // whole graph launches
    hipDeviceProp_t prop;
    hip_check(hipGetDeviceProperties(&prop, device_), "device properties");
    // workload key: device, systems, cells, precision, tension, noise source
    const std::string key = std::string(prop.gcnArchName) + "|" + std::to_string(prop.pciBusID) + "|" +
                            std::to_string(S_) + "|" + std::to_string(N_) + "|" + std::to_string(prec_) + "|" +
                            std::to_string(tension_on_ ? 1 : 0) + "|" + std::to_string(noise_buffer_ ? 1 : 0);
    for (const auto &kv : g_tuned)
        if (kv.first == key) {
            set_shape(kv.second);
            if (out) *out = shape_;
            if (from_cache) *from_cache = 1;
            return CGPU_OK;
        }
    if (from_cache) *from_cache = 0;
    Snapshot sn;
    take_snapshot(sn);
    // best of two timings from the same saved state
    auto trial = [&](const cgpu_launch_shape &sh) {
        set_shape(sh);
        // capture the graph outside the timed region (same buffers and parity
        // after every restore, so it stays valid)
        restore_snapshot(sn);
        time_steps(2);
        double best = 1e30;
        for (int rep = 0; rep < 2; ++rep) {
            restore_snapshot(sn);
            best = std::min(best, time_steps(trial_steps));
        }
        return best;
    };
    cgpu_launch_shape cur{256, 256, 256, 4096, 1024, 8, 0, 0};
    double tcur = trial(cur);
    struct Knob {
        int cgpu_launch_shape::*field;
        std::vector<int> values;
    };
    const std::vector<Knob> knobs = {
        {&cgpu_launch_shape::test_block, {128, 256, 512}},
        {&cgpu_launch_shape::force_block, {128, 256, 512}},
        {&cgpu_launch_shape::force_waves, {0, 5}},
        {&cgpu_launch_shape::star_blocks, {1024, 2048, 4096, 8192, 16384}},
        {&cgpu_launch_shape::repair_blocks, {256, 512, 1024, 2048}},
        {&cgpu_launch_shape::queue_lanes, {4, 8, 16}},
        {&cgpu_launch_shape::cell_block, {128, 256, 512, 1024}},
    };
    // coordinate descent, one pass; a change must gain more than 1 %
    for (const Knob &k : knobs) {
        for (int v : k.values) {
            if (cur.*(k.field) == v) continue;
            cgpu_launch_shape t = cur;
            t.*(k.field) = v;
            const double tt = trial(t);
            if (tt < 0.99 * tcur) {
                cur = t;
                tcur = tt;
            }
        }
    }
    restore_snapshot(sn);
    set_shape(cur);
    g_tuned.push_back({key, cur});
    if (out) *out = shape_;
    return CGPU_OK;
}

void VoronoiBatch::get_active(int *active, int *steps_left) {
    std::vector<cgpu_sys_state> st(S_);
    hip_check(hipMemcpyAsync(st.data(), st_.p, sizeof(cgpu_sys_state) * S_, hipMemcpyDeviceToHost, stream_),
              "get_active");
    sync();
    for (int s = 0; s < S_; ++s) {
        if (active) active[s] = st[s].active;
        if (steps_left) steps_left[s] = st[s].steps_left;
    }
}

size_t VoronoiBatch::device_bytes() const { return device_bytes_; }

}  // namespace cgpu_host

// ================================================================ object-layer C ABI
using cgpu_host::VoronoiBatch;

namespace {
thread_local std::string g_last_error;
}  // namespace

// message of the last failed object-layer call on this thread (every host
// module reports through it)
void cgpu_host::set_last_error(const char *what) { g_last_error = what ? what : ""; }

namespace {
template <typename F> int guarded(F &&f) {
    try {
        return f();
    } catch (const std::invalid_argument &e) {
        g_last_error = e.what();
        return CGPU_ERR_ARGS;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return CGPU_ERR_HIP;
    }
}
}  // namespace

extern "C" {

const char *cgpu_last_error(void) { return g_last_error.c_str(); }

int cgpu_device_synchronize(void) { return hipDeviceSynchronize() == hipSuccess ? CGPU_OK : CGPU_ERR_HIP; }
int cgpu_vor_spv_steps_timed(void *h, int steps, double *ms_out) {
    return guarded([&] { return ((VoronoiBatch *)h)->spv_steps_timed(steps, ms_out); });
}

int cgpu_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

void *cgpu_vor_create(int num_systems, const int *n, const double *L, const uint64_t *seed, const double *x,
                      const double *y, const double *theta, int precision, int device) {
    try {
        return new VoronoiBatch(num_systems, n, L, seed, x, y, theta, precision, device);
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return nullptr;
    }
}
int cgpu_vor_get_external_forces(void *h, double *fx, double *fy) {
    return guarded([&] {
        VoronoiBatch &b = *(VoronoiBatch *)h;
        const int N = b.num_cells();
        std::vector<double> f(2 * (size_t)N, 0.0);
        b.sync();
        if (b.ext_force_device()) b.real_to_host(b.ext_force_device(), f.data(), f.size());
        for (int i = 0; i < N; ++i) {
            fx[i] = f[2 * i];
            fy[i] = f[2 * i + 1];
        }
        return 0;
    });
}
void *cgpu_vor_merge(int count, void *const *batches, const int *systems) {
    try {
        return VoronoiBatch::merge(count, (VoronoiBatch *const *)batches, systems);
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return nullptr;
    }
}
void cgpu_vor_destroy(void *h) { delete (VoronoiBatch *)h; }
int cgpu_vor_set_params(void *h, int s, double KA, double KP, double dt, double mu, double T) {
    return guarded([&] { ((VoronoiBatch *)h)->set_params(s, KA, KP, dt, mu, T); return 0; });
}
int cgpu_vor_set_prefs(void *h, const double *A0, const double *P0) {
    return guarded([&] { ((VoronoiBatch *)h)->set_prefs(A0, P0); return 0; });
}
int cgpu_vor_set_motility(void *h, const double *v0, const double *Dr) {
    return guarded([&] { ((VoronoiBatch *)h)->set_motility(v0, Dr); return 0; });
}
int cgpu_vor_set_exclusions(void *h, const int *ex) {
    return guarded([&] { ((VoronoiBatch *)h)->set_exclusions(ex); return 0; });
}
int cgpu_vor_set_directors(void *h, const double *theta) {
    return guarded([&] { ((VoronoiBatch *)h)->set_directors(theta); return 0; });
}
int cgpu_vor_set_topology(void *h, const int *nn, const int *nbr, int cap, const int *oth) {
    return guarded([&] {
        ((VoronoiBatch *)h)->set_topology(nn, nbr, cap, oth);
        return ((VoronoiBatch *)h)->status();
    });
}
int cgpu_vor_get_del_other(void *h, int *oth, int cap) {
    return guarded([&] {
        ((VoronoiBatch *)h)->get_del_other(oth, cap);
        return 0;
    });
}
int cgpu_vor_set_velocities(void *h, const double *vx, const double *vy) {
    return guarded([&] { ((VoronoiBatch *)h)->set_velocities(vx, vy); return 0; });
}
int cgpu_vor_set_step_counters(void *h, const int *step) {
    return guarded([&] { ((VoronoiBatch *)h)->set_step_counters(step); return 0; });
}
int cgpu_vor_cell_events(void *h, int count, const int *kind, const int *cell, const double *theta,
                         const double *frac, const double *new_director) {
    return guarded([&] {
        int r = ((VoronoiBatch *)h)->cell_events(count, kind, cell, theta, frac, new_director);
        return r ? r : ((VoronoiBatch *)h)->status();
    });
}
int cgpu_vor_get_sizes(void *h, int *n) {
    const VoronoiBatch &b = *(VoronoiBatch *)h;
    if (n)
        for (int s = 0; s < b.num_systems(); ++s) n[s] = b.params()[s].n;
    return b.num_cells();
}
int cgpu_vor_get_tags(void *h, int *tags) {
    return guarded([&] {
        const auto &t = ((VoronoiBatch *)h)->tags();
        std::copy(t.begin(), t.end(), tags);
        return 0;
    });
}
int cgpu_vor_set_cell_types(void *h, const int *types) {
    return guarded([&] { ((VoronoiBatch *)h)->set_cell_types(types); return 0; });
}
int cgpu_vor_set_tension(void *h, int s, int mode, double gamma) {
    return guarded([&] { ((VoronoiBatch *)h)->set_tension(s, mode, gamma); return 0; });
}
int cgpu_vor_set_tension_matrix(void *h, int ntypes, const double *g) {
    return guarded([&] { ((VoronoiBatch *)h)->set_tension_matrix(ntypes, g); return 0; });
}
int cgpu_vor_set_positions(void *h, const double *x, const double *y, int retri) {
    return guarded([&] { ((VoronoiBatch *)h)->set_positions(x, y, retri != 0); return ((VoronoiBatch *)h)->status(); });
}
int cgpu_vor_set_noise_mode(void *h, int use_buffer) {
    return guarded([&] { ((VoronoiBatch *)h)->set_noise_mode(use_buffer != 0); return 0; });
}
int cgpu_vor_upload_noise(void *h, const double *eta) {
    return guarded([&] {
        VoronoiBatch *b = (VoronoiBatch *)h;
        b->upload_noise(eta, 2 * b->num_cells());
        return 0;
    });
}
int cgpu_vor_compute_forces(void *h) { return guarded([&] { return ((VoronoiBatch *)h)->compute_forces(); }); }
int cgpu_vor_enforce_topology(void *h) { return guarded([&] { return ((VoronoiBatch *)h)->enforce_topology(); }); }
int cgpu_vor_spv_steps(void *h, int steps, int use_graph) {
    return guarded([&] { return ((VoronoiBatch *)h)->spv_steps(steps, use_graph != 0); });
}
int cgpu_vor_brownian_steps(void *h, int steps) {
    return guarded([&] { return ((VoronoiBatch *)h)->brownian_steps(steps); });
}
int cgpu_vor_sync(void *h) { return guarded([&] { return ((VoronoiBatch *)h)->sync(); }); }
int cgpu_vor_status(void *h) { return guarded([&] { return ((VoronoiBatch *)h)->status(); }); }
int cgpu_vor_get(void *h, double *x, double *y, double *theta, double *vx, double *vy, double *Fx, double *Fy,
                 double *A, double *P) {
    return guarded([&] { ((VoronoiBatch *)h)->get(x, y, theta, vx, vy, Fx, Fy, A, P); return 0; });
}
int cgpu_vor_get_topology(void *h, int *nn, int *nbr, int cap, int *neigh_max) {
    return guarded([&] { ((VoronoiBatch *)h)->get_topology(nn, nbr, cap, neigh_max); return 0; });
}
int cgpu_vor_get_counters(void *h, int s, long long *out) {
    return guarded([&] { ((VoronoiBatch *)h)->get_counters(s, out); return 0; });
}
int cgpu_vor_num_cells(void *h) { return ((VoronoiBatch *)h)->num_cells(); }
int cgpu_vor_ring_capacity(void *h) { return ((VoronoiBatch *)h)->ring_capacity(); }
double cgpu_vor_energy(void *h, int s) {
    try {
        return ((VoronoiBatch *)h)->energy(s);
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return NAN;
    }
}
int cgpu_vor_set_active(void *h, const int *active, const int *steps_left) {
    return guarded([&] { ((VoronoiBatch *)h)->set_active(active, steps_left); return 0; });
}

int cgpu_vor_get_active(void *h, int *active, int *steps_left) {
    return guarded([&] { ((VoronoiBatch *)h)->get_active(active, steps_left); return 0; });
}

long long cgpu_vor_device_bytes(void *h) {
    if (!h) return -1;
    return (long long)((VoronoiBatch *)h)->device_bytes();
}
int cgpu_vor_autotune(void *h, int trial_steps, cgpu_launch_shape *out, int *from_cache) {
    return guarded([&] { return ((VoronoiBatch *)h)->autotune(trial_steps, out, from_cache); });
}
int cgpu_vor_set_launch_shape(void *h, const cgpu_launch_shape *shape) {
    return guarded([&] {
        if (!shape) throw std::invalid_argument("null launch shape");
        ((VoronoiBatch *)h)->set_shape(*shape);
        return 0;
    });
}
int cgpu_vor_get_launch_shape(void *h, cgpu_launch_shape *shape) {
    return guarded([&] {
        if (!shape) throw std::invalid_argument("null launch shape");
        *shape = ((VoronoiBatch *)h)->shape();
        return 0;
    });
}
int cgpu_vor_spv_steps_profiled(void *h, int steps, double *ms_slots, long long *queue_sums) {
    return guarded([&] { return ((VoronoiBatch *)h)->spv_steps_profiled(steps, ms_slots, queue_sums); });
}
const char *cgpu_profile_slot_name(int slot) {
    static const char *names[CGPU_PROFILE_SLOTS] = {
        "k_force_eom_move", "k_bin_scan", "k_bin_scatter", "k_test_geometry", "k_ld_reduce",
        "k_select_triangles", "k_test_queue", "k_mark_nf", "k_star_local_wave",
        "k_delsets_local", "k_global_fallback", "k_star_global_wave", "k_global_stats",
        "k_finalize"};
    return (slot >= 0 && slot < CGPU_PROFILE_SLOTS) ? names[slot] : "";
}
int cgpu_vor_spatial_sort(void *h, int *perm) {
    return guarded([&] { return ((VoronoiBatch *)h)->spatial_sort(perm); });
}
int cgpu_vor_view(void *h, cgpu_voronoi_batch *out) {
    *out = ((VoronoiBatch *)h)->view();
    return 0;
}
void *cgpu_vor_stream(void *h) { return (void *)((VoronoiBatch *)h)->stream(); }

}  // extern "C"

// host copies of the counter-based generator (same code as the device path)
#include "../cgpu_device.h"
extern "C" void cgpu_philox4x32_10(const uint32_t *ctr, const uint32_t *key, uint32_t *out) {
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    cgpu::philox4x32_10(c, key[0], key[1]);
    for (int q = 0; q < 4; ++q) out[q] = c[q];
}
extern "C" void cgpu_normal_pair(uint64_t seed, uint32_t index, uint32_t step, uint32_t stream, double *out) {
    cgpu::normal_pair(seed, index, step, stream, out[0], out[1]);
}
extern "C" void cgpu_put_in_box(double L, double *x, double *y) {
    cgpu::Box b = cgpu::make_box(L);
    cgpu::dvec2 p{*x, *y};
    cgpu::put_in_box(b, p);
    *x = p.x;
    *y = p.y;
}

namespace cgpu_host {

// Cell death and division (reference voronoiModelBase::cellDeath / cellDivision,
// voronoiModelBase.cpp:1291-1422, with the bookkeeping of Simple2DCell.cpp:775-850
// and Simple2DActiveCell.cpp:118-147). The state of every cell is gathered on the
// host, the events edit each system's segment (a dead cell is removed and the
// cells after it move down, tags above its tag drop by one; a daughter is
// appended with the next tag), the batch is laid out again and retriangulated.
int VoronoiBatch::cell_events(int count, const int *kind, const int *cell, const double *theta, const double *frac,
                              const double *new_director) {
    if (count <= 0) return CGPU_OK;
    if (!kind || !cell) throw std::invalid_argument("cell events need kinds and cells");
    std::vector<int> ev_of_sys(S_, -1);
    bool draw = false;
    for (int e = 0; e < count; ++e) {
        if (cell[e] < 0 || cell[e] >= N_) throw std::invalid_argument("cell index out of range");
        if (kind[e] != 0 && kind[e] != 1) throw std::invalid_argument("event kind is 0 (death) or 1 (division)");
        const int s = cell_sys_h_[cell[e]];
        if (ev_of_sys[s] >= 0) throw std::invalid_argument("at most one event per system per call");
        ev_of_sys[s] = e;
        if (kind[e] == 0 && hp_[s].n <= 4) throw std::invalid_argument("a system keeps at least 4 cells");
        if (kind[e] == 1 && (!theta || !frac || !std::isfinite(theta[e]) || !std::isfinite(frac[e])))
            throw std::invalid_argument("a division needs a finite angle and separation fraction");
        if (kind[e] == 1 && !(new_director && std::isfinite(new_director[e]))) draw = true;
    }
    sync();
    // step counters, for directors drawn from the counter-based stream
    std::vector<cgpu_sys_state> st(S_);
    if (draw) st_.download(st.data(), S_);
    // the division axes, from the current rings on the device: one record of
    // six doubles per dividing cell
    std::vector<int> div_of_event(count, -1), parent;
    std::vector<double> ray;
    for (int e = 0; e < count; ++e)
        if (kind[e] == 1) {
            div_of_event[e] = (int)parent.size();
            parent.push_back(cell[e]);
            ray.push_back(std::cos(theta[e]));
            ray.push_back(std::sin(theta[e]));
        }
    const int nd = (int)parent.size();
    std::vector<double> axis(6 * (size_t)nd);
    if (nd) {
        DevBuf<int> dpar;
        DevBuf<double> dray, dout;
        dpar.alloc(nd);
        dpar.upload(parent.data(), nd);
        dray.alloc(2 * (size_t)nd);
        dray.upload(ray.data(), 2 * (size_t)nd);
        dout.alloc(6 * (size_t)nd);
        cgpu_voronoi_batch b = view();
        int r = cgpu_vor_division_axis(&b, nd, dpar.p, dray.p, dout.p, stream_);
        if (r) return r;
        hip_check(hipStreamSynchronize(stream_), "sync");
        dout.download(axis.data(), axis.size());
    }
    // one record per system; the new positions follow the reference
    // (voronoiModelBase.cpp:1399-1406) from the crossings
    std::vector<cgpu_cell_event> rec(S_);
    std::vector<int> newn(S_);
    int noff = 0;
    for (int s = 0; s < S_; ++s) {
        const cgpu_sys_params &p = hp_[s];
        cgpu_cell_event &r = rec[s];
        std::memset(&r, 0, sizeof(r));
        r.kind = -1;
        r.old_off = p.off;
        r.old_n = p.n;
        r.new_off = noff;
        const int e = ev_of_sys[s];
        if (e >= 0) {
            r.kind = kind[e];
            r.target = cell[e] - p.off;
        }
        if (e >= 0 && kind[e] == 1) {
            const double *ax = &axis[6 * (size_t)div_of_event[e]];
            if (!std::isfinite(ax[2])) throw std::runtime_error("division axis misses the cell's polygon");
            const cgpu::Box box = cgpu::make_box(p.L);
            const cgpu::dvec2 pi{ax[0], ax[1]};
            const double c = ray[2 * div_of_event[e]], sn = ray[2 * div_of_event[e] + 1];
            const double sep = std::max(std::sqrt(ax[2] * ax[2] + ax[3] * ax[3]), std::sqrt(ax[4] * ax[4] + ax[5] * ax[5]));
            const double f = frac[e] * sep;
            cgpu::dvec2 p1{pi.x + f * c, pi.y + f * sn}, p2{pi.x - f * c, pi.y - f * sn};
            cgpu::put_in_box(box, p1);
            cgpu::put_in_box(box, p2);
            double dth = new_director ? new_director[e] : NAN;
            if (!std::isfinite(dth)) {
                uint32_t ctr[4] = {(uint32_t)p.n, (uint32_t)st[s].step, 3u, 0u};
                cgpu::philox4x32_10(ctr, (uint32_t)p.seed, (uint32_t)(p.seed >> 32));
                const uint64_t u = (((uint64_t)ctr[0] << 21) ^ (uint64_t)(ctr[1] >> 11)) & (((uint64_t)1 << 53) - 1);
                dth = 6.283185307179586476925286766559 * ((double)u / 9007199254740992.0);
            }
            r.p1[0] = p1.x;
            r.p1[1] = p1.y;
            r.p2[0] = p2.x;
            r.p2[1] = p2.y;
            r.director = dth;
            r.dcos = std::cos(dth);
            r.dsin = std::sin(dth);
        }
        newn[s] = p.n + (r.kind == 1 ? 1 : 0) - (r.kind == 0 ? 1 : 0);
        noff += newn[s];
    }
    const size_t M = (size_t)noff;
    // the new per-cell arrays, gathered on the device from the old ones
    DevBuf<cgpu_cell_event> drec;
    drec.alloc(S_);
    drec.upload(rec.data(), S_);
    DevBuf<double> npos;
    DevBuf<unsigned char> nth, nvel, npref, nmot;
    DevBuf<int> nex, ntype;
    npos.alloc(2 * M);
    nth.alloc(M * rsz_);
    nvel.alloc(2 * M * rsz_);
    npref.alloc(2 * M * rsz_);
    nmot.alloc(2 * M * rsz_);
    if (excl_.p) nex.alloc(M);
    if (cell_type_.p) ntype.alloc(M);
    {
        cgpu_voronoi_batch b = view();
        int r = cgpu_vor_events_gather(&b, drec.p, (int)M, npos.p, nth.p, nvel.p, npref.p, nmot.p, nex.p, ntype.p,
                                       stream_);
        if (r) return r;
        hip_check(hipStreamSynchronize(stream_), "sync");
    }
    // labels (reference tagToIdx / idxToTag): a death shifts the labels above
    // the dead cell's down by one; a daughter takes the old system size
    std::vector<int> newtag, newtype;
    newtag.reserve(M);
    for (int s = 0; s < S_; ++s) {
        const cgpu_sys_params &p = hp_[s];
        const cgpu_cell_event &r = rec[s];
        const int dead_tag = r.kind == 0 ? tag_[p.off + r.target] : -1;
        for (int i = p.off; i < p.off + p.n; ++i) {
            if (r.kind == 0 && i == p.off + r.target) continue;
            const int t = tag_[i];
            newtag.push_back((r.kind == 0 && t > dead_tag) ? t - 1 : t);
            if (!h_type_.empty()) newtype.push_back(h_type_[i]);
        }
        if (r.kind == 1) {
            newtag.push_back(p.n);
            if (!h_type_.empty()) newtype.push_back(h_type_[p.off + r.target]);
        }
    }
    layout(newn);
    pos0_.swap(npos);
    theta_.swap(nth);
    vel_.swap(nvel);
    APpref_.swap(npref);
    mot_.swap(nmot);
    tag_ = newtag;
    if (nex.p) {
        excl_.swap(nex);
        ext_.alloc(2 * M * rsz_);
    }
    if (ntype.p) {
        h_type_ = newtype;
        cell_type_.swap(ntype);
    }
    if (noise_.p) noise_.alloc(2 * M);
    update_tension_flag();
    return global_triangulation();
}

}  // namespace cgpu_host
