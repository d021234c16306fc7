// EnergyMinimizerFIRE: the per-system minimize loop (FireBase) and its Voronoi
// driver. Reference EnergyMinimizerFIRE2D.cpp:24-41 (defaults), :93-114
// (velocityVerletGPU), :156-207 (fireStepGPU), :269-284 (minimize).
#include "fire.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>

namespace cgpu_host {

FireBase::FireBase(int S, size_t ndof, size_t rsz, size_t nchunks) : S_(S), ndof_(ndof), nchunks_(nchunks), rsz_(rsz) {
    fs_.alloc(S_);
    vel_.alloc(2 * ndof * rsz);
    dots_.alloc(3 * ndof + 4 * nchunks);
    hs_.resize(S_);
    for (auto &f : hs_) {
        std::memset(&f, 0, sizeof(f));
        f.alpha_start = f.alpha = 0.99;
        f.dt = 0.01;
        f.dt_max = 0.1;
        f.dt_inc = 1.05;
        f.dt_dec = 0.95;
        f.alpha_dec = 0.9;
        f.force_cutoff = 1e-7;
        f.nmin = 5;
        f.max_iterations = 1000;
        f.force_max = 100.;
    }
    fs_.upload(hs_.data(), S_);
}

void FireBase::set(int max_it, double fc, double alpha_start, double dt, double dt_max, double dt_inc,
                   double dt_dec, double alpha_dec, int nmin) {
    fs_.download(hs_.data(), S_);
    for (auto &f : hs_) {
        f.max_iterations = max_it;
        f.force_cutoff = fc;
        f.alpha_start = f.alpha = alpha_start;
        f.dt = dt;
        f.dt_max = dt_max;
        f.dt_inc = dt_inc;
        f.dt_dec = dt_dec;
        f.alpha_dec = alpha_dec;
        f.nmin = nmin;
    }
    fs_.upload(hs_.data(), S_);
}

void FireBase::set_param(int which, double v) {
    sync();
    fs_.download(hs_.data(), S_);
    for (auto &f : hs_) {
        switch (which) {
            case CGPU_FIRE_MAX_IT: f.max_iterations = (int)v; break;
            case CGPU_FIRE_FORCE_CUTOFF: f.force_cutoff = v; break;
            case CGPU_FIRE_ALPHA_START: f.alpha_start = f.alpha = v; break;
            case CGPU_FIRE_DT: f.dt = v; break;
            case CGPU_FIRE_DT_MAX: f.dt_max = v; break;
            case CGPU_FIRE_DT_INC: f.dt_inc = v; break;
            case CGPU_FIRE_DT_DEC: f.dt_dec = v; break;
            case CGPU_FIRE_ALPHA_DEC: f.alpha_dec = v; break;
            case CGPU_FIRE_NMIN: f.nmin = (int)v; break;
            default: throw std::invalid_argument("unknown FIRE parameter");
        }
    }
    fs_.upload(hs_.data(), S_);
}

int FireBase::minimize() {
    // cell events change the degrees of freedom: start from rest with buffers of
    // the new size (reference minimize -> initializeFromModel)
    if (dof() != ndof_ || chunks() != nchunks_) {
        sync();
        ndof_ = dof();
        nchunks_ = chunks();
        vel_.alloc(2 * ndof_ * rsz_);
        dots_.alloc(3 * ndof_ + 4 * nchunks_);
    }
    int r = forces();
    if (r) return r;
    sync();
    // converged systems are parked (inactive) so the topology kernels skip them;
    // systems that were inactive before the call are left untouched
    std::vector<cgpu_sys_state> saved(S_);
    hip_check(hipMemcpy(saved.data(), device_state(), sizeof(cgpu_sys_state) * S_, hipMemcpyDeviceToHost), "state");
    fs_.download(hs_.data(), S_);
    for (int s = 0; s < S_; ++s) {
        cgpu_fire_state &f = hs_[s];
        f.force_max = 110.0;  // minimize() re-arms the loop test
        f.converged = (f.iterations < f.max_iterations && std::sqrt(f.force_max) > f.force_cutoff) ? 0 : 1;
        if (!saved[s].active) f.converged = 1;
        f.pending = 0;
    }
    fs_.upload(hs_.data(), S_);
    const int chunk = 16;  // iterations between host checks of the stopping test
    bool done = false;
    while (!done) {
        for (int c = 0; c < chunk; ++c)
            if ((r = iterate())) return r;
        sync();
        fs_.download(hs_.data(), S_);
        if ((r = recover_stalled())) return r;
        fs_.download(hs_.data(), S_);
        done = std::all_of(hs_.begin(), hs_.end(), [](const cgpu_fire_state &f) { return f.converged != 0; });
        if ((r = status())) return r;
    }
    if ((r = apply_mix())) return r;
    // restore only the activity flags: counters advanced during the
    // minimisation stay
    sync();
    std::vector<cgpu_sys_state> now(S_);
    hip_check(hipMemcpy(now.data(), device_state(), sizeof(cgpu_sys_state) * S_, hipMemcpyDeviceToHost), "state");
    for (int s = 0; s < S_; ++s) now[s].active = saved[s].active;
    hip_check(hipMemcpy(device_state(), now.data(), sizeof(cgpu_sys_state) * S_, hipMemcpyHostToDevice), "state");
    return status();
}

void FireBase::get(int s, double *out) {
    sync();
    fs_.download(hs_.data(), S_);
    const cgpu_fire_state &f = hs_[s];
    out[0] = f.force_max;
    out[1] = f.power;
    out[2] = f.alpha;
    out[3] = f.dt;
    out[4] = f.max_f2;
    out[5] = f.iterations;
}

namespace {

// FIRE on a Voronoi batch: the cells are the degrees of freedom and the
// topology repair is enforceTopology
class VoronoiFire : public FireBase {
   public:
    explicit VoronoiFire(VoronoiBatch *vb)
        : FireBase(vb->num_systems(), vb->num_cells(), vb->precision() ? sizeof(float) : sizeof(double),
                   vb->view().num_chunks),
          vb_(vb) {}

   protected:
    hipStream_t stream() override { return vb_->stream(); }
    cgpu_sys_state *device_state() override { return vb_->view().st; }
    int forces() override {
        cgpu_voronoi_batch b = vb_->view();
        return cgpu_voronoi_forces(&b, stream());
    }
    int iterate() override {
        cgpu_voronoi_batch v = vb_->view();
        const int r = cgpu_fire_iteration(&v, fs_.p, vel_.p, dots_.p, stream());
        vb_->swap_parity();
        return r;
    }
    int apply_mix() override {
        cgpu_voronoi_batch v = vb_->view();
        return cgpu_fire_apply_mix(&v, fs_.p, vel_.p, stream());
    }
    void sync() override { vb_->sync(); }
    int status() override { return vb_->status(); }
    int recover_stalled() override { return vb_->finish_stalled_fire(fs_.p, vel_.p, dots_.p); }
    size_t dof() override { return vb_->num_cells(); }
    size_t chunks() override { return vb_->view().num_chunks; }

   private:
    VoronoiBatch *vb_;
};

}  // namespace
}  // namespace cgpu_host

using cgpu_host::FireBase;
using cgpu_host::VoronoiBatch;

extern "C" {
void *cgpu_fire_create(void *vor) {
    try {
        return static_cast<FireBase *>(new cgpu_host::VoronoiFire((VoronoiBatch *)vor));
    } catch (...) {
        return nullptr;
    }
}
void cgpu_fire_destroy(void *h) { delete (FireBase *)h; }
int cgpu_fire_set(void *h, int max_it, double fc, double alpha_start, double dt, double dt_max, double dt_inc,
                  double dt_dec, double alpha_dec, int nmin) {
    try {
        ((FireBase *)h)->set(max_it, fc, alpha_start, dt, dt_max, dt_inc, dt_dec, alpha_dec, nmin);
        return CGPU_OK;
    } catch (...) {
        return CGPU_ERR_HIP;
    }
}
int cgpu_fire_set_param(void *h, int which, double value) {
    try {
        ((FireBase *)h)->set_param(which, value);
        return CGPU_OK;
    } catch (const std::invalid_argument &) {
        return CGPU_ERR_ARGS;
    } catch (...) {
        return CGPU_ERR_HIP;
    }
}
int cgpu_fire_minimize(void *h) {
    try {
        return ((FireBase *)h)->minimize();
    } catch (...) {
        return CGPU_ERR_HIP;
    }
}
int cgpu_fire_get(void *h, int s, double *out) {
    try {
        ((FireBase *)h)->get(s, out);
        return CGPU_OK;
    } catch (...) {
        return CGPU_ERR_HIP;
    }
}
}
