// EnergyMinimizerFIRE on a device-resident Voronoi batch (every system gets its
// own FIRE scalars; systems stop independently). Reference
// EnergyMinimizerFIRE2D.cpp:24-41 (defaults), :93-114 (velocityVerletGPU),
// :156-207 (fireStepGPU), :269-284 (minimize).
#include <algorithm>
#include <cstring>
#include <vector>

#include "voronoi_batch.hpp"

namespace cgpu_host {

class FireMinimizer {
   public:
    explicit FireMinimizer(VoronoiBatch *vb) : vb_(vb) {
        S_ = vb->num_systems();
        N_ = vb->num_cells();
        size_t rsz = vb->precision() ? sizeof(float) : sizeof(double);
        fs_.alloc(S_);
        vel_.alloc(2 * (size_t)N_ * rsz);
        dots_.alloc(3 * (size_t)N_ + 4 * (size_t)vb->view().num_chunks);
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
    void set(int max_it, double fc, double alpha_start, double dt, double dt_max, double dt_inc, double dt_dec,
             double alpha_dec, int nmin) {
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
    int minimize() {
        hipStream_t st = vb_->stream();
        cgpu_voronoi_batch b = vb_->view();
        int r = cgpu_voronoi_forces(&b, st);
        if (r) return r;
        // converged systems are parked (inactive) so the topology kernels skip them;
        // systems that were inactive before the call are left untouched
        std::vector<cgpu_sys_state> saved(S_);
        hip_check(hipMemcpy(saved.data(), b.st, sizeof(cgpu_sys_state) * S_, hipMemcpyDeviceToHost), "state");
        fs_.download(hs_.data(), S_);
        for (int s = 0; s < S_; ++s) {
            cgpu_fire_state &f = hs_[s];
            f.force_max = 110.0;  // minimize() re-arms the loop test
            f.converged = (f.iterations < f.max_iterations && std::sqrt(f.force_max) > f.force_cutoff) ? 0 : 1;
            if (!saved[s].active) f.converged = 1;
            f.pending = 0;
        }
        fs_.upload(hs_.data(), S_);
        const int chunk = 16;
        bool done = false;
        while (!done) {
            for (int c = 0; c < chunk; ++c) {
                cgpu_voronoi_batch v = vb_->view();
                cgpu_voronoi_batch vn = v;
                vn.pos = v.pos_next;
                vn.pos_next = v.pos;
                if ((r = cgpu_step_begin(&v, st))) return r;
                if ((r = cgpu_fire_vv_move(&v, fs_.p, vel_.p, st))) return r;
                if ((r = cgpu_compute_cell_list(&v, 1, st))) return r;
                if ((r = cgpu_test_and_geometry(&vn, 1, st))) return r;
                if ((r = cgpu_repair_topology(&vn, st))) return r;
                vb_->swap_parity();
                if ((r = cgpu_fire_force_reduce(&vn, fs_.p, vel_.p, dots_.p, st))) return r;
                park_converged(st);
            }
            fs_.download(hs_.data(), S_);
            done = std::all_of(hs_.begin(), hs_.end(), [](const cgpu_fire_state &f) { return f.converged != 0; });
            r = vb_->status();
            if (r) return r;
        }
        cgpu_voronoi_batch v = vb_->view();
        if ((r = cgpu_fire_apply_mix(&v, fs_.p, vel_.p, st))) return r;
        // restore only the activity flags: counters and neighMax advanced during
        // the minimisation stay
        std::vector<cgpu_sys_state> now(S_);
        vb_->sync();
        hip_check(hipMemcpy(now.data(), b.st, sizeof(cgpu_sys_state) * S_, hipMemcpyDeviceToHost), "state");
        for (int s = 0; s < S_; ++s) now[s].active = saved[s].active;
        hip_check(hipMemcpy(b.st, now.data(), sizeof(cgpu_sys_state) * S_, hipMemcpyHostToDevice), "state");
        return vb_->status();
    }
    void get(int s, double *out) {
        fs_.download(hs_.data(), S_);
        const cgpu_fire_state &f = hs_[s];
        out[0] = f.force_max;
        out[1] = f.power;
        out[2] = f.alpha;
        out[3] = f.dt;
        out[4] = f.max_f2;
        out[5] = f.iterations;
    }

   private:
    void park_converged(hipStream_t st);
    VoronoiBatch *vb_;
    int S_, N_;
    DevBuf<cgpu_fire_state> fs_;
    DevBuf<unsigned char> vel_;
    DevBuf<double> dots_;
    std::vector<cgpu_fire_state> hs_;
};

}  // namespace cgpu_host



void cgpu_host::FireMinimizer::park_converged(hipStream_t st) {
    cgpu_voronoi_batch v = vb_->view();
    cgpu_fire_park_converged(&v, fs_.p, st);
}

using cgpu_host::FireMinimizer;
using cgpu_host::VoronoiBatch;

extern "C" {
void *cgpu_fire_create(void *vor) {
    try {
        return new FireMinimizer((VoronoiBatch *)vor);
    } catch (...) {
        return nullptr;
    }
}
void cgpu_fire_destroy(void *h) { delete (FireMinimizer *)h; }
int cgpu_fire_set(void *h, int max_it, double fc, double alpha_start, double dt, double dt_max, double dt_inc,
                  double dt_dec, double alpha_dec, int nmin) {
    try {
        ((FireMinimizer *)h)->set(max_it, fc, alpha_start, dt, dt_max, dt_inc, dt_dec, alpha_dec, nmin);
        return CGPU_OK;
    } catch (...) {
        return CGPU_ERR_HIP;
    }
}
int cgpu_fire_minimize(void *h) {
    try {
        return ((FireMinimizer *)h)->minimize();
    } catch (...) {
        return CGPU_ERR_HIP;
    }
}
int cgpu_fire_get(void *h, int s, double *out) {
    try {
        ((FireMinimizer *)h)->get(s, out);
        return CGPU_OK;
    } catch (...) {
        return CGPU_ERR_HIP;
    }
}
}
