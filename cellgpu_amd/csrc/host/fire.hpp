// EnergyMinimizerFIRE on a device-resident batch: every system has its own FIRE
// scalars and stops independently. Reference EnergyMinimizerFIRE2D.cpp:24-41
// (defaults), :93-114 (velocityVerletGPU), :156-207 (fireStepGPU), :269-284
// (minimize). The model-specific parts (forces, one iteration's kernels) are
// supplied by the Voronoi and vertex subclasses.
#pragma once

#include <vector>

#include "voronoi_batch.hpp"

namespace cgpu_host {

class FireBase {
   public:
    // S systems, ndof degrees of freedom in all, real size rsz, nchunks
    // reduction chunks
    FireBase(int S, size_t ndof, size_t rsz, size_t nchunks);
    virtual ~FireBase() = default;
    void set(int max_it, double fc, double alpha_start, double dt, double dt_max, double dt_inc, double dt_dec,
             double alpha_dec, int nmin);
    void set_param(int which, double value);
    int minimize();
    void get(int s, double *out);

   protected:
    virtual hipStream_t stream() = 0;
    virtual cgpu_sys_state *device_state() = 0;
    virtual int forces() = 0;     // forces at the current positions
    virtual int iterate() = 0;    // one iteration of every running system
    virtual int apply_mix() = 0;  // the pending velocity mixing
    virtual void sync() = 0;
    virtual int status() = 0;     // after a sync: first error of the batch
    // systems that stalled on a storage capacity inside an iteration: grown and
    // finished (default: none can stall)
    virtual int recover_stalled() { return 0; }
    virtual size_t dof() = 0;     // degrees of freedom of the model now
    virtual size_t chunks() = 0;  // reduction chunks of the model now
    int S_;
    size_t ndof_ = 0, nchunks_ = 0, rsz_ = 0;
    DevBuf<cgpu_fire_state> fs_;
    DevBuf<unsigned char> vel_;
    DevBuf<double> dots_;
    std::vector<cgpu_fire_state> hs_;
};

}  // namespace cgpu_host
