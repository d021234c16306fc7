// Host-side owner of a device-resident batch of Voronoi systems.
//
// One VoronoiBatch holds S independent periodic monolayers back to back in
// structure-of-arrays device memory (a single reference VoronoiQuadraticEnergy
// is the S = 1 case). It allocates once, runs the step as a fixed kernel
// sequence on its own stream (optionally captured in a hipGraph) and copies
// state to the host only when asked.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../../include/cellgpu_amd.h"

// one profiled step of the kernel layer (voronoi_kernels.hip): ev[0] before
// the first launch, ev[k + 1] after the launches of profile slot k
int cgpu_profiled_spv_step(const cgpu_voronoi_batch *b, void *stream, hipEvent_t *ev);

namespace cgpu_host {

void hip_check(hipError_t e, const char *what);
void set_last_error(const char *what);

template <typename T> struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void alloc(size_t count) {
        release();
        n = count;
        if (count) hip_check(hipMalloc((void **)&p, count * sizeof(T)), "hipMalloc");
        if (count) {
            hip_check(hipMemset(p, 0, count * sizeof(T)), "hipMemset");
            // the fill runs on the null stream, which the batches' non-blocking
            // streams do not wait for: finish it before the buffer is used
            hip_check(hipStreamSynchronize(nullptr), "hipMemset");
        }
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void swap(DevBuf &o) {
        std::swap(p, o.p);
        std::swap(n, o.n);
    }
    void upload(const T *h, size_t count, size_t offset = 0) {
        hip_check(hipMemcpy(p + offset, h, count * sizeof(T), hipMemcpyHostToDevice), "upload");
    }
    void download(T *h, size_t count, size_t offset = 0) const {
        hip_check(hipMemcpy(h, p + offset, count * sizeof(T), hipMemcpyDeviceToHost), "download");
    }
};

// reference initial condition (Simple2DCell.cpp:85-100, Simple2DActiveCell.cpp:55-70):
// std::mt19937(seed), x then y per cell, then the directors, uniform in [0,L) / [0,2pi)
void reference_random_init(int n, uint32_t seed, double L, double *x, double *y, double *theta);

// Per-system constants of a new batch: box side L (L <= 0: sqrt(n), unit
// density), the cell-list grid of the reference cellListGPU::setGridSize(1.25)
// (even number of bins per side, at least 2), moduli 1, dt 0.01, mu 1, T 0.
cgpu_sys_params default_sys_params(int n, double L, uint64_t seed);

// Offsets and work-chunk tables of a batch of systems (the n and xsize of every
// system are read; off, ncells and bin_off are filled in).
struct BatchTables {
    int N = 0, total_bins = 0;
    std::vector<int> cell_sys;         // [N]
    std::vector<int> cs, cb, ce;       // cell chunks: system, begin, end
    std::vector<int> bs, bb, be, sb;   // bin chunks: system, begin, end; first bin chunk per system [S+1]
};
BatchTables batch_tables(std::vector<cgpu_sys_params> &hp);
// capacity (records) of the circumcircle test's region list for N cells
inline int ld_region_capacity(int N) { return N / 8 > 1024 ? N / 8 : 1024; }
// capacity (triangles) of the circumcircle test's triangle queue for N cells
inline int ld_tri_capacity(int N) { return 2 * N + 1024; }

class VoronoiBatch {
   public:
    VoronoiBatch(int num_systems, const int *n, const double *L, const uint64_t *seed, const double *x,
                 const double *y, const double *theta, int precision, int device);
    ~VoronoiBatch();

    int num_systems() const { return S_; }
    int num_cells() const { return N_; }
    int precision() const { return prec_; }
    hipStream_t stream() const { return stream_; }

    void set_params(int s, double KA, double KP, double dt, double mu, double T);
    void set_prefs(const double *A0, const double *P0);
    void set_motility(const double *v0, const double *Dr);
    void set_exclusions(const int *ex);
    // line tension between cell types (VoronoiQuadraticEnergyWithTension)
    void set_cell_types(const int *types);
    void set_tension(int s, int mode, double gamma);
    void set_tension_matrix(int ntypes, const double *g);
    void set_positions(const double *x, const double *y, bool retriangulate);
    void set_directors(const double *theta);
    void set_velocities(const double *vx, const double *vy);
    void set_topology(const int *nn, const int *nbr, int cap, const int *oth);
    void get_del_other(int *oth, int cap);
    void set_step_counters(const int *step);
    // reference tagToIdx / idxToTag (Simple2DCell.h): tag (label within the
    // system at creation) of each cell; spatial sorting permutes it
    const std::vector<int> &tags() const { return tag_; }
    const std::vector<int> &cell_types() const { return h_type_; }
    const int *exclusions_device() const { return excl_.p; }
    const void *ext_force_device() const { return ext_.p; }
    void real_to_host(const void *src, double *dst, size_t n) { real_download(src, dst, n); }
    void set_noise_mode(bool use_buffer);
    void upload_noise(const double *eta, int count);
    void set_active(const int *active, const int *steps_left);
    void get_active(int *active, int *steps_left);
    size_t device_bytes() const;

    int global_triangulation();
    int compute_forces();
    int enforce_topology();
    int spv_steps(int steps, bool use_graph);
    int spv_steps_timed(int steps, double *ms_out);
    // steps with an event after every kernel: summed ms per profile slot
    // (CGPU_PROFILE_SLOTS) and summed queue sizes (flagged, NeedsFixing, global)
    int spv_steps_profiled(int steps, double *ms_slots, long long *queue_sums);
    int brownian_steps(int steps);
    // ring storage (slots per cell) and its growth, outside any graph
    int ring_capacity() const { return ring_cap_; }
    void grow_ring_capacity(int cap);
    // FIRE systems that stalled on the ring capacity: grown, retriangulated, the
    // iteration's force half completed
    int finish_stalled_fire(cgpu_fire_state *fs, void *vel, double *dots);
    int sync();
    int status();

    void get(double *x, double *y, double *theta, double *vx, double *vy, double *Fx, double *Fy, double *A,
             double *P);
    void get_topology(int *nn, int *nbr, int cap, int *neigh_max);
    void get_counters(int s, long long *out);
    double energy(int s);
    cgpu_voronoi_batch view() const;  // current parity
    const std::vector<cgpu_sys_params> &params() const { return hp_; }

    // raw device access for other engines (FIRE, vertex model initialisation)
    double *pos_dev() { return parity_ ? pos1_.p : pos0_.p; }
    void *force_dev() { return force_.p; }
    void swap_parity() { parity_ ^= 1; }
    // launch shapes of this batch's kernels (a zero field: default)
    void set_shape(const cgpu_launch_shape &sh);
    const cgpu_launch_shape &shape() const { return shape_; }
    // Search of the launch shapes on this batch's current state (restored
    // afterwards: tuning does not advance the simulation); the result is cached
    // per device and workload shape and applied. trial_steps: steps per timing.
    int autotune(int trial_steps, cgpu_launch_shape *out, int *from_cache);
    // reference spatialSorting (Simple2DCell.cpp:438-477, voronoiModelBase.cpp:669-680):
    // relabel the cells of every system along a Hilbert curve of their
    // positions, then retriangulate; perm[new] = old (global indices)
    int spatial_sort(int *perm_out);
    // reference voronoiModelBase::cellDeath / cellDivision (voronoiModelBase.cpp:1291-1422):
    // at most one event per system; kind 0 death, 1 division (theta, separation
    // fraction; the daughter, appended at the end of its system, gets director
    // new_director, or a uniform draw when that is NaN). Re-lays the batch out and
    // retriangulates.
    int cell_events(int count, const int *kind, const int *cell, const double *theta, const double *frac,
                    const double *new_director);
    // A new batch holding system sys[k] of batch src[k] as its system k, every
    // piece of its state copied exactly (positions, velocities, directors,
    // preferences, motility, rings and delOther with their indices shifted, the
    // per-system parameters and device state, tags): stepping it gives the same
    // trajectories as the sources would. The sources must share precision and
    // device and use no exclusions, noise buffer or line tension.
    static VoronoiBatch *merge(int count, VoronoiBatch *const *src, const int *sys);

   private:
    struct Empty {};
    // streams and per-batch buffers only; layout() and uploads follow
    VoronoiBatch(Empty, int num_systems, int precision, int device);
    void upload_params();
    void real_upload(void *dst, const double *src2, size_t count);
    void real_download(const void *src, double *dst, size_t count);
    int step_launch(int kind);
    int launch_spv_steps(int steps, bool use_graph);
    void mark_call_start();
    int after_steps(int steps, int kind);
    bool take_stalled(std::vector<int> &stalled);
    void resume(const std::vector<int> &stalled);
    int finish_stalled(int steps, int kind, const std::vector<int> &step_before);
    int retriangulate_active();
    int ring_cap_ = CGPU_RING_CAP;
    void drop_graph();
    void launch_geometry_only(const cgpu_voronoi_batch &b);
    // device copy of every array a step mutates (for the tuner's trials)
    struct Snapshot;
    void take_snapshot(Snapshot &sn);
    void restore_snapshot(const Snapshot &sn);
    double time_steps(int steps);
    cgpu_launch_shape shape_{};

    int S_ = 0, N_ = 0, prec_ = 1, device_ = 0, total_bins_ = 0;
    size_t rsz_ = 4;
    int parity_ = 0;
    bool noise_buffer_ = false;
    hipStream_t stream_ = nullptr;
    hipGraphExec_t graph_exec_ = nullptr;
    hipGraph_t graph_ = nullptr;
    int graph_parity_ = -1;
    std::vector<cgpu_sys_params> hp_;
    std::vector<int> cell_sys_h_;
    DevBuf<cgpu_sys_params> sys_;
    DevBuf<cgpu_sys_state> st_, st_call_;
    DevBuf<int> cell_sys_, nn_, nbr_, oth_, excl_, repair_, mark_, flagged_, nfq_, gq_, counters_;
    DevBuf<int> bin_start_, cell_bin_, cell_rank_, sorted_id_, ld_tri_, bin_mark_wide_;
    DevBuf<float> cell_rmax_;
    DevBuf<unsigned char> ld_regions_, sort_work_;
    DevBuf<unsigned long long> ld_part_;
    int ld_region_cap_ = 0, ld_tri_cap_ = 0;
    DevBuf<int> chunk_sys_, chunk_begin_, chunk_end_, bchunk_sys_, bchunk_begin_, bchunk_end_, sys_bchunk_,
        bchunk_sum_;
    int num_chunks_ = 0, num_bin_chunks_ = 0;
    DevBuf<unsigned int> bin_count_;
    DevBuf<double> pos0_, pos1_, noise_, sorted_pos_;
    DevBuf<unsigned char> AP_, dAP_, APpref_, mot_, vel_, theta_, force_, ext_;
    DevBuf<int> cell_type_;
    std::vector<int> h_type_;  // host copy of the cell types (validation, sorting, energy)
    std::vector<int> tag_;
    DevBuf<double> tension_;
    int ntypes_ = 0;
    bool tension_on_ = false;  // some system has a tension mode and cell types are set
    void update_tension_flag();
    void layout(const std::vector<int> &n);
    size_t device_bytes_ = 0;
};

}  // namespace cgpu_host
