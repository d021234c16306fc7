/*
 * cellgpu_amd — C ABI of the MI355X (gfx950) Voronoi / vertex-model time step.
 *
 * Two layers, both plain C (pointers, sizes, status codes):
 *
 *  1. Kernel-caller layer (cgpu_*). Each entry point replaces one family of the
 *     reference's kernel callers (abaso/cellGPU, declared in the inc/ .cuh headers) and is
 *     cited below. Arguments are device pointers gathered in a batch view struct
 *     (structure-of-arrays, one or many independent systems laid out back to back
 *     with per-system offset tables) plus a hipStream_t passed as void*. Launches
 *     are asynchronous, never allocate, never synchronize the host, and can be
 *     captured into a hipGraph.
 *
 *  2. Object layer (cgpu_vor_*, cgpu_vm_*, cgpu_fire_*, cgpu_server_*). Handles
 *     owned by the C++ classes that keep the reference's Simulation /
 *     voronoiModelBase / vertexModelBase / updater API; used by drivers and by
 *     foreign-language bindings.
 *
 * Status codes: 0 = ok; >0 recoverable / informational; <0 fatal.
 */
#ifndef CELLGPU_AMD_H
#define CELLGPU_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum cgpu_status {
    CGPU_OK = 0,
    CGPU_ERR_ARGS = -1,
    CGPU_ERR_HIP = -2,
    CGPU_ERR_RING_CAPACITY = -3,   /* a Delaunay degree exceeded the ring capacity  */
    CGPU_ERR_BIN_CAPACITY = -4,    /* reserved (the counting-sort cell list has no bin capacity) */
    CGPU_ERR_TOPOLOGY = -5,        /* global triangulation inconsistent / failed    */
    CGPU_ERR_NONFINITE = -6,       /* a position became NaN / inf                   */
    CGPU_ERR_NO_DEVICE = -7,
    CGPU_ERR_SCRATCH_CAPACITY = -8 /* input too large for the static scratch of an allocation-free entry */
};

/* Ring capacity: Delaunay neighbours stored per cell (slot-major storage). A
 * batch starts at CGPU_RING_CAP slots and the object layer grows it, up to
 * CGPU_RING_CAP_MAX (one neighbour per lane of a wavefront in the star
 * kernels), when a star needs more (the reference's neighMax growth,
 * voronoiModelBase.cpp:309-312). */
#define CGPU_RING_CAP 24
#define CGPU_RING_CAP_MAX 64

/* Per-system constants of a batch (host fills, device reads). */
typedef struct cgpu_sys_params {
    int off;            /* first cell (or vertex) of the system            */
    int n;              /* number of cells (degrees of freedom)            */
    int bin_off;        /* first cell-list bin of the system               */
    int xsize;          /* bins per side (reference cellListGPU::xsize)    */
    double L;           /* square box side                                 */
    double boxsize;     /* bin side (L / xsize)                            */
    double KA, KP;      /* quadratic area / perimeter moduli               */
    double dt, mu;      /* integration step, mobility                      */
    double T;           /* Brownian temperature                            */
    uint64_t seed;      /* counter-based noise key                         */
    int ncells;         /* vertex model: number of cells (else = n)        */
    int tension;        /* line tension between cell types: 0 off, 1 one value
                           gamma for every unlike pair, 2 the batch matrix     */
    double gamma;       /* tension of mode 1                               */
} cgpu_sys_params;

/* Device-mutable per-system state of the topology machinery. */
typedef struct cgpu_sys_state {
    int any_fail;       /* a circumcircle test failed this step            */
    int nf_count;       /* |NeedsFixing|                                   */
    int mode;           /* 0 none, 1 local repair, 2 global by count, 3 global fallback */
    int need_global;    /* local repair must be replaced by a global one   */
    int nn_delta;       /* change of sum of degrees by local repairs       */
    int tri_delta;      /* change of owned-triangle count by local repairs */
    int max_deg;        /* max degree seen by the last global pass         */
    int neigh_max;      /* reference neighMax (even rule)                  */
    int step;           /* integrator step counter (noise counter)         */
    int active;         /* 1 = system advances this step                   */
    int steps_left;     /* serving: remaining steps (-1 = unlimited)       */
    int error;          /* sticky cgpu_status for this system              */
    int nflag;          /* cells flagged by the circumcircle test this step */
    int pad_s;
    long long global_fixes, repair_events, local_repairs, t1_events;
    /* Triangulation checks of the circumcircle test's skip rule, per step:
     * wrapping sums of a hash of every ring triangle over all ring slots and
     * over the owned slots (consistent rings: all = 3 x owned), the owned
     * triangles' signed area in units of 2^-30 (one covering of the box: L^2),
     * and a flag that
     * turns the rule off for the system. */
    unsigned long long tri_hash_all, tri_hash_own;
    long long tri_area;          /* fixed point, units of 2^-30 */
    int skip_off, pad_k;
} cgpu_sys_state;

/*
 * Launch shapes of the Voronoi kernels (a zero field takes the built-in
 * default). No field changes a result: every kernel's output is independent of
 * its launch shape (queue order only permutes independent per-cell work).
 * cgpu_vor_autotune searches them per device and workload shape.
 */
typedef struct cgpu_launch_shape {
    int force_block;    /* work-group size (= LDS tile of cells) of the fused force/EOM/move
                           kernel: 128, 256 or 512; default 256                            */
    int test_block;     /* work-group size of the circumcircle test + geometry kernel (its
                           LDS triangle list holds 5 per cell): 128, 256 or 512; default 256 */
    int cell_block;     /* work-group size of the other per-cell kernels: multiple of 64 in
                           [64, 1024]; default 256                                          */
    int star_blocks;    /* 256-thread work-groups of the Delaunay-star kernels, i.e. the repair
                           queue is consumed 4 * star_blocks wavefronts (cells) at a time;
                           default 4096                                                     */
    int repair_blocks;  /* 256-thread work-groups of the grid-stride queue kernels; default 1024 */
    int queue_lanes;    /* threads per queued cell in the queue kernels (64 / queue_lanes cells
                           per wavefront): 4, 8 or 16; default 8                            */
    int force_waves;    /* occupancy target of the fused force kernel: 5 (registers held to
                           5 waves per SIMD, a few bytes of spill) or 0 (the register
                           allocator's choice); default 0                                    */
    int pad_shape;
} cgpu_launch_shape;

/*
 * Batch view of a Voronoi-model batch. All pointers are device pointers.
 * Per-cell arrays have num_cells entries; ring arrays are slot-major:
 * nbr[k * num_cells + i] is the k-th counter-clockwise Delaunay neighbour of cell i.
 * 'real' arrays are float when precision == 1 (FP32 bulk) and double when 0.
 */
typedef struct cgpu_voronoi_batch {
    int num_systems, num_cells, total_bins, precision;
    const cgpu_sys_params *sys;  /* [S] */
    cgpu_sys_state *st;          /* [S] */
    const int *cell_sys;         /* [N] system of each cell */
    double *pos;                 /* [2N] x,y interleaved (current)  */
    double *pos_next;            /* [2N] x,y interleaved (next)     */
    int *nn;                     /* [N] Delaunay degree             */
    int *nbr;                    /* [ring_cap*N] rings              */
    int *oth;                    /* [ring_cap*N] reference delOther */
    void *AP;                    /* [2N] real: area, perimeter      */
    void *dAP;                   /* [2N] real: KA(A-A0), KP(P-P0)   */
    const void *APpref;          /* [2N] real: A0, P0               */
    const void *mot;             /* [2N] real: v0, Dr               */
    void *vel;                   /* [2N] real: velocity             */
    void *theta;                 /* [N]  real: director angle       */
    void *force;                 /* [2N] real: force                */
    const int *excl;             /* [N] or NULL: exclusions         */
    void *ext_force;             /* [2N] real or NULL               */
    unsigned int *bin_count;     /* [total_bins] cells per bin (counting pass)      */
    int *bin_start;              /* [total_bins+1] first bin-ordered slot of a bin  */
    int *cell_bin;               /* [N] bin of each cell                            */
    int *cell_rank;              /* [N] rank of each cell within its bin; after the
                                    scatter pass, its bin-ordered slot              */
    int *sorted_id;              /* [N] cell indices in bin order                   */
    double *sorted_pos;          /* [2N] positions in bin order                     */
    int *repair;                 /* [N] flag (cleared by consumer)  */
    int *mark;                   /* [N] NeedsFixing mark            */
    int *flagged;                /* [N] list of flagged cells       */
    int *nfq;                    /* [N] NeedsFixing queue           */
    int *gq;                     /* [N] global-retriangulation queue */
    int *counters;               /* [16] device counters: 0 flagged, 1 nfq, 2 gq, 3 bin overflow, 4 ring overflow,
                                    5 nonfinite, 6 unused, 7 test epoch,
                                    8 regions to stamp, 9 triangles queued for the circumcircle test */
    const double *noise;         /* [N] per-step N(0,1) buffer or NULL (in-kernel Philox) */
    /* Work chunks: at most CGPU_CHUNK cells (or CGPU_BIN_CHUNK bins) of one
     * system, so per-system passes spread a large system over many work-groups. */
    const int *chunk_sys;        /* [num_chunks] system of each cell chunk         */
    const int *chunk_begin;      /* [num_chunks] first cell                        */
    const int *chunk_end;        /* [num_chunks] one past the last cell            */
    const int *bchunk_sys;       /* [num_bin_chunks] system of each bin chunk      */
    const int *bchunk_begin;     /* [num_bin_chunks] first bin (batch index)       */
    const int *bchunk_end;       /* [num_bin_chunks] one past the last bin         */
    const int *sys_bchunk;       /* [S+1] first bin chunk of each system           */
    int *bchunk_sum;             /* [num_bin_chunks] scratch of the bin scan       */
    int num_chunks, num_bin_chunks;
    /* Line tension (reference VoronoiQuadraticEnergyWithTension): cell types, and
     * the tension matrix g[a*num_types + b] used by systems in mode 2. cell_type
     * NULL: no system uses tension. */
    const int *cell_type;        /* [N] or NULL                                     */
    const double *tension;       /* [num_types^2] or NULL                           */
    int num_types;
    int ring_cap;                /* ring slots allocated per cell (nbr, oth hold ring_cap * N);
                                    0: CGPU_RING_CAP. A star above it stalls its system:
                                    error CGPU_ERR_RING_CAPACITY, active 0, step not advanced */
    /* Circumcircle-test skip rule: bins within 3 bins of an edge that is not
     * locally Delaunay (or of a ring that is not a proper fan), stamped with the
     * test's epoch (counters[7]); each cell's largest circumradius; the queue of
     * triangles the rule leaves to test. */
    int *ld_tri;                 /* [3 ld_tri_cap] (i, j, l) of each queued triangle */
    int *bin_mark_wide;          /* [total_bins] */
    float *cell_rmax;            /* [N] */
    void *ld_regions;            /* [ld_region_cap] regions to stamp (system or -1-cell, bounds) */
    int ld_region_cap, ld_tri_cap;
    unsigned long long *ld_part; /* [3 (N/64 + 1)] sums of each 64-cell group inside one system */
    cgpu_launch_shape shape;     /* launch shapes (zero fields: defaults)           */
} cgpu_voronoi_batch;

#define CGPU_CHUNK 4096
#define CGPU_COUNTERS 16
/* capacity of the circumcircle test's region list: max(1024, N / 8) records of
 * 40 bytes; of its triangle queue: 2 N + 1024 triangles (a triangulation of the
 * torus has 2 N; more, from rings that disagree, are tested where they arise) */
#define CGPU_LD_REGION_BYTES 40
#define CGPU_BIN_CHUNK 8192

/* Vertices stored per vertex-model cell (reference vertexMax = 30,
 * vertexModelBase.cpp:210): a batch starts at CGPU_VERTEX_CAP slots and the
 * object layer grows it when a T1 flip needs more (reference
 * growCellVerticesList, vertexModelBase.cpp:475-500, 815-824). */
#define CGPU_VERTEX_CAP 32

/* Per-system constants of a vertex-model batch. */
typedef struct cgpu_vm_params {
    int voff, nv;       /* first vertex, vertices (= 2 nc)                 */
    int coff, nc;       /* first cell, cells                               */
    double L;           /* square box side                                 */
    double KA, KP;      /* moduli                                          */
    double dt, mu, T;   /* Brownian step                                   */
    double thr;         /* T1 edge-length threshold                        */
    uint64_t seed;      /* counter-based noise key                         */
    int chunk0;         /* first vertex chunk of the system                */
    int pad_i;
} cgpu_vm_params;

/*
 * Batch view of a vertex-model batch (device pointers). Vertex arrays have
 * num_vertices entries (vn, vcn: 3 per vertex, cells counter-clockwise around
 * the vertex, vn[3v+f] across from vcn[3v+f]); cell vertex lists are slot-major
 * cv[k*num_cells + c], counter-clockwise.
 */
typedef struct cgpu_vertex_batch {
    int num_systems, num_vertices, num_cells, precision;
    const cgpu_vm_params *sys;   /* [S] */
    cgpu_sys_state *st;          /* [S] */
    const int *vert_sys;         /* [Nv] */
    const int *cell_sys;         /* [Nc] */
    double *pos, *pos_next;      /* [2Nv] vertex positions */
    int *vn, *vcn;               /* [3Nv] */
    int *cvn, *cv;               /* [Nc], [vertex_cap*Nc] */
    void *AP, *dAP;              /* [2Nc] real */
    const void *APpref;          /* [2Nc] real */
    void *force;                 /* [2Nv] real */
    int *flag_count;             /* [S] */
    int *flag_list;              /* [6Nv] flagged edges (v1, v2) */
    int *flag_cells;             /* [12Nv] */
    int *flag_touched;           /* [12Nv] */
    unsigned long long *cell_min;/* [Nc], all ones when idle */
    const double *noise;         /* [2Nv + Nc] or NULL: x then y per vertex, then
                                    one per cell (director rotation) */
    void *theta;                 /* [Nc] real: cell director angles       */
    const void *motility;        /* [2Nc] real: (v0, Dr) per cell          */
    /* vertex chunks (at most CGPU_CHUNK vertices of one system), for the
     * per-system FIRE reductions */
    const int *chunk_sys, *chunk_begin, *chunk_end;
    int num_chunks;
    int vertex_cap;              /* cell-list slots allocated per cell (cv holds vertex_cap * Nc);
                                    0: CGPU_VERTEX_CAP. A T1 flip that would exceed it stalls its
                                    system: error CGPU_ERR_RING_CAPACITY, active 0 */
} cgpu_vertex_batch;

/* ---------------------------------------------------------------- kernel layer */

/* Replaces gpu_compute_cell_list + gpu_zero_array (reference
 * inc/utility/cellListGPU.cuh:18-38; src/utility/cellListGPU.cu:20-88).
 * Bins cur (0) or next (1) positions. */
int cgpu_compute_cell_list(const cgpu_voronoi_batch *b, int use_next, void *stream);

/* Replaces gpu_test_circumcenters + gpu_compute_voronoi_geometry (reference
 * inc/models/voronoiModelBase.cuh:20-48; voronoiModelBase.cu:24-201): FP64
 * empty-circumcircle test of every Delaunay triangle with the reference's search
 * window, fused with the Voronoi area/perimeter pass. test=0 only computes geometry. */
int cgpu_test_and_geometry(const cgpu_voronoi_batch *b, int test, void *stream);

/* Replaces voronoiModelBase::testAndRepairTriangulation / repairTriangulation /
 * getDelSets / allDelSets and the CGAL calls behind them (reference
 * voronoiModelBase.cpp:429-505, 600-663, 686-739, 1236-1286; DelaunayLoc.cpp:381-403):
 * device-side NeedsFixing, N/6 global rule, exact-predicate Delaunay stars,
 * neighMax and 6N/2N checks, delOther maintenance, geometry of repaired cells. */
int cgpu_repair_topology(const cgpu_voronoi_batch *b, void *stream);

/* Replaces voronoiModelBase::globalTriangulationCGAL (reference
 * voronoiModelBase.cpp:281-351; DelaunayCGAL.cpp:59-65,164-229) for every system
 * of the batch: bins current positions, computes all Delaunay stars, delOther and
 * geometry. */
int cgpu_global_triangulation(const cgpu_voronoi_batch *b, void *stream);
/* The same for the active systems only (the others keep their rings): used to
 * finish a system that stalled on the ring capacity once the caller has grown
 * the ring storage (b->ring_cap) and resumed it. */
int cgpu_global_triangulation_active(const cgpu_voronoi_batch *b, void *stream);

/* Derived topology of rings set by the caller: delOther of every slot (reference
 * allDelSets / getDelSets, voronoiModelBase.cpp:1236-1286; a slot whose two
 * cells do not share the edge marks the system CGPU_ERR_TOPOLOGY) and the
 * Voronoi geometry of every cell. */
int cgpu_derive_topology(const cgpu_voronoi_batch *b, void *stream);

/* Replaces gpu_force_sets + gpu_sum_force_sets(_with_exclusions) (reference
 * inc/models/voronoiQuadraticEnergy.cuh:20-54; voronoiQuadraticEnergy.cu:30-234):
 * per-cell force of the quadratic area-perimeter energy, fused over the ring. */
int cgpu_voronoi_forces(const cgpu_voronoi_batch *b, void *stream);

/* Replaces gpu_spp_eom_integration + gpu_move_degrees_of_freedom (reference
 * inc/updaters/selfPropelledParticleDynamics.cuh:17-27, selfPropelledParticleDynamics.cu:20-57;
 * inc/models/Simple2DCell.cuh:22-40, Simple2DCell.cu:21-58), fused with the
 * force pass: F, then disp = dt(v + mu F), v <- v0 e(theta),
 * theta <- atan2(v) + sqrt(2 dt Dr) eta, x_next = wrap(x + disp). */
int cgpu_spv_force_eom_move(const cgpu_voronoi_batch *b, void *stream);

/* Per-step bookkeeping kernels: reset of the per-system topology state and
 * queues (begin) and the step counter / serving countdown (end). A step is
 * begin, force_eom_move, cell list (next), test_and_geometry and
 * repair_topology on the next positions, end. */
int cgpu_step_begin(const cgpu_voronoi_batch *b, void *stream);
int cgpu_step_end(const cgpu_voronoi_batch *b, void *stream);

/* One complete self-propelled Voronoi step of every active system
 * (Simulation::performTimestep, reference src/Simulation.cpp:141-166). */
int cgpu_spv_step(const cgpu_voronoi_batch *b, void *stream);

/* Brownian step of every active system (brownianParticleDynamics GPU path,
 * reference brownianParticleDynamics.cpp:68-87). */
int cgpu_brownian_step(const cgpu_voronoi_batch *b, void *stream);

/* Default work-group size of the per-cell kernels (multiple of 64, <= 1024). */
int cgpu_set_block_size(int block);

/* Replaces initialize_RNG_array_kernel / curand draws (reference
 * noiseSource.cu:21-29): fills out[i] with the N(0,1) deviate of cell i
 * at (seed of its system, step, stream). */
int cgpu_fill_normal(const cgpu_voronoi_batch *b, int step, int stream_id, double *out, void *stream);

/* Brownian update of cells (reference brownianParticleDynamics.cu:20-39),
 * fused with the force pass and the move. */
int cgpu_brownian_force_eom_move(const cgpu_voronoi_batch *b, void *stream);

/* FIRE pieces (reference EnergyMinimizerFIRE2D.cu:16-68, utilities.cu:28-35, 134-249).
 * fs is a device array cgpu_fire_state[S]; vel a real [2N] velocity array;
 * dots a double [3N] scratch array (f.f, f.v, v.v per cell). */
typedef struct cgpu_fire_state {
    double alpha, alpha_start, alpha_dec, dt, dt_max, dt_inc, dt_dec, force_cutoff;
    double power, force_max, scaling, mix_alpha, max_f2, pad_d;
    int nmin, nsince, iterations, max_iterations;
    int converged, zero_v, pending, pad_i;
} cgpu_fire_state;
/* velocity-Verlet first half + move (gpu_displacement_vv, gpu_update_velocity,
 * moveDegreesOfFreedom), after applying the previous iteration's FIRE mixing */
int cgpu_fire_vv_move(const cgpu_voronoi_batch *b, cgpu_fire_state *fs, void *vel, void *stream);
/* forces on the moved positions, second velocity half step and the three
 * FP64 dot products, then per-system reduction and the FIRE update of alpha,
 * dt and the stopping test (fireStepGPU). dots: 3 N + 4 num_chunks doubles
 * (per-cell products, then per-chunk partial sums). */
int cgpu_fire_force_reduce(const cgpu_voronoi_batch *b, cgpu_fire_state *fs, void *vel, double *dots, void *stream);
/* mark systems whose FIRE loop has stopped as inactive for the topology kernels */
int cgpu_fire_park_converged(const cgpu_voronoi_batch *b, const cgpu_fire_state *fs, void *stream);
/* One FIRE iteration of every running system, fused: the sequence
   cgpu_step_begin, cgpu_fire_vv_move, cgpu_compute_cell_list(use_next = 1),
   then on the moved positions cgpu_test_and_geometry, cgpu_repair_topology,
   cgpu_fire_force_reduce and cgpu_fire_park_converged, with the reset and the
   binning pass inside the move kernel and the pending-mix clear inside the force
   pass. Afterwards b->pos_next holds the new positions (the caller swaps).
   Replaces the per-iteration kernel sequence of EnergyMinimizerFIRE2D.cpp:93-114
   and :156-207. */
int cgpu_fire_iteration(const cgpu_voronoi_batch *b, cgpu_fire_state *fs, void *vel, double *dots, void *stream);
/* The same iteration with no allocation and no workspace argument, for callers
 * that cannot call hipMalloc (code under HIP graph capture, tight serving loops).
 * No counterpart in the reference. The FIRE products and partial sums live in a
 * file-scope device array of cgpu_fire_noalloc_capacity() doubles (8 MiB) per
 * GPU, held until the process ends; a batch needing more than that
 * (3 num_cells + 4 num_chunks doubles) gets CGPU_ERR_SCRATCH_CAPACITY
 * (cgpu_last_error says why) and nothing is launched. Every call on one GPU
 * shares that array, so calls that overlap on one GPU (other streams or host
 * threads) race; nothing in it needs clearing between calls. Replaces the same
 * kernel sequence as cgpu_fire_iteration. */
int cgpu_fire_iteration_noalloc(const cgpu_voronoi_batch *b, cgpu_fire_state *fs, void *vel, void *stream);
size_t cgpu_fire_noalloc_capacity(void);
/* apply a pending FIRE velocity mixing (end of minimize) */
int cgpu_fire_apply_mix(const cgpu_voronoi_batch *b, cgpu_fire_state *fs, void *vel, void *stream);

/* Caller-owned kernel-layer batches (the reference's kernels take arrays their
 * callers own, GPUArray in voronoiModelBase.h:26-107; here one allocation holds
 * them all). cgpu_vor_default_params fills the per-system constants of a new
 * batch (L NULL: sqrt(n); seed NULL: 13377) with its offsets; a caller may edit
 * KA, KP, dt, mu, T, seed, tension, gamma. cgpu_vor_workspace_query gives the
 * array counts and the bytes of one device allocation holding every array of
 * the batch; cgpu_vor_workspace_bind carves such an allocation (256-byte
 * aligned, at least ws->bytes; allocated by the caller) into a batch view:
 * every array zeroed, the layout tables and parameters uploaded, systems active.
 * It allocates nothing and returns after its copies have finished (stream).
 * The caller then writes pos (x, y interleaved), APpref (A0, P0), mot (v0, Dr),
 * vel and theta in the batch precision, builds the rings with
 * cgpu_global_triangulation, and runs the step entry points, swapping pos and
 * pos_next after every step (the step writes pos_next). Optional arrays (excl,
 * ext_force, noise, cell_type, tension) start NULL. */
typedef struct cgpu_vor_workspace {
    int num_systems, num_cells, total_bins, num_chunks, num_bin_chunks, precision;
    long long bytes;
} cgpu_vor_workspace;
int cgpu_vor_default_params(int num_systems, const int *n, const double *L, const uint64_t *seed,
                            cgpu_sys_params *out);
int cgpu_vor_workspace_query(int num_systems, const cgpu_sys_params *params, int precision, cgpu_vor_workspace *ws);
int cgpu_vor_workspace_bind(const cgpu_vor_workspace *ws, const cgpu_sys_params *params, void *mem, size_t mem_bytes,
                            cgpu_voronoi_batch *out, void *stream);

/* Vertex model, replacing gpu_vm_geometry (reference vertexModelBase.cuh:19-30,
 * vertexModelBase.cu:20-77), gpu_avm_force_sets + gpu_avm_sum_force_sets
 * (vertexQuadraticEnergy.cuh:20-33, .cu:21-70), gpu_brownian_eom_integration
 * (brownianParticleDynamics.cuh:17-24), gpu_vm_test_edges_for_T1 +
 * gpu_vm_parse_multiple_flips + gpu_vm_flip_edges (vertexModelBase.cuh:40-79,
 * vertexModelBase.cu:118-544). */
int cgpu_vm_geometry(const cgpu_vertex_batch *b, void *stream);
int cgpu_vm_forces(const cgpu_vertex_batch *b, void *stream);
int cgpu_vm_brownian_step(const cgpu_vertex_batch *b, void *stream);
int cgpu_vm_t1(const cgpu_vertex_batch *b, void *stream);
/* One step of self-propelled cell-vertex dynamics, replacing
 * gpu_spp_cellVertex_eom_integration (reference selfPropelledCellVertexDynamics.cuh,
 * selfPropelledCellVertexDynamics.cu:21-71, .cpp:51-76): forces, vertex
 * displacement dt (mean of v0 e(theta) over the three cells + mu F), director
 * rotation theta += sqrt(2 dt Dr) eta, move, T1 test and flips. */
int cgpu_vm_spp_step(const cgpu_vertex_batch *b, void *stream);
/* FIRE on a vertex batch (the degrees of freedom are the vertices; reference
 * EnergyMinimizerFIRE2D.cpp:93-114, 156-207 with State = vertex model, whose
 * enforceTopology is the T1 pass, vertexModelBase.cpp:93-106). An iteration is
 * vv_move, then cgpu_vm_t1 and force_reduce on the moved positions.
 * dots: 3 Nv + 4 num_chunks doubles. */
int cgpu_vm_fire_vv_move(const cgpu_vertex_batch *b, cgpu_fire_state *fs, void *vel, void *stream);
int cgpu_vm_fire_force_reduce(const cgpu_vertex_batch *b, cgpu_fire_state *fs, void *vel, double *dots,
                              void *stream);
int cgpu_vm_fire_park_converged(const cgpu_vertex_batch *b, const cgpu_fire_state *fs, void *stream);
/* One vertex-model FIRE iteration of every running system, fused: the sequence
   cgpu_vm_fire_vv_move, then on the moved positions cgpu_vm_t1,
   cgpu_vm_fire_force_reduce and cgpu_vm_fire_park_converged, with the
   pending-mix clear inside the force pass. Afterwards b->pos_next holds the new
   positions (the caller swaps). Replaces the per-iteration kernel sequence of
   EnergyMinimizerFIRE2D.cpp:93-114 and :156-207 on a vertex model. */
int cgpu_vm_fire_iteration(const cgpu_vertex_batch *b, cgpu_fire_state *fs, void *vel, double *dots, void *stream);
int cgpu_vm_fire_apply_mix(const cgpu_vertex_batch *b, cgpu_fire_state *fs, void *vel, void *stream);

/* Spatial sorting on the device, replacing Simple2DCell::spatiallySortCells and
 * spatiallySortVertices with the HilbertSorter (reference Simple2DCell.cpp:438-599,
 * inc/utility/HilbertSort.h:27-104). A system of box side L uses the reference's
 * curve order M = min(30, mm + 4), 2^mm >= L (capped at 20), and its pixel
 * (floor(2^M x / L), floor(2^M y / L)); ties keep the current order. Nothing is
 * allocated: work is caller-owned device memory of the queried size.
 *
 * cgpu_vor_hilbert_order: perm[new] = old cell index (global indices; cells
 * stay inside their system's range). cgpu_vor_permute_cells: re-indexes the
 * per-cell arrays of the view in place (pos, theta, vel, APpref, mot, force and,
 * when present, ext_force, excl, cell_type) through tmp (16 num_cells bytes).
 * The rings then refer to old indices: rebuild them with
 * cgpu_global_triangulation, as the reference's spatialSorting does
 * (voronoiModelBase.cpp:669-680). */
size_t cgpu_vor_sort_workspace(int num_cells);
int cgpu_vor_hilbert_order(const cgpu_voronoi_batch *b, int *perm, void *work, size_t work_bytes, void *stream);
int cgpu_vor_permute_cells(const cgpu_voronoi_batch *b, const int *perm, void *tmp, void *stream);
/* Vertex model (vertexModelBase::spatialSorting, vertexModelBase.cpp:236-242):
 * vertices along the curve (vperm[new] = old), then cells in order of first
 * appearance around the sorted vertices (cperm[new] = old). cgpu_vm_permute
 * re-indexes positions, forces, vn / vcn, cvn / cv (ids mapped, list order
 * kept), APpref, AP, dAP, theta and motility in place. */
size_t cgpu_vm_sort_workspace(int num_vertices, int num_cells);
int cgpu_vm_hilbert_order(const cgpu_vertex_batch *b, int *vperm, int *cperm, void *work, size_t work_bytes,
                          void *stream);
size_t cgpu_vm_permute_workspace(int num_vertices, int num_cells);
int cgpu_vm_permute(const cgpu_vertex_batch *b, const int *vperm, const int *cperm, void *work, size_t work_bytes,
                    void *stream);

/* Cell division and death on the device, replacing the array growth and removal
 * of voronoiModelBase::cellDivision / cellDeath (reference
 * voronoiModelBase.cpp:1291-1422; Simple2DActiveCell::cellDivision / cellDeath
 * through growGPUArray / removeGPUArrayElement, functions.h:81-154).
 *
 * cgpu_vor_division_axis: for each dividing cell parent[e] (device array) and
 * axis direction ray[2e], ray[2e+1] = (cos theta, sin theta), writes out[6e..6e+5]
 * = the cell's position, then the first and second crossing of the axis with the
 * cell's Voronoi polygon relative to it (the first twice when only one exists),
 * NaN when the axis misses; scanned in the reference's vertex order. The caller
 * forms the new positions from these (cell +- fraction * max |crossing| * ray,
 * wrapped into the box).
 *
 * cgpu_vor_events_gather: one record per system (device array ev[num_systems],
 * new_off ascending) lays the per-cell arrays out for the new sizes into
 * caller-owned arrays of new_num_cells entries: a death drops its cell (the
 * cells above it move down), a division puts p1 at the parent and appends a
 * daughter at the end of its system, at p2, with the parent's preferences,
 * motility, type, exclusion 0, the director 'director' and the velocity v0 (dcos,
 * dsin). excl / cell_type must be given exactly when the view has them. The
 * rings are then rebuilt by cgpu_global_triangulation on a view of the new
 * layout (the reference's resizeAndReset, voronoiModelBase.cpp:1291-1307). */
typedef struct cgpu_cell_event {
    int kind;             /* -1 none, 0 death, 1 division                  */
    int target;           /* the cell, as an index within its old system   */
    int old_off, old_n;   /* the system in the old layout                  */
    int new_off, pad_e;   /* its first cell in the new layout              */
    double p1[2], p2[2];  /* division: new positions of parent, daughter   */
    double director, dcos, dsin;  /* division: the daughter's director, its cosine and sine */
} cgpu_cell_event;
int cgpu_vor_division_axis(const cgpu_voronoi_batch *b, int count, const int *parent, const double *ray, double *out,
                           void *stream);
int cgpu_vor_events_gather(const cgpu_voronoi_batch *b, const cgpu_cell_event *ev, int new_num_cells, double *pos,
                           void *theta, void *vel, void *APpref, void *mot, int *excl, int *cell_type, void *stream);

/* Vertex-model cell death (T2 of a triangular cell) and division through two
 * edges on the device, replacing vertexModelBase::cellDeath / cellDivision
 * (reference vertexModelBase.cpp:905-1385). One record per system (device array
 * ev[num_systems]); kind -1 leaves the system as it is.
 *
 * cgpu_vm_event_prepare, on the current layout: checks each event and fills the
 * record's local ids (death: the triangle's vertices v[0] < v[1], v[2], the
 * three cells around it and the three vertices beyond it; division: the edges'
 * end vertices a, an, b, bn and the cells na, nb across them) and `grow`, the
 * longest cell list the event makes; status 0 or the reason it is refused
 * (1 a death of a cell that is not a triangle, 2 a system would keep fewer than
 * 4 cells, 3 the triangle is not simple, 4 division positions out of range,
 * 5 an edge without a neighbouring cell).
 *
 * cgpu_vm_events_gather: the caller fills new_voff / new_coff (the new layout:
 * a death drops two vertices and one cell, a division appends two vertices and
 * a daughter cell at the end of its system) and the daughter's director, then
 * this gathers every per-vertex and per-cell array into caller-owned arrays of
 * the new sizes (cell lists with cap slots), ids mapped (the dying triangle's
 * vertices to v[0]). cgpu_vm_events_surgery then makes the local topology
 * change on a view of the new layout: the merged vertex and the lists around it,
 * or the two new vertices (edge midpoints), the split lists and the neighbours'
 * lists, with the slot convention of every vertex it touches restored. The
 * geometry is then stale: cgpu_vm_geometry. */
typedef struct cgpu_vm_event {
    int kind;                    /* -1 none, 0 death, 1 division                       */
    int cell;                    /* the cell, index within its old system             */
    int p1, p2;                  /* division: positions in the cell's vertex list     */
    int old_voff, old_nv, old_coff, old_nc;
    int new_voff, new_coff;      /* set by the caller before the gather                */
    int v[3], around[3], outer[3];            /* death (prepare), old local ids         */
    int a, an, b, bn, na, nb;    /* division (prepare), old local ids                  */
    int status, grow;            /* prepare                                           */
    double director;             /* division: the daughter's director                  */
} cgpu_vm_event;
int cgpu_vm_event_prepare(const cgpu_vertex_batch *b, cgpu_vm_event *ev, void *stream);
int cgpu_vm_events_gather(const cgpu_vertex_batch *b, const cgpu_vm_event *ev, int new_nv, int new_nc, int cap,
                          double *pos, int *vn, int *vcn, int *cvn, int *cv, void *APpref, void *theta, void *mot,
                          void *stream);
int cgpu_vm_events_surgery(const cgpu_vertex_batch *nb, const cgpu_vm_event *ev, void *stream);

/* ---------------------------------------------------------------- object layer */

/* A batch of independent Voronoi systems (one system = the reference's
 * VoronoiQuadraticEnergy; many = an ensemble launched together).
 * n[s], L[s]: cells and box side; x,y,theta: concatenated host arrays (may be
 * NULL: reference initial condition from mt19937(seed[s])). precision 0 = FP64,
 * 1 = FP32 bulk with FP64 predicates. */
void *cgpu_vor_create(int num_systems, const int *n, const double *L, const uint64_t *seed,
                      const double *x, const double *y, const double *theta, int precision, int device);
void cgpu_vor_destroy(void *h);
int cgpu_vor_set_params(void *h, int s, double KA, double KP, double dt, double mu, double T);
int cgpu_vor_set_prefs(void *h, const double *A0, const double *P0);       /* per cell, all systems */
int cgpu_vor_set_motility(void *h, const double *v0, const double *Dr);    /* per cell; resets v   */
int cgpu_vor_set_exclusions(void *h, const int *ex);
/* External force per cell (reference VoronoiQuadraticEnergy::external_forces,
 * voronoiQuadraticEnergy.cu:59-93): minus the summed force sets of an excluded
 * cell, zero elsewhere (all zero before any exclusions are set). */
int cgpu_vor_get_external_forces(void *h, double *fx, double *fy);
/* Line tension between cell types (reference VoronoiQuadraticEnergyWithTension,
 * voronoiQuadraticEnergyWithTension.h:16-60; Simple2DCell::setCellType):
 * cell types per cell; mode for system s (-1: all) 0 off, 1 one value gamma
 * (setSurfaceTension(double)), 2 the matrix (setSurfaceTension(vector), laid out
 * g[a*ntypes + b] = gamma between types a and b). */
int cgpu_vor_set_cell_types(void *h, const int *types);
/* director angles per cell (reference cellDirectors; ReadState restores them) */
int cgpu_vor_set_directors(void *h, const double *theta);
/* stored velocities per cell (reference cellVelocities, Simple2DActiveCell.h:
 * the self-propelled displacement of a step uses the velocity stored by the
 * step before) */
int cgpu_vor_set_velocities(void *h, const double *vx, const double *vy);
/* integrator step counter per system (reference Simulation::setCurrentTimestep,
 * Simulation.h:71; here it also keys the counter-based noise of the next step) */
int cgpu_vor_set_step_counters(void *h, const int *step);
/* tag (label within its system at creation) of every cell; spatial sorting
 * permutes cells, the tags follow them (reference tagToIdx / idxToTag) */
int cgpu_vor_get_tags(void *h, int *tags);
/* Cell death and division (reference voronoiModelBase::cellDeath / cellDivision,
 * voronoiModelBase.cpp:1291-1422; Simple2DCell.cpp:775-850; Simple2DActiveCell.cpp:
 * 118-147), at most one event per system per call. kind[e] 0: death of cell[e]
 * (batch index; the cells after it in its system move down one index). kind[e] 1:
 * division of cell[e] along the axis at angle theta[e], parent and daughter at
 * +-frac[e] times the largest distance from the cell to its Voronoi polygon along
 * that axis; the daughter is appended at the end of its system with the parent's
 * preferences, motility and type and director new_director[e] (NaN, or
 * new_director NULL: a uniform draw of the batch's generator). Systems keep their
 * boxes; the batch is laid out again and retriangulated, so batch views and
 * graphs taken before are invalid. */
int cgpu_vor_cell_events(void *h, int count, const int *kind, const int *cell, const double *theta,
                         const double *frac, const double *new_director);
/* cells per system (n[S]); returns the total */
int cgpu_vor_get_sizes(void *h, int *n);
int cgpu_vor_set_tension(void *h, int s, int mode, double gamma);
int cgpu_vor_set_tension_matrix(void *h, int ntypes, const double *g);
int cgpu_vor_set_positions(void *h, const double *x, const double *y, int retriangulate);
/* Replace the Delaunay rings of every cell (reference cellNeighborNum /
 * cellNeighbors, Simple2DCell.h:200-202, public arrays that callers may
 * overwrite before allDelSets): nn[i] neighbours of cell i in
 * nbr[i * cap + k], counter-clockwise, batch indices inside the cell's system.
 * Checked before upload (degree in [3, CGPU_RING_CAP], neighbours distinct and
 * in the same system, the 6N degree sum per system); geometry and neighMax are
 * derived on the device, and delOther too when oth is NULL (else oth[i * cap + k]
 * is the delOther of slot k, reference voronoiModelBase.h:195-199). */
int cgpu_vor_set_topology(void *h, const int *nn, const int *nbr, int cap, const int *oth);
/* delOther of every slot (oth[i * cap + k]); a local repair keeps a slot's
 * previous value when the neighbour's ring lacks the edge, as getDelSets does
 * (voronoiModelBase.cpp:1236-1275) */
int cgpu_vor_get_del_other(void *h, int *oth, int cap);
int cgpu_vor_set_noise_mode(void *h, int use_buffer);
int cgpu_vor_upload_noise(void *h, const double *eta);                      /* N doubles, this step */
int cgpu_vor_compute_forces(void *h);
/* reference voronoiModelBase::enforceTopology (voronoiModelBase.cpp:686-739) on
 * the current positions (after positions were moved by the caller, e.g. with
 * cgpu_vor_set_positions(..., 0)): cell list, circumcircle test and repair;
 * step counters unchanged */
int cgpu_vor_enforce_topology(void *h);
int cgpu_vor_spv_steps(void *h, int steps, int use_graph);
int cgpu_vor_brownian_steps(void *h, int steps);
int cgpu_vor_sync(void *h);
int cgpu_vor_get(void *h, double *x, double *y, double *theta, double *vx, double *vy, double *Fx,
                 double *Fy, double *A, double *P);
int cgpu_vor_get_topology(void *h, int *nn, int *nbr, int cap, int *neigh_max);
int cgpu_vor_get_counters(void *h, int s, long long *out);
int cgpu_vor_num_cells(void *h);
/* ring slots per cell now allocated (CGPU_RING_CAP, grown up to CGPU_RING_CAP_MAX
 * when a star needs more: the reference's neighMax storage growth) */
int cgpu_vor_ring_capacity(void *h);
double cgpu_vor_energy(void *h, int s);
int cgpu_vor_status(void *h);
int cgpu_vor_set_active(void *h, const int *active, const int *steps_left);
/* Serving layer: a new batch whose system k is system systems[k] of batch
 * batches[k], with every piece of its state copied exactly (positions, stored
 * velocities, directors, preferences, motility, Delaunay rings and delOther,
 * per-system parameters, step counter, remaining steps, repair counters, tags),
 * so each system continues its trajectory bit for bit. Used to drop finished
 * systems and admit new ones at step boundaries. The sources are unchanged;
 * they must share precision and device and use no exclusions, noise buffer or
 * line tension. NULL on error (cgpu_last_error). */
void *cgpu_vor_merge(int count, void *const *batches, const int *systems);
/* Serving layer (no counterpart in the reference, which runs one system per
 * process): per-system active flags and remaining steps after the work queued
 * so far (waits for the batch stream), and the device memory a batch holds, for
 * admission against an HBM budget. */
int cgpu_vor_get_active(void *h, int *active, int *steps_left);
long long cgpu_vor_device_bytes(void *h);
/* Autotuner (no counterpart in the reference, whose launch shapes are fixed:
 * voronoiModelBase.cu:221-223, 128-thread blocks): coordinate search over the
 * launch shapes of cgpu_launch_shape (test and force work-group sizes, star
 * grid = repair-queue depth in wavefronts, queue-kernel grid, threads per queued
 * cell, per-cell work-group size), each candidate timed with HIP events over
 * trial_steps steps from the same saved state. The batch's state is restored
 * afterwards (tuning does not advance the simulation), the best shape is applied
 * and cached for the process per device and workload shape (systems, cells,
 * precision, tension, noise source): *from_cache = 1 when it came from there.
 * Results do not depend on the shape. */
int cgpu_vor_autotune(void *h, int trial_steps, cgpu_launch_shape *out, int *from_cache);
int cgpu_vor_set_launch_shape(void *h, const cgpu_launch_shape *shape);
int cgpu_vor_get_launch_shape(void *h, cgpu_launch_shape *shape);
/* Spatial sorting (reference Simulation::setSortPeriod -> spatialSorting,
 * Simple2DCell.cpp:438-477, voronoiModelBase.cpp:669-680): relabels the cells of
 * every system along a Hilbert curve of their positions (all per-cell state
 * follows) and retriangulates. perm (may be NULL) receives new -> old. */
int cgpu_vor_spatial_sort(void *h, int *perm);
int cgpu_vor_view(void *h, cgpu_voronoi_batch *out);
void *cgpu_vor_stream(void *h);

/* A batch of vertex-model systems built from the Voronoi tessellation of the
 * given cell positions (reference vertexModelBase::setCellsVoronoiTesselation,
 * vertexModelBase.cpp:115-229). Vertex ids follow (smallest cell, ring slot) of
 * the Delaunay triangles. x, y NULL: reference random points from seed[s].
 * spv_initialize: first relax the points by 100 self-propelled Voronoi steps
 * (dt 0.1, P0 3.8, v0 0.1, Dr 1), as the reference's spvInitialize does. */
void *cgpu_vm_create(int num_systems, const int *ncells, const double *L, const uint64_t *seed, const double *x,
                     const double *y, int precision, int device, int spv_initialize);
void cgpu_vm_destroy(void *h);
int cgpu_vm_num_vertices(void *h);
int cgpu_vm_set_params(void *h, int s, double KA, double KP, double dt, double mu, double T, double t1_threshold);
int cgpu_vm_set_prefs(void *h, const double *A0, const double *P0);       /* per cell */
int cgpu_vm_set_positions(void *h, const double *vx, const double *vy);  /* per vertex */
int cgpu_vm_set_noise_mode(void *h, int use_buffer);
int cgpu_vm_upload_noise(void *h, const double *eta);                     /* 2 Nv: x block, y block */
int cgpu_vm_upload_cell_noise(void *h, const double *eta);                /* Nc: director rotation */
int cgpu_vm_compute_forces(void *h);
/* reference vertexModelBase::enforceTopology (vertexModelBase.cpp:93-106): T1
 * test and flips on the current vertex positions, then the geometry */
int cgpu_vm_enforce_topology(void *h);
int cgpu_vm_brownian_steps(void *h, int steps, int use_graph);
/* Self-propelled cell-vertex dynamics (reference selfPropelledCellVertexDynamics):
 * per-cell (v0, Dr) and director angles. The directors start as the reference's
 * setCellDirectorsRandomly draws after the positions (Simple2DActiveCell.cpp:55-70). */
int cgpu_vm_set_motility(void *h, const double *v0, const double *Dr);    /* per cell */
int cgpu_vm_set_directors(void *h, const double *theta);                 /* per cell */
int cgpu_vm_get_directors(void *h, double *theta);
int cgpu_vm_spp_steps(void *h, int steps, int use_graph);
/* per-system activity (inactive systems are skipped by every step and by FIRE) */
int cgpu_vm_set_active(void *h, const int *active);
int cgpu_vm_get_active(void *h, int *active);
/* Replace the topology (vn, vcn [3 Nv]; cvn [Nc]; cv [Nc * cap], counter-clockwise
 * lists) of the whole batch, e.g. when a trajectory record is read back. The
 * lists are checked for consistency (every vertex in the lists of its three
 * cells, cell sizes within the capacity) before anything is uploaded. */
int cgpu_vm_set_topology(void *h, const int *vn, const int *vcn, const int *cvn, const int *cv, int cap);
/* Cell death and division of the vertex model (reference vertexModelBase::cellDeath
 * / cellDivision, vertexModelBase.cpp:905-1110, 1124-1385), at most one event per
 * system per call. kind[e] 0: T2 death of the triangular cell cell[e] (its three
 * vertices merge into the smallest, which keeps its position). kind[e] 1: division
 * of cell[e] through the edges that follow list positions p1[e] and p2[e] of its
 * vertex list (a new vertex at each edge midpoint; the vertices between go to a
 * new cell appended to the system with the parent's preferences and motility and
 * director new_director[e], NaN or NULL: a uniform draw). The batch is laid out
 * again, so views and graphs taken before are invalid. */
int cgpu_vm_cell_events(void *h, int count, const int *kind, const int *cell, const int *p1, const int *p2,
                        const double *new_director);
/* reference vertexModelBase::spatialSorting (vertexModelBase.cpp:236-242,
 * Simple2DCell.cpp:484-599): vertices relabelled along a Hilbert curve, cells in
 * order of first appearance around the sorted vertices, all state re-indexed on
 * the device (cgpu_vm_hilbert_order, cgpu_vm_permute) and the geometry
 * recomputed. vperm / cperm (either may be NULL) receive new -> old. */
int cgpu_vm_spatial_sort(void *h, int *vperm, int *cperm);
/* cell-list slots per cell now allocated (CGPU_VERTEX_CAP, grown when a T1 flip
 * or a cell event needs more: the reference's vertexMax, growCellVerticesList) */
int cgpu_vm_vertex_capacity(void *h);
/* cells and vertices per system (either may be NULL); returns the total vertices */
int cgpu_vm_get_sizes(void *h, int *nc, int *nv);
/* The batch's device arrays as a kernel-layer view, and the stream its work
   runs on (the vertex counterparts of cgpu_vor_view / cgpu_vor_stream; the view
   holds the current position buffer in pos). */
int cgpu_vm_view(void *h, cgpu_vertex_batch *out);
void *cgpu_vm_stream(void *h);

/* ------------------------------------------------------ trajectory databases
 * NetCDF classic (CDF-1) files in the reference's layouts, written by a built-in
 * encoder (no netCDF library):
 *   kind 0: SPVDatabaseNetCDF (DatabaseNetCDFSPV.cpp:29-47): dims rec, Nv, dof,
 *           boxdim, unit; vars time, means0, pos, type, director, BoxMatrix
 *           [, externalForce]; n = cells;
 *   kind 1: AVMDatabaseNetCDF (DatabaseNetCDFAVM.cpp:30-52): dims rec, Nv, Nc,
 *           Nc2, dof, Nvn, boxdim, unit; vars pos, force, VertexCellNeighbors,
 *           Vneighs, cellType, director, cellPositions, BoxMatrix, meanQ, time;
 *           n = vertices.
 * mode (NcFile::FileMode): 0 replace, 1 write (append), 2 read only, 3 new.
 * Arrays are host arrays in tag order; rec < 0 appends. */
void *cgpu_db_open(const char *path, int kind, int n, int mode, int exclusions);
void cgpu_db_close(void *db);
const char *cgpu_db_last_error(void);
int cgpu_db_records(void *db);
int cgpu_db_spv_write(void *db, int rec, double time, double L, double meanq, const double *x, const double *y,
                      const double *director, const int *type, const double *ext_x, const double *ext_y);
int cgpu_db_spv_read(void *db, int rec, double *time, double *box4, double *x, double *y, double *director, int *type,
                     double *ext_x, double *ext_y);
int cgpu_db_avm_write(void *db, int rec, double time, double L, double meanq, const double *vx, const double *vy,
                      const double *fx, const double *fy, const int *vn, const int *vcn, const int *cell_type,
                      const double *director, const double *cell_x, const double *cell_y);
int cgpu_db_avm_read(void *db, int rec, double *time, double *box4, double *vx, double *vy, double *fx, double *fy,
                     int *vn, int *vcn, double *director);
/* One system of a batch to / from a record (reference WriteState / ReadState):
 * Voronoi cells in tag order with type - 5 marking excluded cells and means0 =
 * mean P / sqrt(A); reading sets positions, directors and types and
 * retriangulates. Vertex model: reading rebuilds the cell vertex lists
 * counter-clockwise around each cell's mean vertex position and the vertex
 * neighbour order from the positions. time < 0: step counter x dt. */
int cgpu_db_spv_write_batch(void *db, void *vor, int s, double time, int rec);
int cgpu_db_spv_read_batch(void *db, void *vor, int s, int rec);
int cgpu_db_avm_write_batch(void *db, void *vm, int s, double time, int rec);
int cgpu_db_avm_read_batch(void *db, void *vm, int s, int rec);
int cgpu_vm_get(void *h, double *vx, double *vy, double *Fx, double *Fy, double *A, double *P);
int cgpu_vm_get_topology(void *h, int *vn, int *vcn, int *cvn, int *cv, int cap);
int cgpu_vm_get_counters(void *h, int s, long long *out);
int cgpu_vm_status(void *h);

/* FIRE minimization of every system of a Voronoi batch (EnergyMinimizerFIRE,
 * reference EnergyMinimizerFIRE2D.cpp:269-284, GPU branch). */
void *cgpu_fire_create(void *vor);
void cgpu_fire_destroy(void *h);
int cgpu_fire_set(void *h, int max_it, double force_cutoff, double alpha_start, double dt, double dt_max,
                  double dt_inc, double dt_dec, double alpha_dec, int nmin);
int cgpu_fire_minimize(void *h);
int cgpu_fire_get(void *h, int s, double *out /* force_max, power, alpha, dt, max_f2, iterations */);
/* One parameter of every system, as the reference setters do
 * (EnergyMinimizerFIRE2D.h:37-54): setAlphaStart also sets alpha, setDeltaT sets
 * the current dt; the others change only their own field. */
enum {
    CGPU_FIRE_MAX_IT = 0,
    CGPU_FIRE_FORCE_CUTOFF = 1,
    CGPU_FIRE_ALPHA_START = 2,
    CGPU_FIRE_DT = 3,
    CGPU_FIRE_DT_MAX = 4,
    CGPU_FIRE_DT_INC = 5,
    CGPU_FIRE_DT_DEC = 6,
    CGPU_FIRE_ALPHA_DEC = 7,
    CGPU_FIRE_NMIN = 8
};
int cgpu_fire_set_param(void *h, int which, double value);
/* The same minimizer on a vertex batch (handle used with cgpu_fire_set / _minimize
 * / _get / _destroy); forceMax = sum |F|^2 / Nv per system. */
void *cgpu_vm_fire_create(void *vm);

int cgpu_device_count(void);
int cgpu_device_synchronize(void);
/* Steps with HIP events around each stage (no graph); ms_out[0..3] receive the
 * summed durations of force/EOM/move, cell list, test+geometry, repair chain. */
int cgpu_vor_spv_steps_timed(void *h, int steps, double *ms_out);
/* Steps with a HIP event after every kernel launch (no graph): ms_slots[k]
 * receives the summed duration of profile slot k (the kernels of one step in
 * launch order, named by cgpu_profile_slot_name), queue_sums[0..4] the summed
 * sizes of the flagged, NeedsFixing and global-retriangulation queues, of the
 * circumcircle test's triangle queue and of its list of regions to stamp. */
#define CGPU_PROFILE_SLOTS 14
int cgpu_vor_spv_steps_profiled(void *h, int steps, double *ms_slots, long long *queue_sums);
const char *cgpu_profile_slot_name(int slot);
/* Host-side Philox4x32-10 and the normal pair of (seed, index, step, stream),
 * identical to the device generator; for known-answer tests. */
void cgpu_philox4x32_10(const uint32_t *ctr, const uint32_t *key, uint32_t *out);
void cgpu_normal_pair(uint64_t seed, uint32_t index, uint32_t step, uint32_t stream, double *out);
/* Host copy of the device periodic wrap (reference gpubox::putInBoxReal). */
void cgpu_put_in_box(double L, double *x, double *y);
const char *cgpu_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* CELLGPU_AMD_H */
