// Cell division and death of Voronoi-model batches on the device (reference
// voronoiModelBase::cellDivision / cellDeath, voronoiModelBase.cpp:1291-1422,
// and the array growth / removal of Simple2DActiveCell::cellDivision /
// cellDeath through growGPUArray / removeGPUArrayElement, functions.h:81-154).
//
// Two kernels: the division axis of each dividing cell (where the line through
// the cell at the given angle crosses its Voronoi polygon, from the current
// rings), and one gather that lays every per-cell array out for the new
// sizes, dropping dead cells and appending each daughter at the end of its
// system. The host only sees one small record per event.
#include <hip/hip_runtime.h>

#include "../../include/cellgpu_amd.h"
#include "cgpu_device.h"

using namespace cgpu;

namespace {

inline hipStream_t S(void *s) { return (hipStream_t)s; }
inline int check() { return hipGetLastError() == hipSuccess ? CGPU_OK : CGPU_ERR_HIP; }

// Voronoi polygon of cell i in the reference's CPU vertex order (vertex k =
// circumcenter of ring k-1 and ring k, relative to the cell), then the two
// crossings of the axis (direction ray, normal c) with its edges, scanned as
// the reference does (voronoiModelBase.cpp:1344-1398). out: the cell's
// position, first and second crossing (the first again when only one exists),
// NaN when the axis misses the polygon.
__global__ void k_division_axis(const cgpu_voronoi_batch b, int count, const int *parent, const double *ray,
                                double *out) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= count) return;
    const int i = parent[e], N = b.num_cells;
    const cgpu_sys_params &sp = b.sys[b.cell_sys[i]];
    const Box box = make_box(sp.L);
    const dvec2 pi{b.pos[2 * i], b.pos[2 * i + 1]};
    const int z = b.nn[i];
    double *o = out + 6 * e;
    o[0] = pi.x;
    o[1] = pi.y;
    const int cap = b.ring_cap > 0 ? b.ring_cap : CGPU_RING_CAP;
    if (z < 3 || z > cap) {
        for (int k = 2; k < 6; ++k) o[k] = NAN;
        return;
    }
    auto rel = [&](int k) {
        const int j = b.nbr[k * N + i];
        return min_dist(box, dvec2{b.pos[2 * j], b.pos[2 * j + 1]}, pi);
    };
    const dvec2 c{-ray[2 * e + 1], ray[2 * e]};
    // the polygon's last vertex: circumcenter(ring z-2, ring z-1)
    dvec2 rz2 = rel(z - 2), rz1 = rel(z - 1);
    dvec2 v1 = circ0<double>(rz2, rz1);
    dvec2 rij = rz1, int1{0, 0}, int2{0, 0};
    bool found = false, second = false;
    for (int k = 0; k < z; ++k) {
        const dvec2 rik = rel(k);
        const dvec2 v2 = circ0<double>(rij, rik);
        rij = rik;
        const dvec2 a{0.0 - v2.x, 0.0 - v2.y}, d{v1.x - v2.x, v1.y - v2.y};
        double t2 = -1.0;
        const double dc = d.x * c.x + d.y * c.y;
        if (dc != 0) t2 = (a.x * c.x + a.y * c.y) / dc;
        if (t2 >= 0.0 && t2 <= 1.0) {
            const dvec2 q{v2.x + t2 * (v1.x - v2.x), v2.y + t2 * (v1.y - v2.y)};
            if (found) {
                int2 = q;
                second = true;
            } else {
                int1 = q;
                found = true;
            }
        }
        v1 = v2;
    }
    if (!found) {
        for (int k = 2; k < 6; ++k) o[k] = NAN;
        return;
    }
    if (!second) int2 = int1;
    o[2] = int1.x;
    o[3] = int1.y;
    o[4] = int2.x;
    o[5] = int2.y;
}

// new cell j of system s: the old cell it copies (a daughter copies its parent)
__device__ inline int event_source(const cgpu_cell_event &ev, int local) {
    if (ev.kind == 0) return ev.old_off + local + (local >= ev.target ? 1 : 0);
    if (ev.kind == 1 && local >= ev.old_n) return ev.old_off + ev.target;
    return ev.old_off + local;
}

template <typename R>
__global__ void k_events_gather(const cgpu_voronoi_batch b, const cgpu_cell_event *ev, int num_new, double *pos,
                                R *theta, R *vel, R *pref, R *mot, int *excl, int *type) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= num_new) return;
    // system of new cell j: the last s with new_off <= j
    int lo = 0, hi = b.num_systems - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (ev[mid].new_off <= j) lo = mid; else hi = mid - 1;
    }
    const cgpu_cell_event &e = ev[lo];
    const int local = j - e.new_off;
    const int i = event_source(e, local);
    const bool divides = e.kind == 1 && local == e.target, daughter = e.kind == 1 && local >= e.old_n;
    const R *th0 = (const R *)b.theta, *v0 = (const R *)b.vel, *p0 = (const R *)b.APpref, *m0 = (const R *)b.mot;
    if (divides) {
        pos[2 * j] = e.p1[0];
        pos[2 * j + 1] = e.p1[1];
    } else if (daughter) {
        pos[2 * j] = e.p2[0];
        pos[2 * j + 1] = e.p2[1];
    } else {
        pos[2 * j] = b.pos[2 * i];
        pos[2 * j + 1] = b.pos[2 * i + 1];
    }
    pref[2 * j] = p0[2 * i];
    pref[2 * j + 1] = p0[2 * i + 1];
    mot[2 * j] = m0[2 * i];
    mot[2 * j + 1] = m0[2 * i + 1];
    if (daughter) {
        // v = v0 (cos, sin) of the daughter's director, the cosine and sine
        // supplied with the event
        theta[j] = (R)e.director;
        vel[2 * j] = (R)((double)m0[2 * i] * e.dcos);
        vel[2 * j + 1] = (R)((double)m0[2 * i] * e.dsin);
    } else {
        theta[j] = th0[i];
        vel[2 * j] = v0[2 * i];
        vel[2 * j + 1] = v0[2 * i + 1];
    }
    if (excl) excl[j] = daughter ? 0 : b.excl[i];
    if (type) type[j] = b.cell_type[i];
}

inline dim3 grid(long long n) { return dim3((unsigned)((n + 255) / 256)); }

}  // namespace

extern "C" {

int cgpu_vor_division_axis(const cgpu_voronoi_batch *b, int count, const int *parent, const double *ray, double *out,
                           void *stream) {
    if (!b || count < 0 || (count > 0 && (!parent || !ray || !out))) return CGPU_ERR_ARGS;
    if (count == 0) return CGPU_OK;
    hipLaunchKernelGGL(k_division_axis, grid(count), dim3(256), 0, S(stream), *b, count, parent, ray, out);
    return check();
}

int cgpu_vor_events_gather(const cgpu_voronoi_batch *b, const cgpu_cell_event *ev, int new_num_cells, double *pos,
                           void *theta, void *vel, void *APpref, void *mot, int *excl, int *cell_type, void *stream) {
    if (!b || !ev || new_num_cells <= 0 || !pos || !theta || !vel || !APpref || !mot) return CGPU_ERR_ARGS;
    if ((excl != nullptr) != (b->excl != nullptr) || (cell_type != nullptr) != (b->cell_type != nullptr))
        return CGPU_ERR_ARGS;
    if (b->precision)
        hipLaunchKernelGGL(k_events_gather<float>, grid(new_num_cells), dim3(256), 0, S(stream), *b, ev,
                           new_num_cells, pos, (float *)theta, (float *)vel, (float *)APpref, (float *)mot, excl,
                           cell_type);
    else
        hipLaunchKernelGGL(k_events_gather<double>, grid(new_num_cells), dim3(256), 0, S(stream), *b, ev,
                           new_num_cells, pos, (double *)theta, (double *)vel, (double *)APpref, (double *)mot, excl,
                           cell_type);
    return check();
}

}  // extern "C"
