// On-device spatial sorting along a Hilbert curve (reference spatialSorting:
// Simple2DCell::spatiallySortCells / spatiallySortVertices, Simple2DCell.cpp:
// 438-599; HilbertSorter, inc/utility/HilbertSort.h:27-104), for the Voronoi and
// vertex-model batches. Keys, the sort (rocPRIM radix sort through hipCUB) and
// the re-indexing of every per-cell and per-vertex array all run on the device;
// only the permutation itself may be copied to the host by the caller.
//
// Order: a system of box side L uses a Hilbert curve of order
// M = min(30, mm + 4) with 2^mm >= L, pixel (floor(2^M x / L), floor(2^M y / L))
// computed as the reference's general-box inverse transform does. Points are
// ordered by (system, curve index, current index), which is the reference's
// std::sort of (index, i) pairs within each system.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <climits>
#include <cstdint>

#include "../../include/cellgpu_amd.h"

namespace {

constexpr int KEY_BITS = 40;   // curve index bits below the system index
constexpr int MAX_ORDER = 20;  // 2 * MAX_ORDER <= KEY_BITS

inline hipStream_t S(void *s) { return (hipStream_t)s; }
inline int check() { return hipGetLastError() == hipSuccess ? CGPU_OK : CGPU_ERR_HIP; }
inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// curve order of the reference HilbertSorter for a square box of side L
// (HilbertSort.h:33-44); orders above MAX_ORDER (L > 65536, where the
// reference's int index overflows) are capped
__device__ inline int curve_order(double L) {
    int mm = 1;
    double t = 2.0;
    while (t < L && mm < 40) {
        t *= 2.0;
        ++mm;
    }
    const int M = mm + 4 < 30 ? mm + 4 : 30;
    return M < MAX_ORDER ? M : MAX_ORDER;
}

// Index of pixel (x, y) along the Hilbert curve filling a 2^M x 2^M grid: the
// classical iterative rotate-and-reflect construction (Hilbert 1891), the same
// curve the reference's xy2d call walks.
__device__ inline uint64_t curve_index(int M, uint32_t x, uint32_t y) {
    uint64_t d = 0;
    for (uint32_t s = 1u << (M - 1); s > 0; s >>= 1) {
        const uint32_t rx = (x & s) ? 1u : 0u, ry = (y & s) ? 1u : 0u;
        d += (uint64_t)s * s * ((3u * rx) ^ ry);
        if (ry == 0) {
            if (rx == 1) {
                x = s - 1 - x;
                y = s - 1 - y;
            }
            const uint32_t t = x;
            x = y;
            y = t;
        }
    }
    return d;
}

// (system << KEY_BITS) | curve index of point p in a box of side L; the pixel
// follows the reference's general-box inverse transform: 1 / (L L) times L,
// times the coordinate (the off-diagonal terms add -0)
__device__ inline uint64_t point_key(int s, double L, double px, double py) {
    const int M = curve_order(L);
    const double inv = (1.0 / (L * L)) * L;
    const double n = (double)(1u << M);
    const int x = (int)floor(n * (inv * px));
    const int y = (int)floor(n * (inv * py));
    return ((uint64_t)s << KEY_BITS) | curve_index(M, (uint32_t)x, (uint32_t)y);
}

__global__ void k_cell_keys(const cgpu_voronoi_batch b, const double *pos, uint64_t *key, int *idx) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= b.num_cells) return;
    const int s = b.cell_sys[i];
    key[i] = point_key(s, b.sys[s].L, pos[2 * i], pos[2 * i + 1]);
    idx[i] = i;
}

__global__ void k_vertex_keys(const cgpu_vertex_batch b, uint64_t *key, int *idx) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= b.num_vertices) return;
    const int s = b.vert_sys[v];
    key[v] = point_key(s, b.sys[s].L, b.pos[2 * v], b.pos[2 * v + 1]);
    idx[v] = v;
}

__global__ void k_fill(unsigned int *a, int n, unsigned int v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ void k_iota(int *a, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = i;
}

// Cell order of the reference spatiallySortVertices (Simple2DCell.cpp:524-543):
// cells numbered by their first appearance walking the sorted vertices and
// each vertex's three cells in order, i.e. by the smallest 3 vv + f at which
// they appear
__global__ void k_cell_first_seen(const cgpu_vertex_batch b, const int *vperm, unsigned int *first) {
    const int vv = blockIdx.x * blockDim.x + threadIdx.x;
    if (vv >= b.num_vertices) return;
    const int v = vperm[vv];
    for (int f = 0; f < 3; ++f) atomicMin(&first[b.vcn[3 * v + f]], (unsigned int)(3 * vv + f));
}

__global__ void k_invert(const int *perm, int *inv, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) inv[perm[i]] = i;
}

// dst[i] = src[perm[i]] for elements of W 4-byte words
template <int W> struct Words { unsigned int w[W]; };
template <int W>
__global__ void k_gather(const Words<W> *__restrict__ src, Words<W> *__restrict__ dst, const int *__restrict__ perm,
                         int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[perm[i]];
}

// new index lists: dst[W i + k] = map[src[W perm[i] + k]]
template <int W>
__global__ void k_gather_remap(const int *__restrict__ src, int *__restrict__ dst, const int *__restrict__ perm,
                               const int *__restrict__ map, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int o = perm[i];
#pragma unroll
    for (int k = 0; k < W; ++k) dst[W * i + k] = map[src[W * o + k]];
}

// slot-major cell vertex lists: new cell c takes old cell cperm[c]'s list with
// vertex ids mapped, unused slots -1
__global__ void k_gather_cell_lists(const int *__restrict__ cv_old, const int *__restrict__ cvn_new,
                                    int *__restrict__ cv_new, const int *__restrict__ cperm,
                                    const int *__restrict__ vinv, int nc) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nc * CGPU_VERTEX_CAP) return;
    const int k = t / nc, c = t - k * nc;
    cv_new[t] = k < cvn_new[c] ? vinv[cv_old[(size_t)k * nc + cperm[c]]] : -1;
}

inline dim3 grid(long long n, int bs = 256) { return dim3((unsigned)((n + bs - 1) / bs)); }

// gather one per-element array of `bytes` bytes per element through tmp, in place
int gather_inplace(void *a, size_t bytes, const int *perm, int n, void *tmp, hipStream_t st) {
    if (!a || n <= 0) return CGPU_OK;
    switch (bytes) {
        case 4:
            hipLaunchKernelGGL(k_gather<1>, grid(n), dim3(256), 0, st, (const Words<1> *)a, (Words<1> *)tmp, perm, n);
            break;
        case 8:
            hipLaunchKernelGGL(k_gather<2>, grid(n), dim3(256), 0, st, (const Words<2> *)a, (Words<2> *)tmp, perm, n);
            break;
        case 16:
            hipLaunchKernelGGL(k_gather<4>, grid(n), dim3(256), 0, st, (const Words<4> *)a, (Words<4> *)tmp, perm, n);
            break;
        default:
            return CGPU_ERR_ARGS;
    }
    if (hipMemcpyAsync(a, tmp, bytes * (size_t)n, hipMemcpyDeviceToDevice, st) != hipSuccess) return CGPU_ERR_HIP;
    return check();
}

size_t radix_temp_u64(int n) {
    size_t bytes = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const uint64_t *)nullptr, (uint64_t *)nullptr,
                                             (const int *)nullptr, (int *)nullptr, n);
    return bytes;
}
size_t radix_temp_u32(int n) {
    size_t bytes = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const unsigned int *)nullptr, (unsigned int *)nullptr,
                                             (const int *)nullptr, (int *)nullptr, n);
    return bytes;
}

// workspace: keys in / out (8 n each), indices in (4 n), then the sort's temp
struct SortWork {
    uint64_t *kin, *kout;
    int *iin;
    void *temp;
    size_t temp_bytes;
};
size_t sort_work_bytes(int n, size_t temp) { return 3 * align256(8 * (size_t)n) + temp + 256; }
SortWork carve(void *work, int n, size_t temp) {
    char *p = (char *)(((uintptr_t)work + 255) & ~(uintptr_t)255);
    SortWork w;
    w.kin = (uint64_t *)p;
    p += align256(8 * (size_t)n);
    w.kout = (uint64_t *)p;
    p += align256(8 * (size_t)n);
    w.iin = (int *)p;
    p += align256(8 * (size_t)n);
    w.temp = p;
    w.temp_bytes = temp;
    return w;
}

}  // namespace

extern "C" {

size_t cgpu_vor_sort_workspace(int num_cells) {
    if (num_cells <= 0) return 0;
    return sort_work_bytes(num_cells, radix_temp_u64(num_cells));
}

int cgpu_vor_hilbert_order(const cgpu_voronoi_batch *b, int *perm, void *work, size_t work_bytes, void *stream) {
    if (!b || !perm || !work) return CGPU_ERR_ARGS;
    const int n = b->num_cells;
    if (n <= 0) return CGPU_OK;
    const size_t temp = radix_temp_u64(n);
    if (work_bytes < sort_work_bytes(n, temp)) return CGPU_ERR_ARGS;
    const SortWork w = carve(work, n, temp);
    hipStream_t st = S(stream);
    hipLaunchKernelGGL(k_cell_keys, grid(n), dim3(256), 0, st, *b, (const double *)b->pos, w.kin, w.iin);
    size_t tb = w.temp_bytes;
    if (hipcub::DeviceRadixSort::SortPairs(w.temp, tb, w.kin, w.kout, w.iin, perm, n, 0, 64, st) != hipSuccess)
        return CGPU_ERR_HIP;
    return check();
}

int cgpu_vor_permute_cells(const cgpu_voronoi_batch *b, const int *perm, void *tmp, void *stream) {
    if (!b || !perm || !tmp) return CGPU_ERR_ARGS;
    const int n = b->num_cells;
    const size_t r = b->precision ? 4 : 8;
    hipStream_t st = S(stream);
    int rc = gather_inplace(b->pos, 16, perm, n, tmp, st);
    if (!rc) rc = gather_inplace(b->theta, r, perm, n, tmp, st);
    if (!rc) rc = gather_inplace(b->vel, 2 * r, perm, n, tmp, st);
    if (!rc) rc = gather_inplace((void *)b->APpref, 2 * r, perm, n, tmp, st);
    if (!rc) rc = gather_inplace((void *)b->mot, 2 * r, perm, n, tmp, st);
    if (!rc) rc = gather_inplace(b->force, 2 * r, perm, n, tmp, st);
    if (!rc) rc = gather_inplace(b->ext_force, 2 * r, perm, n, tmp, st);
    if (!rc) rc = gather_inplace((void *)b->excl, 4, perm, n, tmp, st);
    if (!rc) rc = gather_inplace((void *)b->cell_type, 4, perm, n, tmp, st);
    return rc;
}

size_t cgpu_vm_sort_workspace(int num_vertices, int num_cells) {
    if (num_vertices <= 0 || num_cells <= 0) return 0;
    const int n = num_vertices > num_cells ? num_vertices : num_cells;
    size_t t = radix_temp_u64(n), t32 = radix_temp_u32(n);
    return sort_work_bytes(n, t > t32 ? t : t32);
}

int cgpu_vm_hilbert_order(const cgpu_vertex_batch *b, int *vperm, int *cperm, void *work, size_t work_bytes,
                          void *stream) {
    if (!b || !vperm || !cperm || !work) return CGPU_ERR_ARGS;
    const int nv = b->num_vertices, nc = b->num_cells;
    if (nv <= 0 || nc <= 0) return CGPU_OK;
    if (3LL * nv >= (long long)UINT_MAX) return CGPU_ERR_ARGS;
    const int n = nv > nc ? nv : nc;
    size_t temp = radix_temp_u64(n), t32 = radix_temp_u32(n);
    temp = temp > t32 ? temp : t32;
    if (work_bytes < sort_work_bytes(n, temp)) return CGPU_ERR_ARGS;
    const SortWork w = carve(work, n, temp);
    hipStream_t st = S(stream);
    // vertices along the curve
    hipLaunchKernelGGL(k_vertex_keys, grid(nv), dim3(256), 0, st, *b, w.kin, w.iin);
    size_t tb = w.temp_bytes;
    if (hipcub::DeviceRadixSort::SortPairs(w.temp, tb, w.kin, w.kout, w.iin, vperm, nv, 0, 64, st) != hipSuccess)
        return CGPU_ERR_HIP;
    // cells by first appearance; a cell's smallest 3 vv + f lies in its own
    // system's vertex range, so one sort over the batch keeps the systems apart
    unsigned int *first = (unsigned int *)w.kin, *fout = (unsigned int *)w.kout;
    hipLaunchKernelGGL(k_fill, grid(nc), dim3(256), 0, st, first, nc, UINT_MAX);
    hipLaunchKernelGGL(k_cell_first_seen, grid(nv), dim3(256), 0, st, *b, (const int *)vperm, first);
    hipLaunchKernelGGL(k_iota, grid(nc), dim3(256), 0, st, w.iin, nc);
    tb = w.temp_bytes;
    if (hipcub::DeviceRadixSort::SortPairs(w.temp, tb, (const unsigned int *)first, fout, (const int *)w.iin, cperm,
                                           nc, 0, 32, st) != hipSuccess)
        return CGPU_ERR_HIP;
    return check();
}

size_t cgpu_vm_permute_workspace(int num_vertices, int num_cells) {
    if (num_vertices <= 0 || num_cells <= 0) return 0;
    size_t tmp = 16 * (size_t)num_vertices;
    const size_t lists = 4 * (size_t)CGPU_VERTEX_CAP * num_cells;
    if (lists > tmp) tmp = lists;
    return align256(4 * (size_t)num_vertices) + align256(4 * (size_t)num_cells) + tmp + 256;
}

int cgpu_vm_permute(const cgpu_vertex_batch *b, const int *vperm, const int *cperm, void *work, size_t work_bytes,
                    void *stream) {
    if (!b || !vperm || !cperm || !work) return CGPU_ERR_ARGS;
    const int nv = b->num_vertices, nc = b->num_cells;
    if (nv <= 0 || nc <= 0) return CGPU_OK;
    if (work_bytes < cgpu_vm_permute_workspace(nv, nc)) return CGPU_ERR_ARGS;
    char *p = (char *)(((uintptr_t)work + 255) & ~(uintptr_t)255);
    int *vinv = (int *)p;
    p += align256(4 * (size_t)nv);
    int *cinv = (int *)p;
    p += align256(4 * (size_t)nc);
    void *tmp = p;
    const size_t r = b->precision ? 4 : 8;
    hipStream_t st = S(stream);
    hipLaunchKernelGGL(k_invert, grid(nv), dim3(256), 0, st, vperm, vinv, nv);
    hipLaunchKernelGGL(k_invert, grid(nc), dim3(256), 0, st, cperm, cinv, nc);
    // per-vertex: positions, forces, and the neighbour / cell index triples
    int rc = gather_inplace(b->pos, 16, vperm, nv, tmp, st);
    if (!rc) rc = gather_inplace(b->force, 2 * r, vperm, nv, tmp, st);
    if (!rc) {
        hipLaunchKernelGGL(k_gather_remap<3>, grid(nv), dim3(256), 0, st, (const int *)b->vn, (int *)tmp, vperm,
                           (const int *)vinv, nv);
        if (hipMemcpyAsync(b->vn, tmp, 12 * (size_t)nv, hipMemcpyDeviceToDevice, st) != hipSuccess) return CGPU_ERR_HIP;
        hipLaunchKernelGGL(k_gather_remap<3>, grid(nv), dim3(256), 0, st, (const int *)b->vcn, (int *)tmp, vperm,
                           (const int *)cinv, nv);
        if (hipMemcpyAsync(b->vcn, tmp, 12 * (size_t)nv, hipMemcpyDeviceToDevice, st) != hipSuccess) return CGPU_ERR_HIP;
        rc = check();
    }
    // per-cell: vertex counts first (the list gather reads the new counts)
    if (!rc) rc = gather_inplace(b->cvn, 4, cperm, nc, tmp, st);
    if (!rc) {
        hipLaunchKernelGGL(k_gather_cell_lists, grid((long long)nc * CGPU_VERTEX_CAP), dim3(256), 0, st,
                           (const int *)b->cv, (const int *)b->cvn, (int *)tmp, cperm, (const int *)vinv, nc);
        if (hipMemcpyAsync(b->cv, tmp, 4 * (size_t)CGPU_VERTEX_CAP * nc, hipMemcpyDeviceToDevice, st) != hipSuccess)
            return CGPU_ERR_HIP;
        rc = check();
    }
    if (!rc) rc = gather_inplace((void *)b->APpref, 2 * r, cperm, nc, tmp, st);
    if (!rc) rc = gather_inplace(b->AP, 2 * r, cperm, nc, tmp, st);
    if (!rc) rc = gather_inplace(b->dAP, 2 * r, cperm, nc, tmp, st);
    if (!rc) rc = gather_inplace(b->theta, r, cperm, nc, tmp, st);
    if (!rc) rc = gather_inplace((void *)b->motility, 2 * r, cperm, nc, tmp, st);
    return rc;
}

}  // extern "C"
