// Caller-owned batches of the kernel layer: the size of one device allocation
// that holds every array of a Voronoi batch, and the carve-up of such an
// allocation into a cgpu_voronoi_batch view. Nothing here allocates device
// memory; the kernel layer then runs on the view as on a handle's view.
#include <cstring>
#include <stdexcept>
#include <vector>

#include "voronoi_batch.hpp"

namespace cgpu_host {
namespace {

constexpr size_t ALIGN = 256;

// Assigns consecutive 256-byte aligned ranges of a base address (or of address
// 0, to measure the total).
struct Carver {
    char *base;
    size_t used = 0;
    template <typename T> T *take(size_t count) {
        used = (used + ALIGN - 1) / ALIGN * ALIGN;
        T *p = base ? (T *)(base + used) : nullptr;
        used += count * sizeof(T);
        return p;
    }
};

// The arrays of a batch in one fixed order; b receives the pointers.
void carve(Carver &c, const cgpu_vor_workspace &ws, cgpu_voronoi_batch &b) {
    const size_t S = ws.num_systems, N = ws.num_cells, rs = ws.precision ? sizeof(float) : sizeof(double);
    b.sys = c.take<cgpu_sys_params>(S);
    b.st = c.take<cgpu_sys_state>(S);
    b.cell_sys = c.take<int>(N);
    b.pos = c.take<double>(2 * N);
    b.pos_next = c.take<double>(2 * N);
    b.nn = c.take<int>(N);
    b.nbr = c.take<int>((size_t)CGPU_RING_CAP * N);
    b.oth = c.take<int>((size_t)CGPU_RING_CAP * N);
    b.AP = c.take<char>(2 * N * rs);
    b.dAP = c.take<char>(2 * N * rs);
    b.APpref = c.take<char>(2 * N * rs);
    b.mot = c.take<char>(2 * N * rs);
    b.vel = c.take<char>(2 * N * rs);
    b.theta = c.take<char>(N * rs);
    b.force = c.take<char>(2 * N * rs);
    b.bin_count = c.take<unsigned int>(ws.total_bins);
    b.bin_start = c.take<int>((size_t)ws.total_bins + 1);
    b.cell_bin = c.take<int>(N);
    b.cell_rank = c.take<int>(N);
    b.sorted_id = c.take<int>(N);
    b.sorted_pos = c.take<double>(2 * N);
    b.repair = c.take<int>(N);
    b.mark = c.take<int>(N);
    b.flagged = c.take<int>(N);
    b.nfq = c.take<int>(N);
    b.gq = c.take<int>(N);
    b.counters = c.take<int>(CGPU_COUNTERS);
    b.chunk_sys = c.take<int>(ws.num_chunks);
    b.chunk_begin = c.take<int>(ws.num_chunks);
    b.chunk_end = c.take<int>(ws.num_chunks);
    b.bchunk_sys = c.take<int>(ws.num_bin_chunks);
    b.bchunk_begin = c.take<int>(ws.num_bin_chunks);
    b.bchunk_end = c.take<int>(ws.num_bin_chunks);
    b.sys_bchunk = c.take<int>(S + 1);
    b.bchunk_sum = c.take<int>(ws.num_bin_chunks);
    b.ld_tri = c.take<int>(3 * (size_t)ld_tri_capacity((int)N));
    b.ld_tri_cap = ld_tri_capacity((int)N);
    b.bin_mark_wide = c.take<int>(ws.total_bins);
    b.cell_rmax = c.take<float>(N);
    b.ld_regions = c.take<char>((size_t)ld_region_capacity((int)N) * CGPU_LD_REGION_BYTES);
    b.ld_region_cap = ld_region_capacity((int)N);
    b.ld_part = c.take<unsigned long long>(3 * (N / 64 + 1));
}

template <typename F> int ws_guard(F &&f) {
    try {
        return f();
    } catch (const std::invalid_argument &e) {
        set_last_error(e.what());
        return CGPU_ERR_ARGS;
    } catch (const std::exception &e) {
        set_last_error(e.what());
        return CGPU_ERR_HIP;
    }
}

}  // namespace
}  // namespace cgpu_host

using namespace cgpu_host;

extern "C" {

int cgpu_vor_default_params(int num_systems, const int *n, const double *L, const uint64_t *seed,
                            cgpu_sys_params *out) {
    if (num_systems <= 0 || !n || !out) return CGPU_ERR_ARGS;
    return ws_guard([&] {
        std::vector<cgpu_sys_params> hp(num_systems);
        for (int s = 0; s < num_systems; ++s)
            hp[s] = default_sys_params(n[s], L ? L[s] : 0.0, seed ? seed[s] : 13377ull);
        batch_tables(hp);  // fills the offsets
        std::memcpy(out, hp.data(), sizeof(cgpu_sys_params) * num_systems);
        return CGPU_OK;
    });
}

int cgpu_vor_workspace_query(int num_systems, const cgpu_sys_params *params, int precision, cgpu_vor_workspace *ws) {
    if (num_systems <= 0 || !params || !ws) return CGPU_ERR_ARGS;
    return ws_guard([&] {
        std::vector<cgpu_sys_params> hp(params, params + num_systems);
        const BatchTables t = batch_tables(hp);
        std::memset(ws, 0, sizeof(*ws));
        ws->num_systems = num_systems;
        ws->num_cells = t.N;
        ws->total_bins = t.total_bins;
        ws->num_chunks = (int)t.cs.size();
        ws->num_bin_chunks = (int)t.bs.size();
        ws->precision = precision ? 1 : 0;
        Carver c{nullptr};
        cgpu_voronoi_batch b;
        carve(c, *ws, b);
        ws->bytes = (long long)c.used;
        return CGPU_OK;
    });
}

int cgpu_vor_workspace_bind(const cgpu_vor_workspace *ws, const cgpu_sys_params *params, void *mem, size_t mem_bytes,
                            cgpu_voronoi_batch *out, void *stream) {
    if (!ws || !params || !mem || !out) return CGPU_ERR_ARGS;
    if ((uintptr_t)mem % ALIGN) {
        set_last_error("the workspace must be 256-byte aligned");
        return CGPU_ERR_ARGS;
    }
    if (mem_bytes < (size_t)ws->bytes) {
        set_last_error("the workspace is smaller than cgpu_vor_workspace_query's bytes");
        return CGPU_ERR_SCRATCH_CAPACITY;
    }
    return ws_guard([&] {
        std::vector<cgpu_sys_params> hp(params, params + ws->num_systems);
        const BatchTables t = batch_tables(hp);
        if (t.N != ws->num_cells || t.total_bins != ws->total_bins || (int)t.cs.size() != ws->num_chunks ||
            (int)t.bs.size() != ws->num_bin_chunks)
            throw std::invalid_argument("the parameters do not describe the queried workspace");
        cgpu_voronoi_batch b;
        std::memset(&b, 0, sizeof(b));
        b.num_systems = ws->num_systems;
        b.num_cells = ws->num_cells;
        b.total_bins = ws->total_bins;
        b.precision = ws->precision;
        b.num_chunks = ws->num_chunks;
        b.num_bin_chunks = ws->num_bin_chunks;
        b.ring_cap = CGPU_RING_CAP;
        Carver c{(char *)mem};
        carve(c, *ws, b);
        // every array starts at zero; systems start active with unlimited steps
        std::vector<cgpu_sys_state> st(ws->num_systems);
        std::memset(st.data(), 0, sizeof(cgpu_sys_state) * st.size());
        for (auto &q : st) {
            q.active = 1;
            q.steps_left = -1;
        }
        hipStream_t sm = (hipStream_t)stream;
        hip_check(hipMemsetAsync(mem, 0, (size_t)ws->bytes, sm), "workspace clear");
        auto put = [&](const void *dst, const void *src, size_t bytes) {
            hip_check(hipMemcpyAsync((void *)dst, src, bytes, hipMemcpyHostToDevice, sm), "workspace upload");
        };
        put(b.sys, hp.data(), sizeof(cgpu_sys_params) * hp.size());
        put(b.st, st.data(), sizeof(cgpu_sys_state) * st.size());
        put(b.cell_sys, t.cell_sys.data(), sizeof(int) * t.N);
        put(b.chunk_sys, t.cs.data(), sizeof(int) * t.cs.size());
        put(b.chunk_begin, t.cb.data(), sizeof(int) * t.cb.size());
        put(b.chunk_end, t.ce.data(), sizeof(int) * t.ce.size());
        put(b.bchunk_sys, t.bs.data(), sizeof(int) * t.bs.size());
        put(b.bchunk_begin, t.bb.data(), sizeof(int) * t.bb.size());
        put(b.bchunk_end, t.be.data(), sizeof(int) * t.be.size());
        put(b.sys_bchunk, t.sb.data(), sizeof(int) * t.sb.size());
        // the host tables are freed on return
        hip_check(hipStreamSynchronize(sm), "workspace bind");
        *out = b;
        return CGPU_OK;
    });
}

}  // extern "C"
