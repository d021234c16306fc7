// Trajectory databases in the reference's NetCDF layouts, written and read as
// NetCDF classic files (netcdf_classic.cpp). Host arrays in, host arrays out;
// the Python and C++ front ends gather them from a batch.
//   SPV: DatabaseNetCDFSPV.cpp:29-47 (layout), :65-133 (WriteState), :135-217 (ReadState)
//   AVM: DatabaseNetCDFAVM.cpp:30-52 (layout), :192-271 (WriteState), :78-190 (ReadState)
#include <sys/stat.h>

#include <cmath>
#include <memory>
#include <stdexcept>
#include <string>

#include "../../../include/cellgpu_amd.h"
#include "netcdf_classic.hpp"
#include "voronoi_batch.hpp"

namespace cgpu_host {

std::string &db_last_error();

namespace {

using Type = NcClassic::Type;
using VarSpec = std::pair<std::string, std::pair<Type, std::vector<int>>>;

struct Db {
    int kind;     // 0 SPV, 1 AVM
    int n;        // SPV: cells (the reference's "Nv"); AVM: vertices
    bool excl;    // SPV: externalForce variable present
    NcClassic nc;
    int var(const char *name) const {
        int v = nc.var_id(name);
        if (v < 0) throw std::runtime_error(std::string("netcdf: no variable ") + name);
        return v;
    }
};

bool exists(const char *path) {
    struct stat st;
    return stat(path, &st) == 0;
}

// mode: 0 replace, 1 write (append to an existing file), 2 read only, 3 new
// (fails when the file exists) -- the reference's NcFile::FileMode values
Db *open_db(const char *path, int kind, int n, int mode, int exclusions) {
    if (!path || n <= 0) throw std::invalid_argument("database: path and size required");
    std::unique_ptr<Db> db(new Db{kind, n, exclusions != 0, NcClassic()});
    if (mode == 1 || mode == 2) {
        db->nc = NcClassic::open(path, mode == 1);
        const int64_t want = db->nc.dim_len("Nv");
        if (want != n) throw std::invalid_argument("database: file holds a different number of degrees of freedom");
        if (kind == 0) db->excl = db->nc.vars().size() == 7;  // reference ReadState's test
        return db.release();
    }
    if (mode == 3 && exists(path)) throw std::invalid_argument("database: file exists");
    if (mode != 0 && mode != 3) throw std::invalid_argument("database: mode 0..3");
    if (kind == 0) {
        // dims: rec, Nv, dof, boxdim, unit
        std::vector<NcClassic::Dim> d = {{"rec", 0}, {"Nv", n}, {"dof", 2 * (int64_t)n}, {"boxdim", 4}, {"unit", 1}};
        std::vector<VarSpec> v = {
            {"time", {NcClassic::NC_DOUBLE, {0, 4}}},     {"means0", {NcClassic::NC_DOUBLE, {0, 4}}},
            {"pos", {NcClassic::NC_DOUBLE, {0, 2}}},      {"type", {NcClassic::NC_INT, {0, 1}}},
            {"director", {NcClassic::NC_DOUBLE, {0, 1}}}, {"BoxMatrix", {NcClassic::NC_DOUBLE, {0, 3}}}};
        if (db->excl) v.push_back({"externalForce", {NcClassic::NC_DOUBLE, {0, 2}}});
        db->nc = NcClassic::create(path, d, v);
    } else {
        if (n % 2) throw std::invalid_argument("database: a vertex model has an even number of vertices");
        const int64_t nc = n / 2;
        // dims: rec, Nv, Nc, Nc2, dof, Nvn, boxdim, unit
        std::vector<NcClassic::Dim> d = {{"rec", 0},          {"Nv", n},     {"Nc", nc},       {"Nc2", 2 * nc},
                                         {"dof", 2 * (int64_t)n}, {"Nvn", 3 * (int64_t)n}, {"boxdim", 4}, {"unit", 1}};
        std::vector<VarSpec> v = {{"pos", {NcClassic::NC_DOUBLE, {0, 4}}},
                                  {"force", {NcClassic::NC_DOUBLE, {0, 4}}},
                                  {"VertexCellNeighbors", {NcClassic::NC_INT, {0, 5}}},
                                  {"Vneighs", {NcClassic::NC_INT, {0, 5}}},
                                  {"cellType", {NcClassic::NC_DOUBLE, {0, 2}}},
                                  {"director", {NcClassic::NC_DOUBLE, {0, 2}}},
                                  {"cellPositions", {NcClassic::NC_DOUBLE, {0, 3}}},
                                  {"BoxMatrix", {NcClassic::NC_DOUBLE, {0, 6}}},
                                  {"meanQ", {NcClassic::NC_DOUBLE, {0, 7}}},
                                  {"time", {NcClassic::NC_DOUBLE, {0, 7}}}};
        db->nc = NcClassic::create(path, d, v);
    }
    return db.release();
}

std::vector<double> interleave(const double *x, const double *y, int n) {
    std::vector<double> v(2 * (size_t)n);
    for (int i = 0; i < n; ++i) {
        v[2 * i] = x[i];
        v[2 * i + 1] = y[i];
    }
    return v;
}
void split(const std::vector<double> &v, double *x, double *y, int n) {
    for (int i = 0; i < n; ++i) {
        if (x) x[i] = v[2 * i];
        if (y) y[i] = v[2 * i + 1];
    }
}

template <typename F> int db_guard(F &&f) {
    try {
        return f();
    } catch (const std::invalid_argument &e) {
        db_last_error() = e.what();
        set_last_error(e.what());
        return CGPU_ERR_ARGS;
    } catch (const std::exception &e) {
        db_last_error() = e.what();
        set_last_error(e.what());
        return CGPU_ERR_HIP;
    }
}

}  // namespace

std::string &db_last_error() {
    static std::string e;
    return e;
}

}  // namespace cgpu_host

using cgpu_host::Db;
using cgpu_host::interleave;
using cgpu_host::split;

extern "C" {

void *cgpu_db_open(const char *path, int kind, int n, int mode, int exclusions) {
    try {
        return cgpu_host::open_db(path, kind, n, mode, exclusions);
    } catch (const std::exception &e) {
        cgpu_host::db_last_error() = e.what();
        return nullptr;
    }
}
void cgpu_db_close(void *db) { delete (Db *)db; }
const char *cgpu_db_last_error(void) { return cgpu_host::db_last_error().c_str(); }
int cgpu_db_records(void *db) { return db ? (int)((Db *)db)->nc.num_records() : -1; }

int cgpu_db_spv_write(void *h, int rec, double time, double L, double meanq, const double *x, const double *y,
                      const double *director, const int *type, const double *ext_x, const double *ext_y) {
    return cgpu_host::db_guard([&] {
        Db &db = *(Db *)h;
        if (db.kind != 0) throw std::invalid_argument("not a Voronoi database");
        const int n = db.n;
        if (rec < 0) rec = (int)db.nc.num_records();
        const double box[4] = {L, 0.0, 0.0, L};  // square box: x11, x12, x21, x22
        std::vector<int> t(n, 0);
        if (type)
            for (int i = 0; i < n; ++i) t[i] = type[i];
        std::vector<double> d(n, 0.0);
        if (director)
            for (int i = 0; i < n; ++i) d[i] = director[i];
        db.nc.put_record(db.var("means0"), rec, &meanq);
        db.nc.put_record(db.var("time"), rec, &time);
        db.nc.put_record(db.var("pos"), rec, interleave(x, y, n).data());
        db.nc.put_record(db.var("type"), rec, t.data());
        db.nc.put_record(db.var("director"), rec, d.data());
        db.nc.put_record(db.var("BoxMatrix"), rec, box);
        if (db.excl) {
            std::vector<double> z(n, 0.0);
            db.nc.put_record(db.var("externalForce"), rec,
                             interleave(ext_x ? ext_x : z.data(), ext_y ? ext_y : z.data(), n).data());
        }
        db.nc.sync();
        return CGPU_OK;
    });
}

int cgpu_db_spv_read(void *h, int rec, double *time, double *box, double *x, double *y, double *director, int *type,
                     double *ext_x, double *ext_y) {
    return cgpu_host::db_guard([&] {
        const Db &db = *(Db *)h;
        if (db.kind != 0) throw std::invalid_argument("not a Voronoi database");
        const int n = db.n;
        if (time) db.nc.get_record(db.var("time"), rec, time);
        if (box) db.nc.get_record(db.var("BoxMatrix"), rec, box);
        std::vector<double> v(2 * (size_t)n);
        if (x || y) {
            db.nc.get_record(db.var("pos"), rec, v.data());
            split(v, x, y, n);
        }
        if (director) db.nc.get_record(db.var("director"), rec, director);
        if (type) db.nc.get_record(db.var("type"), rec, type);
        if ((ext_x || ext_y) && db.excl) {
            db.nc.get_record(db.var("externalForce"), rec, v.data());
            split(v, ext_x, ext_y, n);
        }
        return CGPU_OK;
    });
}

int cgpu_db_avm_write(void *h, int rec, double time, double L, double meanq, const double *vx, const double *vy,
                      const double *fx, const double *fy, const int *vn, const int *vcn, const int *cell_type,
                      const double *director, const double *cx, const double *cy) {
    return cgpu_host::db_guard([&] {
        Db &db = *(Db *)h;
        if (db.kind != 1) throw std::invalid_argument("not a vertex-model database");
        const int nv = db.n, nc = nv / 2;
        if (rec < 0) rec = (int)db.nc.num_records();
        const double box[4] = {L, 0.0, 0.0, L};
        std::vector<double> zero(std::max(nv, nc), 0.0), t(nc, 0.0), d(nc, 0.0);
        if (cell_type)
            for (int i = 0; i < nc; ++i) t[i] = cell_type[i];
        if (director)
            for (int i = 0; i < nc; ++i) d[i] = director[i];
        db.nc.put_record(db.var("time"), rec, &time);
        db.nc.put_record(db.var("meanQ"), rec, &meanq);
        db.nc.put_record(db.var("pos"), rec, interleave(vx, vy, nv).data());
        db.nc.put_record(db.var("force"), rec,
                         interleave(fx ? fx : zero.data(), fy ? fy : zero.data(), nv).data());
        db.nc.put_record(db.var("Vneighs"), rec, vn);
        db.nc.put_record(db.var("VertexCellNeighbors"), rec, vcn);
        db.nc.put_record(db.var("director"), rec, d.data());
        db.nc.put_record(db.var("BoxMatrix"), rec, box);
        db.nc.put_record(db.var("cellPositions"), rec,
                         interleave(cx ? cx : zero.data(), cy ? cy : zero.data(), nc).data());
        db.nc.put_record(db.var("cellType"), rec, t.data());
        db.nc.sync();
        return CGPU_OK;
    });
}

int cgpu_db_avm_read(void *h, int rec, double *time, double *box, double *vx, double *vy, double *fx, double *fy,
                     int *vn, int *vcn, double *director) {
    return cgpu_host::db_guard([&] {
        const Db &db = *(Db *)h;
        if (db.kind != 1) throw std::invalid_argument("not a vertex-model database");
        const int nv = db.n;
        if (time) db.nc.get_record(db.var("time"), rec, time);
        if (box) db.nc.get_record(db.var("BoxMatrix"), rec, box);
        std::vector<double> v(2 * (size_t)nv);
        if (vx || vy) {
            db.nc.get_record(db.var("pos"), rec, v.data());
            split(v, vx, vy, nv);
        }
        if (fx || fy) {
            db.nc.get_record(db.var("force"), rec, v.data());
            split(v, fx, fy, nv);
        }
        if (vn) db.nc.get_record(db.var("Vneighs"), rec, vn);
        if (vcn) db.nc.get_record(db.var("VertexCellNeighbors"), rec, vcn);
        if (director) db.nc.get_record(db.var("director"), rec, director);
        return CGPU_OK;
    });
}

}  // extern "C"

// ------------------------------------------------------------- batch glue
// WriteState / ReadState of one system of a Voronoi batch (reference
// SPVDatabaseNetCDF::WriteState :65-133, ReadState :135-217): cells in tag
// order, type - 5 for excluded cells, means0 = mean P / sqrt(A).

extern "C" {

int cgpu_db_spv_write_batch(void *h, void *vor, int s, double time, int rec) {
    return cgpu_host::db_guard([&] {
        Db &db = *(Db *)h;
        cgpu_host::VoronoiBatch &b = *(cgpu_host::VoronoiBatch *)vor;
        if (s < 0 || s >= b.num_systems()) throw std::invalid_argument("no such system");
        const cgpu_sys_params &p = b.params()[s];
        if (db.kind != 0 || db.n != p.n) throw std::invalid_argument("database size differs from the system's");
        const int N = b.num_cells();
        std::vector<double> x(N), y(N), th(N), A(N), P(N), ef;
        b.get(x.data(), y.data(), th.data(), nullptr, nullptr, nullptr, nullptr, A.data(), P.data());
        std::vector<int> ex;
        if (b.exclusions_device()) {
            ex.resize(N);
            cgpu_host::hip_check(hipMemcpy(ex.data(), b.exclusions_device(), N * sizeof(int), hipMemcpyDeviceToHost), "ex");
            if (b.ext_force_device()) {
                ef.resize(2 * (size_t)N);
                b.real_to_host(b.ext_force_device(), ef.data(), ef.size());
            }
        }
        const std::vector<int> &tags = b.tags();
        const std::vector<int> &types = b.cell_types();
        const int n = p.n;
        std::vector<double> tx(n), ty(n), td(n), tex(n, 0.0), tey(n, 0.0);
        std::vector<int> tt(n, 0);
        double q = 0.0;
        for (int i = p.off; i < p.off + n; ++i) {
            const int t = tags[i];
            tx[t] = x[i];
            ty[t] = y[i];
            td[t] = th[i];
            tt[t] = (types.empty() ? 0 : types[i]) - (!ex.empty() && ex[i] ? 5 : 0);
            if (!ef.empty()) {
                tex[t] = ef[2 * (size_t)i];
                tey[t] = ef[2 * (size_t)i + 1];
            }
            q += P[i] / std::sqrt(A[i]);
        }
        q /= n;
        if (time < 0) {
            long long c[8];
            b.get_counters(s, c);
            time = (double)c[3] * p.dt;
        }
        return cgpu_db_spv_write(h, rec, time, p.L, q, tx.data(), ty.data(), td.data(), tt.data(), tex.data(),
                                 tey.data());
    });
}

int cgpu_db_spv_read_batch(void *h, void *vor, int s, int rec) {
    return cgpu_host::db_guard([&] {
        Db &db = *(Db *)h;
        cgpu_host::VoronoiBatch &b = *(cgpu_host::VoronoiBatch *)vor;
        if (s < 0 || s >= b.num_systems()) throw std::invalid_argument("no such system");
        const cgpu_sys_params &p = b.params()[s];
        if (db.kind != 0 || db.n != p.n) throw std::invalid_argument("database size differs from the system's");
        const int n = p.n, N = b.num_cells();
        std::vector<double> tx(n), ty(n), td(n), box(4);
        std::vector<int> tt(n);
        int r = cgpu_db_spv_read(h, rec, nullptr, box.data(), tx.data(), ty.data(), td.data(), tt.data(), nullptr,
                                 nullptr);
        if (r) return r;
        if (std::fabs(box[0] - p.L) > 1e-12 * p.L || box[1] != 0.0 || box[2] != 0.0 || box[3] != box[0])
            throw std::invalid_argument("the record's box differs from the system's square box");
        std::vector<double> x(N), y(N), th(N);
        b.get(x.data(), y.data(), th.data(), nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
        std::vector<int> types(b.cell_types());
        bool any_type = false;
        for (int t = 0; t < n; ++t) any_type = any_type || tt[t] != 0;
        if (any_type && types.empty()) types.assign(N, 0);
        const std::vector<int> &tags = b.tags();
        for (int i = p.off; i < p.off + n; ++i) {
            const int t = tags[i];
            x[i] = tx[t];
            y[i] = ty[t];
            th[i] = td[t];
            if (!types.empty()) types[i] = tt[t] < 0 ? tt[t] + 5 : tt[t];
        }
        b.set_directors(th.data());
        if (!types.empty()) b.set_cell_types(types.data());
        b.set_positions(x.data(), y.data(), true);  // retriangulates and recomputes the geometry
        return b.status();
    });
}

}  // extern "C"
