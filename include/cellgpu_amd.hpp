// C++ drop-in layer over the cellgpu_amd C ABI, keeping the reference's
// plugin API (abaso/cellGPU): Simulation (inc/Simulation.h:18-87), the model
// hierarchy Simple2DModel (inc/models/Simple2DModel.h:26-67) -> Simple2DCell
// (Simple2DCell.h:22-126, 325-385) -> Simple2DActiveCell
// (Simple2DActiveCell.h:18-60) -> voronoiModelBase (voronoiModelBase.h:26-107)
// / vertexModelBase (vertexModelBase.h:26-109) -> VoronoiQuadraticEnergy /
// VertexQuadraticEnergy, and the updaters updater (inc/updaters/updater.h:17-103)
// -> simpleEquationOfMotion -> selfPropelledParticleDynamics /
// brownianParticleDynamics / selfPropelledCellVertexDynamics /
// EnergyMinimizerFIRE.
//
// The state lives on the GPU. The reference's GPUArray getters return a
// GPUArray here too; it mirrors device state: an ArrayHandle with
// access_location::host copies the current device values to the host when it
// is made (read, readwrite) and, for arrays with a setter, stores them back
// when it goes out of scope (readwrite, overwrite). Device access through
// ArrayHandle is refused: the device layout is structure-of-arrays over many
// systems; device code uses the kernel layer of cellgpu_amd.h. setCPU /
// setCPUOperation(true) throw: there is no CPU path. A model is a one-system
// batch (cgpu_vor_create runs many systems in one launch).
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <functional>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "cellgpu_amd.h"

namespace cellgpu {

using std::shared_ptr;
using std::vector;

typedef double Dscalar;
struct double2v {
    double x, y;
};
typedef double2v Dscalar2;
struct int2v {
    int x, y;
};

inline void check(int status, const char *what) {
    if (status != CGPU_OK)
        throw std::runtime_error(std::string(what) + " failed, status " + std::to_string(status) + ": " +
                                 cgpu_last_error());
}

// ------------------------------------------------------------ host mirrors
//! reference access_location / access_mode (gpuarray.h:41-75)
namespace access_location {
enum Enum { host, device };
}
namespace access_mode {
enum Enum { read, readwrite, overwrite };
}

template <class T> class ArrayHandle;

//! A host copy of one array of device state. bind() attaches a pull (device ->
//! host) and, for writable state, a push (host -> device); an unbound GPUArray
//! is a plain host array (e.g. the displacements a caller passes to
//! moveDegreesOfFreedom).
template <class T> class GPUArray {
   public:
    GPUArray() = default;
    explicit GPUArray(size_t n) : h_(n) {}
    size_t getNumElements() const { return size_ ? size_() : h_.size(); }
    void resize(size_t n) { h_.resize(n); }
    void swap(GPUArray &o) {
        h_.swap(o.h_);
        std::swap(pull_, o.pull_);
        std::swap(push_, o.push_);
    }
    //! size: the current number of elements of the device state
    void bind(std::function<size_t()> size, std::function<void(vector<T> &)> pull,
              std::function<void(const vector<T> &)> push = nullptr) {
        size_ = std::move(size);
        pull_ = std::move(pull);
        push_ = std::move(push);
    }
    bool writable() const { return !pull_ || (bool)push_; }

   private:
    friend class ArrayHandle<T>;
    vector<T> h_;
    std::function<size_t()> size_;
    std::function<void(vector<T> &)> pull_;
    std::function<void(const vector<T> &)> push_;
};

//! reference ArrayHandle (gpuarray.h:91-101, 180-190), host access only
template <class T> class ArrayHandle {
   public:
    ArrayHandle(GPUArray<T> &a, access_location::Enum loc = access_location::host,
                access_mode::Enum mode = access_mode::readwrite)
        : data(prepare(a, loc, mode)), a_(a), mode_(mode) {}
    ~ArrayHandle() {
        if (mode_ == access_mode::read || !a_.push_) return;
        try {
            a_.push_(a_.h_);
        } catch (const std::exception &e) {
            std::fprintf(stderr, "cellgpu: storing an ArrayHandle back to the device failed: %s\n", e.what());
        }
    }
    ArrayHandle(const ArrayHandle &) = delete;
    ArrayHandle &operator=(const ArrayHandle &) = delete;
    T *const data;

   private:
    static T *prepare(GPUArray<T> &a, access_location::Enum loc, access_mode::Enum mode) {
        if (loc != access_location::host)
            throw std::runtime_error("cellgpu: ArrayHandle gives host access only (device kernels use the C ABI)");
        if (mode != access_mode::read && !a.writable())
            throw std::runtime_error("cellgpu: this array mirrors device state that has no setter (read only)");
        if (mode != access_mode::overwrite && a.pull_)
            a.pull_(a.h_);
        else if (a.size_)
            a.h_.resize(a.size_());
        return a.h_.data();
    }
    GPUArray<T> &a_;
    access_mode::Enum mode_;
};

// ------------------------------------------------------------------ models
//! reference inc/models/Simple2DModel.h:26-67
class Simple2DModel {
   public:
    virtual ~Simple2DModel() = default;
    virtual void setGPU() = 0;
    virtual void setCPU() = 0;
    virtual int getNumberOfDegreesOfFreedom() = 0;
    virtual void computeForces() = 0;
    //! dynamical-matrix entries are not provided by this library
    virtual void getDynMatEntries(vector<int2v> &, vector<Dscalar> &, Dscalar = 1.0, Dscalar = 1.0) {
        throw std::runtime_error("cellgpu: dynamical matrix entries are not provided");
    }
    virtual void spatialSorting() {}
    virtual void enforceTopology() {}
    virtual void getForces(GPUArray<Dscalar2> &forces) = 0;
    virtual Dscalar getMaxForce() { return 0.; }
    virtual GPUArray<Dscalar2> &returnPositions() = 0;
    virtual GPUArray<Dscalar> &returnMasses() = 0;
    virtual GPUArray<Dscalar> &returnOtherData() = 0;
    virtual GPUArray<Dscalar2> &returnVelocities() = 0;
    virtual GPUArray<Dscalar2> &returnForces() = 0;
    virtual void moveDegreesOfFreedom(GPUArray<Dscalar2> &displacements, Dscalar scale = 1.) = 0;
    virtual Dscalar reportq() = 0;
    virtual void reportMeanCellForce(bool verbose) = 0;
    Dscalar currentTime = 0;
    virtual void setTime(Dscalar time) { currentTime = time; }
};

//! reference inc/models/Simple2DCell.h:22-126, 325-385: cells with area and
//! perimeter preferences; the concrete models bind the arrays to device state
class Simple2DCell : public Simple2DModel {
   public:
    void setGPU() override {}
    void setCPU() override { throw std::runtime_error("cellgpu_amd has no CPU path"); }
    void setCPU(bool cpu) {
        if (cpu) setCPU();
    }
    virtual void setv0Dr(Dscalar v0, Dscalar Dr) = 0;
    int getNumberOfDegreesOfFreedom() override { return Ncells; }
    virtual void computeGeometry() = 0;
    virtual Dscalar computeEnergy() = 0;

    //! A0, P0 of every cell (reference :64-67)
    void setCellPreferencesUniform(Dscalar A0, Dscalar P0) {
        vector<Dscalar2> ap(Ncells, Dscalar2{A0, P0});
        setCellPreferences(ap);
    }
    void setCellPreferences(vector<Dscalar2> &AreaPeriPreferences) {
        ArrayHandle<Dscalar2> h(this->AreaPeriPreferences, access_location::host, access_mode::overwrite);
        for (int i = 0; i < Ncells; ++i) h.data[i] = AreaPeriPreferences.at(i);
    }
    //! cell types (reference :91-93; line tension uses them)
    void setCellTypeUniform(int t) {
        vector<int> types(Ncells, t);
        setCellType(types);
    }
    virtual void setCellType(vector<int> &types) = 0;
    virtual void setModuliUniform(Dscalar KA, Dscalar KP) = 0;
    void setDeltaT(Dscalar dt) {
        deltaT = dt;
        pushParameters();
    }

    GPUArray<Dscalar2> &returnModuli() { return Moduli; }
    GPUArray<Dscalar2> &returnAreaPeri() { return AreaPeri; }
    GPUArray<Dscalar2> &returnAreaPeriPreferences() { return AreaPeriPreferences; }
    GPUArray<Dscalar2> &returnVelocities() override { return cellVelocities; }
    GPUArray<Dscalar2> &returnPositions() override { return cellPositions; }
    GPUArray<Dscalar2> &returnForces() override { return cellForces; }
    GPUArray<Dscalar> &returnMasses() override { return cellMasses; }
    GPUArray<Dscalar> &returnOtherData() override { return cellMasses; }

    //! reference getMaxForce (:325-340): the largest |component| over the forces
    Dscalar getMaxForce() override {
        ArrayHandle<Dscalar2> f(returnForces(), access_location::host, access_mode::read);
        Dscalar m = 0;
        const int n = (int)returnForces().getNumElements();
        for (int i = 0; i < n; ++i) m = std::max(m, std::max(std::fabs(f.data[i].x), std::fabs(f.data[i].y)));
        return m;
    }
    void reportMeanCellForce(bool verbose) override {
        const Dscalar2 m = meanCellForce(verbose);
        std::printf("mean force = (%g,%g)\n", m.x, m.y);
    }
    Dscalar2 meanCellForce(bool verbose = false) {
        ArrayHandle<Dscalar2> f(cellForces, access_location::host, access_mode::read);
        Dscalar2 m{0, 0};
        for (int i = 0; i < Ncells; ++i) {
            if (verbose) std::printf("cell %i force = (%g,%g)\n", i, f.data[i].x, f.data[i].y);
            m.x += f.data[i].x / Ncells;
            m.y += f.data[i].y / Ncells;
        }
        return m;
    }
    void reportAP(bool verbose = false) {
        ArrayHandle<Dscalar2> ap(AreaPeri, access_location::host, access_mode::read);
        Dscalar vtot = 0.0;
        for (int i = 0; i < Ncells; ++i) {
            if (verbose) std::printf("%i: (%f,%f)\n", i, ap.data[i].x, ap.data[i].y);
            vtot += ap.data[i].x;
        }
        std::printf("total area = %f\n", vtot);
    }
    //! mean P / sqrt(A) (reference Simple2DCell.cpp reportq)
    Dscalar reportq() override {
        ArrayHandle<Dscalar2> ap(AreaPeri, access_location::host, access_mode::read);
        Dscalar q = 0;
        for (int i = 0; i < Ncells; ++i) q += ap.data[i].y / std::sqrt(ap.data[i].x);
        return q / Ncells;
    }
    Dscalar reportVarq() {
        ArrayHandle<Dscalar2> ap(AreaPeri, access_location::host, access_mode::read);
        Dscalar m = 0, m2 = 0;
        for (int i = 0; i < Ncells; ++i) {
            const Dscalar q = ap.data[i].y / std::sqrt(ap.data[i].x);
            m += q / Ncells;
            m2 += q * q / Ncells;
        }
        return m2 - m * m;
    }
    Dscalar2 reportVarAP() {
        ArrayHandle<Dscalar2> ap(AreaPeri, access_location::host, access_mode::read);
        Dscalar mA = 0, mP = 0, vA = 0, vP = 0;
        for (int i = 0; i < Ncells; ++i) {
            mA += ap.data[i].x / Ncells;
            mP += ap.data[i].y / Ncells;
        }
        for (int i = 0; i < Ncells; ++i) {
            vA += (ap.data[i].x - mA) * (ap.data[i].x - mA) / Ncells;
            vP += (ap.data[i].y - mP) * (ap.data[i].y - mP) / Ncells;
        }
        return Dscalar2{vA, vP};
    }
    Dscalar reportMeanP() {
        ArrayHandle<Dscalar2> ap(AreaPeri, access_location::host, access_mode::read);
        Dscalar m = 0;
        for (int i = 0; i < Ncells; ++i) m += ap.data[i].y / Ncells;
        return m;
    }

    int Ncells = 0, Nvertices = 0;
    //! side of the square periodic box (sqrt of the initial cell count: unit density)
    Dscalar boxSide = 1;
    Dscalar KA = 1, KP = 1, deltaT = 0.01;
    GPUArray<Dscalar2> cellPositions, vertexPositions, cellVelocities, vertexVelocities, cellForces, vertexForces;
    GPUArray<Dscalar2> AreaPeri, AreaPeriPreferences, Moduli;
    GPUArray<Dscalar> cellMasses, vertexMasses;
    GPUArray<int> cellType;
    //! index maps of the last spatial sorting (reference :241-253): current
    //! index i held old index itt[i]; idxToTag / tagToIdx: label at creation
    vector<int> itt, tti, tagToIdx, idxToTag;

   protected:
    virtual void pushParameters() = 0;
    //! unit masses; count: the member holding the number of elements
    void bindMasses(const int &count, GPUArray<Dscalar> &m) {
        m.bind([&count] { return (size_t)count; }, [&count](vector<Dscalar> &h) { h.assign(count, 1.0); });
    }
};

//! reference inc/models/Simple2DActiveCell.h:18-60: self-propulsion
//! directors and (v0, Dr) per cell
class Simple2DActiveCell : public Simple2DCell {
   public:
    //! reference setCellMotility (Simple2DActiveCell.cpp:98-112)
    virtual void setCellMotility(vector<Dscalar> &v0s, vector<Dscalar> &drs) = 0;
    void setv0Dr(Dscalar v0, Dscalar Dr) override {
        vector<Dscalar> a(Ncells, v0), b(Ncells, Dr);
        setCellMotility(a, b);
    }
    //! uniform directors in [0, 2 pi) from std::mt19937 (reference
    //! setCellDirectorsRandomly, Simple2DActiveCell.cpp:55-70; reproducible
    //! seed 13377 as there)
    void setCellDirectorsRandomly(bool reproducible = true) {
        std::mt19937 gen(reproducible ? 13377u : std::random_device{}());
        std::uniform_real_distribution<double> u(0.0, 2.0 * 3.14159265358979323846);
        ArrayHandle<Dscalar> d(cellDirectors, access_location::host, access_mode::overwrite);
        for (int i = 0; i < Ncells; ++i) d.data[i] = u(gen);
    }
    GPUArray<Dscalar> cellDirectors;
    GPUArray<Dscalar2> Motility;
};

//! reference voronoiModelBase (voronoiModelBase.h:26-107, GPU branch): cells
//! are the Voronoi tessellation of their positions; topology = the Delaunay
//! triangulation, maintained on the device
class voronoiModelBase : public Simple2DActiveCell {
   public:
    voronoiModelBase(int n, bool reprod, int precision, unsigned seed) {
        Ncells = n;
        boxSide = std::sqrt((double)n);
        uint64_t s = reprod ? seed : (uint64_t)std::random_device{}();
        h_ = cgpu_vor_create(1, &n, nullptr, &s, nullptr, nullptr, nullptr, precision, 0);
        if (!h_) throw std::runtime_error(std::string("cgpu_vor_create: ") + cgpu_last_error());
        bindArrays();
        refreshTags();
    }
    ~voronoiModelBase() override { cgpu_vor_destroy(h_); }
    voronoiModelBase(const voronoiModelBase &) = delete;
    voronoiModelBase &operator=(const voronoiModelBase &) = delete;

    int getNumberOfDegreesOfFreedom() override { return Ncells; }
    //! x += scale * displacement, wrapped into the box (reference movePoints,
    //! voronoiModelBase.cpp:170-200); the topology is enforced separately
    void moveDegreesOfFreedom(GPUArray<Dscalar2> &displacements, Dscalar scale = 1.) override {
        vector<double> x(Ncells), y(Ncells);
        check(cgpu_vor_get(h_, x.data(), y.data(), 0, 0, 0, 0, 0, 0, 0), "positions");
        ArrayHandle<Dscalar2> d(displacements, access_location::host, access_mode::read);
        const double L = boxSide;
        for (int i = 0; i < Ncells; ++i) {
            x[i] += scale * d.data[i].x;
            y[i] += scale * d.data[i].y;
            cgpu_put_in_box(L, &x[i], &y[i]);
        }
        check(cgpu_vor_set_positions(h_, x.data(), y.data(), 0), "moveDegreesOfFreedom");
    }
    //! test and repair of the triangulation at the current positions
    //! (voronoiModelBase.cpp:686-739)
    void enforceTopology() override { check(cgpu_vor_enforce_topology(h_), "enforceTopology"); }
    void getForces(GPUArray<Dscalar2> &forces) override { copyOut(cellForces, forces); }
    void computeGeometry() override { check(cgpu_vor_set_params(h_, 0, KA, KP, deltaT, mu, T_), "geometry"); }
    //! reference setExclusions (voronoiModelBase.cpp:89-107)
    void setExclusions(vector<int> &exes) { check(cgpu_vor_set_exclusions(h_, exes.data()), "setExclusions"); }
    //! Hilbert relabelling of the cells, then retriangulation (reference
    //! spatialSorting, voronoiModelBase.cpp:669-680)
    void spatialSorting() override {
        vector<int> perm(Ncells);
        check(cgpu_vor_spatial_sort(h_, perm.data()), "spatialSorting");
        itt = perm;
        tti.assign(Ncells, 0);
        for (int i = 0; i < Ncells; ++i) tti[itt[i]] = i;
        refreshTags();
        reIndex(prefs_);
        reIndex(motility_);
    }
    //! reference reIndexCellArray (Simple2DCell.h): a host copy follows the sort
    template <typename T> void reIndex(vector<T> &a) const {
        if (a.size() != itt.size()) return;
        vector<T> b(a.size());
        for (size_t i = 0; i < a.size(); ++i) b[i] = a[itt[i]];
        a.swap(b);
    }
    void setCellPositions(vector<Dscalar2> newCellPositions) {
        vector<double> x(Ncells), y(Ncells);
        for (int i = 0; i < Ncells; ++i) {
            x[i] = newCellPositions.at(i).x;
            y[i] = newCellPositions.at(i).y;
        }
        check(cgpu_vor_set_positions(h_, x.data(), y.data(), 1), "setCellPositions");
    }
    void setCellMotility(vector<Dscalar> &v0s, vector<Dscalar> &drs) override {
        check(cgpu_vor_set_motility(h_, v0s.data(), drs.data()), "setCellMotility");
    }
    void setCellType(vector<int> &types) override {
        check(cgpu_vor_set_cell_types(h_, types.data()), "setCellType");
    }
    void setModuliUniform(Dscalar ka, Dscalar kp) override {
        KA = ka;
        KP = kp;
        pushParameters();
    }
    void setMobility(Dscalar m) {
        mu = m;
        pushParameters();
    }
    void setTemperature(Dscalar t) {
        T_ = t;
        pushParameters();
    }
    //! reference getCellNeighs (Simple2DCell.h:193-201)
    void getCellNeighs(int idx, int &nNeighs, vector<int> &neighs) {
        ArrayHandle<int> nn(cellNeighborNum, access_location::host, access_mode::read);
        ArrayHandle<int> n(cellNeighbors, access_location::host, access_mode::read);
        nNeighs = nn.data[idx];
        neighs.assign(n.data + (size_t)idx * neighMax(), n.data + (size_t)idx * neighMax() + nNeighs);
    }
    //! reference neighMax (even number above the largest degree)
    int neighMax() {
        long long c[8];
        check(cgpu_vor_get_counters(h_, 0, c), "counters");
        return (int)c[6];
    }
    //! NeedsFixing cells repaired locally so far (reference localTopologyUpdates)
    long long localTopologyUpdates() {
        long long c[8];
        check(cgpu_vor_get_counters(h_, 0, c), "counters");
        return c[1];
    }
    //! reference cellDeath / cellDivision (voronoiModelBase.cpp:1291-1422):
    //! parameters = {cell}, dParams = {theta, separation fraction}
    virtual void cellDeath(int cellIndex) {
        int kind = 0;
        check(cgpu_vor_cell_events(h_, 1, &kind, &cellIndex, nullptr, nullptr, nullptr), "cellDeath");
        Ncells = cgpu_vor_get_sizes(h_, nullptr);
        refreshTags();
    }
    virtual void cellDivision(const vector<int> &parameters, const vector<Dscalar> &dParams) {
        int kind = 1;
        check(cgpu_vor_cell_events(h_, 1, &kind, &parameters.at(0), &dParams.at(0), &dParams.at(1), nullptr),
              "cellDivision");
        Ncells = cgpu_vor_get_sizes(h_, nullptr);
        refreshTags();
    }
    void *handle() { return h_; }

    //! Delaunay rings: cellNeighbors[i * neighMax() + k] (counter-clockwise)
    GPUArray<int> cellNeighborNum, cellNeighbors, delOther;
    Dscalar mu = 1;

   protected:
    void pushParameters() override { check(cgpu_vor_set_params(h_, 0, KA, KP, deltaT, mu, T_), "set params"); }
    void *h_;
    Dscalar T_ = 0;

   private:
    void fetch2(int what, vector<Dscalar2> &out) {
        vector<double> a(Ncells), b(Ncells);
        int r;
        if (what == 0)
            r = cgpu_vor_get(h_, a.data(), b.data(), 0, 0, 0, 0, 0, 0, 0);
        else if (what == 1)
            r = cgpu_vor_get(h_, 0, 0, 0, a.data(), b.data(), 0, 0, 0, 0);
        else if (what == 2)
            r = cgpu_vor_get(h_, 0, 0, 0, 0, 0, a.data(), b.data(), 0, 0);
        else
            r = cgpu_vor_get(h_, 0, 0, 0, 0, 0, 0, 0, a.data(), b.data());
        check(r, "get");
        out.resize(Ncells);
        for (int i = 0; i < Ncells; ++i) out[i] = Dscalar2{a[i], b[i]};
    }
    void topo(vector<int> *nn, vector<int> *nbr, vector<int> *oth) {
        const int cap = CGPU_RING_CAP_MAX;
        vector<int> n(Ncells), r((size_t)Ncells * cap, -1);
        check(cgpu_vor_get_topology(h_, n.data(), r.data(), cap, nullptr), "topology");
        if (nn) *nn = n;
        const int nm = neighMax();
        auto repack = [&](const vector<int> &src, vector<int> *dst) {
            dst->assign((size_t)Ncells * nm, -1);
            for (int i = 0; i < Ncells; ++i)
                for (int k = 0; k < n[i] && k < nm; ++k) (*dst)[(size_t)i * nm + k] = src[(size_t)i * cap + k];
        };
        if (nbr) repack(r, nbr);
        if (oth) {
            vector<int> o((size_t)Ncells * cap, -1);
            check(cgpu_vor_get_del_other(h_, o.data(), cap), "delOther");
            repack(o, oth);
        }
    }
    void bindArrays() {
        cellPositions.bind(
            [this] { return (size_t)Ncells; }, [this](vector<Dscalar2> &h) { fetch2(0, h); },
                           [this](const vector<Dscalar2> &h) { setCellPositions(h); });
        cellVelocities.bind(
            [this] { return (size_t)Ncells; }, [this](vector<Dscalar2> &h) { fetch2(1, h); },
                            [this](const vector<Dscalar2> &h) {
                                vector<double> a(Ncells), b(Ncells);
                                for (int i = 0; i < Ncells; ++i) {
                                    a[i] = h[i].x;
                                    b[i] = h[i].y;
                                }
                                check(cgpu_vor_set_velocities(h_, a.data(), b.data()), "velocities");
                            });
        cellForces.bind(
            [this] { return (size_t)Ncells; }, [this](vector<Dscalar2> &h) { fetch2(2, h); });
        AreaPeri.bind(
            [this] { return (size_t)Ncells; }, [this](vector<Dscalar2> &h) { fetch2(3, h); });
        AreaPeriPreferences.bind(
            [this] { return (size_t)Ncells; }, [this](vector<Dscalar2> &h) { h = prefs_; },
                                 [this](const vector<Dscalar2> &h) {
                                     vector<double> a(Ncells), p(Ncells);
                                     for (int i = 0; i < Ncells; ++i) {
                                         a[i] = h[i].x;
                                         p[i] = h[i].y;
                                     }
                                     check(cgpu_vor_set_prefs(h_, a.data(), p.data()), "preferences");
                                     prefs_ = h;
                                 });
        Moduli.bind(
            [this] { return (size_t)Ncells; }, [this](vector<Dscalar2> &h) { h.assign(Ncells, Dscalar2{KA, KP}); });
        bindMasses(Ncells, cellMasses);
        cellDirectors.bind(
            [this] { return (size_t)Ncells; },
            [this](vector<Dscalar> &h) {
                h.resize(Ncells);
                check(cgpu_vor_get(h_, 0, 0, h.data(), 0, 0, 0, 0, 0, 0), "directors");
            },
            [this](const vector<Dscalar> &h) { check(cgpu_vor_set_directors(h_, h.data()), "directors"); });
        Motility.bind(
            [this] { return (size_t)Ncells; }, [this](vector<Dscalar2> &h) { h = motility_; });
        cellNeighborNum.bind(
            [this] { return (size_t)Ncells; }, [this](vector<int> &h) { topo(&h, nullptr, nullptr); });
        cellNeighbors.bind(
            [this] { return (size_t)Ncells * neighMax(); }, [this](vector<int> &h) { topo(nullptr, &h, nullptr); });
        delOther.bind(
            [this] { return (size_t)Ncells * neighMax(); }, [this](vector<int> &h) { topo(nullptr, nullptr, &h); });
    }
    void refreshTags() {
        idxToTag.resize(Ncells);
        check(cgpu_vor_get_tags(h_, idxToTag.data()), "tags");
        tagToIdx.assign(Ncells, 0);
        for (int i = 0; i < Ncells; ++i) tagToIdx[idxToTag[i]] = i;
    }
    static void copyOut(GPUArray<Dscalar2> &src, GPUArray<Dscalar2> &dst) {
        ArrayHandle<Dscalar2> s(src, access_location::host, access_mode::read);
        dst.resize(src.getNumElements());
        ArrayHandle<Dscalar2> d(dst, access_location::host, access_mode::overwrite);
        std::copy(s.data, s.data + src.getNumElements(), d.data);
    }

   protected:
    vector<Dscalar2> prefs_, motility_;
};

//! reference VoronoiQuadraticEnergy (voronoiQuadraticEnergy.h:19-64, GPU branch)
class VoronoiQuadraticEnergy : public voronoiModelBase {
   public:
    VoronoiQuadraticEnergy(int n, Dscalar A0, Dscalar P0, bool reprod, int precision = 0, unsigned seed = 13377)
        : voronoiModelBase(n, reprod, precision, seed) {
        setCellPreferencesUniform(A0, P0);
        setv0Dr(0.05, 1.0);
    }
    void setCellMotility(vector<Dscalar> &v0s, vector<Dscalar> &drs) override {
        voronoiModelBase::setCellMotility(v0s, drs);
        motility_.resize(Ncells);
        for (int i = 0; i < Ncells; ++i) motility_[i] = Dscalar2{v0s[i], drs[i]};
    }
    void computeForces() override { check(cgpu_vor_compute_forces(h_), "computeForces"); }
    //! reference computeEnergy (voronoiQuadraticEnergy.cpp:384-397)
    Dscalar computeEnergy() override { return cgpu_vor_energy(h_, 0); }
    void setCellPreferencesUniform(Dscalar A0, Dscalar P0) {
        Simple2DCell::setCellPreferencesUniform(A0, P0);
    }
    //! the mean force, returned (reportMeanCellForce prints it)
    Dscalar2 reportMeanCellForceValue() { return meanCellForce(false); }
};

//! reference voronoiQuadraticEnergyWithTension.h:16-60 (GPU branch): line
//! tension between cells of different types
class VoronoiQuadraticEnergyWithTension : public VoronoiQuadraticEnergy {
   public:
    VoronoiQuadraticEnergyWithTension(int n, Dscalar A0, Dscalar P0, bool reprod = false)
        : VoronoiQuadraticEnergy(n, A0, P0, reprod) {}
    void setUseSurfaceTension(bool use_tension) {
        Tension = use_tension;
        push_tension();
    }
    //! one tension value for every pair of unlike cells
    void setSurfaceTension(Dscalar g) {
        gamma = g;
        simpleTension = true;
        push_tension();
    }
    //! gammas[a*n + b]: tension between types a and b (n*n values)
    void setSurfaceTension(const vector<Dscalar> &gammas) {
        int n = (int)std::lround(std::sqrt((double)gammas.size()));
        if (n * n != (int)gammas.size()) throw std::invalid_argument("the tension matrix has n*n entries");
        check(cgpu_vor_set_tension_matrix(h_, n, gammas.data()), "setSurfaceTension");
        simpleTension = false;
        push_tension();
    }
    Dscalar getSurfaceTension() { return gamma; }

   protected:
    Dscalar gamma = 0.0;
    bool Tension = false, simpleTension = true;

   private:
    void push_tension() {
        check(cgpu_vor_set_tension(h_, -1, Tension ? (simpleTension ? 1 : 2) : 0, gamma), "set tension");
    }
};

//! reference vertexModelBase (vertexModelBase.h:26-109, GPU branch): the
//! vertices are the degrees of freedom; topology changes are T1 flips
class vertexModelBase : public Simple2DActiveCell {
   public:
    //! runSPV: relax the random points with 100 self-propelled Voronoi steps
    //! first (reference setCellsVoronoiTesselation(spvInitialize))
    vertexModelBase(int n, bool reprod, bool runSPV, int precision, unsigned seed) {
        Ncells = n;
        boxSide = std::sqrt((double)n);
        uint64_t s = reprod ? seed : (uint64_t)std::random_device{}();
        h_ = cgpu_vm_create(1, &n, nullptr, &s, nullptr, nullptr, precision, 0, runSPV ? 1 : 0);
        if (!h_) throw std::runtime_error(std::string("cgpu_vm_create: ") + cgpu_last_error());
        Nvertices = cgpu_vm_num_vertices(h_);
        deltaT = 0.005;
        bindArrays();
    }
    ~vertexModelBase() override { cgpu_vm_destroy(h_); }
    vertexModelBase(const vertexModelBase &) = delete;
    vertexModelBase &operator=(const vertexModelBase &) = delete;

    int getNumberOfDegreesOfFreedom() override { return Nvertices; }
    GPUArray<Dscalar2> &returnForces() override { return vertexForces; }
    GPUArray<Dscalar2> &returnVelocities() override { return vertexVelocities; }
    GPUArray<Dscalar2> &returnPositions() override { return vertexPositions; }
    GPUArray<Dscalar> &returnMasses() override { return vertexMasses; }
    void getForces(GPUArray<Dscalar2> &forces) override {
        ArrayHandle<Dscalar2> s(vertexForces, access_location::host, access_mode::read);
        forces.resize(Nvertices);
        ArrayHandle<Dscalar2> d(forces, access_location::host, access_mode::overwrite);
        std::copy(s.data, s.data + Nvertices, d.data);
    }
    //! vertex x += scale * displacement, wrapped (reference
    //! vertexModelBase::moveDegreesOfFreedom, vertexModelBase.cpp:11-47)
    void moveDegreesOfFreedom(GPUArray<Dscalar2> &displacements, Dscalar scale = 1.) override {
        vector<double> x(Nvertices), y(Nvertices);
        check(cgpu_vm_get(h_, x.data(), y.data(), 0, 0, 0, 0), "vertex positions");
        ArrayHandle<Dscalar2> d(displacements, access_location::host, access_mode::read);
        const double L = boxSide;
        for (int i = 0; i < Nvertices; ++i) {
            x[i] += scale * d.data[i].x;
            y[i] += scale * d.data[i].y;
            cgpu_put_in_box(L, &x[i], &y[i]);
        }
        check(cgpu_vm_set_positions(h_, x.data(), y.data()), "moveDegreesOfFreedom");
    }
    //! T1 test and flips at the current positions (vertexModelBase.cpp:93-106)
    void enforceTopology() override { check(cgpu_vm_enforce_topology(h_), "enforceTopology"); }
    void computeGeometry() override { check(cgpu_vm_compute_forces(h_), "geometry"); }
    void setT1Threshold(Dscalar t) {
        T1Threshold = t;
        pushParameters();
    }
    void setTemperature(Dscalar t) {
        T = t;
        pushParameters();
    }
    void setMu(Dscalar m) {
        mu = m;
        pushParameters();
    }
    void setModuliUniform(Dscalar ka, Dscalar kp) override {
        KA = ka;
        KP = kp;
        pushParameters();
    }
    void setCellType(vector<int> &) override {
        throw std::runtime_error("cellgpu: the vertex model here has no type-dependent forces");
    }
    void setCellMotility(vector<Dscalar> &v0s, vector<Dscalar> &drs) override {
        check(cgpu_vm_set_motility(h_, v0s.data(), drs.data()), "setCellMotility");
        motility_.resize(Ncells);
        for (int i = 0; i < Ncells; ++i) motility_[i] = Dscalar2{v0s[i], drs[i]};
    }
    //! reference reportMeanVertexForce (Simple2DCell.h:358-371)
    void reportMeanVertexForce(bool verbose = false) {
        const Dscalar2 m = meanVertexForce();
        if (verbose) std::printf("%d vertices\n", Nvertices);
        std::printf("mean force = (%g,%g)\n", m.x, m.y);
    }
    Dscalar2 meanVertexForce() {
        ArrayHandle<Dscalar2> f(vertexForces, access_location::host, access_mode::read);
        Dscalar2 m{0, 0};
        for (int i = 0; i < Nvertices; ++i) {
            m.x += f.data[i].x / Nvertices;
            m.y += f.data[i].y / Nvertices;
        }
        return m;
    }
    void reportMeanCellForce(bool) override { reportMeanVertexForce(false); }
    //! reference vertexModelBase::spatialSorting (vertexModelBase.cpp:236-242):
    //! vertices along a Hilbert curve, cells by first appearance around them,
    //! re-indexed on the device; ittVertex / itt hold new -> old of this sort
    void spatialSorting() override {
        vector<int> vp(Nvertices), cp(Ncells);
        check(cgpu_vm_spatial_sort(h_, vp.data(), cp.data()), "spatialSorting");
        compose(vp, ittVertex, ttiVertex, idxToTagVertex, tagToIdxVertex);
        compose(cp, itt, tti, idxToTag, tagToIdx);
        for (vector<Dscalar2> *a : {&prefs_, &motility_}) {
            if (a->size() != cp.size()) continue;
            vector<Dscalar2> b(a->size());
            for (size_t i = 0; i < b.size(); ++i) b[i] = (*a)[cp[i]];
            a->swap(b);
        }
    }
    //! index maps of the last vertex sort and vertex labels at creation
    //! (reference Simple2DCell.h:253-262)
    vector<int> ittVertex, ttiVertex, tagToIdxVertex, idxToTagVertex;
    //! reference cellDeath (T2, vertexModelBase.cpp:905-1110) / cellDivision
    //! (:1124-1385): parameters = {cell, vertex-list position 1, position 2}
    virtual void cellDeath(int cellIndex) {
        int kind = 0, z = 0;
        check(cgpu_vm_cell_events(h_, 1, &kind, &cellIndex, &z, &z, nullptr), "cellDeath");
        Nvertices = cgpu_vm_get_sizes(h_, &Ncells, nullptr);
    }
    virtual void cellDivision(const vector<int> &parameters, const vector<Dscalar> &dParams = {}) {
        (void)dParams;
        int kind = 1;
        check(cgpu_vm_cell_events(h_, 1, &kind, &parameters.at(0), &parameters.at(1), &parameters.at(2), nullptr),
              "cellDivision");
        Nvertices = cgpu_vm_get_sizes(h_, &Ncells, nullptr);
    }
    void *handle() { return h_; }

   private:
    static void compose(const vector<int> &perm, vector<int> &it, vector<int> &ti, vector<int> &i2t,
                        vector<int> &t2i) {
        const int n = (int)perm.size();
        if ((int)i2t.size() != n) {
            i2t.resize(n);
            for (int i = 0; i < n; ++i) i2t[i] = i;
        }
        it = perm;
        ti.assign(n, 0);
        vector<int> old = i2t;
        t2i.assign(n, 0);
        for (int i = 0; i < n; ++i) {
            ti[perm[i]] = i;
            i2t[i] = old[perm[i]];
            t2i[i2t[i]] = i;
        }
    }

   public:
    //! vertexNeighbors[3 v + k], vertexCellNeighbors[3 v + k] (reference
    //! Simple2DCell.h:168-189); cellVertexNum[c], cellVertices[c * vertexMax + k]
    GPUArray<int> vertexNeighbors, vertexCellNeighbors, cellVertexNum, cellVertices;
    int vertexMax = CGPU_VERTEX_CAP;
    Dscalar mu = 1, T = 0, T1Threshold = 0.01;

   protected:
    void pushParameters() override {
        check(cgpu_vm_set_params(h_, 0, KA, KP, deltaT, mu, T, T1Threshold), "set params");
    }
    void *h_;
    vector<Dscalar2> prefs_, motility_;

   private:
    void fetchVertices(int what, vector<Dscalar2> &out) {
        vector<double> a(Nvertices), b(Nvertices);
        if (what == 0)
            check(cgpu_vm_get(h_, a.data(), b.data(), 0, 0, 0, 0), "get");
        else
            check(cgpu_vm_get(h_, 0, 0, a.data(), b.data(), 0, 0), "get");
        out.resize(Nvertices);
        for (int i = 0; i < Nvertices; ++i) out[i] = Dscalar2{a[i], b[i]};
    }
    void topo(vector<int> *vn, vector<int> *vcn, vector<int> *cvn, vector<int> *cv) {
        vertexMax = cgpu_vm_vertex_capacity(h_);  // the reference's vertexMax grows with the lists
        vector<int> a(3 * (size_t)Nvertices), b(3 * (size_t)Nvertices), c(Ncells), d((size_t)Ncells * vertexMax, -1);
        check(cgpu_vm_get_topology(h_, a.data(), b.data(), c.data(), d.data(), vertexMax), "topology");
        if (vn) *vn = a;
        if (vcn) *vcn = b;
        if (cvn) *cvn = c;
        if (cv) *cv = d;
    }
    void bindArrays() {
        vertexPositions.bind(
            [this] { return (size_t)Nvertices; }, [this](vector<Dscalar2> &h) { fetchVertices(0, h); },
                             [this](const vector<Dscalar2> &h) {
                                 vector<double> a(Nvertices), b(Nvertices);
                                 for (int i = 0; i < Nvertices; ++i) {
                                     a[i] = h[i].x;
                                     b[i] = h[i].y;
                                 }
                                 check(cgpu_vm_set_positions(h_, a.data(), b.data()), "vertex positions");
                             });
        vertexForces.bind(
            [this] { return (size_t)Nvertices; }, [this](vector<Dscalar2> &h) { fetchVertices(1, h); });
        AreaPeri.bind(
            [this] { return (size_t)Ncells; }, [this](vector<Dscalar2> &h) {
            vector<double> A(Ncells), P(Ncells);
            check(cgpu_vm_get(h_, 0, 0, 0, 0, A.data(), P.data()), "get");
            h.resize(Ncells);
            for (int c = 0; c < Ncells; ++c) h[c] = Dscalar2{A[c], P[c]};
        });
        AreaPeriPreferences.bind(
            [this] { return (size_t)Ncells; }, [this](vector<Dscalar2> &h) { h = prefs_; },
                                 [this](const vector<Dscalar2> &h) {
                                     vector<double> a(Ncells), p(Ncells);
                                     for (int i = 0; i < Ncells; ++i) {
                                         a[i] = h[i].x;
                                         p[i] = h[i].y;
                                     }
                                     check(cgpu_vm_set_prefs(h_, a.data(), p.data()), "preferences");
                                     prefs_ = h;
                                 });
        Moduli.bind(
            [this] { return (size_t)Ncells; }, [this](vector<Dscalar2> &h) { h.assign(Ncells, Dscalar2{KA, KP}); });
        bindMasses(Nvertices, vertexMasses);
        bindMasses(Ncells, cellMasses);
        cellDirectors.bind(
            [this] { return (size_t)Ncells; },
            [this](vector<Dscalar> &h) {
                h.resize(Ncells);
                check(cgpu_vm_get_directors(h_, h.data()), "directors");
            },
            [this](const vector<Dscalar> &h) { check(cgpu_vm_set_directors(h_, h.data()), "directors"); });
        Motility.bind(
            [this] { return (size_t)Ncells; }, [this](vector<Dscalar2> &h) { h = motility_; });
        vertexNeighbors.bind(
            [this] { return 3 * (size_t)Nvertices; }, [this](vector<int> &h) { topo(&h, nullptr, nullptr, nullptr); });
        vertexCellNeighbors.bind(
            [this] { return 3 * (size_t)Nvertices; }, [this](vector<int> &h) { topo(nullptr, &h, nullptr, nullptr); });
        cellVertexNum.bind(
            [this] { return (size_t)Ncells; }, [this](vector<int> &h) { topo(nullptr, nullptr, &h, nullptr); });
        cellVertices.bind(
            [this] { return (size_t)Ncells * (vertexMax = cgpu_vm_vertex_capacity(h_)); },
            [this](vector<int> &h) { topo(nullptr, nullptr, nullptr, &h); });
        // cell positions: mean of each cell's vertices (minimum image about the
        // first vertex), as the reference's getCellPositions
        cellPositions.bind(
            [this] { return (size_t)Ncells; }, [this](vector<Dscalar2> &h) {
            vector<Dscalar2> v;
            fetchVertices(0, v);
            vector<int> cvn, cv;
            topo(nullptr, nullptr, &cvn, &cv);
            const double L = boxSide;
            h.assign(Ncells, Dscalar2{0, 0});
            for (int c = 0; c < Ncells; ++c) {
                const Dscalar2 p0 = v[cv[(size_t)c * vertexMax]];
                double sx = 0, sy = 0;
                for (int k = 0; k < cvn[c]; ++k) {
                    const Dscalar2 p = v[cv[(size_t)c * vertexMax + k]];
                    double dx = p.x - p0.x, dy = p.y - p0.y;
                    dx -= L * std::round(dx / L);
                    dy -= L * std::round(dy / L);
                    sx += dx;
                    sy += dy;
                }
                double x = p0.x + sx / cvn[c], y = p0.y + sy / cvn[c];
                cgpu_put_in_box(L, &x, &y);
                h[c] = Dscalar2{x, y};
            }
        });
    }
};

//! reference VertexQuadraticEnergy (vertexQuadraticEnergy.h:16-36, GPU branch)
class VertexQuadraticEnergy : public vertexModelBase {
   public:
    VertexQuadraticEnergy(int n, Dscalar A0, Dscalar P0, bool reprod, bool runSPV = false, int precision = 0,
                          unsigned seed = 13377)
        : vertexModelBase(n, reprod, runSPV, precision, seed) {
        setCellPreferencesUniform(A0, P0);
    }
    void setCellPreferencesUniform(Dscalar A0, Dscalar P0) {
        A0_ = A0;
        P0_ = P0;
        Simple2DCell::setCellPreferencesUniform(A0, P0);
    }
    void computeForces() override { check(cgpu_vm_compute_forces(h_), "computeForces"); }
    //! sum over cells of KA (A - A0)^2 + KP (P - P0)^2 (FP64 on the host)
    Dscalar computeEnergy() override {
        computeForces();
        ArrayHandle<Dscalar2> ap(AreaPeri, access_location::host, access_mode::read);
        ArrayHandle<Dscalar2> pref(AreaPeriPreferences, access_location::host, access_mode::read);
        Dscalar E = 0;
        for (int c = 0; c < Ncells; ++c) {
            const Dscalar a = ap.data[c].x - pref.data[c].x, p = ap.data[c].y - pref.data[c].y;
            E += KA * a * a + KP * p * p;
        }
        return E;
    }
    Dscalar A0_ = 1.0, P0_ = 3.9;
};

// ---------------------------------------------------------------- updaters
//! reference inc/updaters/updater.h:17-103
class updater {
   public:
    updater() = default;
    explicit updater(int p) : Period(p) {}
    virtual ~updater() = default;
    virtual void Update(int timestep) {
        if (Period > 0 && (timestep + Phase) % Period == 0) performUpdate();
    }
    virtual void performUpdate() {}
    virtual void set2DModel(shared_ptr<Simple2DModel> m) { model = m; }
    void setPeriod(int p) { Period = p; }
    void setPhase(int p) { Phase = p; }
    virtual void spatialSorting() {}
    virtual void setReproducible(bool) {}
    virtual void setGPU() {}
    virtual void setCPU() { throw std::runtime_error("cellgpu_amd has no CPU path"); }
    int getNdof() { return Ndof; }
    void setNdof(int n) { Ndof = n; }
    virtual void setDeltaT(Dscalar dt) { deltaT = dt; }
    shared_ptr<Simple2DModel> model;

   protected:
    int Period = -1, Phase = 0, Ndof = 0;
    Dscalar deltaT = 0.01;
};

//! reference simpleEquationOfMotion.h:26-56 (updates every step)
class simpleEquationOfMotion : public updater {
   public:
    simpleEquationOfMotion() : updater(1) {}
    void performUpdate() override { integrateEquationsOfMotion(); }
    virtual void integrateEquationsOfMotion() = 0;
    void set2DModel(shared_ptr<Simple2DModel> m) override {
        updater::set2DModel(m);
        Ndof = m->getNumberOfDegreesOfFreedom();
    }
};

//! reference selfPropelledParticleDynamics.h:13-44 (GPU semantics: the fused
//! force, update and move of one step, then the topology)
class selfPropelledParticleDynamics : public simpleEquationOfMotion {
   public:
    explicit selfPropelledParticleDynamics(int n = 0) { Ndof = n; }
    void setMu(Dscalar m) { mu = m; }
    void integrateEquationsOfMotion() override {
        auto v = std::dynamic_pointer_cast<voronoiModelBase>(model);
        if (!v) throw std::runtime_error("selfPropelledParticleDynamics drives a Voronoi model");
        if (v->mu != mu) v->setMobility(mu);
        check(cgpu_vor_spv_steps(v->handle(), 1, 0), "self-propelled step");
    }
    Dscalar mu = 1.0;
};

//! reference brownianParticleDynamics.h:14-51
class brownianParticleDynamics : public simpleEquationOfMotion {
   public:
    explicit brownianParticleDynamics(int n = 0) { Ndof = n; }
    void setT(Dscalar t) { Temperature = t; }
    void setMu(Dscalar m) { mu = m; }
    void integrateEquationsOfMotion() override {
        if (auto vm = std::dynamic_pointer_cast<vertexModelBase>(model)) {
            if (vm->T != Temperature || vm->mu != mu) {
                vm->mu = mu;
                vm->setTemperature(Temperature);
            }
            check(cgpu_vm_brownian_steps(vm->handle(), 1, 0), "vertex Brownian step");
        } else if (auto v = std::dynamic_pointer_cast<voronoiModelBase>(model)) {
            v->mu = mu;
            v->setTemperature(Temperature);
            check(cgpu_vor_brownian_steps(v->handle(), 1), "Voronoi Brownian step");
        } else {
            throw std::runtime_error("brownianParticleDynamics drives a Voronoi or vertex model");
        }
    }
    Dscalar Temperature = 0.0, mu = 1.0;
};

//! reference selfPropelledCellVertexDynamics.h (GPU branch,
//! selfPropelledCellVertexDynamics.cpp:51-76)
class selfPropelledCellVertexDynamics : public simpleEquationOfMotion {
   public:
    selfPropelledCellVertexDynamics(int Ncells = 0, int Nvertices = 0) { Ndof = Nvertices ? Nvertices : Ncells; }
    void integrateEquationsOfMotion() override {
        auto vm = std::dynamic_pointer_cast<vertexModelBase>(model);
        if (!vm) throw std::runtime_error("selfPropelledCellVertexDynamics drives a vertex model");
        if (vm->mu != mu) vm->setMu(mu);
        check(cgpu_vm_spp_steps(vm->handle(), 1, 0), "self-propelled cell-vertex step");
    }
    Dscalar mu = 1.0;
};

//! reference EnergyMinimizerFIRE2D.h:20-132 (GPU branch), Voronoi and vertex models
class EnergyMinimizerFIRE : public simpleEquationOfMotion {
   public:
    explicit EnergyMinimizerFIRE(shared_ptr<Simple2DModel> system) {
        set2DModel(system);
        if (auto v = std::dynamic_pointer_cast<voronoiModelBase>(system))
            h_ = cgpu_fire_create(v->handle());
        else if (auto vm = std::dynamic_pointer_cast<vertexModelBase>(system))
            h_ = cgpu_vm_fire_create(vm->handle());
        else
            throw std::runtime_error("EnergyMinimizerFIRE drives a Voronoi or vertex model");
        if (!h_) throw std::runtime_error("FIRE create failed");
        push();
    }
    ~EnergyMinimizerFIRE() override { cgpu_fire_destroy(h_); }
    // one parameter at a time, as the reference setters (EnergyMinimizerFIRE2D.h:37-54)
    void setMaximumIterations(int m) {
        maxIt = m;
        one(CGPU_FIRE_MAX_IT, m);
    }
    void setForceCutoff(Dscalar f) {
        fc = f;
        one(CGPU_FIRE_FORCE_CUTOFF, f);
    }
    void setAlphaStart(Dscalar a) {
        a0 = a;
        one(CGPU_FIRE_ALPHA_START, a);
    }
    void setDeltaT(Dscalar d) override {
        dt = d;
        if (h_) one(CGPU_FIRE_DT, d);
    }
    void setDeltaTMax(Dscalar d) {
        dtMax = d;
        one(CGPU_FIRE_DT_MAX, d);
    }
    void setDeltaTInc(Dscalar d) {
        dtInc = d;
        one(CGPU_FIRE_DT_INC, d);
    }
    void setDeltaTDec(Dscalar d) {
        dtDec = d;
        one(CGPU_FIRE_DT_DEC, d);
    }
    void setAlphaDec(Dscalar a) {
        aDec = a;
        one(CGPU_FIRE_ALPHA_DEC, a);
    }
    void setNMin(int n) {
        nMin = n;
        one(CGPU_FIRE_NMIN, n);
    }
    void minimize() { check(cgpu_fire_minimize(h_), "minimize"); }
    void integrateEquationsOfMotion() override { minimize(); }
    //! forceMax = sum |F|^2 / N of the last iteration (the loop test is its
    //! square root, GPU branch: EnergyMinimizerFIRE2D.cpp:184, 277)
    Dscalar getMaxForce() {
        double out[6];
        check(cgpu_fire_get(h_, 0, out), "fire get");
        return out[0];
    }
    int getIterations() {
        double out[6];
        check(cgpu_fire_get(h_, 0, out), "fire get");
        return (int)out[5];
    }

   private:
    void push() { check(cgpu_fire_set(h_, maxIt, fc, a0, dt, dtMax, dtInc, dtDec, aDec, nMin), "fire set"); }
    void one(int which, double v) { check(cgpu_fire_set_param(h_, which, v), "fire set"); }
    void *h_ = nullptr;
    int maxIt = 1000, nMin = 5;
    Dscalar fc = 1e-7, a0 = 0.99, dt = 0.01, dtMax = 0.1, dtInc = 1.05, dtDec = 0.95, aDec = 0.9;
};

// --------------------------------------------------------------- Simulation
//! reference inc/Simulation.h:18-87, src/Simulation.cpp:100-166
class Simulation {
   public:
    void setConfiguration(shared_ptr<Simple2DModel> c) { cellConfiguration = c; }
    void addUpdater(shared_ptr<updater> u) { updaters.push_back(u); }
    void addUpdater(shared_ptr<updater> u, shared_ptr<Simple2DModel> c) {
        u->set2DModel(c);
        updaters.push_back(u);
    }
    void clearUpdaters() { updaters.clear(); }
    void setIntegrationTimestep(Dscalar dt) {
        integrationTimestep = dt;
        if (auto c = std::dynamic_pointer_cast<Simple2DCell>(cellConfiguration)) c->setDeltaT(dt);
        for (auto &u : updaters) u->setDeltaT(dt);
    }
    void setCPUOperation(bool setcpu) {
        if (setcpu) throw std::runtime_error("cellgpu_amd has no CPU path");
        USE_GPU = true;
    }
    void setReproducible(bool r) {
        for (auto &u : updaters) u->setReproducible(r);
    }
    //! spatial sorting every sp steps (sp <= 0: never)
    void setSortPeriod(int sp) { sortPeriod = sp; }
    virtual void setCurrentTime(Dscalar t) {
        Time = t;
        if (cellConfiguration) cellConfiguration->setTime(t);
    }
    virtual void setCurrentTimestep(int t) { integerTimestep = t; }
    //! the configuration's forces, copied into forces
    void computeForces(GPUArray<Dscalar2> &forces) {
        cellConfiguration->computeForces();
        cellConfiguration->getForces(forces);
    }
    void moveDegreesOfFreedom(GPUArray<Dscalar2> &displacements) {
        cellConfiguration->moveDegreesOfFreedom(displacements);
    }
    //! every updater in insertion order, then the spatial sorting every
    //! sortPeriod steps, then the configuration's time (Simulation.cpp:141-166)
    void performTimestep() {
        integerTimestep += 1;
        Time += integrationTimestep;
        for (auto &u : updaters) u->Update(integerTimestep);
        if (sortPeriod > 0 && integerTimestep % sortPeriod == 0) {
            cellConfiguration->spatialSorting();
            for (auto &u : updaters) u->spatialSorting();
        }
        cellConfiguration->setTime(Time);
    }
    int integerTimestep = 0;
    Dscalar Time = 0, integrationTimestep = 0.01;
    bool USE_GPU = true;
    shared_ptr<Simple2DModel> cellConfiguration;
    vector<shared_ptr<updater>> updaters;

   protected:
    int sortPeriod = -1;
};

// ---------------------------------------------------------------- databases
//! reference NcFile::FileMode values used by the databases
namespace NcFile {
enum FileMode { Replace = 0, Write = 1, ReadOnly = 2, New = 3 };
}

//! reference SPVDatabaseNetCDF (DatabaseNetCDFSPV.h): NetCDF classic file,
//! one record per WriteState
class SPVDatabaseNetCDF {
   public:
    SPVDatabaseNetCDF(int np, const std::string &fn, int mode = NcFile::ReadOnly, bool exclude = false)
        : filename(fn), Mode(mode) {
        h_ = cgpu_db_open(fn.c_str(), 0, np, mode, exclude ? 1 : 0);
        if (!h_) throw std::runtime_error(std::string("SPVDatabaseNetCDF: ") + cgpu_db_last_error());
    }
    ~SPVDatabaseNetCDF() { cgpu_db_close(h_); }
    SPVDatabaseNetCDF(const SPVDatabaseNetCDF &) = delete;
    SPVDatabaseNetCDF &operator=(const SPVDatabaseNetCDF &) = delete;
    void WriteState(shared_ptr<voronoiModelBase> s, Dscalar time = -1.0, int rec = -1) {
        check(cgpu_db_spv_write_batch(h_, s->handle(), 0, time, rec), "WriteState");
    }
    void ReadState(shared_ptr<voronoiModelBase> s, int rec, bool geometry = true) {
        (void)geometry;  // positions are always retriangulated
        check(cgpu_db_spv_read_batch(h_, s->handle(), 0, rec), "ReadState");
    }
    int Records() const { return cgpu_db_records(h_); }
    std::string filename;
    const int Mode;

   private:
    void *h_;
};

//! reference AVMDatabaseNetCDF (DatabaseNetCDFAVM.h)
class AVMDatabaseNetCDF {
   public:
    AVMDatabaseNetCDF(int np, const std::string &fn, int mode = NcFile::ReadOnly) : filename(fn), Mode(mode) {
        h_ = cgpu_db_open(fn.c_str(), 1, np, mode, 0);
        if (!h_) throw std::runtime_error(std::string("AVMDatabaseNetCDF: ") + cgpu_db_last_error());
    }
    ~AVMDatabaseNetCDF() { cgpu_db_close(h_); }
    AVMDatabaseNetCDF(const AVMDatabaseNetCDF &) = delete;
    AVMDatabaseNetCDF &operator=(const AVMDatabaseNetCDF &) = delete;
    void WriteState(shared_ptr<vertexModelBase> s, Dscalar time = -1.0, int rec = -1) {
        check(cgpu_db_avm_write_batch(h_, s->handle(), 0, time, rec), "WriteState");
    }
    void ReadState(shared_ptr<vertexModelBase> s, int rec, bool geometry = true) {
        (void)geometry;
        check(cgpu_db_avm_read_batch(h_, s->handle(), 0, rec), "ReadState");
    }
    int Records() const { return cgpu_db_records(h_); }
    std::string filename;
    const int Mode;

   private:
    void *h_;
};

//! reference smart-pointer typedefs (Simple2DCell.h:387-388,
//! simpleEquationOfMotion.h:58, updater.h:105, Simulation.h:88)
typedef shared_ptr<Simple2DCell> ForcePtr;
typedef shared_ptr<simpleEquationOfMotion> EOMPtr;
typedef shared_ptr<updater> UpdaterPtr;
typedef shared_ptr<Simulation> SimulationPtr;

}  // namespace cellgpu
