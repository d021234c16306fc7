// C++ drop-in layer over the cellgpu_amd C ABI, keeping the reference's
// plugin API (abaso/cellGPU): Simulation (inc/Simulation.h:18-87), the model
// hierarchy Simple2DModel -> Simple2DCell -> Simple2DActiveCell ->
// voronoiModelBase / vertexModelBase -> VoronoiQuadraticEnergy /
// VertexQuadraticEnergy (inc/models/*.h), and the updaters updater ->
// simpleEquationOfMotion -> selfPropelledParticleDynamics /
// brownianParticleDynamics / EnergyMinimizerFIRE (inc/updaters/*.h).
//
// Differences a caller sees: state lives on the GPU and host accessors return
// copies (std::vector) instead of GPUArray handles; setCPU / setCPUOperation(true)
// throw, since there is no CPU path; a model is a one-system batch (see
// cgpu_vor_create for many systems in one launch).
#pragma once
#include <cmath>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "cellgpu_amd.h"

namespace cellgpu {

using std::shared_ptr;
using std::vector;

struct double2v { double x, y; };

inline void check(int status, const char *what) {
    if (status != CGPU_OK)
        throw std::runtime_error(std::string(what) + " failed, status " + std::to_string(status) + ": " +
                                 cgpu_last_error());
}

//! reference inc/models/Simple2DModel.h:26-67
class Simple2DModel {
   public:
    virtual ~Simple2DModel() = default;
    virtual int getNumberOfDegreesOfFreedom() = 0;
    virtual void computeForces() = 0;
    virtual double computeEnergy() = 0;
    virtual void setGPU() {}
    virtual void setCPU(bool = true) { throw std::runtime_error("cellgpu_amd has no CPU path"); }
    virtual void setDeltaT(double dt) = 0;
    virtual vector<double2v> returnForces() = 0;
    virtual vector<double2v> returnPositions() = 0;
};

//! reference voronoiModelBase + VoronoiQuadraticEnergy (GPU branch)
class VoronoiQuadraticEnergy : public Simple2DModel {
   public:
    VoronoiQuadraticEnergy(int n, double A0, double P0, bool reprod, int precision = 0, unsigned seed = 13377)
        : Ncells(n) {
        uint64_t s = reprod ? seed : (uint64_t)std::random_device{}();
        h_ = cgpu_vor_create(1, &n, nullptr, &s, nullptr, nullptr, nullptr, precision, 0);
        if (!h_) throw std::runtime_error(std::string("cgpu_vor_create: ") + cgpu_last_error());
        setCellPreferencesUniform(A0, P0);
        setv0Dr(0.05, 1.0);
    }
    ~VoronoiQuadraticEnergy() override { cgpu_vor_destroy(h_); }
    void setCellPreferencesUniform(double A0, double P0) {
        vector<double> a(Ncells, A0), p(Ncells, P0);
        check(cgpu_vor_set_prefs(h_, a.data(), p.data()), "setCellPreferencesUniform");
    }
    void setCellPreferences(const vector<double2v> &ap) {
        vector<double> a(Ncells), p(Ncells);
        for (int i = 0; i < Ncells; ++i) { a[i] = ap[i].x; p[i] = ap[i].y; }
        check(cgpu_vor_set_prefs(h_, a.data(), p.data()), "setCellPreferences");
    }
    void setv0Dr(double v0, double Dr) {
        vector<double> a(Ncells, v0), b(Ncells, Dr);
        check(cgpu_vor_set_motility(h_, a.data(), b.data()), "setv0Dr");
    }
    void setModuliUniform(double ka, double kp) {
        KA = ka;
        KP = kp;
        push();
    }
    void setExclusions(const vector<int> &ex) { check(cgpu_vor_set_exclusions(h_, ex.data()), "setExclusions"); }
    void setDeltaT(double dt) override {
        deltaT = dt;
        push();
    }
    void setMobility(double m) {
        mu = m;
        push();
    }
    int getNumberOfDegreesOfFreedom() override { return Ncells; }
    void computeForces() override { check(cgpu_vor_compute_forces(h_), "computeForces"); }
    double computeEnergy() override { return cgpu_vor_energy(h_, 0); }
    vector<double2v> returnForces() override { return fetch(2); }
    vector<double2v> returnPositions() override { return fetch(0); }
    vector<double2v> returnAreaPeri() { return fetch(3); }
    double reportq() {
        auto ap = returnAreaPeri();
        double q = 0;
        for (auto &v : ap) q += v.y / std::sqrt(v.x);
        return q / Ncells;
    }
    double2v reportMeanCellForce() {
        auto f = returnForces();
        double2v m{0, 0};
        for (auto &v : f) { m.x += v.x / Ncells; m.y += v.y / Ncells; }
        return m;
    }
    void *handle() { return h_; }
    int Ncells;
    double KA = 1, KP = 1, deltaT = 0.01, mu = 1;

   private:
    void push() { check(cgpu_vor_set_params(h_, 0, KA, KP, deltaT, mu, 0.0), "set params"); }
    vector<double2v> fetch(int what) {
        vector<double> a(Ncells), b(Ncells);
        int r;
        if (what == 0) r = cgpu_vor_get(h_, a.data(), b.data(), 0, 0, 0, 0, 0, 0, 0);
        else if (what == 2) r = cgpu_vor_get(h_, 0, 0, 0, 0, 0, a.data(), b.data(), 0, 0);
        else r = cgpu_vor_get(h_, 0, 0, 0, 0, 0, 0, 0, a.data(), b.data());
        check(r, "get");
        vector<double2v> out(Ncells);
        for (int i = 0; i < Ncells; ++i) out[i] = double2v{a[i], b[i]};
        return out;
    }

   protected:
    void *h_;

   public:
    //! reference Simple2DCell::setCellType
    void setCellType(const vector<int> &types) { check(cgpu_vor_set_cell_types(h_, types.data()), "setCellType"); }
    //! reference voronoiModelBase::cellDeath
    void cellDeath(int cellIndex) {
        int kind = 0;
        check(cgpu_vor_cell_events(h_, 1, &kind, &cellIndex, nullptr, nullptr, nullptr), "cellDeath");
        Ncells = cgpu_vor_get_sizes(h_, nullptr);
    }
    //! reference voronoiModelBase::cellDivision: parameters = {cell}, dParams = {theta, fraction}
    void cellDivision(const vector<int> &parameters, const vector<double> &dParams) {
        int kind = 1;
        check(cgpu_vor_cell_events(h_, 1, &kind, &parameters.at(0), &dParams.at(0), &dParams.at(1), nullptr),
              "cellDivision");
        Ncells = cgpu_vor_get_sizes(h_, nullptr);
    }
};

//! reference voronoiQuadraticEnergyWithTension.h:16-60 (GPU branch): line
//! tension between cells of different types
class VoronoiQuadraticEnergyWithTension : public VoronoiQuadraticEnergy {
   public:
    VoronoiQuadraticEnergyWithTension(int n, double A0, double P0, bool reprod = false)
        : VoronoiQuadraticEnergy(n, A0, P0, reprod) {}
    void setUseSurfaceTension(bool use_tension) {
        Tension = use_tension;
        push_tension();
    }
    //! one tension value for every pair of unlike cells
    void setSurfaceTension(double g) {
        gamma = g;
        simpleTension = true;
        push_tension();
    }
    //! gammas[a*n + b]: tension between types a and b (n*n values)
    void setSurfaceTension(const vector<double> &gammas) {
        int n = (int)std::lround(std::sqrt((double)gammas.size()));
        if (n * n != (int)gammas.size()) throw std::invalid_argument("the tension matrix has n*n entries");
        check(cgpu_vor_set_tension_matrix(h_, n, gammas.data()), "setSurfaceTension");
        simpleTension = false;
        push_tension();
    }
    double getSurfaceTension() { return gamma; }

   protected:
    double gamma = 0.0;
    bool Tension = false, simpleTension = true;

   private:
    void push_tension() {
        check(cgpu_vor_set_tension(h_, -1, Tension ? (simpleTension ? 1 : 2) : 0, gamma), "set tension");
    }
};

//! reference vertexModelBase + VertexQuadraticEnergy (GPU branch)
class VertexQuadraticEnergy : public Simple2DModel {
   public:
    //! runSPV: relax the random points with 100 self-propelled Voronoi steps first
    //! (reference vertexModelBase::setCellsVoronoiTesselation(spvInitialize))
    VertexQuadraticEnergy(int n, double A0, double P0, bool reprod, bool runSPV = false, int precision = 0,
                          unsigned seed = 13377)
        : Ncells(n) {
        uint64_t s = reprod ? seed : (uint64_t)std::random_device{}();
        h_ = cgpu_vm_create(1, &n, nullptr, &s, nullptr, nullptr, precision, 0, runSPV ? 1 : 0);
        if (!h_) throw std::runtime_error("cgpu_vm_create failed");
        Nvertices = cgpu_vm_num_vertices(h_);
        setCellPreferencesUniform(A0, P0);
    }
    ~VertexQuadraticEnergy() override { cgpu_vm_destroy(h_); }
    void setCellPreferencesUniform(double A0, double P0) {
        A0_ = A0;
        P0_ = P0;
        vector<double> a(Ncells, A0), p(Ncells, P0);
        check(cgpu_vm_set_prefs(h_, a.data(), p.data()), "setCellPreferencesUniform");
    }
    void setT1Threshold(double t) {
        T1Threshold = t;
        push();
    }
    void setDeltaT(double dt) override {
        deltaT = dt;
        push();
    }
    void setTemperature(double t) {
        T = t;
        push();
    }
    void setMu(double m) {
        mu = m;
        push();
    }
    //! reference Simple2DActiveCell::setv0Dr (Simple2DActiveCell.cpp:76-92)
    void setv0Dr(double v0, double Dr) {
        vector<double> a(Ncells, v0), b(Ncells, Dr);
        check(cgpu_vm_set_motility(h_, a.data(), b.data()), "setv0Dr");
    }
    vector<double> returnCellDirectors() {
        vector<double> t(Ncells);
        check(cgpu_vm_get_directors(h_, t.data()), "cellDirectors");
        return t;
    }
    //! reference vertexModelBase::cellDeath: T2 transition of a triangular cell
    void cellDeath(int cellIndex) {
        int kind = 0, z = 0;
        check(cgpu_vm_cell_events(h_, 1, &kind, &cellIndex, &z, &z, nullptr), "cellDeath");
        refresh_sizes();
    }
    //! reference vertexModelBase::cellDivision: parameters = {cell, list position 1, list position 2}
    void cellDivision(const vector<int> &parameters, const vector<double> &dParams = {}) {
        (void)dParams;
        int kind = 1;
        check(cgpu_vm_cell_events(h_, 1, &kind, &parameters.at(0), &parameters.at(1), &parameters.at(2), nullptr),
              "cellDivision");
        refresh_sizes();
    }
    //! reference vertexModelBase::reportMeanVertexForce: mean force on the vertices
    double2v reportMeanVertexForce() {
        vector<double> a(Nvertices), b(Nvertices);
        check(cgpu_vm_get(h_, 0, 0, a.data(), b.data(), 0, 0), "get");
        double2v m{0, 0};
        for (int i = 0; i < Nvertices; ++i) {
            m.x += a[i];
            m.y += b[i];
        }
        m.x /= Nvertices;
        m.y /= Nvertices;
        return m;
    }
    //! reference Simple2DCell::reportq: mean P / sqrt(A) over the cells
    double reportq() {
        vector<double> A(Ncells), P(Ncells);
        check(cgpu_vm_get(h_, 0, 0, 0, 0, A.data(), P.data()), "get");
        double q = 0;
        for (int c = 0; c < Ncells; ++c) q += P[c] / std::sqrt(A[c]);
        return q / Ncells;
    }
    int getNumberOfDegreesOfFreedom() override { return Nvertices; }
    void computeForces() override { check(cgpu_vm_compute_forces(h_), "computeForces"); }
    double computeEnergy() override {
        computeForces();
        vector<double> A(Ncells), P(Ncells);
        check(cgpu_vm_get(h_, 0, 0, 0, 0, A.data(), P.data()), "get");
        double E = 0;
        for (int c = 0; c < Ncells; ++c) E += KA * (A[c] - A0_) * (A[c] - A0_) + KP * (P[c] - P0_) * (P[c] - P0_);
        return E;
    }
    vector<double2v> returnForces() override {
        vector<double> a(Nvertices), b(Nvertices);
        check(cgpu_vm_get(h_, 0, 0, a.data(), b.data(), 0, 0), "get");
        vector<double2v> out(Nvertices);
        for (int i = 0; i < Nvertices; ++i) out[i] = double2v{a[i], b[i]};
        return out;
    }
    vector<double2v> returnPositions() override {
        vector<double> a(Nvertices), b(Nvertices);
        check(cgpu_vm_get(h_, a.data(), b.data(), 0, 0, 0, 0), "get");
        vector<double2v> out(Nvertices);
        for (int i = 0; i < Nvertices; ++i) out[i] = double2v{a[i], b[i]};
        return out;
    }
    void *handle() { return h_; }
    int Ncells, Nvertices;
    double KA = 1, KP = 1, deltaT = 0.005, mu = 1, T = 0, T1Threshold = 0.01, A0_ = 1.0, P0_ = 3.9;

   private:
    void push() { check(cgpu_vm_set_params(h_, 0, KA, KP, deltaT, mu, T, T1Threshold), "set params"); }
    void refresh_sizes() {
        Nvertices = cgpu_vm_get_sizes(h_, &Ncells, nullptr);
    }
    void *h_;
};

//! reference inc/updaters/updater.h:17-103
class updater {
   public:
    virtual ~updater() = default;
    virtual void Update(int timestep) {
        if (Period > 0 && (timestep + Phase) % Period == 0) performUpdate();
    }
    virtual void performUpdate() {}
    virtual void set2DModel(shared_ptr<Simple2DModel> m) { model = m; }
    virtual void setGPU() {}
    virtual void setCPU() { throw std::runtime_error("cellgpu_amd has no CPU path"); }
    virtual void setReproducible(bool) {}
    virtual void setDeltaT(double dt) { deltaT = dt; }
    void setPeriod(int p) { Period = p; }
    void setPhase(int p) { Phase = p; }
    shared_ptr<Simple2DModel> model;

   protected:
    int Period = 1, Phase = 0;
    double deltaT = 0.01;
};

//! reference simpleEquationOfMotion.h:26-56
class simpleEquationOfMotion : public updater {
   public:
    void performUpdate() override { integrateEquationsOfMotion(); }
    virtual void integrateEquationsOfMotion() = 0;
};

//! reference selfPropelledParticleDynamics.h:13-44 (GPU semantics)
class selfPropelledParticleDynamics : public simpleEquationOfMotion {
   public:
    explicit selfPropelledParticleDynamics(int n = 0) { (void)n; }
    void integrateEquationsOfMotion() override {
        auto v = std::dynamic_pointer_cast<VoronoiQuadraticEnergy>(model);
        if (!v) throw std::runtime_error("selfPropelledParticleDynamics drives a Voronoi model");
        check(cgpu_vor_spv_steps(v->handle(), 1, 0), "self-propelled step");
    }
};

//! reference brownianParticleDynamics.h:14-51
class brownianParticleDynamics : public simpleEquationOfMotion {
   public:
    explicit brownianParticleDynamics(int n = 0) { (void)n; }
    void setT(double t) { Temperature = t; }
    void integrateEquationsOfMotion() override {
        if (auto vm = std::dynamic_pointer_cast<VertexQuadraticEnergy>(model)) {
            vm->setTemperature(Temperature);
            check(cgpu_vm_brownian_steps(vm->handle(), 1, 0), "vertex Brownian step");
        } else if (auto v = std::dynamic_pointer_cast<VoronoiQuadraticEnergy>(model)) {
            check(cgpu_vor_set_params(v->handle(), 0, v->KA, v->KP, v->deltaT, v->mu, Temperature), "params");
            check(cgpu_vor_brownian_steps(v->handle(), 1), "Voronoi Brownian step");
        }
    }
    double Temperature = 0.0;
};

//! reference selfPropelledCellVertexDynamics.h (GPU branch,
//! selfPropelledCellVertexDynamics.cpp:51-76)
class selfPropelledCellVertexDynamics : public simpleEquationOfMotion {
   public:
    selfPropelledCellVertexDynamics(int Ncells = 0, int Nvertices = 0) {
        (void)Ncells;
        (void)Nvertices;
    }
    void integrateEquationsOfMotion() override {
        auto vm = std::dynamic_pointer_cast<VertexQuadraticEnergy>(model);
        if (!vm) throw std::runtime_error("selfPropelledCellVertexDynamics drives a vertex model");
        vm->setMu(mu);
        check(cgpu_vm_spp_steps(vm->handle(), 1, 0), "self-propelled cell-vertex step");
    }
    double mu = 1.0;
};

//! reference EnergyMinimizerFIRE2D.h:20-132 (GPU branch), Voronoi and vertex models
class EnergyMinimizerFIRE : public simpleEquationOfMotion {
   public:
    explicit EnergyMinimizerFIRE(shared_ptr<Simple2DModel> system) {
        set2DModel(system);
        if (auto v = std::dynamic_pointer_cast<VoronoiQuadraticEnergy>(system))
            h_ = cgpu_fire_create(v->handle());
        else if (auto vm = std::dynamic_pointer_cast<VertexQuadraticEnergy>(system))
            h_ = cgpu_vm_fire_create(vm->handle());
        else
            throw std::runtime_error("EnergyMinimizerFIRE drives a Voronoi or vertex model");
        if (!h_) throw std::runtime_error("FIRE create failed");
        push();
    }
    ~EnergyMinimizerFIRE() override { cgpu_fire_destroy(h_); }
    // one parameter at a time, as the reference setters (EnergyMinimizerFIRE2D.h:37-54)
    void setMaximumIterations(int m) { maxIt = m; one(CGPU_FIRE_MAX_IT, m); }
    void setForceCutoff(double f) { fc = f; one(CGPU_FIRE_FORCE_CUTOFF, f); }
    void setAlphaStart(double a) { a0 = a; one(CGPU_FIRE_ALPHA_START, a); }
    void setDeltaT(double d) override { dt = d; if (h_) one(CGPU_FIRE_DT, d); }
    void setDeltaTMax(double d) { dtMax = d; one(CGPU_FIRE_DT_MAX, d); }
    void setDeltaTInc(double d) { dtInc = d; one(CGPU_FIRE_DT_INC, d); }
    void setDeltaTDec(double d) { dtDec = d; one(CGPU_FIRE_DT_DEC, d); }
    void setAlphaDec(double a) { aDec = a; one(CGPU_FIRE_ALPHA_DEC, a); }
    void setNMin(int n) { nMin = n; one(CGPU_FIRE_NMIN, n); }
    void minimize() { check(cgpu_fire_minimize(h_), "minimize"); }
    void integrateEquationsOfMotion() override { minimize(); }
    double getMaxForce() {
        double out[6];
        check(cgpu_fire_get(h_, 0, out), "fire get");
        return out[0];
    }

   private:
    void push() { check(cgpu_fire_set(h_, maxIt, fc, a0, dt, dtMax, dtInc, dtDec, aDec, nMin), "fire set"); }
    void one(int which, double v) { check(cgpu_fire_set_param(h_, which, v), "fire set"); }
    void *h_ = nullptr;
    int maxIt = 1000, nMin = 5;
    double fc = 1e-7, a0 = 0.99, dt = 0.01, dtMax = 0.1, dtInc = 1.05, dtDec = 0.95, aDec = 0.9;
};

//! reference inc/Simulation.h:18-87, src/Simulation.cpp:141-166
class Simulation {
   public:
    void setConfiguration(shared_ptr<Simple2DModel> c) { config = c; }
    void addUpdater(shared_ptr<updater> u, shared_ptr<Simple2DModel> c) {
        u->set2DModel(c);
        updaters.push_back(u);
    }
    void setIntegrationTimestep(double dt) {
        integrationTimestep = dt;
        config->setDeltaT(dt);
        for (auto &u : updaters) u->setDeltaT(dt);
    }
    void setCPUOperation(bool setcpu) {
        if (setcpu) throw std::runtime_error("cellgpu_amd has no CPU path");
    }
    void setReproducible(bool r) {
        for (auto &u : updaters) u->setReproducible(r);
    }
    void performTimestep() {
        integerTimestep += 1;
        Time += integrationTimestep;
        for (auto &u : updaters) u->Update(integerTimestep);
    }
    int integerTimestep = 0;
    double Time = 0, integrationTimestep = 0.01;
    shared_ptr<Simple2DModel> config;
    vector<shared_ptr<updater>> updaters;
};

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
    void WriteState(shared_ptr<VoronoiQuadraticEnergy> s, double time = -1.0, int rec = -1) {
        check(cgpu_db_spv_write_batch(h_, s->handle(), 0, time, rec), "WriteState");
    }
    void ReadState(shared_ptr<VoronoiQuadraticEnergy> s, int rec, bool geometry = true) {
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
    void WriteState(shared_ptr<VertexQuadraticEnergy> s, double time = -1.0, int rec = -1) {
        check(cgpu_db_avm_write_batch(h_, s->handle(), 0, time, rec), "WriteState");
    }
    void ReadState(shared_ptr<VertexQuadraticEnergy> s, int rec, bool geometry = true) {
        (void)geometry;
        check(cgpu_db_avm_read_batch(h_, s->handle(), 0, rec), "ReadState");
    }
    int Records() const { return cgpu_db_records(h_); }
    std::string filename;
    const int Mode;

   private:
    void *h_;
};

typedef shared_ptr<Simple2DModel> ForcePtr;
typedef shared_ptr<simpleEquationOfMotion> EOMPtr;
typedef shared_ptr<updater> UpdaterPtr;
typedef shared_ptr<Simulation> SimulationPtr;

}  // namespace cellgpu
