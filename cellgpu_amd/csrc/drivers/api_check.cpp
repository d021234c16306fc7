// Exercise of the C++ drop-in layer (include/cellgpu_amd.hpp) for the parity
// tests: it drives the models through the reference's API (ArrayHandle reads
// and writes, getForces, moveDegreesOfFreedom + enforceTopology, Simulation
// with a sort period, EnergyMinimizerFIRE) and writes every array it sees to a
// text file, one line per array: "name count v0 v1 ...". The test compares the
// arrays with the oracle; this program checks only what needs no oracle.
//
//   api_check <out.txt>
#include <cmath>
#include <cstdio>
#include <memory>
#include <string>

#include "../../../include/cellgpu_amd.hpp"

using namespace cellgpu;

static FILE *g_out = nullptr;
static int g_fail = 0;

static void dump(const char *name, const double *v, size_t n) {
    std::fprintf(g_out, "%s %zu", name, n);
    for (size_t i = 0; i < n; ++i) std::fprintf(g_out, " %.17g", v[i]);
    std::fprintf(g_out, "\n");
}
static void dump(const char *name, const int *v, size_t n) {
    std::fprintf(g_out, "%s %zu", name, n);
    for (size_t i = 0; i < n; ++i) std::fprintf(g_out, " %d", v[i]);
    std::fprintf(g_out, "\n");
}
static void dump(const char *name, double v) { dump(name, &v, 1); }
static void dump2(const char *name, GPUArray<Dscalar2> &a) {
    ArrayHandle<Dscalar2> h(a, access_location::host, access_mode::read);
    dump(name, &h.data[0].x, 2 * a.getNumElements());
}
static void dumpi(const char *name, GPUArray<int> &a) {
    ArrayHandle<int> h(a, access_location::host, access_mode::read);
    dump(name, h.data, a.getNumElements());
}
static void expect(bool ok, const char *what) {
    std::printf("%s: %s\n", ok ? "ok" : "FAILED", what);
    if (!ok) ++g_fail;
}

// deterministic displacements of size <= amp
static GPUArray<Dscalar2> displacements(int n, double amp) {
    GPUArray<Dscalar2> d(n);
    ArrayHandle<Dscalar2> h(d, access_location::host, access_mode::overwrite);
    for (int i = 0; i < n; ++i) h.data[i] = Dscalar2{amp * std::sin(1.7 * i + 0.3), amp * std::cos(2.3 * i + 0.1)};
    return d;
}

static void voronoi_part() {
    const int n = 300;
    auto spv = std::make_shared<VoronoiQuadraticEnergy>(n, 1.0, 3.8, true, 0, 13377);
    spv->setCellPreferencesUniform(1.0, 3.75);
    dump2("v_pos0", spv->returnPositions());
    spv->computeForces();
    GPUArray<Dscalar2> f;
    spv->getForces(f);
    dump2("v_force0", f);
    dump2("v_ap0", spv->returnAreaPeri());
    dump("v_energy0", spv->computeEnergy());
    expect(spv->getMaxForce() > 0, "getMaxForce of a random start is positive");

    // move + enforceTopology (reference Simulation::moveDegreesOfFreedom, then the
    // model's enforceTopology)
    GPUArray<Dscalar2> d = displacements(n, 0.04);
    dump2("v_disp", d);
    spv->moveDegreesOfFreedom(d, 1.0);
    spv->enforceTopology();
    dump2("v_pos1", spv->returnPositions());
    dumpi("v_nn1", spv->cellNeighborNum);
    dumpi("v_nbr1", spv->cellNeighbors);
    dump("v_nm1", (double)spv->neighMax());
    spv->computeForces();
    dump2("v_force1", spv->returnForces());

    // writing through a host handle stores the positions back (retriangulated)
    {
        ArrayHandle<Dscalar2> p(spv->returnPositions(), access_location::host, access_mode::readwrite);
        const double L = spv->boxSide;
        for (int i = 0; i < 10; ++i) {
            p.data[i].x = std::fmod(p.data[i].x + 0.3, L);
            p.data[i].y = std::fmod(p.data[i].y + 0.2, L);
        }
    }
    dump2("v_pos2", spv->returnPositions());
    dumpi("v_nn2", spv->cellNeighborNum);
    dumpi("v_nbr2", spv->cellNeighbors);
    dump("v_nm2", (double)spv->neighMax());

    // directors round trip
    {
        ArrayHandle<Dscalar> t(spv->cellDirectors, access_location::host, access_mode::overwrite);
        for (int i = 0; i < n; ++i) t.data[i] = 0.01 * i;
    }
    {
        ArrayHandle<Dscalar> t(spv->cellDirectors, access_location::host, access_mode::read);
        bool same = true;
        for (int i = 0; i < n; ++i) same = same && t.data[i] == 0.01 * i;
        expect(same, "directors written through a handle read back");
    }
    bool refused = false;
    try {
        ArrayHandle<Dscalar2> bad(spv->returnForces(), access_location::host, access_mode::readwrite);
    } catch (const std::runtime_error &) {
        refused = true;
    }
    expect(refused, "forces are read only");
    refused = false;
    try {
        ArrayHandle<Dscalar2> bad(spv->returnPositions(), access_location::device, access_mode::read);
    } catch (const std::runtime_error &) {
        refused = true;
    }
    expect(refused, "device access through ArrayHandle is refused");
}

// Simulation with a sort period: the relabelled system follows the unsorted one
// cell by cell (no noise: v0 = 0)
static void sorting_part() {
    const int n = 200;
    auto a = std::make_shared<VoronoiQuadraticEnergy>(n, 1.0, 3.8, true, 0, 99);
    auto b = std::make_shared<VoronoiQuadraticEnergy>(n, 1.0, 3.8, true, 0, 99);
    SimulationPtr sa = std::make_shared<Simulation>(), sb = std::make_shared<Simulation>();
    for (auto &pr : {std::make_pair(a, sa), std::make_pair(b, sb)}) {
        pr.first->setv0Dr(0.0, 1.0);
        pr.second->setConfiguration(pr.first);
        pr.second->addUpdater(std::make_shared<selfPropelledParticleDynamics>(n), pr.first);
        pr.second->setIntegrationTimestep(0.01);
    }
    sa->setSortPeriod(3);
    for (int t = 0; t < 7; ++t) {
        sa->performTimestep();
        sb->performTimestep();
    }
    dump2("s_pos_sorted", a->returnPositions());
    dump2("s_pos_plain", b->returnPositions());
    dump("s_tags", a->idxToTag.data(), a->idxToTag.size());
    dump("s_itt", a->itt.data(), a->itt.size());
    dump("s_time", sa->Time);
    dump("s_model_time", a->currentTime);
}

static void fire_part() {
    const int n = 100;
    auto spv = std::make_shared<VoronoiQuadraticEnergy>(n, 1.0, 3.7, true, 0, 17);
    spv->setv0Dr(0.0, 0.0);
    const double e0 = spv->computeEnergy();
    auto fire = std::make_shared<EnergyMinimizerFIRE>(spv);
    fire->setMaximumIterations(20000);
    fire->setForceCutoff(1e-8);
    fire->setAlphaStart(0.1);
    fire->setDeltaT(0.01);
    SimulationPtr sim = std::make_shared<Simulation>();
    sim->setConfiguration(spv);
    sim->addUpdater(fire, spv);
    sim->performTimestep();
    dump("f_forcemax", fire->getMaxForce());
    dump("f_iterations", (double)fire->getIterations());
    dump("f_e0", e0);
    dump("f_e1", spv->computeEnergy());
    spv->computeForces();
    dump2("f_force", spv->returnForces());
}

static void vertex_part() {
    const int n = 64;
    auto avm = std::make_shared<VertexQuadraticEnergy>(n, 1.0, 3.9, true, false, 0, 5);
    avm->setT1Threshold(0.04);
    dump("m_nv", (double)avm->Nvertices);
    dump2("m_pos0", avm->returnPositions());
    avm->computeForces();
    GPUArray<Dscalar2> f;
    avm->getForces(f);
    dump2("m_force0", f);
    dump2("m_ap0", avm->returnAreaPeri());
    dumpi("m_vn0", avm->vertexNeighbors);
    dumpi("m_vcn0", avm->vertexCellNeighbors);
    dumpi("m_cvn0", avm->cellVertexNum);
    dumpi("m_cv0", avm->cellVertices);
    dump("m_vmax", (double)avm->vertexMax);
    dump2("m_cellpos0", avm->cellPositions);

    GPUArray<Dscalar2> d = displacements(avm->Nvertices, 0.06);
    dump2("m_disp", d);
    avm->moveDegreesOfFreedom(d, 1.0);
    // a T1 flip rotates its edge's vertices: the positions before it are the oracle's input
    dump2("m_pos_moved", avm->returnPositions());
    avm->enforceTopology();
    dump2("m_pos1", avm->returnPositions());
    dumpi("m_vn1", avm->vertexNeighbors);
    dumpi("m_vcn1", avm->vertexCellNeighbors);
    dumpi("m_cvn1", avm->cellVertexNum);
    dumpi("m_cv1", avm->cellVertices);
    avm->computeForces();
    dump2("m_force1", avm->returnForces());
    dump("m_energy1", avm->computeEnergy());

    // a Brownian step at T = 0 through the Simulation
    auto bd = std::make_shared<brownianParticleDynamics>(avm->Nvertices);
    bd->setT(0.0);
    SimulationPtr sim = std::make_shared<Simulation>();
    sim->setConfiguration(avm);
    sim->addUpdater(bd, avm);
    sim->setIntegrationTimestep(0.005);
    sim->performTimestep();
    dump2("m_pos2", avm->returnPositions());
    dumpi("m_cvn2", avm->cellVertexNum);
    dumpi("m_cv2", avm->cellVertices);
}

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: api_check <out.txt>\n");
        return 2;
    }
    g_out = std::fopen(argv[1], "w");
    if (!g_out) {
        std::perror(argv[1]);
        return 2;
    }
    try {
        std::printf("voronoi model\n");
        std::fflush(stdout);
        voronoi_part();
        std::printf("simulation with spatial sorting\n");
        std::fflush(stdout);
        sorting_part();
        std::printf("FIRE\n");
        std::fflush(stdout);
        fire_part();
        std::printf("vertex model\n");
        std::fflush(stdout);
        vertex_part();
    } catch (const std::exception &e) {
        std::fprintf(stderr, "api_check: %s\n", e.what());
        std::fclose(g_out);
        return 1;
    }
    std::fclose(g_out);
    std::printf("%s\n", g_fail ? "api_check FAILED" : "api_check done");
    return g_fail ? 1 : 0;
}
