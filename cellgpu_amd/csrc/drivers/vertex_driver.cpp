// Timing driver of the vertex model with the reference's command-line flags
// (abaso/cellGPU Vertex.cpp:27-52: -n cells, -t steps, -i initialization steps,
// -e dt, -p p0, -a a0, -v v0, -d Dr; -z 1 selects self-propelled cell-vertex
// dynamics instead of Brownian vertices at T = v0; -g is accepted and must name
// a GPU). Cells start from 100 smoothing Voronoi steps (runSPV), the T1
// threshold is 0.1, as in the reference driver. -z -1 also saves the state to
// ./data/test_p<p0>.nc (AVM NetCDF layout) at the start, after the initial
// steps and every 100/dt steps. Written against the C++ drop-in layer.
#include <getopt.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <memory>
#include <string>

#include "../../../include/cellgpu_amd.hpp"

using namespace cellgpu;

int main(int argc, char *argv[]) {
    int numpts = 200, USE_GPU = 0, tSteps = 5, initSteps = 1, program_switch = 0, c;
    double dt = 0.01, p0 = 4.0, a0 = 1.0, v0 = 0.05, Dr = 1.0;
    while ((c = getopt(argc, argv, "n:g:m:s:r:a:i:v:b:x:y:z:p:t:e:d:")) != -1) switch (c) {
            case 'n': numpts = atoi(optarg); break;
            case 't': tSteps = atoi(optarg); break;
            case 'g': USE_GPU = atoi(optarg); break;
            case 'i': initSteps = atoi(optarg); break;
            case 'z': program_switch = atoi(optarg); break;
            case 'e': dt = atof(optarg); break;
            case 'p': p0 = atof(optarg); break;
            case 'a': a0 = atof(optarg); break;
            case 'v': v0 = atof(optarg); break;
            case 'd': Dr = atof(optarg); break;
            default: break;
        }
    if (USE_GPU < 0) {
        fprintf(stderr, "cellgpu_amd runs on the GPU only\n");
        return 1;
    }
    auto avm = std::make_shared<VertexQuadraticEnergy>(numpts, 1.0, 4.0, true, true);
    avm->setCellPreferencesUniform(a0, p0);
    avm->setv0Dr(v0, Dr);
    avm->setT1Threshold(0.1);
    shared_ptr<updater> eom;
    if (program_switch == 1) {
        eom = std::make_shared<selfPropelledCellVertexDynamics>(numpts, avm->Nvertices);
    } else {
        auto bd = std::make_shared<brownianParticleDynamics>(avm->Nvertices);
        bd->setT(v0);
        eom = bd;
    }
    SimulationPtr sim = std::make_shared<Simulation>();
    sim->setConfiguration(avm);
    sim->addUpdater(eom, avm);
    sim->setIntegrationTimestep(dt);
    sim->setCPUOperation(false);
    sim->setReproducible(true);
    std::unique_ptr<AVMDatabaseNetCDF> ncdat;
    if (program_switch < 0) {
        char dataname[256];
        snprintf(dataname, sizeof(dataname), "./data/test_p%.3f.nc", p0);
        ncdat.reset(new AVMDatabaseNetCDF(avm->Nvertices, dataname, NcFile::Replace));
        ncdat->WriteState(avm);
    }
    const int saveEvery = std::max(1, (int)(100 / dt));
    double2v mf = avm->meanVertexForce();
    printf("mean vertex force = (%g, %g)\n", mf.x, mf.y);
    for (int ii = 0; ii < initSteps + 1; ++ii) sim->performTimestep();
    mf = avm->meanVertexForce();
    printf("mean vertex force = (%g, %g)\n", mf.x, mf.y);
    if (ncdat) ncdat->WriteState(avm);
    auto t1 = std::chrono::steady_clock::now();
    for (int ii = 0; ii < tSteps; ++ii) {
        sim->performTimestep();
        if (ncdat && ii % saveEvery == 0) ncdat->WriteState(avm);
    }
    cgpu_device_synchronize();
    auto t2 = std::chrono::steady_clock::now();
    printf("timestep time per iteration currently at %g\n\n",
           std::chrono::duration<double>(t2 - t1).count() / tSteps);
    mf = avm->meanVertexForce();
    printf("mean vertex force = (%g, %g)\n", mf.x, mf.y);
    printf("Mean q = %f\n", avm->reportq());
    return 0;
}
