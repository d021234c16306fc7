// Timing driver of the self-propelled Voronoi model with the reference's
// command-line flags (abaso/cellGPU voronoi.cpp:37-58: -n cells, -t steps,
// -i initialization steps, -e dt, -p p0, -a a0, -v v0; -g is accepted and must
// name a GPU). Written against the C++ drop-in layer.
#include <getopt.h>

#include <chrono>
#include <cstdio>
#include <memory>

#include "../../../include/cellgpu_amd.hpp"

using namespace cellgpu;

int main(int argc, char *argv[]) {
    int numpts = 200, USE_GPU = 0, tSteps = 5, initSteps = 1, c;
    double dt = 0.01, p0 = 3.8, a0 = 1.0, v0 = 0.1;
    while ((c = getopt(argc, argv, "n:g:m:s:r:a:i:v:b:x:y:z:p:t:e:")) != -1) switch (c) {
            case 'n': numpts = atoi(optarg); break;
            case 't': tSteps = atoi(optarg); break;
            case 'g': USE_GPU = atoi(optarg); break;
            case 'i': initSteps = atoi(optarg); break;
            case 'e': dt = atof(optarg); break;
            case 'p': p0 = atof(optarg); break;
            case 'a': a0 = atof(optarg); break;
            case 'v': v0 = atof(optarg); break;
            default: break;
        }
    if (USE_GPU < 0) {
        fprintf(stderr, "cellgpu_amd runs on the GPU only\n");
        return 1;
    }
    EOMPtr spp = std::make_shared<selfPropelledParticleDynamics>(numpts);
    auto spv = std::make_shared<VoronoiQuadraticEnergy>(numpts, 1.0, 4.0, true);
    spv->setCellPreferencesUniform(1.0, p0);
    spv->setv0Dr(v0, 1.0);
    SimulationPtr sim = std::make_shared<Simulation>();
    sim->setConfiguration(spv);
    sim->addUpdater(spp, spv);
    sim->setIntegrationTimestep(dt);
    sim->setCPUOperation(false);
    sim->setReproducible(true);
    for (int ii = 0; ii < initSteps; ++ii) sim->performTimestep();
    printf("current q = %f\n", spv->reportq());
    // the mean force should be numerically close to zero
    spv->reportMeanCellForce(false);
    auto t1 = std::chrono::steady_clock::now();
    for (int ii = 0; ii < tSteps; ++ii) sim->performTimestep();
    cgpu_device_synchronize();
    printf("energy %f, local topology updates %lld\n", spv->computeEnergy(), spv->localTopologyUpdates());
    auto t2 = std::chrono::steady_clock::now();
    double steptime = std::chrono::duration<double>(t2 - t1).count() / tSteps;
    printf("timestep ~ %g per frame; \n%f\n", steptime, spv->reportq());
    return 0;
}
