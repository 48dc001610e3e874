// Reference-style access to the lattice state through GPUArray / ArrayHandle
// (src/utilities/gpuarray.h) and force::computeForces(GPUArray<dVec>&, bool,
// int) (src/forces/baseForce.h:23), as code written for open-Qmin does it:
//
//   array-views-example L seed
//
// sets a seeded random state, changes one site's Q through a read-write
// handle, computes the forces into the model's force array and into an array
// of its own (replacing, then adding), and prints one line per check plus the
// forces of the first sites, which the test compares with the CPU oracle.
#include <cstdio>
#include <cstdlib>

#include "openqmin_hip.h"

int main(int argc, char** argv) {
    const int L = argc > 1 ? std::atoi(argv[1]) : 12;
    const int seed = argc > 2 ? std::atoi(argv[2]) : 1234;
    try {
        auto Configuration = std::make_shared<multirankQTensorLatticeModel>(L, L, L);
        auto force = std::make_shared<landauDeGennesLC>(-1.0, -12.325581395348838, 10.058139534883722, 2.32, 0.5,
                                                        0.3, 0.0, 0.0);
        force->setModel(Configuration);
        noiseSource noise(true);
        noise.setReproducibleSeed(seed);
        const scalar S0 = 0.532864615469017;
        Configuration->setNematicQTensorRandomly(noise, S0);
        boundaryObject sphere(boundaryType::homeotropic, 5.3, S0);
        auto sim = std::make_shared<multirankSimulation>();
        sim->setConfiguration(Configuration);
        sim->addForce(force);
        sim->createSphericalColloid(make_scalar3(L / 2.0, L / 2.0, L / 2.0), L / 4.0, sphere);
        {
            // a read-write handle: the change reaches the device when it closes
            ArrayHandle<dVec> q(Configuration->returnPositions());
            q.data[1][0] += 0.125;
            q.data[1][4] -= 0.0625;
        }
        {
            ArrayHandle<dVec> q(Configuration->returnPositions(), access_location::host, access_mode::read);
            std::printf("q1 %.17g %.17g\n", q.data[1][0], q.data[1][4]);
        }
        ArrayHandle<int> t(Configuration->returnTypes(), access_location::host, access_mode::read);
        int objectSites = 0;
        for (int i = 0; i < Configuration->getNumberOfParticles(); ++i) objectSites += t.data[i] > 0;
        std::printf("objects %d\n", objectSites);
        // into the model's own array, then twice into an array of the caller's
        force->computeForces(Configuration->returnForces());
        GPUArray<dVec> mine;
        force->computeForces(mine, true, 0);
        force->computeForces(mine, false, 1);   // the CPU path's boundary pass: adds nothing
        force->computeForces(mine, false, 0);   // adds the forces once more
        ArrayHandle<dVec> f(Configuration->returnForces(), access_location::host, access_mode::read);
        ArrayHandle<dVec> g(mine, access_location::host, access_mode::read);
        double worst = 0;
        for (int i = 0; i < Configuration->getNumberOfParticles(); ++i)
            for (int k = 0; k < 5; ++k) worst = std::fmax(worst, std::fabs(g.data[i][k] - 2 * f.data[i][k]));
        std::printf("doubled %.3g\n", worst);
        for (int i = 0; i < 4; ++i)
            std::printf("f %d %.17g %.17g %.17g %.17g %.17g\n", i, f.data[i][0], f.data[i][1], f.data[i][2],
                        f.data[i][3], f.data[i][4]);
        std::printf("size %.6g\n", Configuration->getClassSize());
        bool threw = false;
        try {
            ArrayHandle<dVec> d(Configuration->returnPositions(), access_location::device);
        } catch (const qminError&) {
            threw = true;
        }
        std::printf("device-handle-throws %d\n", threw ? 1 : 0);
    } catch (const qminError& e) {
        std::fprintf(stderr, "qmin error: %s\n", e.what());
        return 1;
    }
    return 0;
}
