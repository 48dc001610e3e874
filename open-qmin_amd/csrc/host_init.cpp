// Host-side initial state and object rasterization that must reproduce the
// reference bit for bit.  Compiled with g++ (as the reference's own CPU build
// is) rather than hipcc: the two compilers lower some libm calls differently,
// which changes the seeded Q in the last bit.
#include "host_init.h"

#include <cmath>
#include <random>

namespace qmin {

// qTensorFromDirector, src/../inc/qTensorFunctions.h:34-51, same operation order
void director_q(double nx, double ny, double nz, double S0, double* q) {
    double scale = std::sqrt(nx * nx + ny * ny + nz * nz);
    if (scale == 0) {
        q[0] = -0.5 * S0;
        q[3] = -0.5 * S0;
    } else if (scale != 1) {
        nx /= scale;
        ny /= scale;
        nz /= scale;
    }
    q[0] = 1.5 * S0 * (nx * nx - 1.0 / 3.0);
    q[1] = 1.5 * S0 * nx * ny;
    q[2] = 1.5 * S0 * nx * nz;
    q[3] = 1.5 * S0 * (ny * ny - 1.0 / 3.0);
    q[4] = 1.5 * S0 * ny * nz;
}

// noiseSource::setReproducibleSeed + getRealUniform (src/utilities/noiseSource.cpp:16-47)
// and qTensorLatticeModel::setNematicQTensorRandomly (src/model/qTensorLatticeModel.cpp:134-177):
// two global draws, then theta = acos(2u - 1), phi = 2 pi u per site, applied where t <= 0.
void random_nematic(long long n, const int* types, int seed, double S0, double* q) {
    const double pi = 3.14159265358979323846;
    std::mt19937 gen(seed);
    auto uniform = [&gen]() {
        std::uniform_real_distribution<double> d(0.0, 1.0);
        return d(gen);
    };
    double gTheta = std::acos(2.0 * uniform() - 1);
    double gPhi = 2.0 * pi * uniform();
    (void)gTheta;
    (void)gPhi;
    for (long long pp = 0; pp < n; ++pp) {
        double theta = std::acos(2.0 * uniform() - 1);
        double phi = 2.0 * pi * uniform();
        if (types[pp] <= 0) {
            double nx = std::cos(phi) * std::sin(theta);
            double ny = std::sin(phi) * std::sin(theta);
            double nz = std::cos(theta);
            director_q(nx, ny, nz, S0, q + 5 * pp);
        }
    }
}

// The globally aligned case of the same functions (and of
// multirankQTensorLatticeModel::setRandomDirectors, multirankQTensorLatticeModel.cpp:52-77):
// the director of the two global draws of the reproducible stream.
void random_director(int seed, double* n) {
    const double pi = 3.14159265358979323846;
    std::mt19937 gen(seed);
    std::uniform_real_distribution<double> d(0.0, 1.0);
    const double theta = std::acos(2.0 * d(gen) - 1);
    const double phi = 2.0 * pi * d(gen);
    n[0] = std::cos(phi) * std::sin(theta);
    n[1] = std::sin(phi) * std::sin(theta);
    n[2] = std::cos(theta);
}

// normalizeDirector (inc/functions.h:193-205): unit vector, x for a zero vector
void normalize_director(double* n) {
    const double a = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (a == 0) {
        n[0] = 1;
        n[1] = 0;
        n[2] = 0;
    } else {
        n[0] = n[0] / a;
        n[1] = n[1] / a;
        n[2] = n[2] / a;
    }
}

// multirankSimulation::createSphericalColloid (src/simulation/multirankSimulationBoundaries.cpp:169-215):
// sites with |r - c|^2 < R^2 over the half-open box [ceil(c - R), floor(c + R))
void sphere_sites(double cx, double cy, double cz, double radius, int btype, double S0, std::vector<int>& xyz,
                  std::vector<double>& qb) {
    for (int xx = (int)std::ceil(cx - radius); xx < std::floor(cx + radius); ++xx)
        for (int yy = (int)std::ceil(cy - radius); yy < std::floor(cy + radius); ++yy)
            for (int zz = (int)std::ceil(cz - radius); zz < std::floor(cz + radius); ++zz) {
                double dx = xx - cx, dy = yy - cy, dz = zz - cz;
                if ((dx * dx + dy * dy + dz * dz) < radius * radius) {
                    xyz.push_back(xx);
                    xyz.push_back(yy);
                    xyz.push_back(zz);
                    double Qt[5] = {0, 0, 0, 0, 0};
                    if (btype == 0) {
                        director_q(dx, dy, dz, S0, Qt);
                    } else {
                        double nrm = std::sqrt(dx * dx + dy * dy + dz * dz);
                        Qt[0] = dx / nrm;
                        Qt[1] = dy / nrm;
                        Qt[2] = dz / nrm;
                    }
                    qb.insert(qb.end(), Qt, Qt + 5);
                }
            }
}

}  // namespace qmin
