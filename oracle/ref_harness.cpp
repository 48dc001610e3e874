// Test-harness entry points over the reference's own CPU sources.
//
// TEST INFRASTRUCTURE ONLY.  This file is compiled together with unmodified
// source files of open-Qmin (read in place under the reference checkout by
// oracle/Makefile) into oracle/_ref/libqmref.so.  Only the tests, the golden
// fixture scripts and bench.py's CPU-baseline leg load that library; the
// product never does.
//
// What is exercised here is the reference's code itself:
//   * per-site stencils     src/forces/lcForces.h:13-1043
//   * anchoring             src/forces/landauDeGennesLCBoundary.h:26-124
//   * Q algebra             inc/qTensorFunctions.h:34-241, inc/functions.h:131-159
//   * lattice model         src/model/cubicLattice.cpp:143-318,
//                           src/model/qTensorLatticeModel.cpp:95-238
//   * RNG                   src/utilities/noiseSource.cpp:16-47
//   * minimizers            src/updaters/energyMinimizerFIRE.cpp:77-170,
//                           velocityVerlet.cpp:8-40, energyMinimizerAdam.cpp:25-75,
//                           energyMinimizerNesterovAG.cpp:75-120,
//                           energyMinimizerLoLBFGS.cpp:105-226, baseUpdater.cpp:10-28
// The landauDeGennesLC class and multirankSimulation cannot be compiled here
// (their headers need mpi.h, which this image does not ship), so the force
// callback below is supplied by the caller (the C++ restatement in
// oracle/qmin_oracle.cpp).  The harness simulation only forwards calls.

#include <unistd.h>
#include <fcntl.h>
#include <cstdio>
#include <memory>
#include <vector>

#include "std_include.h"
#include "functions.h"
#include "qTensorFunctions.h"
#include "lcForces.h"
#include "latticeBoundaries.h"
#include "landauDeGennesLCBoundary.h"
#include "qTensorLatticeModel.h"
#include "basicSimulation.h"
#include "energyMinimizerFIRE.h"
#include "energyMinimizerAdam.h"
#include "energyMinimizerNesterovAG.h"
#include "energyMinimizerLoLBFGS.h"
#include "noiseSource.h"

namespace {

// Silence the reference's progress printf/cout during a call.
struct Quiet {
    int saved = -1;
    Quiet() {
        fflush(stdout);
        std::cout.flush();
        saved = dup(1);
        int devnull = open("/dev/null", O_WRONLY);
        if (devnull >= 0) { dup2(devnull, 1); close(devnull); }
    }
    ~Quiet() {
        fflush(stdout);
        std::cout.flush();
        if (saved >= 0) { dup2(saved, 1); close(saved); }
    }
};

inline dVec toDVec(const double* p) {
    dVec v;
    for (int i = 0; i < 5; ++i) v[i] = p[i];
    return v;
}
inline void fromDVec(const dVec& v, double* p) {
    for (int i = 0; i < 5; ++i) p[i] = v[i];
}
inline cubicLatticeDerivativeVector toDeriv(const double* p) {
    cubicLatticeDerivativeVector v;
    for (int i = 0; i < 15; ++i) v[i] = p[i];
    return v;
}

typedef void (*force_fn)(void* ctx, int n, const double* q_aos, double* f_aos);
typedef double (*energy_fn)(void* ctx, int n, const double* q_aos);

// A single-rank simulation that forwards computeForces to a caller-supplied
// function.  sumUpdaterData is the base no-op, i.e. one rank.
class harnessSimulation : public basicSimulation {
  public:
    shared_ptr<qTensorLatticeModel> model;
    force_fn force = nullptr;
    energy_fn energy = nullptr;
    void* ctx = nullptr;
    void computeForces() override {
        int n = model->getNumberOfParticles();
        std::vector<double> q(5 * (size_t)n), f(5 * (size_t)n);
        {
            ArrayHandle<dVec> pos(model->returnPositions(), access_location::host, access_mode::read);
            for (int i = 0; i < n; ++i) fromDVec(pos.data[i], &q[5 * (size_t)i]);
        }
        force(ctx, n, q.data(), f.data());
        ArrayHandle<dVec> frc(model->returnForces(), access_location::host, access_mode::overwrite);
        for (int i = 0; i < n; ++i) frc.data[i] = toDVec(&f[5 * (size_t)i]);
    }
    void moveParticles(GPUArray<dVec>& disp, scalar scale = 1.0) override {
        model->moveParticles(disp, scale);
    }
    scalar computePotentialEnergy(bool) override {
        if (!energy) return 0.0;
        int n = model->getNumberOfParticles();
        std::vector<double> q(5 * (size_t)n);
        ArrayHandle<dVec> pos(model->returnPositions(), access_location::host, access_mode::read);
        for (int i = 0; i < n; ++i) fromDVec(pos.data[i], &q[5 * (size_t)i]);
        return energy(ctx, n, q.data());
    }
};

// Subclasses that expose the protected FIRE scalars for per-iteration traces.
class tracedFIRE : public energyMinimizerFIRE {
  public:
    using energyMinimizerFIRE::energyMinimizerFIRE;
    double power() const { return Power; }
    double alphaValue() const { return alpha; }
    double dt() const { return deltaT; }
    double scale() const { return scaling; }
    int nPositive() const { return NSinceNegativePower; }
};

struct RefModel {
    shared_ptr<qTensorLatticeModel> model;
};

}  // namespace

extern "C" {

// ---------------------------------------------------------------- per site
int qmref_boundary_case(const int* t) {
    return lcForce::getBoundaryCase(t[0], t[1], t[2], t[3], t[4], t[5]);
}

void qmref_bulk_l1(double L1, const double* q, const double* nb, double* out) {
    dVec s(0.0);
    lcForce::bulkL1Force(L1, toDVec(q), toDVec(nb), toDVec(nb + 5), toDVec(nb + 10), toDVec(nb + 15),
                         toDVec(nb + 20), toDVec(nb + 25), s);
    fromDVec(s, out);
}

void qmref_boundary_l1(double L1, const double* q, const double* nb, const int* t, double* out) {
    dVec s(0.0);
    lcForce::boundaryL1Force(L1, toDVec(q), toDVec(nb), toDVec(nb + 5), toDVec(nb + 10), toDVec(nb + 15),
                             toDVec(nb + 20), toDVec(nb + 25), t[0], t[1], t[2], t[3], t[4], t[5], s);
    fromDVec(s, out);
}

// k in {2,3,4,6}; bcase < 0 selects the bulk function.
void qmref_lk(int k, double L, int bcase, const double* q, const double* nb, const double* nd, double* out) {
    dVec s(0.0);
    dVec qc = toDVec(q), xd = toDVec(nb), xu = toDVec(nb + 5), yd = toDVec(nb + 10), yu = toDVec(nb + 15),
         zd = toDVec(nb + 20), zu = toDVec(nb + 25);
    cubicLatticeDerivativeVector dxd = toDeriv(nd), dxu = toDeriv(nd + 15), dyd = toDeriv(nd + 30),
                                 dyu = toDeriv(nd + 45), dzd = toDeriv(nd + 60), dzu = toDeriv(nd + 75);
    if (bcase < 0) {
        switch (k) {
            case 2: lcForce::bulkL2Force(L, qc, xd, xu, yd, yu, zd, zu, dxd, dxu, dyd, dyu, dzd, dzu, s); break;
            case 3: lcForce::bulkL3Force(L, qc, xd, xu, yd, yu, zd, zu, dxd, dxu, dyd, dyu, dzd, dzu, s); break;
            case 4: lcForce::bulkL4Force(L, qc, xd, xu, yd, yu, zd, zu, dxd, dxu, dyd, dyu, dzd, dzu, s); break;
            case 6: lcForce::bulkL6Force(L, qc, xd, xu, yd, yu, zd, zu, dxd, dxu, dyd, dyu, dzd, dzu, s); break;
        }
    } else {
        switch (k) {
            case 2: lcForce::boundaryL2Force(L, bcase, qc, xd, xu, yd, yu, zd, zu, dxd, dxu, dyd, dyu, dzd, dzu, s); break;
            case 3: lcForce::boundaryL3Force(L, bcase, qc, xd, xu, yd, yu, zd, zu, dxd, dxu, dyd, dyu, dzd, dzu, s); break;
            case 4: lcForce::boundaryL4Force(L, bcase, qc, xd, xu, yd, yu, zd, zu, dxd, dxu, dyd, dyu, dzd, dzu, s); break;
            case 6: lcForce::boundaryL6Force(L, bcase, qc, xd, xu, yd, yu, zd, zu, dxd, dxu, dyd, dyu, dzd, dzu, s); break;
        }
    }
    fromDVec(s, out);
}

// btype 0 = homeotropic, 1 = degenerate planar (latticeBoundaries.h:8)
void qmref_anchoring_force(int btype, double P1, double P2, const double* q, const double* qb, double* out) {
    boundaryObject b(btype == 0 ? boundaryType::homeotropic : boundaryType::degeneratePlanar, P1, P2);
    dVec f(0.0);
    computeBoundaryForce(toDVec(q), toDVec(qb), b, f);
    fromDVec(f, out);
}

double qmref_anchoring_energy(int btype, double P1, double P2, const double* q, const double* qb) {
    boundaryObject b(btype == 0 ? boundaryType::homeotropic : boundaryType::degeneratePlanar, P1, P2);
    return computeBoundaryEnergy(toDVec(q), toDVec(qb), b);
}

// out[0..4] dTrQ2, out[5..9] dTrQ3, out[10..14] d(TrQ2)^2; tr = {TrQ2, TrQ3, TrQ2^2}
void qmref_phase(const double* q, double* out, double* tr) {
    dVec qc = toDVec(q);
    fromDVec(derivativeTrQ2(qc), out);
    fromDVec(derivativeTrQ3(qc), out + 5);
    fromDVec(derivativeTrQ2Squared(qc), out + 10);
    tr[0] = TrQ2(qc);
    tr[1] = TrQ3(qc);
    tr[2] = TrQ2Squared(qc);
}

void qmref_eigenvalues(const double* q, double* ev) {
    dVec qc = toDVec(q);
    double a, b, c;
    eigenvaluesOfQ(qc, a, b, c);
    ev[0] = a; ev[1] = b; ev[2] = c;
}

double qmref_determinant(const double* q) {
    dVec qc = toDVec(q);
    return determinantOfQ(qc);
}

void qmref_q_from_director(const double* n, double S0, double* q) {
    scalar3 v; v.x = n[0]; v.y = n[1]; v.z = n[2];
    dVec out(0.0);
    qTensorFromDirector(v, S0, out);
    fromDVec(out, q);
}

double qmref_dotvec(const double* a, const double* b) { return dotVec(toDVec(a), toDVec(b)); }
double qmref_dot(const double* a, const double* b) { return dot(toDVec(a), toDVec(b)); }

// Raw mt19937 stream and the doubles getRealUniform() draws from it.
void qmref_uniform_stream(int seed, int n, double* out) {
    noiseSource noise(true);
    noise.setReproducibleSeed(seed);
    for (int i = 0; i < n; ++i) out[i] = noise.getRealUniform();
}

// ---------------------------------------------------------------- model
void* qmref_model_create(int lx, int ly, int lz) {
    RefModel* m = new RefModel;
    m->model = make_shared<qTensorLatticeModel>(lx, ly, lz, false, true);
    return m;
}

void qmref_model_destroy(void* h) { delete static_cast<RefModel*>(h); }

void qmref_model_neighbors(void* h, int* out) {
    auto& m = static_cast<RefModel*>(h)->model;
    Quiet q;
    m->fillNeighborLists(0);
    ArrayHandle<int> nb(m->neighboringSites, access_location::host, access_mode::read);
    int n = m->getNumberOfParticles();
    for (int i = 0; i < 6 * n; ++i) out[i] = nb.data[i];
}

void qmref_model_random_init(void* h, int seed, double S0) {
    auto& m = static_cast<RefModel*>(h)->model;
    noiseSource noise(true);
    noise.setReproducibleSeed(seed);
    m->setNematicQTensorRandomly(noise, S0);
}

void qmref_model_random_init_aligned(void* h, int seed, double S0) {
    auto& m = static_cast<RefModel*>(h)->model;
    noiseSource noise(true);
    noise.setReproducibleSeed(seed);
    m->setNematicQTensorRandomly(noise, S0, true);
}

// cubicLattice::boundarySites / surfaceSites of object obj; returns the count
int qmref_model_object_sites(void* h, int obj, int surface, int* out) {
    auto& m = static_cast<RefModel*>(h)->model;
    GPUArray<int>& a = surface ? m->surfaceSites[obj] : m->boundarySites[obj];
    const int count = a.getNumElements();
    if (out) {
        ArrayHandle<int> v(a, access_location::host, access_mode::read);
        for (int i = 0; i < count; ++i) out[i] = v.data[i];
    }
    return count;
}

void qmref_model_displace_object(void* h, int obj, int direction, int magnitude) {
    auto& m = static_cast<RefModel*>(h)->model;
    {
        Quiet q;
        m->fillNeighborLists(0);   // the move reads the neighbour table
    }
    m->displaceBoundaryObject(obj, direction, magnitude);
}

// normalizeDirector (inc/functions.h), in place on n[3]
void qmref_normalize_director(double* n) {
    scalar3 v;
    v.x = n[0];
    v.y = n[1];
    v.z = n[2];
    normalizeDirector(v);
    n[0] = v.x;
    n[1] = v.y;
    n[2] = v.z;
}

// sites: lattice indices; btype as in qmref_anchoring_force
void qmref_model_add_object(void* h, const int* sites, int n, int btype, double P1, double P2) {
    auto& m = static_cast<RefModel*>(h)->model;
    std::vector<int> s(sites, sites + n);
    Quiet q;
    m->createBoundaryObject(s, btype == 0 ? boundaryType::homeotropic : boundaryType::degeneratePlanar, P1, P2);
}

void qmref_model_get_q(void* h, double* out) {
    auto& m = static_cast<RefModel*>(h)->model;
    ArrayHandle<dVec> pos(m->returnPositions(), access_location::host, access_mode::read);
    for (int i = 0; i < m->getNumberOfParticles(); ++i) fromDVec(pos.data[i], out + 5 * (size_t)i);
}

void qmref_model_set_q(void* h, const double* in) {
    auto& m = static_cast<RefModel*>(h)->model;
    ArrayHandle<dVec> pos(m->returnPositions(), access_location::host, access_mode::overwrite);
    for (int i = 0; i < m->getNumberOfParticles(); ++i) pos.data[i] = toDVec(in + 5 * (size_t)i);
}

void qmref_model_get_types(void* h, int* out) {
    auto& m = static_cast<RefModel*>(h)->model;
    ArrayHandle<int> t(m->returnTypes(), access_location::host, access_mode::read);
    for (int i = 0; i < m->getNumberOfParticles(); ++i) out[i] = t.data[i];
}

void qmref_model_defects(void* h, int defectType, double* out) {
    auto& m = static_cast<RefModel*>(h)->model;
    int n = m->getNumberOfParticles();
    if (m->defectMeasures.getNumElements() < (unsigned)n) m->defectMeasures.resize(n);
    m->computeDefectMeasures(defectType);
    ArrayHandle<double> d(m->defectMeasures, access_location::host, access_mode::read);
    for (int i = 0; i < n; ++i) out[i] = d.data[i];
}

// ---------------------------------------------------------------- minimizers
// All minimizers run on the model's current Q (velocities start at zero), with
// forces from `force(ctx, ...)`.  Per-iteration traces are recorded by running
// one iteration per call of the reference's own loop body.

// trace: per iteration {forceMax, Power, alpha, deltaT, scaling, NSinceNegativePower}
int qmref_fire_run(void* h, force_fn force, void* ctx, double dt, double alphaStart, double dtMax,
                   double dtInc, double dtDec, double alphaDec, int nMin, double fc, double alphaMin,
                   int maxIter, double* trace) {
    auto& m = static_cast<RefModel*>(h)->model;
    Quiet quiet;
    auto sim = make_shared<harnessSimulation>();
    sim->model = m;
    sim->force = force;
    sim->ctx = ctx;
    auto fire = make_shared<tracedFIRE>(m);
    fire->setSimulation(sim);
    fire->setMaximumIterations(maxIter);
    fire->setFIREParameters(dt, alphaStart, dtMax, dtInc, dtDec, alphaDec, nMin, fc, alphaMin);
    fire->getNTotal();
    // energyMinimizerFIRE::minimize (energyMinimizerFIRE.cpp:151-170), one
    // loop body at a time so the scalars can be recorded.
    sim->computeForces();
    int it0 = fire->iterations;
    double fmax = 100.0;
    while ((fire->iterations < maxIter && fmax > fc) || fire->iterations == it0) {
        fire->iterations += 1;
        fire->integrateEquationOfMotion();
        fire->fireStep();
        fmax = fire->getMaxForce();
        if (trace) {
            double* t = trace + 6 * (size_t)(fire->iterations - it0 - 1);
            t[0] = fmax; t[1] = fire->power(); t[2] = fire->alphaValue(); t[3] = fire->dt();
            t[4] = fire->scale(); t[5] = fire->nPositive();
        }
    }
    return fire->iterations - it0;
}

// The reference's own FIRE minimize() loop; returns iterations performed.
int qmref_fire_minimize(void* h, force_fn force, void* ctx, double dt, double alphaStart, double dtMax,
                        double dtInc, double dtDec, double alphaDec, int nMin, double fc, double alphaMin,
                        int maxIter, double* forceMaxOut) {
    auto& m = static_cast<RefModel*>(h)->model;
    Quiet quiet;
    auto sim = make_shared<harnessSimulation>();
    sim->model = m;
    sim->force = force;
    sim->ctx = ctx;
    auto fire = make_shared<energyMinimizerFIRE>(m);
    fire->setSimulation(sim);
    fire->setMaximumIterations(maxIter);
    fire->setFIREParameters(dt, alphaStart, dtMax, dtInc, dtDec, alphaDec, nMin, fc, alphaMin);
    fire->getNTotal();
    fire->minimize();
    if (forceMaxOut) *forceMaxOut = fire->getMaxForce();
    return fire->iterations;
}

// trace: forceMax per iteration
int qmref_adam_run(void* h, force_fn force, void* ctx, double b1, double b2, double eps, double dt, double fc,
                   int maxIter, double* trace) {
    auto& m = static_cast<RefModel*>(h)->model;
    Quiet quiet;
    auto sim = make_shared<harnessSimulation>();
    sim->model = m;
    sim->force = force;
    sim->ctx = ctx;
    auto adam = make_shared<energyMinimizerAdam>();
    adam->setModel(m);
    adam->initializeFromModel();
    adam->setSimulation(sim);
    adam->setAdamParameters(b1, b2, eps, dt, fc);
    int done = 0;
    for (int k = 1; k <= maxIter; ++k) {
        adam->setMaximumIterations(k);
        int before = adam->iterations;
        adam->minimize();
        if (adam->iterations == before) break;
        if (trace) trace[k - 1] = adam->getMaxForce();
        done = k;
        if (adam->getMaxForce() <= fc) break;
    }
    return done;
}

int qmref_nesterov_run(void* h, force_fn force, void* ctx, double dt, double mu, double fc, int scheduled,
                       int maxIter, double* trace) {
    auto& m = static_cast<RefModel*>(h)->model;
    Quiet quiet;
    auto sim = make_shared<harnessSimulation>();
    sim->model = m;
    sim->force = force;
    sim->ctx = ctx;
    auto nag = make_shared<energyMinimizerNesterovAG>(m);
    nag->setSimulation(sim);
    nag->setNesterovAGParameters(dt, mu, fc);
    nag->scheduledMomentum = scheduled != 0;
    int done = 0;
    for (int k = 1; k <= maxIter; ++k) {
        nag->setMaximumIterations(k);
        int before = nag->iterations;
        nag->minimize();
        if (nag->iterations == before) break;
        if (trace) trace[k - 1] = nag->getMaxForce();
        done = k;
        if (nag->getMaxForce() <= fc) break;
    }
    return done;
}

int qmref_lolbfgs_run(void* h, force_fn force, void* ctx, int mem, double dt, double c, double fc, double tau,
                      int scheduled, int maxIter, double* trace) {
    auto& m = static_cast<RefModel*>(h)->model;
    Quiet quiet;
    auto sim = make_shared<harnessSimulation>();
    sim->model = m;
    sim->force = force;
    sim->ctx = ctx;
    auto lb = make_shared<energyMinimizerLoLBFGS>(m);
    lb->setSimulation(sim);
    lb->setLoLBFGSParameters(mem, dt, c, fc, tau);
    lb->scheduledMomentum = scheduled != 0;
    int done = 0;
    for (int k = 1; k <= maxIter; ++k) {
        lb->setMaximumIterations(k);
        int before = lb->iterations;
        lb->minimize();
        if (lb->iterations == before) break;
        if (trace) trace[k - 1] = lb->getMaxForce();
        done = k;
        if (lb->getMaxForce() <= fc) break;
    }
    return done;
}

}  // extern "C"
