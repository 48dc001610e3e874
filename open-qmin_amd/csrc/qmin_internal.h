// Internal definitions shared by the host translation units of libqmin.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/qmin.h"
#include "qmin_device.h"

using qmin::Ctrl;
using qmin::DevObject;
using qmin::StateSet;
using qmin::MarchArgs;
using qmin::BoundaryArgs;
using qmin::FireConst;

struct HostObject {
    int type;
    double W, S0;
    std::vector<int64_t> sites;     // cubicLattice::boundarySites: creation order
    std::vector<int64_t> surface;   // cubicLattice::surfaceSites: order of discovery
    std::vector<double> qb;         // FP64 Q of each entry of sites (5 per site), as last set from the host
};

struct qm_lattice {
    uint64_t serial = 0;   // unique per lattice of the process (addresses can be reused)
    int lx = 0, ly = 0, lz = 0, precision = QM_FP64;
    long long n = 0;
    size_t realSize = 8;
    int device = 0;                  // the GPU that holds the lattice (current at creation)
    hipStream_t stream = nullptr;    // the stream all of the lattice's work runs on
    hipStream_t ownStream = nullptr; // the lattice's own (stream unless the caller set one)
    std::vector<int32_t> types;      // host copy, reference encoding
    std::vector<HostObject> objects;
    qmin::SiteType* dTypes = nullptr;
    int8_t* dClass = nullptr;             // 0 composed bulk formula, > 0 object, < 0 boundary pass
    void* dW = nullptr;                   // composed bulk operator weights (Real)
    DevObject* dObjects = nullptr;
    void* dSet[2] = {nullptr, nullptr};   // each 15 n reals: Q, U, F planes
    int32_t* dSites = nullptr;            // compacted sites finished by the boundary pass
    long long nBoundary = 0;
    Ctrl* dCtrl = nullptr;
    Ctrl hCtrl{};                          // host mirror, valid after syncCtrl
    double* dPartials = nullptr;
    size_t partialsCap = 0;
    double* dTrace = nullptr;
    int traceCap = 0;
    bool typesDirty = true;
    int64_t nActive = 0;      // sites with type <= 0, counted when the types change
    int zchunk = 16;          // z extent one work-group marches through
    int zchunkOverride = 0;   // fixed z chunk (0 = choose from occupancy)
    int tileOverride = -1;    // tile shape of the fused pass (-1 = by precision)
    int terms = 0;            // distortion terms of the force last prepared on this lattice
    double* dHField = nullptr;  // spatially varying H: planes hx, hy, hz of n (FP64)
    int numCUs = 0;
};

struct qm_ldg {
    double A = -1, B = 0, C = 0;
    double L1 = 2.32, L2 = 0, L3 = 0, L4 = 0, L6 = 0;
    int multi = 0;
    int semantics = QM_SEMANTICS_REF_GPU;
    int eOn = 0, hOn = 0;
    double E[3] = {0, 0, 0}, eps = 1, eps0 = 1, deps = 0.5;
    double H[3] = {0, 0, 0}, chi = 1, mu0 = 1, dchi = 0.5;
    int hvarOn = 0;   // use the lattice's spatially varying H field
};

struct qm_fire {
    // energyMinimizerFIRE::initializeParameters defaults (energyMinimizerFIRE.cpp:24-43)
    double dt = 0.01, dtMin = 0.0001, alphaStart = 0.99, alpha = 0.99, dtMax = 0.1, dtInc = 1.05, dtDec = 0.95,
           alphaDec = 0.9, alphaMin = 0.0, fc = 1e-7;
    int nMin = 5, maxIter = 1000;
    int iterations = 0, nPositive = 0;
    double forceMax = 100.0, power = 0, scaling = 0;
    int converged = 0, nan = 0;
    int traceCap = 0;
    std::vector<double> trace;
};


namespace qmin {
int fail(int code, const std::string& msg);
// type upload, partial-sum storage and weights before a pass
int prepare(qm_lattice* L, const qm_ldg* f);
// F = forces(Q) into the current state set; gated passes are skipped on the
// device once the control block reports done
int force_pass(qm_lattice* L, const qm_ldg* f, bool gated);
// One Adam or Nesterov iteration's lattice passes (mode kAdamStep or
// kNesterovStep), skipped on the device once the control block reports done:
// the force of every site and the minimizer's update, written with that force
// to the other state set; the partial sums of |F|^2 go to the lattice's
// partials buffer, three doubles per slot, and *slots receives their count.
// The caller's finalize sums them, advances the minimizer and flips the set.
int step_pass(qm_lattice* L, const qm_ldg* f, int mode, const StepArgs& step, long long* slots);
// copies U of the current state set into the other one, so that passes that
// alternate the sets without touching U keep the FIRE velocity
int share_velocity(qm_lattice* L);
int sync_ctrl_to_host(qm_lattice* L);
// makes a pending FIRE velocity mix concrete in U of the current state set
int materialize(qm_lattice* L);
int push_ctrl(qm_lattice* L);
// synchronous copy ordered after the work queued on the lattice's stream
hipError_t copy_on(qm_lattice* L, void* dst, const void* src, size_t bytes, hipMemcpyKind kind);
int ensure_partials(qm_lattice* L, size_t count);
int64_t n_active(const qm_lattice* L);
// stencil variant of a force computer (kOneConstant .. kMultiL6)
int terms_of(const qm_ldg* f);
// tile shape of the fused pass for this lattice, and its tile columns
int active_tile(const qm_lattice* L, int mode);
long long tile_columns(const qm_lattice* L, int tile);
// kernel arguments of one lattice's passes (z chunk taken from L->zchunk)
template <typename R>
void fill_args(qm_lattice* L, const qm_ldg* f, MarchArgs<R>& ma, BoundaryArgs<R>& ba, const FireConst& fire,
               int tile);
FireConst fire_const(const qm_fire* m, const qm_lattice* L);
int boundary_blocks(long long nsites);
template <typename R>
inline StateSet<R> state_set(const qm_lattice* L, int s) {
    R* base = static_cast<R*>(L->dSet[s]);
    return StateSet<R>{base, base + 5 * L->n, base + 10 * L->n};
}
}  // namespace qmin

#define HIP_TRY(expr)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return qmin::fail(QM_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)
