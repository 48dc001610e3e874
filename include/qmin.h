/*
 * qmin: MI355X (gfx950) lattice Landau-de Gennes energy minimization.
 *
 * C ABI of the HIP library libqmin.so.  Plain pointers and sizes only; every
 * function returns QM_OK (0) or a negative QM_E* code, with a message from
 * qm_last_error().  Host arrays use the reference's layouts: Q, v, F are AoS
 * double[N][5] (Qxx, Qxy, Qxz, Qyy, Qyz), site = x + Lx*(y + Ly*z), types are
 * int32 (0 bulk, -1 next to an object, j >= 1 object j).  On the device the
 * library keeps structure-of-arrays planes in FP64 or FP32.
 *
 * Each entry point names the reference interface it replaces (open-Qmin,
 * paths relative to the reference checkout).  The C++ classes in
 * open-qmin_amd/host/ wrap these calls under the reference's class and method
 * names.
 */
#ifndef QMIN_H
#define QMIN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    QM_OK = 0,
    QM_EINVAL = -1,   /* bad argument */
    QM_EHIP = -2,     /* HIP runtime error */
    QM_ENOMEM = -3,   /* allocation failed */
    QM_EIO = -4,      /* file error */
    QM_ESTATE = -5,   /* call not valid in the current state */
    QM_ENAN = -6      /* the minimizer produced NaN forces */
};

enum { QM_FP64 = 0, QM_FP32 = 1 };                     /* device storage and stencil precision */
enum { QM_HOMEOTROPIC = 0, QM_DEGENERATE_PLANAR = 1 }; /* src/model/latticeBoundaries.h:8 */
enum { QM_SEMANTICS_REF_CPU = 0, QM_SEMANTICS_REF_GPU = 1 };
enum { QM_FIELD_Q = 0, QM_FIELD_V = 1, QM_FIELD_F = 2 };

typedef struct qm_lattice qm_lattice;
typedef struct qm_ldg qm_ldg;
typedef struct qm_fire qm_fire;
typedef struct qm_adam qm_adam;
typedef struct qm_nesterov qm_nesterov;
typedef struct qm_lolbfgs qm_lolbfgs;
typedef struct qm_ensemble qm_ensemble;

const char* qm_last_error(void);
int qm_device_count(int* count);

/* ---- lattice model --------------------------------------------------------
 * Replaces multirankQTensorLatticeModel(lx,ly,lz,...) on one rank without halos
 * (src/model/multirankQTensorLatticeModel.cpp:8-50) with its cubicLattice and
 * qTensorLatticeModel bases (src/model/cubicLattice.cpp:22-36,
 * src/model/qTensorLatticeModel.h:25-62).  One lattice is one periodic domain
 * resident in HBM. */
int qm_lattice_create(int lx, int ly, int lz, int precision, qm_lattice** out);
int qm_lattice_destroy(qm_lattice* lat);
int qm_lattice_dims(const qm_lattice* lat, int* lx, int* ly, int* lz, int* precision);
/* the HIP stream (hipStream_t) the lattice's work runs on from now on, e.g. the
 * caller's framework stream; NULL returns to the lattice's own stream.  It must
 * have the default priority (QM_EINVAL otherwise); work queued on the previous
 * stream completes first, and calls still return once their results are on the
 * host.  No counterpart in the reference, whose GPU path runs on the default
 * stream throughout. */
int qm_lattice_set_stream(qm_lattice* lat, void* stream);
/* z extent each work-group of the fused lattice pass marches through (>= 4);
 * 0 (default) sizes it so that the whole grid is resident at once.  A tuning
 * knob of this implementation; it has no counterpart in the reference. */
int qm_lattice_set_zchunk(qm_lattice* lat, int zchunk);
/* tile shape of the fused lattice pass: -1 (default) picks by precision, 0 is
 * 32x8 sites per work-group, 1 is 64x4, 2 is 32x16, 3 is 64x8.  A tuning knob of this
 * implementation. */
int qm_lattice_set_tile(qm_lattice* lat, int tile);
/* tile autotuner, the role of kernelTuner (src/utilities/kernelTuner.cpp:5-125):
 * times the fused FIRE pass with HIP events for every tile shape and z chunks
 * {automatic, 8, 16, 32, 64} on the lattice's current state (which it leaves
 * unchanged), keeps the fastest on the lattice and returns it (any output may be
 * NULL).  Explicit, so that a run's reduction order does not depend on timing. */
int qm_lattice_autotune(qm_lattice* lat, qm_ldg* f, int* tile, int* zchunk, double* ms);
/* converts the device state (Q, v, F) to the other storage precision, e.g. for an
 * FP64 polish after an FP32 minimization; minimizers bound to the lattice other
 * than FIRE restart their history.  Converting to FP64 writes back the object
 * sites' Q in FP64 (as last given by the host), so the fixed boundary values carry
 * no float rounding.  No reference counterpart (the reference is FP64
 * throughout). */
int qm_lattice_set_precision(qm_lattice* lat, int precision);
/* AoS host <-> device state; which = QM_FIELD_Q/V/F (simpleModel::returnPositions/
 * returnVelocities/returnForces, src/model/simpleModel.h:63-73) */
int qm_lattice_set_field(qm_lattice* lat, int which, const double* aos);
int qm_lattice_get_field(qm_lattice* lat, int which, double* aos);
/* qTensorLatticeModel::setNematicQTensorRandomly after noiseSource::setReproducibleSeed
 * (src/model/qTensorLatticeModel.cpp:134-177, src/utilities/noiseSource.cpp:38-47) */
int qm_lattice_random_init(qm_lattice* lat, int seed, double S0);
/* multirankQTensorLatticeModel::setUniformDirectors (multirankQTensorLatticeModel.cpp:79-88;
 * openQmin -z 2): qTensorFromDirector of one director on every site with t <= 0 */
int qm_lattice_uniform_init(qm_lattice* lat, double nx, double ny, double nz, double S0);
/* setNematicQTensorRandomly / setRandomDirectors with globallyAligned = true
 * (qTensorLatticeModel.cpp:134-177, multirankQTensorLatticeModel.cpp:52-77;
 * openQmin -z 1): one director from the first two draws of mt19937(seed)
 * (theta = acos(2u - 1), phi = 2 pi u) on every site with t <= 0 */
int qm_lattice_random_uniform_init(qm_lattice* lat, int seed, double S0);
/* multirankQTensorLatticeModel::setDirectorFromFunction
 * (multirankQTensorLatticeModel.cpp:90-110; openQmin -z 3): directors holds the
 * function's (nx, ny, nz, S0) for every site in site order (4 n doubles);
 * qTensorFromDirector is applied on the sites with t <= 0 */
int qm_lattice_director_init(qm_lattice* lat, const double* directors);
/* multirankSimulation::createMultirankBoundaryObject + cubicLattice::createBoundaryObject
 * (src/simulation/multirankSimulationBoundaries.cpp:512-543, src/model/cubicLattice.cpp:267-318);
 * xyz is [n][3] (wrapped periodically), qb is [n][5]; returns the object index via *obj */
int qm_lattice_add_object(qm_lattice* lat, const int32_t* xyz, int64_t n, const double* qb, int btype, double W,
                          double S0, int* obj);
/* multirankSimulation::createWall (multirankSimulationBoundaries.cpp:101-167) */
int qm_lattice_wall(qm_lattice* lat, int xyz, int plane, int btype, double W, double S0, int* obj);
/* multirankSimulation::createSphericalColloid (multirankSimulationBoundaries.cpp:169-215) */
int qm_lattice_sphere(qm_lattice* lat, double cx, double cy, double cz, double radius, int btype, double W,
                      double S0, int* obj);
/* multirankSimulation::createBoundaryFromFile (multirankSimulationBoundaries.cpp:41-96) */
int qm_lattice_boundary_file(qm_lattice* lat, const char* path);
int qm_lattice_get_types(const qm_lattice* lat, int32_t* types);
/* neighbour table in the reference order {x-,x+,y-,y+,z-,z+}, int32[N][6]
 * (cubicLattice::fillNeighborLists, src/model/cubicLattice.cpp:143-166) */
int qm_lattice_get_neighbors(const qm_lattice* lat, int32_t* nbrs);
/* updater::getNTotal (src/updaters/baseUpdater.cpp:10-28): number of sites with t <= 0 */
int qm_lattice_n_active(const qm_lattice* lat, int64_t* n);
int qm_lattice_n_boundary(const qm_lattice* lat, int64_t* n);
/* qTensorLatticeModel::computeDefectMeasures (src/model/qTensorLatticeModel.cpp:95-133):
 * kind 0 largest eigenvalue, 1 det Q, 2 TrQ2^3 - 54 det^2; out is double[N] */
int qm_lattice_defects(qm_lattice* lat, int kind, double* out);
/* multirankSimulation::saveState / loadState text format, one rank
 * (src/simulation/multirankSimulation.cpp:573-665): lines "x y z Qxx Qxy Qxz Qyy Qyz type Smax";
 * path is the full file name */
/* cubicLattice::boundarySites / surfaceSites (cubicLattice.h:78-81) of object
 * `object` (0-based, creation order): its sites in creation order (surface = 0)
 * or its surface sites in order of discovery (surface = 1); *count is always
 * set, sites (may be NULL) receives *count indices */
int qm_lattice_object_sites(const qm_lattice* lat, int object, int surface, int64_t* sites, int64_t* count);
/* cubicLattice::displaceBoundaryObject (cubicLattice.cpp:329-372): moves object
 * `object` and its surface sites, with their Q, by `magnitude` sites along
 * direction 0..5 (x-, x+, y-, y+, z-, z+) */
int qm_lattice_displace_object(qm_lattice* lat, int object, int direction, int magnitude);
int qm_lattice_save_state(qm_lattice* lat, const char* path);
/* multirankSimulation::saveState(fname, latticeSkip, defectType)
 * (multirankSimulation.cpp:630-665): only sites whose x, y and z are multiples
 * of latticeSkip; the last column is computeDefectMeasures(defectType) */
int qm_lattice_save_state_ex(qm_lattice* lat, const char* path, int latticeSkip, int defectType);
int qm_lattice_load_state(qm_lattice* lat, const char* path);
/* copy the device state Q of one lattice into another of the same size */
int qm_lattice_copy_q(qm_lattice* dst, const qm_lattice* src);
/* spatially varying H field of the lattice, AoS double[N][3] (NULL removes it);
 * the reference keeps it in landauDeGennesLC::spatiallyVaryingField
 * (landauDeGennesLC.h:93, .cpp:81-123) */
int qm_lattice_set_hfield(qm_lattice* lat, const double* h3);
/* the reference's field file ("x y z Hx Hy Hz" per line, landauDeGennesLC.cpp:69-123);
 * path is the full file name (the reference appends _x%iy%iz%i.txt per rank) */
int qm_lattice_load_hfield(qm_lattice* lat, const char* path);

/* ---- force computer: landauDeGennesLC (src/forces/landauDeGennesLC.h:12-173) */
int qm_ldg_create(qm_ldg** out);
int qm_ldg_destroy(qm_ldg* f);
/* setPhaseConstants(A,B,C) (landauDeGennesLC.h:34): energy A/2 TrQ^2 + B/3 TrQ^3 + C/4 (TrQ^2)^2 */
int qm_ldg_set_phase(qm_ldg* f, double A, double B, double C);
/* setElasticConstants(L1,L2,L3,L4,L6) + setNumberOfConstants (landauDeGennesLC.h:35-36);
 * multi = 0 selects the one-constant (L1) path */
int qm_ldg_set_elastic(qm_ldg* f, double L1, double L2, double L3, double L4, double L6, int multi);
/* setEField(E, eps, eps0, deltaEps) / setHField(H, chi, mu0, deltaChi) (landauDeGennesLC.h:76-91);
 * on = 0 removes the field */
int qm_ldg_set_efield(qm_ldg* f, int on, double Ex, double Ey, double Ez, double eps, double eps0, double deps);
int qm_ldg_set_hfield(qm_ldg* f, int on, double Hx, double Hy, double Hz, double chi, double mu0, double dchi);
/* QM_SEMANTICS_REF_GPU (default): one metric correction after all terms, as the
 * reference GPU path (landauDeGennesLC.cpp:128-134, 357-398).
 * QM_SEMANTICS_REF_CPU: the reference CPU path's two passes, which correct bulk
 * sites twice (multirankSimulation.cpp:462-466). */
int qm_ldg_set_semantics(qm_ldg* f, int semantics);
/* setSpatiallyVaryingField's constants (landauDeGennesLC.cpp:81-87): on = 1 adds
 * the lattice's varying H field (computeSpatiallyVaryingFieldCPU,
 * landauDeGennesLCOtherForces.cpp:254-283; energy landauDeGennesLC.cpp:256-264);
 * chi, mu0, deltaChi are shared with the uniform H field, as in the reference */
int qm_ldg_set_varying_hfield(qm_ldg* f, int on, double chi, double mu0, double dchi);
/* force::computeForces (src/forces/baseForce.h:15-77): F = forces(Q) into the lattice */
int qm_ldg_compute_forces(qm_ldg* f, qm_lattice* lat);
/* landauDeGennesLC::computeEnergy (landauDeGennesLC.cpp:204-353): total energy and
 * components {phase, distortion, anchoring, E field, H field}; the reference's
 * computePotentialEnergy reports total / n_active (multirankSimulation.cpp:475-490) */
/* landauDeGennesLC::computeObjectForces with computeStressTensors
 * (landauDeGennesLCOtherForces.cpp:147-251): force[3] on object `object`
 * (0-based) from the stress tensor over its surface sites; this call's force
 * (the reference adds it to lattice->boundaryForce).  One-constant only, as in
 * the reference (QM_EINVAL otherwise). */
int qm_ldg_object_force(qm_ldg* f, qm_lattice* lat, int object, double* force);
int qm_ldg_compute_energy(qm_ldg* f, qm_lattice* lat, double* total, double* components5);
/* landauDeGennesLC::energyDensity (landauDeGennesLC.h:96) as computeEnergy fills
 * it (landauDeGennesLC.cpp:204-353): density[n] in site order, 0 on object sites,
 * else the site's phase, field, anchoring and distortion energy; the values that
 * qm_ldg_compute_energy sums */
int qm_ldg_energy_density(qm_ldg* f, qm_lattice* lat, double* density);

/* ---- FIRE: energyMinimizerFIRE (src/updaters/energyMinimizerFIRE.h:20-110) */
typedef struct {
    int iterations;      /* cumulative, as updater::iterations */
    int nPositive;       /* NSinceNegativePower */
    int converged;       /* last minimize stopped on the force tolerance */
    int nanDetected;     /* forceMax became NaN (the reference would report convergence) */
    double forceMax;     /* sqrt(sum dotVec(F,F)) / nTotal (energyMinimizerFIRE.cpp:110) */
    double power, alpha, deltaT, scaling;
} qm_fire_state;

int qm_fire_create(qm_fire** out);
int qm_fire_destroy(qm_fire* m);
/* setFIREParameters(dt, alphaStart, dtMax, dtInc, dtDec, alphaDec, nMin, forceCutoff, alphaMin)
 * (energyMinimizerFIRE.cpp:172-183); also resets alpha and dtMin = dt/100 */
int qm_fire_set_parameters(qm_fire* m, double dt, double alphaStart, double dtMax, double dtInc, double dtDec,
                           double alphaDec, int nMin, double forceCutoff, double alphaMin);
/* updater::setMaximumIterations (baseUpdater.h:77) */
int qm_fire_set_max_iterations(qm_fire* m, int maxIterations);
/* energyMinimizerFIRE::setForceCutoff (energyMinimizerFIRE.h:29): the other
 * FIRE state (dt, alpha, N+) is kept, e.g. for an FP64 polish to a tighter cutoff */
int qm_fire_set_force_cutoff(qm_fire* m, double forceCutoff);
/* energyMinimizerFIRE::minimize (energyMinimizerFIRE.cpp:151-170) on the device:
 * forces, then fused velocity-Verlet + force + FIRE iterations until the tolerance
 * or the iteration cap.  The velocity field persists in the lattice. */
int qm_fire_minimize(qm_fire* m, qm_ldg* f, qm_lattice* lat);
/* run exactly n iterations (no tolerance test): throughput measurement */
int qm_fire_run_iterations(qm_fire* m, qm_ldg* f, qm_lattice* lat, int n);
/* qm_fire_run_iterations without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The passes'
 * work-group partial sums go to a fixed 1.5 MiB array in the library's device
 * image (65536 slots, one per work-group of a pass: QM_EINVAL on a lattice
 * that needs more), one copy per GPU, shared by every call on that GPU: calls
 * that overlap on one GPU race.  No trace is recorded.  Results equal
 * qm_fire_run_iterations'. */
int qm_fire_run_iterations_noalloc(qm_fire* m, qm_ldg* f, qm_lattice* lat, int n);
int qm_fire_get_state(const qm_fire* m, qm_fire_state* out);
/* per-iteration trace {forceMax, power, alpha, deltaT, scaling, nPositive}: capacity in
 * iterations (0 disables); qm_fire_get_trace copies up to cap rows, returns the count in *n */
int qm_fire_set_trace(qm_fire* m, int capacity);
int qm_fire_get_trace(qm_fire* m, double* rows6, int cap, int* n);

/* ---- Adam: energyMinimizerAdam (src/updaters/energyMinimizerAdam.h:7-50) */
int qm_adam_create(qm_adam** out);
int qm_adam_destroy(qm_adam* m);
/* setAdamParameters(b1, b2, eps, dt, fc) (energyMinimizerAdam.h:16-25) */
int qm_adam_set_parameters(qm_adam* m, double beta1, double beta2, double eps, double dt, double forceCutoff);
int qm_adam_set_max_iterations(qm_adam* m, int maxIterations);
/* energyMinimizerAdam::minimize (energyMinimizerAdam.cpp:56-75); forceMax = sqrt(F.F)/N */
int qm_adam_minimize(qm_adam* m, qm_ldg* f, qm_lattice* lat);
int qm_adam_get_state(const qm_adam* m, int* iterations, double* forceMax);

/* ---- Nesterov: energyMinimizerNesterovAG (src/updaters/energyMinimizerNesterovAG.h:8-61) */
int qm_nesterov_create(qm_nesterov** out);
int qm_nesterov_destroy(qm_nesterov* m);
/* setNesterovAGParameters(dt, mu, fc) + scheduledMomentum (energyMinimizerNesterovAG.h:22-31) */
int qm_nesterov_set_parameters(qm_nesterov* m, double dt, double mu, double forceCutoff, int scheduledMomentum);
int qm_nesterov_set_max_iterations(qm_nesterov* m, int maxIterations);
/* energyMinimizerNesterovAG::minimize (energyMinimizerNesterovAG.cpp:94-120) */
int qm_nesterov_minimize(qm_nesterov* m, qm_ldg* f, qm_lattice* lat);
int qm_nesterov_get_state(const qm_nesterov* m, int* iterations, double* forceMax);

/* ---- online L-BFGS: energyMinimizerLoLBFGS (src/updaters/energyMinimizerLoLBFGS.h:13-110) */
int qm_lolbfgs_create(qm_lolbfgs** out);
int qm_lolbfgs_destroy(qm_lolbfgs* m);
/* setLoLBFGSParameters(m, dt, c, fc, tau) + scheduledMomentum (energyMinimizerLoLBFGS.h:29-45);
 * the memory never shrinks below 5, as in the reference */
int qm_lolbfgs_set_parameters(qm_lolbfgs* m, int memory, double dt, double c, double forceCutoff, double tau,
                              int scheduledMomentum);
int qm_lolbfgs_set_max_iterations(qm_lolbfgs* m, int maxIterations);
/* energyMinimizerLoLBFGS::minimize (energyMinimizerLoLBFGS.cpp:195-226) */
int qm_lolbfgs_minimize(qm_lolbfgs* m, qm_ldg* f, qm_lattice* lat);
int qm_lolbfgs_get_state(const qm_lolbfgs* m, int* iterations, double* forceMax);

/* per-iteration forceMax traces of the three minimizers above (the quantity their
 * minimize() loops print every 1000 steps): capacity in iterations (0 disables);
 * get copies up to cap values and returns the count in *n */
int qm_adam_set_trace(qm_adam* m, int capacity);
int qm_adam_get_trace(const qm_adam* m, double* forceMax, int cap, int* n);
int qm_nesterov_set_trace(qm_nesterov* m, int capacity);
int qm_nesterov_get_trace(const qm_nesterov* m, double* forceMax, int cap, int* n);
int qm_lolbfgs_set_trace(qm_lolbfgs* m, int capacity);
int qm_lolbfgs_get_trace(const qm_lolbfgs* m, double* forceMax, int cap, int* n);

/* ---- ensembles: many lattices minimized together --------------------------
 * No reference counterpart: the reference runs one lattice per process.  The
 * lattices (any sizes, one precision) stay owned by the caller and must outlive
 * the ensemble.  Each FIRE kernel is one launch over every lattice, every
 * lattice keeps its own FIRE scalars and stops on its own; batches of
 * graphIterations iterations are captured once in a hipGraph and replayed. */
int qm_ensemble_create(qm_lattice* const* lattices, int count, qm_ensemble** out);
int qm_ensemble_destroy(qm_ensemble* e);
/* energyMinimizerFIRE::minimize on every lattice (fresh FIRE scalars from params'
 * parameters, cutoff and iteration cap); per-lattice iterations and forceMax out
 * (arrays of count, may be NULL) */
int qm_ensemble_fire_minimize(qm_ensemble* e, const qm_fire* params, qm_ldg* f, int graphIterations,
                              int* iterations, double* forceMax);
/* tile autotuner of the batched pass, the role of kernelTuner
 * (src/utilities/kernelTuner.cpp:5-125) for ensembles: times the batched fused
 * FIRE pass with HIP events for every tile shape, and for QM_TILE_PER_LATTICE
 * (each lattice the 8-row tile, 32 or 64 sites wide, that pads its x edge less;
 * one launch per shape), with z chunks {8, 16, 32, 64} on the lattices' current
 * states (which it leaves unchanged) and keeps the fastest for later
 * qm_ensemble_fire_minimize calls (outputs may be NULL). */
#define QM_TILE_PER_LATTICE (-2)
/* the batched pass's tile for later qm_ensemble_fire_minimize calls: -1 the first
 * lattice's default, QM_TILE_PER_LATTICE, or a fused-pass tile 0..3 (as
 * qm_lattice_set_tile); zchunk 0 (whole columns up to 32 planes) or at least 4.
 * The setting the autotuner leaves, made by hand.  A tuning knob of this
 * implementation. */
int qm_ensemble_set_tile(qm_ensemble* e, int tile, int zchunk);
int qm_ensemble_autotune(qm_ensemble* e, qm_ldg* f, int* tile, int* zchunk, double* ms);

/* ---- learned warm start -----------------------------------------------------
 * No reference counterpart.  Building blocks of a small 3-D convolutional
 * network that predicts an initial Q from the geometry; random-start FIRE stays
 * the fallback.  Activations are channels-last [n][C] BF16 (raw 16-bit words),
 * C in {16, 32, 64}; weights are [27][cout][cin] (tap = (dz+1)*9 + (dy+1)*3 +
 * (dx+1), offsets times the dilation, periodic); all pointers are device
 * memory, work is queued on `stream` (a hipStream_t, NULL = default). */
/* y[s][co] = bias[co] + sum_{tap,ci} x'[s + dil*off(tap)][ci] wt[tap][co][ci],
 * x' = max(x, 0) if reluIn; y zeroed where mask[s][co] <= 0 (mask may be NULL);
 * output to ybf16 (BF16) or yf32 (FP32), exactly one non-NULL */
int qm_ws_conv_forward(const void* x, int cin, const void* wt, const float* bias, int cout, const void* mask,
                       void* ybf16, float* yf32, int lx, int ly, int lz, int dil, int reluIn, void* stream);
/* dw[tap][co][ci] = sum_s g[s][co] x'[s + dil*off(tap)][ci]; part holds
 * ceil(n/chunk) * 27 * cout * cin floats of per-chunk sums (chunk a multiple of 16) */
int qm_ws_conv_wgrad(const void* g, int cout, const void* x, int cin, int reluIn, int lx, int ly, int lz, int dil,
                     float* part, long long chunk, float* dw, void* stream);
/* db[co] = sum_s g[s][co]; part holds blocks * cout floats of per-block sums,
 * added in a fixed order */
int qm_ws_bias_grad(const void* g, long long n, int cout, float* part, int blocks, float* db, void* stream);
/* Adam step `step` (1-based) on FP32 weights; refreshes the BF16 copy and, when
 * wbd16 is not NULL, the mirrored transpose wbd[26-tap][ci][co] used for input gradients */
int qm_ws_adam(float* w, const float* g, float* m, float* v, long long count, float lr, float b1, float b2, float eps,
               int step, void* wbf16, void* wbd16, int cout, int cin, void* stream);
int qm_ws_weights_to_bf16(const float* w, long long count, void* wbf16, void* wbd16, int cout, int cin, void* stream);
/* g = 2 (y + offset - target) scale on the first five of cy channels at sites
 * with type <= 0, zero elsewhere (offset5: host double[5], may be NULL);
 * lossPart[blocks] gets per-block sums of squared differences */
int qm_ws_mse_grad(const float* y, int cy, const float* target, const void* types, long long n, float scale,
                   const double* offset5, void* g, double* lossPart, int blocks, void* stream);
/* energy term of the training loss: with the prediction in the lattice
 * (qm_ws_apply), computes the forces F of force computer f there and adds
 * -scale F to the first five of cy channels of g (BF16, device) at sites with
 * type <= 0; F is minus the energy gradient shaped by the metric.  The
 * caller's stream is drained first; the call returns when g is updated. */
int qm_ws_energy_grad(qm_ldg* f, qm_lattice* lat, float scale, void* g, int cy, void* stream);
/* [n][16] BF16 input features of the lattice's geometry and far-field director d */
int qm_ws_features(qm_lattice* lat, double dx, double dy, double dz, void* out);
int qm_ws_types(qm_lattice* lat, void* out);
/* write the first five of cy FP32 channels of y plus offset5 (host double[5],
 * may be NULL) into Q of the lattice's liquid-crystal sites; velocities cleared */
int qm_ws_apply(qm_lattice* lat, const float* y, int cy, const double* offset5);
int qm_ws_extract_q(qm_lattice* lat, float* out);

/* ---- kernel timing (HIP events on the library's stream) -------------------
 * When enabled, every launch of the two FIRE kernels is bracketed by events;
 * qm_timing_get reports launches and total milliseconds per kernel:
 * kernel 0 = fused lattice pass, 1 = boundary-site pass (with the FIRE scalar step). */
int qm_timing_enable(int on);
int qm_timing_reset(void);
int qm_timing_get(int kernel, int64_t* launches, double* total_ms);

#ifdef __cplusplus
}
#endif
#endif /* QMIN_H */
