// CPU oracle for the lattice Landau-de Gennes minimization path.
//
// TEST INFRASTRUCTURE.  This is a plain restatement, in C++, of the reference
// CPU algorithm (open-Qmin, src/forces, src/updaters, src/model,
// src/simulation); every function cites the reference lines it follows.  It
// is the checker for the HIP implementation: only tests/, the smoke check and
// bench.py's CPU-baseline leg load it.  Nothing in the product links to it.
//
// Layout: Q, v, F are AoS double[N][5] in the reference's site order
// site = x + Lx*(y + Ly*z) (inc/indexer.h:92-100); types are int32 with the
// reference encoding (0 bulk, -1 next to an object, j>=1 object j).
//
// Parity pins (see tests/): per-site stencils and anchoring against the
// compiled reference (oracle/_ref), the minimizer loops against the compiled
// reference minimizers driven with this oracle's forces, lattice/types/init
// against the compiled reference model, energy and stress polynomials against
// the reference's expressions read from its source, and closed forms and
// invariances (plane wave, lattice rotations, locality, first-derivative rules
// on a quadratic field) for the rest.
// orc_object_force (the reference's force class that computes it needs MPI and
// is not compiled here) is pinned in closed form on flat walls, at corners and
// between two walls one site apart (objects on both sides of one axis).
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_system orc_system;

// --- lattice, objects, initial state
orc_system* orc_create(int lx, int ly, int lz);
void orc_destroy(orc_system* s);
void orc_set_threads(orc_system* s, int nthreads);
void orc_neighbors(const orc_system* s, int32_t* out6n);
void orc_get_types(const orc_system* s, int32_t* out);
void orc_set_types(orc_system* s, const int32_t* in);
void orc_get_q(const orc_system* s, double* out);
void orc_set_q(orc_system* s, const double* in);
void orc_random_init(orc_system* s, int seed, double S0);
void orc_random_uniform_init(orc_system* s, int seed, double S0);
void orc_director_init(orc_system* s, const double* n4);
int64_t orc_object_sites(orc_system* s, int obj, int surface, int64_t* out);
int orc_displace_object(orc_system* s, int obj, int direction, int magnitude);
int orc_object_force(orc_system* s, int obj, double* out);
// btype 0 homeotropic, 1 degenerate planar; sites are (x,y,z) triples, qb is [n][5]
int orc_add_object(orc_system* s, const int32_t* xyz, int n, const double* qb, int btype, double W, double S0);
int orc_wall(orc_system* s, int xyz, int plane, int btype, double W, double S0);
int orc_sphere(orc_system* s, double cx, double cy, double cz, double radius, int btype, double W, double S0);
int64_t orc_n_active(const orc_system* s);
int orc_n_objects(const orc_system* s);

// --- force / energy parameters
// semantics: 0 = reference CPU path (two passes, metric applied twice to bulk
// sites), 1 = reference GPU path (single metric correction)
void orc_set_params(orc_system* s, double A, double B, double C, double L1, double L2, double L3,
                    double L4, double L6, int multiConstant, int semantics);
void orc_set_efield(orc_system* s, int on, double Ex, double Ey, double Ez, double eps, double eps0, double deps);
void orc_set_hfield(orc_system* s, int on, double Hx, double Hy, double Hz, double chi, double mu0, double dchi);
void orc_set_hfield_varying(orc_system* s, int on, const double* h3, double chi, double mu0, double dchi);
// stencil tables: n entries, each (term slot 0..3 = L2,L3,L4,L6; case slot 0..27;
// component; coefficient; input a; input b or -1)
void orc_load_tables(int n, const int32_t* term, const int32_t* cslot, const int32_t* comp, const double* coef,
                     const int32_t* a, const int32_t* b);

void orc_forces(orc_system* s, const double* q, double* f);
void orc_first_derivatives(orc_system* s, const double* q, double* d15);
// returns the total energy; comps = {phase, distortion, anchoring, E, H}
double orc_energy(orc_system* s, const double* q, double* comps);
// per-site energy density (landauDeGennesLC::energyDensity), n values; 0 on object sites
void orc_energy_density(orc_system* s, const double* q, double* out);
// one-constant stress tensor of derivatives fd[15] (computeStressTensors), row-major m[9], e on the diagonal
void orc_stress_tensor(const double* fd, double L1, double e, double* m);
// force callback usable by the reference harness (ctx = orc_system*)
void orc_force_cb(void* ctx, int n, const double* q, double* f);
double orc_energy_cb(void* ctx, int n, const double* q);

// --- minimizers, all on the system's Q (and v for FIRE), returning iterations done
// FIRE trace per iteration: {forceMax, Power, alpha, dt, scaling, NSinceNegativePower}
int orc_fire(orc_system* s, double dt, double alphaStart, double dtMax, double dtInc, double dtDec,
             double alphaDec, int nMin, double fc, double alphaMin, int maxIter, double* trace);
void orc_get_v(const orc_system* s, double* out);
void orc_set_v(orc_system* s, const double* in);
void orc_get_f(const orc_system* s, double* out);
int orc_adam(orc_system* s, double b1, double b2, double eps, double dt, double fc, int maxIter, double* trace);
int orc_nesterov(orc_system* s, double dt, double mu, double fc, int scheduled, int maxIter, double* trace);
int orc_lolbfgs(orc_system* s, int m, double dt, double c, double fc, double tau, int scheduled, int maxIter,
                double* trace);

// --- post-processing: defectType 0 largest eigenvalue, 1 det Q, 2 TrQ2^3 - 54 det^2
void orc_defects(const orc_system* s, int defectType, double* out);
void orc_eigenvalues(const double* q, double* ev);

// --- per-site pieces (unit checks)
// distortion energy densities {L1, L2, L3, L4, L6 parts} per unit constant
void orc_distortion_density(const double* q, const double* dX, const double* dY, const double* dZ, double* e5);
double orc_anchoring_energy(int btype, double P1, double P2, const double* q, const double* qb);
void orc_anchoring_force(int btype, double P1, double P2, const double* q, const double* qb, double* f);
void orc_phase_parts(const double* q, double* g15, double* tr3);
void orc_l1_site(double L1, int boundary, const double* q, const double* nb, const int32_t* t, double* out);

#ifdef __cplusplus
}
#endif
