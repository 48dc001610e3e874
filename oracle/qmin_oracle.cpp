// CPU oracle: C++ restatement of the reference CPU path (TEST INFRASTRUCTURE;
// see qmin_oracle.h).  Each function names the reference lines it follows.
// Compiled with -ffp-contract=off so that expression order, not FMA
// contraction, decides rounding (as in the reference's own host build).
#include "qmin_oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "stencil_eval.h"

namespace {

const double kPi = 3.14159265358979323846;  // inc/std_include.h:39
oracle::StencilTables g_tables;

struct BoundaryObj {
    int type;  // 0 homeotropic, 1 degenerate planar (src/model/latticeBoundaries.h:8-40)
    double P1, P2;
};

typedef double Vec5[5];

}  // namespace

struct orc_system {
    int lx, ly, lz;
    int64_t n;
    std::vector<int32_t> types, nbr;
    std::vector<double> q, v, f;
    std::vector<BoundaryObj> objects;
    // per object, as cubicLattice keeps them (cubicLattice.h:78-86): its sites in
    // creation order, and its surface sites in order of discovery
    std::vector<std::vector<int64_t>> objSites, objSurface;
    double A = -1, B = 0, C = 0, L1 = 0, L2 = 0, L3 = 0, L4 = 0, L6 = 0;
    int multi = 0, semantics = 0;
    int eOn = 0, hOn = 0;
    double E[3] = {0, 0, 0}, eps = 1, eps0 = 1, deps = 0;
    double H[3] = {0, 0, 0}, chi = 1, mu0 = 1, dchi = 0;
    int hvarOn = 0;
    std::vector<double> hvar;   // spatially varying H, [n][3]
    int nthreads = 1;
    std::vector<double> deriv;  // forceCalculationAssist, [n][15]
};

namespace {

inline int64_t siteIndex(const orc_system* s, int x, int y, int z) {
    return (int64_t)x + (int64_t)s->lx * ((int64_t)y + (int64_t)s->ly * z);
}

// Periodic neighbour table {x-, x+, y-, y+, z-, z+}: cubicLattice::getNeighbors
// (src/model/cubicLattice.cpp:191-243) and, for one rank without halos,
// multirankQTensorLatticeModel::positionToIndex (multirankQTensorLatticeModel.cpp:281-295).
void buildNeighbors(orc_system* s) {
    s->nbr.assign(6 * s->n, 0);
    for (int z = 0; z < s->lz; ++z)
        for (int y = 0; y < s->ly; ++y)
            for (int x = 0; x < s->lx; ++x) {
                int64_t i = siteIndex(s, x, y, z);
                int xm = x == 0 ? s->lx - 1 : x - 1, xp = x == s->lx - 1 ? 0 : x + 1;
                int ym = y == 0 ? s->ly - 1 : y - 1, yp = y == s->ly - 1 ? 0 : y + 1;
                int zm = z == 0 ? s->lz - 1 : z - 1, zp = z == s->lz - 1 ? 0 : z + 1;
                int32_t* o = &s->nbr[6 * i];
                o[0] = (int32_t)siteIndex(s, xm, y, z);
                o[1] = (int32_t)siteIndex(s, xp, y, z);
                o[2] = (int32_t)siteIndex(s, x, ym, z);
                o[3] = (int32_t)siteIndex(s, x, yp, z);
                o[4] = (int32_t)siteIndex(s, x, y, zm);
                o[5] = (int32_t)siteIndex(s, x, y, zp);
            }
}

// qTensorFromDirector, inc/qTensorFunctions.h:34-51 (same operation order)
void qFromDirector(double nx, double ny, double nz, double S0, double* q) {
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

// Tr(Q^2) (qTensorFunctions.h:54-57), in the QT inner product form
inline double trQ2(const double* q) {
    return 2.0 * (q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3] + q[4] * q[4] + q[0] * q[3]);
}

// 3x3 form of q (qTensorFunctions.h:8-14)
inline void full(const double* q, double M[3][3]) {
    M[0][0] = q[0]; M[0][1] = q[1]; M[0][2] = q[2];
    M[1][0] = q[1]; M[1][1] = q[3]; M[1][2] = q[4];
    M[2][0] = q[2]; M[2][1] = q[4]; M[2][2] = -q[0] - q[3];
}

inline double det3(const double M[3][3]) {
    return M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) - M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
           M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
}

// Tr(Q^3) = 3 det Q for traceless symmetric Q (qTensorFunctions.h:60-63)
inline double trQ3(const double* q) {
    double M[3][3];
    full(q, M);
    return 3.0 * det3(M);
}

// Gradient of a scalar with respect to q0..q4 from its gradient G with respect
// to the symmetric matrix entries, using Qzz = -q0-q3.
inline void matGradToQ(const double G[3][3], double* g) {
    g[0] = G[0][0] - G[2][2];
    g[1] = G[0][1] + G[1][0];
    g[2] = G[0][2] + G[2][0];
    g[3] = G[1][1] - G[2][2];
    g[4] = G[1][2] + G[2][1];
}

// d TrQ2 / dq (qTensorFunctions.h:73-82)
inline void dTrQ2(const double* q, double* g) {
    g[0] = 2.0 * (2.0 * q[0] + q[3]);
    g[1] = 4.0 * q[1];
    g[2] = 4.0 * q[2];
    g[3] = 2.0 * (q[0] + 2.0 * q[3]);
    g[4] = 4.0 * q[4];
}

// d TrQ3 / dq = 3 d det/dq, det's matrix gradient is the cofactor matrix
// (qTensorFunctions.h:85-94)
inline void dTrQ3(const double* q, double* g) {
    double M[3][3], Cf[3][3];
    full(q, M);
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            int a1 = (a + 1) % 3, a2 = (a + 2) % 3, b1 = (b + 1) % 3, b2 = (b + 2) % 3;
            Cf[a][b] = 3.0 * (M[a1][b1] * M[a2][b2] - M[a1][b2] * M[a2][b1]);
        }
    matGradToQ(Cf, g);
}

// d (TrQ2)^2 / dq (qTensorFunctions.h:97-107)
inline void dTrQ2Sq(const double* q, double* g) {
    double t = trQ2(q);
    dTrQ2(q, g);
    for (int i = 0; i < 5; ++i) g[i] *= 2.0 * t;
}

// Phase part of the force: F -= a dTrQ2 + b dTrQ3 + c d(TrQ2)^2 with
// a = A/2, b = B/3, c = C/4 (landauDeGennesLCDistortionTerms.cpp:24-38)
inline void phaseForce(const orc_system* s, const double* q, double* force) {
    double a = 0.5 * s->A, b = s->B / 3.0, c = 0.25 * s->C;
    double g2[5], g3[5], g4[5];
    dTrQ2(q, g2);
    dTrQ3(q, g3);
    dTrQ2Sq(q, g4);
    for (int i = 0; i < 5; ++i) force[i] -= a * g2[i];
    for (int i = 0; i < 5; ++i) force[i] -= b * g3[i];
    for (int i = 0; i < 5; ++i) force[i] -= c * g4[i];
}

// Shape of the L1 operator in the Q basis (lcForces.h:31-36, 57-62):
// s0 <- 2 s0 + s3, s3 <- s0 + 2 s3, off-diagonals doubled
inline void l1Shape(double* s) {
    double d = s[0] + s[3];
    s[0] += d;
    s[1] *= 2.0;
    s[2] *= 2.0;
    s[3] += d;
    s[4] *= 2.0;
}

// bulkL1Force, lcForces.h:26-37: L1 (6 q - sum of neighbours), shaped
inline void bulkL1(double L1, const double* q, const double* nb, double* out) {
    for (int i = 0; i < 5; ++i) {
        double t = 6.0 * q[i] - nb[i] - nb[5 + i] - nb[10 + i] - nb[15 + i] - nb[20 + i] - nb[25 + i];
        out[i] = L1 * t;
    }
    l1Shape(out);
}

// boundaryL1Force, lcForces.h:39-63: only object-facing directions contribute,
// each as -(opposite neighbour - q)
inline void boundaryL1(double L1, const double* q, const double* nb, const int* t, double* out) {
    static const int opposite[6] = {1, 0, 3, 2, 5, 4};
    for (int i = 0; i < 5; ++i) out[i] = 0.0;
    for (int d = 0; d < 6; ++d)
        if (t[d] > 0)
            for (int i = 0; i < 5; ++i) out[i] -= (nb[5 * opposite[d] + i] - q[i]);
    for (int i = 0; i < 5; ++i) out[i] *= L1;
    l1Shape(out);
}

// getBoundaryCase, lcForces.h:13-25 (the down neighbour wins)
inline int boundaryCase(const int* t) {
    int xx = 0, yy = 0, zz = 0;
    if (t[1] > 0) xx = 1;
    if (t[0] > 0) xx = 2;
    if (t[3] > 0) yy = 1;
    if (t[2] > 0) yy = 2;
    if (t[5] > 0) zz = 1;
    if (t[4] > 0) zz = 2;
    return xx + 4 * yy + 16 * zz;
}

// computeBoundaryForce, landauDeGennesLCBoundary.h:26-74.
// homeotropic: E = W Tr((Q-Q^B)^2); planar (Fournier-Galatola, W1 = W2 = P1,
// nu = Q^B[0..2]): E = W1 Tr((Qt - P Qt P)^2) + W1 (Tr(Qt^2) - (3 S0/2)^2)^2,
// Qt = Q + S0/2 I, P = I - nu nu.  The force is -dE/dq.
void anchoringForce(const BoundaryObj& o, const double* q, const double* qb, double* f) {
    if (o.type == 0) {
        double d[5], g[5];
        for (int i = 0; i < 5; ++i) d[i] = q[i] - qb[i];
        dTrQ2(d, g);
        for (int i = 0; i < 5; ++i) f[i] = -o.P1 * g[i];
        return;
    }
    double W = o.P1, S0 = o.P2;
    double nu[3] = {qb[0], qb[1], qb[2]};
    double Qt[3][3], P[3][3], D[3][3], PDP[3][3], G[3][3], T[3][3];
    full(q, Qt);
    for (int a = 0; a < 3; ++a) Qt[a][a] += 0.5 * S0;
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) P[a][b] = (a == b ? 1.0 : 0.0) - nu[a] * nu[b];
    // D = Qt - P Qt P
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            double acc = 0;
            for (int c = 0; c < 3; ++c) acc += Qt[a][c] * P[c][b];
            T[a][b] = acc;
        }
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            double acc = 0;
            for (int c = 0; c < 3; ++c) acc += P[a][c] * T[c][b];
            D[a][b] = Qt[a][b] - acc;
        }
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            double acc = 0;
            for (int c = 0; c < 3; ++c) acc += D[a][c] * P[c][b];
            T[a][b] = acc;
        }
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            double acc = 0;
            for (int c = 0; c < 3; ++c) acc += P[a][c] * T[c][b];
            PDP[a][b] = acc;
        }
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) G[a][b] = 2.0 * W * (D[a][b] - PDP[a][b]);
    double g1[5], g2[5];
    matGradToQ(G, g1);
    dTrQ2(q, g2);
    double w2 = 2.0 * W * (trQ2(q) - 1.5 * S0 * S0);
    for (int i = 0; i < 5; ++i) f[i] = -(g1[i] + w2 * g2[i]);
}

// computeBoundaryEnergy, landauDeGennesLCBoundary.h:77-124 (same energies as above)
double anchoringEnergy(const BoundaryObj& o, const double* q, const double* qb) {
    if (o.type == 0) {
        double d[5];
        for (int i = 0; i < 5; ++i) d[i] = q[i] - qb[i];
        return o.P1 * trQ2(d);
    }
    double W = o.P1, S0 = o.P2;
    double nu[3] = {qb[0], qb[1], qb[2]};
    double Qt[3][3], P[3][3], T[3][3], D[3][3];
    full(q, Qt);
    for (int a = 0; a < 3; ++a) Qt[a][a] += 0.5 * S0;
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) P[a][b] = (a == b ? 1.0 : 0.0) - nu[a] * nu[b];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            double acc = 0;
            for (int c = 0; c < 3; ++c) acc += Qt[a][c] * P[c][b];
            T[a][b] = acc;
        }
    double e1 = 0, tq = 0;
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            double acc = 0;
            for (int c = 0; c < 3; ++c) acc += P[a][c] * T[c][b];
            D[a][b] = Qt[a][b] - acc;
        }
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            e1 += D[a][b] * D[b][a];
            tq += Qt[a][b] * Qt[b][a];
        }
    double e2 = tq - 2.25 * S0 * S0;
    return W * e1 + W * e2 * e2;
}

// Uniform field force on t <= 0 sites: F -= fieldForce with
// fieldForce = (-P(Fx^2-Fz^2)/2, -P FxFy, -P FxFz, -P(Fy^2-Fz^2)/2, -P FyFz),
// P = anisotropy * permeability (landauDeGennesLCOtherForces.cpp:285-312)
inline void fieldForce(const double* F, double anis, double perm, double* out) {
    double p = anis * perm;
    out[0] = -0.5 * p * (F[0] * F[0] - F[2] * F[2]);
    out[1] = -p * F[0] * F[1];
    out[2] = -p * F[0] * F[2];
    out[3] = -0.5 * p * (F[1] * F[1] - F[2] * F[2]);
    out[4] = -p * F[1] * F[2];
}

// correctForceFromMetric, landauDeGennesLC.cpp:137-160
inline void metric(double* f) {
    double xx = f[0], yy = f[3], t = 2. / 3.;
    f[0] = t * (2. * xx - yy);
    f[3] = t * (2. * yy - xx);
}

// computeFirstDerivatives, landauDeGennesLCOtherForces.cpp:15-145: central
// differences for t == 0; for t < 0 per axis central if neither neighbour is an
// object, else one-sided (an object above takes precedence); objects untouched
// (zero from allocation, gpuarray.cpp:343-360).
void firstDerivatives(orc_system* s, const double* q) {
    s->deriv.assign(15 * s->n, 0.0);
#pragma omp parallel for num_threads(s->nthreads) schedule(static)
    for (int64_t i = 0; i < s->n; ++i) {
        int t = s->types[i];
        if (t > 0) continue;
        const int32_t* nb = &s->nbr[6 * i];
        const double* qc = q + 5 * i;
        double* d = &s->deriv[15 * i];
        for (int ax = 0; ax < 3; ++ax) {
            int64_t dn = nb[2 * ax], up = nb[2 * ax + 1];
            const double* qd = q + 5 * dn;
            const double* qu = q + 5 * up;
            if (t == 0 || (s->types[dn] <= 0 && s->types[up] <= 0)) {
                for (int c = 0; c < 5; ++c) d[5 * ax + c] = 0.5 * (qu[c] - qd[c]);
            } else if (s->types[up] > 0) {
                for (int c = 0; c < 5; ++c) d[5 * ax + c] = (qc[c] - qd[c]);
            } else {
                for (int c = 0; c < 5; ++c) d[5 * ax + c] = (qu[c] - qc[c]);
            }
        }
    }
}

// Raw distortion ("spatial") term of one site for the multi-constant path:
// computeAllDistortionTermsBulkCPU / BoundaryCPU
// (landauDeGennesLCDistortionTerms.cpp:113-330).
void multiConstantSpatial(const orc_system* s, int64_t i, const double* q, double* spatial) {
    const int32_t* nb = &s->nbr[6 * i];
    double x[125];
    int t[6];
    for (int c = 0; c < 5; ++c) x[c] = q[5 * i + c];
    for (int d = 0; d < 6; ++d) {
        t[d] = s->types[nb[d]];
        for (int c = 0; c < 5; ++c) x[5 + 5 * d + c] = q[5 * (int64_t)nb[d] + c];
        for (int c = 0; c < 15; ++c) x[35 + 15 * d + c] = s->deriv[15 * (int64_t)nb[d] + c];
    }
    int self = s->types[i];
    int slot = 0;
    double tmp[5];
    for (int c = 0; c < 5; ++c) spatial[c] = 0.0;
    bool boundaryForm = self == -1;  // -2 uses the bulk forms (:261-307)
    if (boundaryForm) slot = oracle::caseSlot(boundaryCase(t));
    if (s->L1 != 0) {
        if (boundaryForm)
            boundaryL1(s->L1, x, x + 5, t, tmp);
        else
            bulkL1(s->L1, x, x + 5, tmp);
        for (int c = 0; c < 5; ++c) spatial[c] += tmp[c];
    }
    const double Ls[4] = {s->L2, s->L3, s->L4, s->L6};
    for (int k = 0; k < 4; ++k) {
        if (Ls[k] == 0) continue;
        oracle::evalStencil(g_tables.t[k][slot], Ls[k], x, tmp);
        for (int c = 0; c < 5; ++c) spatial[c] += tmp[c];
    }
}

// computeL1BulkCPU / computeL1BoundaryCPU (landauDeGennesLCDistortionTerms.cpp:16-111)
void oneConstantSpatial(const orc_system* s, int64_t i, const double* q, double* spatial) {
    const int32_t* nb = &s->nbr[6 * i];
    double x[35];
    int t[6];
    for (int c = 0; c < 5; ++c) x[c] = q[5 * i + c];
    for (int d = 0; d < 6; ++d) {
        t[d] = s->types[nb[d]];
        for (int c = 0; c < 5; ++c) x[5 + 5 * d + c] = q[5 * (int64_t)nb[d] + c];
    }
    if (s->types[i] == -1)
        boundaryL1(s->L1, x, x + 5, t, spatial);
    else
        bulkL1(s->L1, x, x + 5, spatial);
}

// Phase + distortion force of one LC site (t <= 0), before metric correction.
void rawForce(const orc_system* s, int64_t i, const double* q, double* f) {
    double spatial[5];
    for (int c = 0; c < 5; ++c) f[c] = 0.0;
    phaseForce(s, q + 5 * i, f);
    if (s->multi)
        multiConstantSpatial(s, i, q, spatial);
    else
        oneConstantSpatial(s, i, q, spatial);
    for (int c = 0; c < 5; ++c) f[c] -= spatial[c];
}

// Anchoring force summed over object neighbours (computeBoundaryForcesCPU,
// landauDeGennesLCOtherForces.cpp:314-373, neighbour order xd,xu,yd,yu,zd,zu)
void anchoringSum(const orc_system* s, int64_t i, const double* q, double* f) {
    const int32_t* nb = &s->nbr[6 * i];
    double tmp[5];
    for (int d = 0; d < 6; ++d) {
        int tn = s->types[nb[d]];
        if (tn > 0) {
            anchoringForce(s->objects[tn - 1], q + 5 * i, q + 5 * (int64_t)nb[d], tmp);
            for (int c = 0; c < 5; ++c) f[c] += tmp[c];
        }
    }
}

// uniform E, uniform H, then the spatially varying H of site i
// (landauDeGennesLC.cpp:190-200; computeSpatiallyVaryingFieldCPU,
// landauDeGennesLCOtherForces.cpp:254-283)
void fieldSum(const orc_system* s, int64_t i, double* f) {
    double tmp[5];
    if (s->eOn) {
        fieldForce(s->E, s->deps, s->eps0, tmp);
        for (int c = 0; c < 5; ++c) f[c] -= tmp[c];
    }
    if (s->hOn) {
        fieldForce(s->H, s->dchi, s->mu0, tmp);
        for (int c = 0; c < 5; ++c) f[c] -= tmp[c];
    }
    if (s->hvarOn) {
        fieldForce(&s->hvar[3 * i], s->dchi, s->mu0, tmp);
        for (int c = 0; c < 5; ++c) f[c] -= tmp[c];
    }
}

// Full force evaluation.
// semantics 0, the CPU branch of multirankSimulation::computeForces
// (multirankSimulation.cpp:462-466): pass 1 (type 0, zero) writes the bulk
// force and applies the metric to every site; pass 2 (type 1, accumulate) adds
// the boundary-site force, anchoring and fields, then applies the metric again
// (landauDeGennesLC.cpp:125-135, 162-202).
// semantics 1, the GPU branch (landauDeGennesLC.cpp:357-398): one pass over all
// LC sites, then one metric correction.
void computeForces(orc_system* s, const double* q, double* f) {
    if (s->multi) firstDerivatives(s, q);
#pragma omp parallel for num_threads(s->nthreads) schedule(static)
    for (int64_t i = 0; i < s->n; ++i) {
        double* fi = f + 5 * i;
        int t = s->types[i];
        for (int c = 0; c < 5; ++c) fi[c] = 0.0;
        if (t > 0) continue;
        double raw[5];
        rawForce(s, i, q, raw);
        if (s->semantics == 0) {
            if (t == 0) {
                for (int c = 0; c < 5; ++c) fi[c] = raw[c];
                metric(fi);
                fieldSum(s, i, fi);
                metric(fi);
            } else {
                // pass 1 left zero (metric of zero is zero); pass 2 adds
                for (int c = 0; c < 5; ++c) fi[c] += raw[c];
                anchoringSum(s, i, q, fi);
                fieldSum(s, i, fi);
                metric(fi);
            }
        } else {
            for (int c = 0; c < 5; ++c) fi[c] = raw[c];
            if (t < 0) anchoringSum(s, i, q, fi);
            fieldSum(s, i, fi);
            metric(fi);
        }
    }
}

// Distortion energy density with first derivatives dX, dY, dZ at a site (the
// L1..L6 terms of landauDeGennesLC::computeEnergyCPU, landauDeGennesLC.cpp:268-335),
// written from the invariants:
//   L1/2 d_k Q_ij d_k Q_ij,   L2/2 (d_j Q_ij)^2,   L3/2 d_k Q_ij d_j Q_ik,
//   L4/2 eps_lik Q_lj d_k Q_ij (CPU form),   L6/2 Q_lk d_l Q_ij d_k Q_ij.
void distortionTerms(const double* q, const double* dX, const double* dY, const double* dZ, double* e) {
    double Q[3][3], dQ[3][3][3];  // dQ[k][i][j] = d_k Q_ij
    full(q, Q);
    full(dX, dQ[0]);
    full(dY, dQ[1]);
    full(dZ, dQ[2]);
    double l1 = 0, l3 = 0, l4 = 0, l6 = 0, l2 = 0;
    for (int k = 0; k < 3; ++k)
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                l1 += dQ[k][i][j] * dQ[k][i][j];
                l3 += dQ[k][i][j] * dQ[j][i][k];
            }
    for (int i = 0; i < 3; ++i) {
        double div = 0;
        for (int j = 0; j < 3; ++j) div += dQ[j][i][j];
        l2 += div * div;
    }
    static const int eps[3][3][3] = {{{0, 0, 0}, {0, 0, 1}, {0, -1, 0}},
                                     {{0, 0, -1}, {0, 0, 0}, {1, 0, 0}},
                                     {{0, 1, 0}, {-1, 0, 0}, {0, 0, 0}}};
    for (int l = 0; l < 3; ++l)
        for (int i = 0; i < 3; ++i)
            for (int k = 0; k < 3; ++k) {
                if (!eps[l][i][k]) continue;
                for (int j = 0; j < 3; ++j) l4 += eps[l][i][k] * Q[l][j] * dQ[k][i][j];
            }
    for (int l = 0; l < 3; ++l)
        for (int k = 0; k < 3; ++k)
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) l6 += Q[l][k] * dQ[l][i][j] * dQ[k][i][j];
    e[0] = 0.5 * l1;
    e[1] = 0.5 * l2;
    e[2] = 0.5 * l3;
    e[3] = 0.5 * l4;
    e[4] = 0.5 * l6;
}

// Field energy -(perm/2) F.(iso I + anis Q).F (landauDeGennesLC.cpp:238-265)
inline double fieldEnergy(const double* F, double iso, double anis, double perm, const double* q) {
    double FF = F[0] * F[0] + F[1] * F[1] + F[2] * F[2];
    double FQF = F[0] * F[0] * q[0] + F[1] * F[1] * q[3] + F[2] * F[2] * (-q[0] - q[3]) +
                 2.0 * (F[0] * F[1] * q[1] + F[0] * F[2] * q[2] + F[1] * F[2] * q[4]);
    return -0.5 * perm * (iso * FF + anis * FQF);
}

}  // namespace

extern "C" {

orc_system* orc_create(int lx, int ly, int lz) {
    orc_system* s = new orc_system;
    s->lx = lx;
    s->ly = ly;
    s->lz = lz;
    s->n = (int64_t)lx * ly * lz;
    s->types.assign(s->n, 0);  // simpleModel::initializeSimpleModel (simpleModel.cpp:23-43)
    s->q.assign(5 * s->n, 0.0);
    s->v.assign(5 * s->n, 0.0);
    s->f.assign(5 * s->n, 0.0);
    buildNeighbors(s);
    return s;
}

void orc_destroy(orc_system* s) { delete s; }
void orc_set_threads(orc_system* s, int n) { s->nthreads = n < 1 ? 1 : n; }
void orc_neighbors(const orc_system* s, int32_t* out) { std::memcpy(out, s->nbr.data(), 6 * s->n * sizeof(int32_t)); }
void orc_get_types(const orc_system* s, int32_t* out) { std::memcpy(out, s->types.data(), s->n * sizeof(int32_t)); }
void orc_set_types(orc_system* s, const int32_t* in) { std::memcpy(s->types.data(), in, s->n * sizeof(int32_t)); }
void orc_get_q(const orc_system* s, double* out) { std::memcpy(out, s->q.data(), 5 * s->n * sizeof(double)); }
void orc_set_q(orc_system* s, const double* in) { std::memcpy(s->q.data(), in, 5 * s->n * sizeof(double)); }
void orc_get_v(const orc_system* s, double* out) { std::memcpy(out, s->v.data(), 5 * s->n * sizeof(double)); }
void orc_set_v(orc_system* s, const double* in) { std::memcpy(s->v.data(), in, 5 * s->n * sizeof(double)); }
void orc_get_f(const orc_system* s, double* out) { std::memcpy(out, s->f.data(), 5 * s->n * sizeof(double)); }
int orc_n_objects(const orc_system* s) { return (int)s->objects.size(); }

// noiseSource(reproducible) + setReproducibleSeed (noiseSource.cpp:38-47) and
// qTensorLatticeModel::setNematicQTensorRandomly (qTensorLatticeModel.cpp:134-177):
// two global draws, then per site theta = acos(2u-1), phi = 2 pi u, drawn for
// every site but applied only where t <= 0.
void orc_random_init(orc_system* s, int seed, double S0) {
    std::mt19937 gen(seed);
    auto uniform = [&gen]() {
        std::uniform_real_distribution<double> d(0.0, 1.0);  // noiseSource.cpp:16-25
        return d(gen);
    };
    double gTheta = std::acos(2.0 * uniform() - 1);
    double gPhi = 2.0 * kPi * uniform();
    (void)gTheta;
    (void)gPhi;
    for (int64_t pp = 0; pp < s->n; ++pp) {
        double theta = std::acos(2.0 * uniform() - 1);
        double phi = 2.0 * kPi * uniform();
        if (s->types[pp] <= 0) {
            double nx = std::cos(phi) * std::sin(theta);
            double ny = std::sin(phi) * std::sin(theta);
            double nz = std::cos(theta);
            qFromDirector(nx, ny, nz, S0, &s->q[5 * pp]);
        }
    }
}

// setNematicQTensorRandomly(noise, S0, globallyAligned = true)
// (qTensorLatticeModel.cpp:134-177; multirankQTensorLatticeModel.cpp:52-77 draws the
// same way): every t <= 0 site gets the director of the two global draws; the
// per-site draws are made and discarded, which does not change the result.
void orc_random_uniform_init(orc_system* s, int seed, double S0) {
    std::mt19937 gen(seed);
    std::uniform_real_distribution<double> d(0.0, 1.0);
    const double theta = std::acos(2.0 * d(gen) - 1);
    const double phi = 2.0 * kPi * d(gen);
    const double nx = std::cos(phi) * std::sin(theta), ny = std::sin(phi) * std::sin(theta), nz = std::cos(theta);
    for (int64_t pp = 0; pp < s->n; ++pp)
        if (s->types[pp] <= 0) qFromDirector(nx, ny, nz, S0, &s->q[5 * pp]);
}

// multirankQTensorLatticeModel::setDirectorFromFunction (multirankQTensorLatticeModel.cpp:90-110)
// with the function's values given: n4 = (nx, ny, nz, S0) per site, site order
void orc_director_init(orc_system* s, const double* n4) {
    for (int64_t pp = 0; pp < s->n; ++pp)
        if (s->types[pp] <= 0) {
            const double* d = n4 + 4 * pp;
            qFromDirector(d[0], d[1], d[2], d[3], &s->q[5 * pp]);
        }
}

// multirankSimulation::createMultirankBoundaryObject (multirankSimulationBoundaries.cpp:512-543)
// followed by cubicLattice::createBoundaryObject (cubicLattice.cpp:267-318).
int orc_add_object(orc_system* s, const int32_t* xyz, int n, const double* qb, int btype, double W, double S0) {
    std::vector<int64_t> sites;
    sites.reserve(n);
    for (int ii = 0; ii < n; ++ii) {
        int x = xyz[3 * ii], y = xyz[3 * ii + 1], z = xyz[3 * ii + 2];
        // wrap (inc/functions.h:299-317)
        auto wrap = [](int v, int m) {
            int a = v;
            if (v >= m) a = v % m;
            while (a < 0) a += m;
            return a;
        };
        x = wrap(x, s->lx);
        y = wrap(y, s->ly);
        z = wrap(z, s->lz);
        int64_t idx = siteIndex(s, x, y, z);
        sites.push_back(idx);
        std::memcpy(&s->q[5 * idx], qb + 5 * ii, 5 * sizeof(double));
    }
    s->objects.push_back({btype, W, S0});
    int j = (int)s->objects.size();
    for (int64_t site : sites) s->types[site] = j;
    // cubicLattice::createBoundaryObject (cubicLattice.cpp:282-306): every
    // non-object neighbour becomes a surface site, duplicates removed keeping
    // the first occurrence (removeDuplicateVectorElements, inc/functions.h:23-35)
    std::vector<int64_t> surface;
    std::vector<char> seen(s->n, 0);
    for (int64_t site : sites)
        for (int d = 0; d < 6; ++d) {
            int32_t nb = s->nbr[6 * site + d];
            if (s->types[nb] < 1) {
                s->types[nb] = -1;
                if (!seen[nb]) {
                    seen[nb] = 1;
                    surface.push_back(nb);
                }
            }
        }
    s->objSites.push_back(sites);
    s->objSurface.push_back(surface);
    return j;
}

// cubicLattice::boundarySites / surfaceSites of object obj (0-based); returns the count
int64_t orc_object_sites(orc_system* s, int obj, int surface, int64_t* out) {
    if (obj < 0 || obj >= (int)s->objSites.size()) return -1;
    const std::vector<int64_t>& v = surface ? s->objSurface[obj] : s->objSites[obj];
    if (out) std::memcpy(out, v.data(), v.size() * sizeof(int64_t));
    return (int64_t)v.size();
}

// cubicLattice::displaceBoundaryObject (cubicLattice.cpp:329-372), CPU branch:
// per unit step, the object's and its surface's Q move one site along
// direction (neighbour order x-, x+, y-, y+, z-, z+); the old surface sites
// become bulk (type 0) first, then the moved sites take the object's type and -1.
int orc_displace_object(orc_system* s, int obj, int direction, int magnitude) {
    if (obj < 0 || obj >= (int)s->objSites.size() || direction < 0 || direction > 5) return -1;
    std::vector<int64_t>& bs = s->objSites[obj];
    std::vector<int64_t>& ss = s->objSurface[obj];
    for (int mm = 0; mm < magnitude; ++mm) {
        std::vector<std::pair<int64_t, std::array<double, 5>>> b1(bs.size()), b2(ss.size());
        for (size_t k = 0; k < bs.size(); ++k) {
            b1[k].first = s->nbr[6 * bs[k] + direction];
            std::memcpy(b1[k].second.data(), &s->q[5 * bs[k]], 5 * sizeof(double));
        }
        for (size_t k = 0; k < ss.size(); ++k) {
            b2[k].first = s->nbr[6 * ss[k] + direction];
            std::memcpy(b2[k].second.data(), &s->q[5 * ss[k]], 5 * sizeof(double));
            s->types[ss[k]] = 0;
        }
        for (size_t k = 0; k < bs.size(); ++k) {
            bs[k] = b1[k].first;
            std::memcpy(&s->q[5 * bs[k]], b1[k].second.data(), 5 * sizeof(double));
            s->types[bs[k]] = obj + 1;
        }
        for (size_t k = 0; k < ss.size(); ++k) {
            ss[k] = b2[k].first;
            std::memcpy(&s->q[5 * ss[k]], b2[k].second.data(), 5 * sizeof(double));
            s->types[ss[k]] = -1;
        }
    }
    return 0;
}

// multirankSimulation::createWall (multirankSimulationBoundaries.cpp:101-167)
int orc_wall(orc_system* s, int xyz, int plane, int btype, double W, double S0) {
    if (xyz < 0 || xyz > 2) return -2;
    double Qt[5] = {0, 0, 0, 0, 0};
    if (btype == 0) {
        if (xyz == 0) { Qt[0] = S0; Qt[3] = -0.5 * S0; }
        else if (xyz == 1) { Qt[0] = -0.5 * S0; Qt[3] = S0; }
        else { Qt[0] = -0.5 * S0; Qt[3] = -0.5 * S0; }
    } else {
        Qt[xyz] = 1.0;
    }
    int size1, size2;
    if (xyz == 0) { size1 = s->ly; size2 = s->lz; }
    else if (xyz == 1) { size1 = s->lx; size2 = s->lz; }
    else { size1 = s->lx; size2 = s->ly; }
    std::vector<int32_t> pts;
    std::vector<double> qs;
    for (int a = 0; a < size1; ++a)
        for (int b = 0; b < size2; ++b) {
            if (xyz == 0) { pts.push_back(plane); pts.push_back(a); pts.push_back(b); }
            else if (xyz == 1) { pts.push_back(a); pts.push_back(plane); pts.push_back(b); }
            else { pts.push_back(a); pts.push_back(b); pts.push_back(plane); }
            qs.insert(qs.end(), Qt, Qt + 5);
        }
    return orc_add_object(s, pts.data(), (int)(pts.size() / 3), qs.data(), btype, W, S0);
}

// multirankSimulation::createSphericalColloid (multirankSimulationBoundaries.cpp:169-215)
int orc_sphere(orc_system* s, double cx, double cy, double cz, double radius, int btype, double W, double S0) {
    std::vector<int32_t> pts;
    std::vector<double> qs;
    for (int xx = (int)std::ceil(cx - radius); xx < std::floor(cx + radius); ++xx)
        for (int yy = (int)std::ceil(cy - radius); yy < std::floor(cy + radius); ++yy)
            for (int zz = (int)std::ceil(cz - radius); zz < std::floor(cz + radius); ++zz) {
                double dx = xx - cx, dy = yy - cy, dz = zz - cz;
                if ((dx * dx + dy * dy + dz * dz) < radius * radius) {
                    pts.push_back(xx); pts.push_back(yy); pts.push_back(zz);
                    double Qt[5] = {0, 0, 0, 0, 0};
                    if (btype == 0) {
                        qFromDirector(dx, dy, dz, S0, Qt);
                    } else {
                        double nrm = std::sqrt(dx * dx + dy * dy + dz * dz);
                        Qt[0] = dx / nrm; Qt[1] = dy / nrm; Qt[2] = dz / nrm;
                    }
                    qs.insert(qs.end(), Qt, Qt + 5);
                }
            }
    return orc_add_object(s, pts.data(), (int)(pts.size() / 3), qs.data(), btype, W, S0);
}

// updater::getNTotal (baseUpdater.cpp:10-28): number of sites with t <= 0
int64_t orc_n_active(const orc_system* s) {
    int64_t n = 0;
    for (int64_t i = 0; i < s->n; ++i)
        if (s->types[i] <= 0) ++n;
    return n;
}

void orc_set_params(orc_system* s, double A, double B, double C, double L1, double L2, double L3, double L4,
                    double L6, int multiConstant, int semantics) {
    s->A = A; s->B = B; s->C = C;
    s->L1 = L1; s->L2 = L2; s->L3 = L3; s->L4 = L4; s->L6 = L6;
    s->multi = multiConstant;
    s->semantics = semantics;
}

void orc_set_efield(orc_system* s, int on, double Ex, double Ey, double Ez, double eps, double eps0, double deps) {
    s->eOn = on;
    s->E[0] = Ex; s->E[1] = Ey; s->E[2] = Ez;
    s->eps = eps; s->eps0 = eps0; s->deps = deps;
}

// landauDeGennesLC::setSpatiallyVaryingField (landauDeGennesLC.cpp:81-123): the
// per-site field, with chi, mu0, deltaChi shared with the uniform H field
void orc_set_hfield_varying(orc_system* s, int on, const double* h3, double chi, double mu0, double dchi) {
    s->hvarOn = on;
    if (on) s->hvar.assign(h3, h3 + 3 * s->n);
    s->chi = chi;
    s->mu0 = mu0;
    s->dchi = dchi;
}

void orc_set_hfield(orc_system* s, int on, double Hx, double Hy, double Hz, double chi, double mu0, double dchi) {
    s->hOn = on;
    s->H[0] = Hx; s->H[1] = Hy; s->H[2] = Hz;
    s->chi = chi; s->mu0 = mu0; s->dchi = dchi;
}

void orc_load_tables(int n, const int32_t* term, const int32_t* cslot, const int32_t* comp, const double* coef,
                     const int32_t* a, const int32_t* b) {
    for (int k = 0; k < 4; ++k)
        for (int c = 0; c < 28; ++c) g_tables.t[k][c].clear();
    for (int i = 0; i < n; ++i) g_tables.t[term[i]][cslot[i]].push_back({comp[i], coef[i], a[i], b[i]});
    g_tables.loaded = true;
}

void orc_forces(orc_system* s, const double* q, double* f) { computeForces(s, q, f); }

void orc_first_derivatives(orc_system* s, const double* q, double* d15) {
    firstDerivatives(s, q);
    std::memcpy(d15, s->deriv.data(), 15 * s->n * sizeof(double));
}

void orc_force_cb(void* ctx, int n, const double* q, double* f) {
    (void)n;
    computeForces(static_cast<orc_system*>(ctx), q, f);
}

// landauDeGennesLC::computeEnergyCPU (landauDeGennesLC.cpp:204-353); returns the
// total (the simulation divides it by NActive, multirankSimulation.cpp:475-490)
double orc_energy(orc_system* s, const double* q, double* comps) {
    double phase = 0, distortion = 0, anchoring = 0, eF = 0, hF = 0;
    double a = 0.5 * s->A, b = s->B / 3.0, c = 0.25 * s->C;
    const double Ls[5] = {s->L1, s->L2, s->L3, s->L4, s->L6};
    for (int64_t i = 0; i < s->n; ++i) {
        int t = s->types[i];
        if (t > 0) continue;
        const double* qc = q + 5 * i;
        double t2 = trQ2(qc);
        phase += a * t2 + b * trQ3(qc) + c * t2 * t2;
        if (s->eOn) eF += fieldEnergy(s->E, s->eps, s->deps, s->eps0, qc);
        if (s->hOn) hF += fieldEnergy(s->H, s->chi, s->dchi, s->mu0, qc);
        if (s->hvarOn) hF += fieldEnergy(&s->hvar[3 * i], s->chi, s->dchi, s->mu0, qc);
        const int32_t* nb = &s->nbr[6 * i];
        double d[3][5];
        for (int ax = 0; ax < 3; ++ax)
            for (int k = 0; k < 5; ++k) d[ax][k] = 0.5 * (q[5 * (int64_t)nb[2 * ax + 1] + k] - q[5 * (int64_t)nb[2 * ax] + k]);
        if (t < 0) {
            for (int dd = 0; dd < 6; ++dd) {
                int tn = s->types[nb[dd]];
                if (tn <= 0) continue;
                const double* qn = q + 5 * (int64_t)nb[dd];
                anchoring += anchoringEnergy(s->objects[tn - 1], qc, qn);
                int ax = dd / 2;
                const double* other = q + 5 * (int64_t)nb[dd ^ 1];
                for (int k = 0; k < 5; ++k) d[ax][k] = (dd & 1) ? (qc[k] - other[k]) : (other[k] - qc[k]);
            }
        }
        double e[5];
        distortionTerms(qc, d[0], d[1], d[2], e);
        for (int k = 0; k < 5; ++k)
            if (Ls[k] != 0) distortion += Ls[k] * e[k];
    }
    if (comps) {
        comps[0] = phase; comps[1] = distortion; comps[2] = anchoring; comps[3] = eF; comps[4] = hF;
    }
    return phase + distortion + anchoring + eF + hF;
}

// landauDeGennesLC::energyDensity (landauDeGennesLC.h:96) as computeEnergyCPU
// fills it (landauDeGennesLC.cpp:204-353): per site, 0 for object sites, else
// phase, then E, uniform H and spatially varying H, anchoring and distortion,
// added in that order.
void orc_energy_density(orc_system* s, const double* q, double* out) {
    const double a = 0.5 * s->A, b = s->B / 3.0, c = 0.25 * s->C;
    const double Ls[5] = {s->L1, s->L2, s->L3, s->L4, s->L6};
    for (int64_t i = 0; i < s->n; ++i) {
        out[i] = 0.0;
        const int t = s->types[i];
        if (t > 0) continue;
        const double* qc = q + 5 * i;
        const double t2 = trQ2(qc);
        out[i] += a * t2 + b * trQ3(qc) + c * t2 * t2;
        if (s->eOn) out[i] += fieldEnergy(s->E, s->eps, s->deps, s->eps0, qc);
        if (s->hOn) out[i] += fieldEnergy(s->H, s->chi, s->dchi, s->mu0, qc);
        if (s->hvarOn) out[i] += fieldEnergy(&s->hvar[3 * i], s->chi, s->dchi, s->mu0, qc);
        const int32_t* nb = &s->nbr[6 * i];
        double d[3][5];
        for (int ax = 0; ax < 3; ++ax)
            for (int k = 0; k < 5; ++k) d[ax][k] = 0.5 * (q[5 * (int64_t)nb[2 * ax + 1] + k] - q[5 * (int64_t)nb[2 * ax] + k]);
        if (t < 0) {
            double anch = 0;
            for (int dd = 0; dd < 6; ++dd) {
                const int tn = s->types[nb[dd]];
                if (tn <= 0) continue;
                anch += anchoringEnergy(s->objects[tn - 1], qc, q + 5 * (int64_t)nb[dd]);
                const int ax = dd / 2;
                const double* other = q + 5 * (int64_t)nb[dd ^ 1];
                for (int k = 0; k < 5; ++k) d[ax][k] = (dd & 1) ? (qc[k] - other[k]) : (other[k] - qc[k]);
            }
            out[i] += anch;
        }
        double e[5];
        distortionTerms(qc, d[0], d[1], d[2], e);
        double dist = 0;
        for (int k = 0; k < 5; ++k)
            if (Ls[k] != 0) dist += Ls[k] * e[k];
        out[i] += dist;
    }
}

// computeStressTensors, one-constant form (landauDeGennesLCOtherForces.cpp:214-251):
// the stress of site derivatives fd = [dx q0..4, dy q0..4, dz q0..4] at elastic
// constant L1, row-major m[9], with the site's energy density e added on the
// diagonal.
void orc_stress_tensor(const double* fd, double L1, double e, double* m) {
    const double* X = fd;
    const double* Y = fd + 5;
    const double* Z = fd + 10;
    auto diag = [&](const double* D) {
        return -2 * L1 * (D[0] * D[0] + 2 * (D[1] * D[1]) + 2 * (D[2] * D[2]) + D[3] * D[3] + 2 * (D[4] * D[4]) +
                          D[0] * D[3]);
    };
    auto off = [&](const double* P, const double* R) {
        return L1 * (-(R[0] * (2 * P[0] + P[3])) - R[3] * (P[0] + 2 * P[3]) -
                     4 * (R[1] * P[1] + R[2] * P[2] + R[4] * P[4]));
    };
    m[0] = diag(X) + e;
    m[1] = off(X, Y);
    m[2] = off(X, Z);
    m[3] = off(X, Y);
    m[4] = diag(Y) + e;
    m[5] = off(Y, Z);
    m[6] = off(X, Z);
    m[7] = off(Y, Z);
    m[8] = diag(Z) + e;
}

// landauDeGennesLC::computeObjectForces (landauDeGennesLCOtherForces.cpp:147-208)
// with computeStressTensors (:214-251; one-constant only, as there): over the
// object's surface sites, the stress tensor from the first derivatives
// (computeFirstDerivatives rules) and L1, plus the site's energy density
// (computeEnergyCPU: phase, fields, anchoring, distortion) on the diagonal,
// applied to the site's surface-area vector (+-1 per axis towards an object
// neighbour, the upper neighbour taking precedence); summed in surface-site
// order.  Returns -1 unless the model is one-constant.
int orc_object_force(orc_system* s, int obj, double* out) {
    if (obj < 0 || obj >= (int)s->objSurface.size() || s->multi) return -1;
    const double* q = s->q.data();
    const double a = 0.5 * s->A, b = s->B / 3.0, c = 0.25 * s->C, L1 = s->L1;
    double fx = 0, fy = 0, fz = 0;
    for (int64_t i : s->objSurface[obj]) {
        const int t = s->types[i];
        if (t > 0) continue;   // a surface site taken by a later object
        const double* qc = q + 5 * i;
        const int32_t* nb = &s->nbr[6 * i];
        // energy density of the site (computeEnergyCPU)
        double e = 0;
        const double t2 = trQ2(qc);
        e += a * t2 + b * trQ3(qc) + c * t2 * t2;
        if (s->eOn) e += fieldEnergy(s->E, s->eps, s->deps, s->eps0, qc);
        if (s->hOn) e += fieldEnergy(s->H, s->chi, s->dchi, s->mu0, qc);
        if (s->hvarOn) e += fieldEnergy(&s->hvar[3 * i], s->chi, s->dchi, s->mu0, qc);
        double d[3][5];
        for (int ax = 0; ax < 3; ++ax)
            for (int k = 0; k < 5; ++k) d[ax][k] = 0.5 * (q[5 * (int64_t)nb[2 * ax + 1] + k] - q[5 * (int64_t)nb[2 * ax] + k]);
        double anch = 0;
        if (t < 0)
            for (int dd = 0; dd < 6; ++dd) {
                const int tn = s->types[nb[dd]];
                if (tn <= 0) continue;
                anch += anchoringEnergy(s->objects[tn - 1], qc, q + 5 * (int64_t)nb[dd]);
                const int ax = dd / 2;
                const double* other = q + 5 * (int64_t)nb[dd ^ 1];
                for (int k = 0; k < 5; ++k) d[ax][k] = (dd & 1) ? (qc[k] - other[k]) : (other[k] - qc[k]);
            }
        e += anch;
        double et[5];
        distortionTerms(qc, d[0], d[1], d[2], et);
        e += L1 * et[0];
        double fd[15], m[9];
        for (int ax = 0; ax < 3; ++ax)
            for (int k = 0; k < 5; ++k) fd[5 * ax + k] = d[ax][k];
        orc_stress_tensor(fd, L1, e, m);
        const double m11 = m[0], m12 = m[1], m13 = m[2], m21 = m[3], m22 = m[4], m23 = m[5], m31 = m[6], m32 = m[7],
                     m33 = m[8];
        double vx = 0, vy = 0, vz = 0;
        if (s->types[nb[0]] > 0) vx = -1.0;
        if (s->types[nb[1]] > 0) vx = 1.0;
        if (s->types[nb[2]] > 0) vy = -1.0;
        if (s->types[nb[3]] > 0) vy = 1.0;
        if (s->types[nb[4]] > 0) vz = -1.0;
        if (s->types[nb[5]] > 0) vz = 1.0;
        // scalar3 * Matrix3x3 (inc/matrix.h:281-288)
        fx = fx + (m11 * vx + m12 * vy + m13 * vz);
        fy = fy + (m21 * vx + m22 * vy + m23 * vz);
        fz = fz + (m31 * vx + m32 * vy + m33 * vz);
    }
    out[0] = fx;
    out[1] = fy;
    out[2] = fz;
    return 0;
}

double orc_energy_cb(void* ctx, int n, const double* q) {
    (void)n;
    return orc_energy(static_cast<orc_system*>(ctx), q, nullptr);
}

// energyMinimizerFIRE: constructor/defaults (energyMinimizerFIRE.cpp:14-43,
// energyMinimizerFIRE.h:36-53), minimize (:151-170), velocity Verlet
// (velocityVerlet.cpp:8-40), fireStepCPU (:77-145).  Uses and updates the
// system's Q, v, F.
int orc_fire(orc_system* s, double dt, double alphaStart, double dtMax, double dtInc, double dtDec,
             double alphaDec, int nMin, double fc, double alphaMin, int maxIter, double* trace) {
    double deltaT = dt, deltaTMin = dt * .01, alpha = alphaStart;
    int NSinceNegativePower = 0, iterations = 0;
    double forceMax = 100.0, Power = 0, scaling = 0;
    double nTotal = (double)orc_n_active(s);
    const int64_t N = s->n;
    double* Q = s->q.data();
    double* V = s->v.data();
    double* F = s->f.data();
    computeForces(s, Q, F);
    int it0 = iterations;
    while ((iterations < maxIter && forceMax > fc) || iterations == it0) {
        iterations += 1;
        // velocityVerlet::integrateEOMCPU
        for (int64_t i = 0; i < 5 * N; ++i) {
            double disp = deltaT * V[i] + (0.5 * deltaT * deltaT) * F[i];
            V[i] += (0.5) * deltaT * F[i];
            Q[i] += disp;  // qTensorLatticeModel::moveParticles, scale 1 (:179-200)
        }
        computeForces(s, Q, F);
        for (int64_t i = 0; i < 5 * N; ++i) V[i] += (0.5) * deltaT * F[i];
        // fireStepCPU: QT dot products (inc/functions.h:140-147)
        double forceNorm = 0, velocityNorm = 0;
        Power = 0;
        for (int64_t i = 0; i < N; ++i) {
            const double* f = F + 5 * i;
            const double* v = V + 5 * i;
            double fd = 0, vd = 0, pd = 0;
            for (int k = 0; k < 5; ++k) fd += f[k] * f[k];
            fd += f[0] * f[3];
            for (int k = 0; k < 5; ++k) vd += v[k] * v[k];
            vd += v[0] * v[3];
            for (int k = 0; k < 5; ++k) pd += f[k] * v[k];
            pd += f[0] * v[3];
            forceNorm += fd;
            velocityNorm += vd;
            Power += pd;
        }
        forceMax = std::sqrt(forceNorm) / nTotal;
        scaling = 0.0;
        if (forceNorm > 0.) scaling = std::sqrt(velocityNorm / forceNorm);
        for (int64_t i = 0; i < 5 * N; ++i) V[i] = (1.0 - alpha) * V[i] + alpha * scaling * F[i];
        if (Power > 0 && iterations % 500 != 0) {
            if (NSinceNegativePower > nMin) {
                deltaT = std::min(deltaT * dtInc, dtMax);
                alpha = alpha * alphaDec;
                alpha = std::max(alpha, alphaMin);
            }
            NSinceNegativePower += 1;
        } else {
            NSinceNegativePower = 0;
            deltaT = deltaT * dtDec;
            deltaT = std::max(deltaT, deltaTMin);
            alpha = alphaStart;
            std::fill(V, V + 5 * N, 0.0);
        }
        if (trace) {
            double* t = trace + 6 * (int64_t)(iterations - it0 - 1);
            t[0] = forceMax; t[1] = Power; t[2] = alpha; t[3] = deltaT; t[4] = scaling;
            t[5] = NSinceNegativePower;
        }
        if (std::isnan(forceMax)) break;  // the reference would exit as "converged"
    }
    return iterations - it0;
}

// energyMinimizerAdam::adamStepCPU / minimize (energyMinimizerAdam.cpp:25-75);
// moments start at zero, beta^t at beta (energyMinimizerAdam.h:16-25).
int orc_adam(orc_system* s, double b1, double b2, double eps, double dt, double fc, int maxIter, double* trace) {
    const int64_t N = s->n;
    std::vector<double> m(5 * N, 0.0), v(5 * N, 0.0), disp(5 * N, 0.0);
    double b1t = b1, b2t = b2;
    int iterations = 0;
    double forceMax = 110.0;
    double* Q = s->q.data();
    double* F = s->f.data();
    while (iterations < maxIter && forceMax > fc) {
        iterations += 1;
        computeForces(s, Q, F);
        double forceNorm = 0.0;
        for (int64_t nn = 0; nn < N; ++nn) {
            double dotf = 0;
            for (int k = 0; k < 5; ++k) dotf += F[5 * nn + k] * F[5 * nn + k];
            forceNorm += dotf;
            for (int k = 0; k < 5; ++k) {
                int64_t i = 5 * nn + k;
                m[i] = b1 * m[i] + (b1 - 1) * F[i];
                v[i] = b2 * v[i] + (1 - b2) * (F[i] * F[i]);
                double mc = m[i] * (1.0 / (1.0 - b1t));
                double vc = v[i] * (1.0 / (1.0 - b2t));
                double rootvc = std::sqrt(vc);
                if (rootvc == 0) rootvc = eps;
                disp[i] = -dt * mc / (rootvc);
            }
        }
        for (int64_t i = 0; i < 5 * N; ++i) Q[i] += disp[i];
        forceMax = std::sqrt(forceNorm) / N;
        b1t *= b1;
        b2t *= b2;
        if (trace) trace[iterations - 1] = forceMax;
    }
    return iterations;
}

// energyMinimizerNesterovAG (energyMinimizerNesterovAG.cpp:9-120): y starts as
// a copy of Q, lambda = 1; optional scheduled momentum mu = 1 + gamma.
int orc_nesterov(orc_system* s, double dt, double mu, double fc, int scheduled, int maxIter, double* trace) {
    const int64_t N = s->n;
    double* Q = s->q.data();
    double* F = s->f.data();
    std::vector<double> alt(s->q);
    double lambda = 1.0, gamma = 0.0, forceMax = 110.0;
    int iterations = 0;
    while (iterations < maxIter && forceMax > fc) {
        double oldLambda = lambda;
        lambda = 0.5 * (1.0 + std::sqrt(1.0 + 4.0 * oldLambda * oldLambda));
        gamma = (1 - oldLambda) / lambda;
        if (scheduled) mu = 1. + gamma;
        iterations += 1;
        computeForces(s, Q, F);
        double forceNorm = 0.0;
        for (int64_t nn = 0; nn < N; ++nn) {
            double dotf = 0;
            for (int k = 0; k < 5; ++k) dotf += F[5 * nn + k] * F[5 * nn + k];
            forceNorm += dotf;
            for (int k = 0; k < 5; ++k) {
                int64_t i = 5 * nn + k;
                double old = alt[i];
                alt[i] = Q[i] + dt * F[i];
                Q[i] = alt[i] + mu * (alt[i] - old);
            }
        }
        forceMax = std::sqrt(forceNorm) / N;
        if (trace) trace[iterations - 1] = forceMax;
    }
    return iterations;
}

namespace {
// utilities.cpp:4-23
double plainDot(const double* a, const double* b, int64_t n5) {
    double ans = 0.0;
    for (int64_t i = 0; i < n5; ++i) ans += a[i] * b[i];
    return ans;
}
void axpy(double* y, const double* x, double factor, int64_t n5) {
    for (int64_t i = 0; i < n5; ++i) y[i] = y[i] + factor * x[i];
}
}  // namespace

// energyMinimizerLoLBFGS (energyMinimizerLoLBFGS.cpp:20-226, .h:29-45): memory
// is max(5, m) (the constructor already sizes it for 5); no line search; step
// 3 adds (s.y / y.y) y; s = eta p, y = F_old - F_new.
int orc_lolbfgs(orc_system* s, int mreq, double dt, double c, double fc, double tau, int scheduled, int maxIter,
                double* trace) {
    (void)c;
    const int64_t N = s->n, n5 = 5 * N;
    int m = std::max(5, mreq);
    double* Q = s->q.data();
    double* F = s->f.data();
    std::vector<std::vector<double>> y(m, std::vector<double>(n5, 0.0)), sv(m, std::vector<double>(n5, 0.0));
    std::vector<double> p(n5, 0.0), sy(m, 0.0), al(m, 0.0);
    int iterations = 0, cur = 0;
    double forceMax = 100.0;
    int it0 = iterations;
    while (((iterations < maxIter) && (forceMax > fc)) || iterations == it0) {
        double eta = scheduled ? tau / (tau + iterations) * dt : dt;
        // step 1
        if (iterations == 0) {
            computeForces(s, Q, F);
            p.assign(F, F + n5);
        }
        // step 2, newest first
        for (int ii = 0; ii < m; ++ii) {
            int tMinusI = (cur - ii - 1 + m) % m;
            sy[ii] = plainDot(sv[tMinusI].data(), y[tMinusI].data(), n5);
            al[ii] = 0;
            if (sy[ii] != 0) al[ii] = plainDot(sv[tMinusI].data(), p.data(), n5) / sy[ii];
            axpy(p.data(), y[tMinusI].data(), -al[ii], n5);
        }
        // step 3
        int lastM = (cur - 1 + m) % m;
        double val1 = sy[0];
        double val2 = plainDot(y[lastM].data(), y[lastM].data(), n5);
        if (val2 != 0) axpy(p.data(), y[lastM].data(), val1 / val2, n5);
        // step 4, oldest first
        for (int ii = m - 1; ii >= 0; --ii) {
            int tMinusI = (cur - ii - 1 + m) % m;
            if (sy[ii] != 0) {
                double beta = plainDot(y[tMinusI].data(), p.data(), n5) / sy[ii];
                axpy(p.data(), sv[tMinusI].data(), al[ii] - beta, n5);
            }
        }
        for (int64_t i = 0; i < n5; ++i) sv[cur][i] = eta * p[i];
        y[cur].assign(F, F + n5);
        for (int64_t i = 0; i < n5; ++i) Q[i] += sv[cur][i];
        computeForces(s, Q, F);
        p.assign(F, F + n5);
        axpy(y[cur].data(), p.data(), -1.0, n5);
        forceMax = std::sqrt(plainDot(p.data(), p.data(), n5)) / N;
        iterations += 1;
        cur = (cur + 1) % m;
        if (trace) trace[iterations - it0 - 1] = forceMax;
    }
    return iterations - it0;
}

// eigenvaluesOfQ (qTensorFunctions.h:215-241): trigonometric closed form
void orc_distortion_density(const double* q, const double* dX, const double* dY, const double* dZ, double* e5) {
    distortionTerms(q, dX, dY, dZ, e5);
}

double orc_anchoring_energy(int btype, double P1, double P2, const double* q, const double* qb) {
    BoundaryObj o{btype, P1, P2};
    return anchoringEnergy(o, q, qb);
}

void orc_anchoring_force(int btype, double P1, double P2, const double* q, const double* qb, double* f) {
    BoundaryObj o{btype, P1, P2};
    anchoringForce(o, q, qb, f);
}

// per-site pieces for the unit checks: phase gradient parts and L1 stencils
void orc_phase_parts(const double* q, double* g15, double* tr3) {
    dTrQ2(q, g15);
    dTrQ3(q, g15 + 5);
    dTrQ2Sq(q, g15 + 10);
    double t = trQ2(q);
    tr3[0] = t;
    tr3[1] = trQ3(q);
    tr3[2] = t * t;
}

void orc_l1_site(double L1, int boundary, const double* q, const double* nb, const int32_t* t, double* out) {
    int tt[6];
    for (int d = 0; d < 6; ++d) tt[d] = t[d];
    if (boundary)
        boundaryL1(L1, q, nb, tt, out);
    else
        bulkL1(L1, q, nb, out);
}

void orc_eigenvalues(const double* q, double* ev) {
    double p1 = q[1] * q[1] + q[2] * q[2] + q[4] * q[4];
    if (p1 == 0) {
        ev[0] = q[0]; ev[1] = q[3]; ev[2] = -q[0] - q[3];
        return;
    }
    double p2 = 2 * (q[0] * q[0] + q[3] * q[3] + q[0] * q[3]) + 2 * p1;
    double p = std::sqrt(p2 / 6.0);
    double Bq[5];
    for (int i = 0; i < 5; ++i) Bq[i] = (1.0 / p) * q[i];
    double M[3][3];
    full(Bq, M);
    double r = 0.5 * det3(M);
    double phi;
    if (r <= -1) phi = kPi / 3.0;
    else if (r >= 1) phi = 0.0;
    else phi = std::acos(r) / 3.0;
    double a = 2.0 * p * std::cos(phi);
    double cc = 2.0 * p * std::cos(phi + 2.0 * kPi / 3.0);
    ev[0] = a; ev[1] = -a - cc; ev[2] = cc;
}

// qTensorLatticeModel::computeDefectMeasures (qTensorLatticeModel.cpp:95-133);
// object sites are left at zero.
void orc_defects(const orc_system* s, int defectType, double* out) {
    for (int64_t pp = 0; pp < s->n; ++pp) {
        out[pp] = 0.0;
        if (s->types[pp] > 0) continue;
        const double* q = &s->q[5 * pp];
        double M[3][3];
        full(q, M);
        if (defectType == 0) {
            double ev[3];
            orc_eigenvalues(q, ev);
            out[pp] = std::max(std::max(ev[0], ev[1]), ev[2]);
        } else if (defectType == 1) {
            out[pp] = det3(M);
        } else {
            double t = trQ2(q), d = det3(M);
            out[pp] = t * t * t - 54.0 * d * d;
        }
    }
}

}  // extern "C"
