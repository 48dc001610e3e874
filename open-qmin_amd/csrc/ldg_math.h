// Per-site Landau-de Gennes algebra shared by the gfx950 kernels.
//
// q = (Qxx, Qxy, Qxz, Qyy, Qyz) with Qzz = -Qxx - Qyy.  The force on q is
// F = -(phase gradient) - (distortion term), then shaped by the metric of this
// non-orthonormal basis.  Reference behaviour followed (open-Qmin):
//   phase           src/forces/landauDeGennesLCDistortionTerms.cpp:24-38
//   L1 stencils     src/forces/lcForces.h:26-63
//   anchoring       src/forces/landauDeGennesLCBoundary.h:26-124
//   metric          src/forces/landauDeGennesLC.cpp:137-160
//   fields          src/forces/landauDeGennesLCOtherForces.cpp:285-312
//   energy density  src/forces/landauDeGennesLC.cpp:204-353
#pragma once
#include "qmin_device.h"

namespace qmin {

template <typename R>
__host__ __device__ __forceinline__ R trQ2(const R* q) {
    return R(2) * (q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3] + q[4] * q[4] + q[0] * q[3]);
}

template <typename R>
__host__ __device__ __forceinline__ void dTrQ2(const R* q, R* g) {
    g[0] = R(2) * (R(2) * q[0] + q[3]);
    g[1] = R(4) * q[1];
    g[2] = R(4) * q[2];
    g[3] = R(2) * (q[0] + R(2) * q[3]);
    g[4] = R(4) * q[4];
}

// Tr(Q^3) = 3 det Q for traceless Q; its gradient is 3x the cofactor matrix
// mapped onto (q0..q4).
template <typename R>
__host__ __device__ __forceinline__ void dTrQ3(const R* q, R* g) {
    const R w = -q[0] - q[3];
    const R cxx = q[3] * w - q[4] * q[4];
    const R cyy = q[0] * w - q[2] * q[2];
    const R czz = q[0] * q[3] - q[1] * q[1];
    const R cxy = q[2] * q[4] - q[1] * w;
    const R cxz = q[1] * q[4] - q[2] * q[3];
    const R cyz = q[1] * q[2] - q[0] * q[4];
    g[0] = R(3) * (cxx - czz);
    g[1] = R(6) * cxy;
    g[2] = R(6) * cxz;
    g[3] = R(3) * (cyy - czz);
    g[4] = R(6) * cyz;
}

template <typename R>
__host__ __device__ __forceinline__ R trQ3(const R* q) {
    const R w = -q[0] - q[3];
    const R det = q[0] * (q[3] * w - q[4] * q[4]) - q[1] * (q[1] * w - q[4] * q[2]) + q[2] * (q[1] * q[4] - q[3] * q[2]);
    return R(3) * det;
}

// F -= a dTrQ2 + b dTrQ3 + c d(TrQ2)^2
template <typename R>
__device__ __forceinline__ void phase_force(R a, R b, R c, const R* q, R* F) {
    R g2[5], g3[5];
    dTrQ2(q, g2);
    dTrQ3(q, g3);
    const R t2 = trQ2(q);
    const R c2 = R(2) * c * t2;
#pragma unroll
    for (int i = 0; i < 5; ++i) F[i] -= a * g2[i];
#pragma unroll
    for (int i = 0; i < 5; ++i) F[i] -= b * g3[i];
#pragma unroll
    for (int i = 0; i < 5; ++i) F[i] -= c2 * g2[i];
}

// s0 <- 2 s0 + s3, s3 <- s0 + 2 s3, off-diagonal components doubled
template <typename R>
__device__ __forceinline__ void l1_shape(R* s) {
    const R d = s[0] + s[3];
    s[0] += d;
    s[1] *= R(2);
    s[2] *= R(2);
    s[3] += d;
    s[4] *= R(2);
}

template <typename R>
__host__ __device__ __forceinline__ void metric(R* f) {
    const R xx = f[0], yy = f[3];
    const R t = R(2) / R(3);
    f[0] = t * (R(2) * xx - yy);
    f[3] = t * (R(2) * yy - xx);
}

template <typename R>
__device__ __forceinline__ void full3(const R* q, R M[3][3]) {
    M[0][0] = q[0]; M[0][1] = q[1]; M[0][2] = q[2];
    M[1][0] = q[1]; M[1][1] = q[3]; M[1][2] = q[4];
    M[2][0] = q[2]; M[2][1] = q[4]; M[2][2] = -q[0] - q[3];
}

template <typename R>
__device__ __forceinline__ void mat_grad_to_q(const R G[3][3], R* g) {
    g[0] = G[0][0] - G[2][2];
    g[1] = G[0][1] + G[1][0];
    g[2] = G[0][2] + G[2][0];
    g[3] = G[1][1] - G[2][2];
    g[4] = G[1][2] + G[2][1];
}

// X -> P X P with P = I - nu nu
template <typename R>
__device__ __forceinline__ void project(const R P[3][3], const R X[3][3], R out[3][3]) {
    R T[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) T[a][b] = X[a][0] * P[0][b] + X[a][1] * P[1][b] + X[a][2] * P[2][b];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) out[a][b] = P[a][0] * T[0][b] + P[a][1] * T[1][b] + P[a][2] * T[2][b];
}

// Anchoring force of one object neighbour with state qb on a site q.
// homeotropic: E = W Tr((Q - Q^B)^2).  degenerate planar (Fournier-Galatola,
// W1 = W2 = W, nu = qb[0..2]): E = W Tr((Qt - P Qt P)^2) + W (Tr Qt^2 - (3 S0/2)^2)^2
// with Qt = Q + S0/2 I, P = I - nu nu.  Returns -dE/dq added into F.
template <typename R>
__device__ __forceinline__ void add_anchoring_force(const DevObject& o, const R* q, const R* qb, R* F) {
    if (o.type == 0) {
        R d[5], g[5];
#pragma unroll
        for (int i = 0; i < 5; ++i) d[i] = q[i] - qb[i];
        dTrQ2(d, g);
        const R W = R(o.P1);
#pragma unroll
        for (int i = 0; i < 5; ++i) F[i] -= W * g[i];
        return;
    }
    const R W = R(o.P1), S0 = R(o.P2);
    const R nu[3] = {qb[0], qb[1], qb[2]};
    R Qt[3][3], P[3][3], D[3][3], PDP[3][3], G[3][3];
    full3(q, Qt);
#pragma unroll
    for (int a = 0; a < 3; ++a) Qt[a][a] += R(0.5) * S0;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) P[a][b] = (a == b ? R(1) : R(0)) - nu[a] * nu[b];
    project(P, Qt, D);
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) D[a][b] = Qt[a][b] - D[a][b];
    project(P, D, PDP);
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) G[a][b] = R(2) * W * (D[a][b] - PDP[a][b]);
    R g1[5], g2[5];
    mat_grad_to_q(G, g1);
    dTrQ2(q, g2);
    const R w2 = R(2) * W * (trQ2(q) - R(1.5) * S0 * S0);
#pragma unroll
    for (int i = 0; i < 5; ++i) F[i] -= g1[i] + w2 * g2[i];
}

template <typename R>
__device__ __forceinline__ R anchoring_energy(const DevObject& o, const R* q, const R* qb) {
    if (o.type == 0) {
        R d[5];
#pragma unroll
        for (int i = 0; i < 5; ++i) d[i] = q[i] - qb[i];
        return R(o.P1) * trQ2(d);
    }
    const R W = R(o.P1), S0 = R(o.P2);
    const R nu[3] = {qb[0], qb[1], qb[2]};
    R Qt[3][3], P[3][3], D[3][3];
    full3(q, Qt);
#pragma unroll
    for (int a = 0; a < 3; ++a) Qt[a][a] += R(0.5) * S0;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) P[a][b] = (a == b ? R(1) : R(0)) - nu[a] * nu[b];
    project(P, Qt, D);
    R e1 = 0, tq = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const R dd = Qt[a][b] - D[a][b];
            e1 += dd * dd;
            tq += Qt[a][b] * Qt[a][b];
        }
    const R e2 = tq - R(2.25) * S0 * S0;
    return W * e1 + W * e2 * e2;
}

// Distortion energy densities per unit constant from first derivatives:
// e = {1/2 dQ:dQ, 1/2 (div Q)^2, 1/2 d_k Q_ij d_j Q_ik, 1/2 eps_lik Q_lj d_k Q_ij,
//      1/2 Q_lk d_l Q_ij d_k Q_ij}
template <typename R>
__device__ __forceinline__ void distortion_densities(const R* q, const R* dX, const R* dY, const R* dZ, R* e) {
    R Q[3][3], dQ[3][3][3];
    full3(q, Q);
    full3(dX, dQ[0]);
    full3(dY, dQ[1]);
    full3(dZ, dQ[2]);
    R l1 = 0, l2 = 0, l3 = 0, l4 = 0, l6 = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                l1 += dQ[k][i][j] * dQ[k][i][j];
                l3 += dQ[k][i][j] * dQ[j][i][k];
            }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const R div = dQ[0][i][0] + dQ[1][i][1] + dQ[2][i][2];
        l2 += div * div;
    }
    // eps_lik Q_lj d_k Q_ij over the six nonzero Levi-Civita entries
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        l4 += Q[0][j] * (dQ[2][1][j] - dQ[1][2][j]);
        l4 += Q[1][j] * (dQ[0][2][j] - dQ[2][0][j]);
        l4 += Q[2][j] * (dQ[1][0][j] - dQ[0][1][j]);
    }
#pragma unroll
    for (int l = 0; l < 3; ++l)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            R s = 0;
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) s += dQ[l][i][j] * dQ[k][i][j];
            l6 += Q[l][k] * s;
        }
    e[0] = R(0.5) * l1;
    e[1] = R(0.5) * l2;
    e[2] = R(0.5) * l3;
    e[3] = R(0.5) * l4;
    e[4] = R(0.5) * l6;
}

// -(perm/2) F.(iso I + anis Q).F
template <typename R>
__device__ __forceinline__ R field_energy(const double* Fv, double iso, double anis, double perm, const R* q) {
    const R fx = R(Fv[0]), fy = R(Fv[1]), fz = R(Fv[2]);
    const R ff = fx * fx + fy * fy + fz * fz;
    const R fqf = fx * fx * q[0] + fy * fy * q[3] + fz * fz * (-q[0] - q[3]) +
                  R(2) * (fx * fy * q[1] + fx * fz * q[2] + fy * fz * q[4]);
    return R(-0.5) * R(perm) * (R(iso) * ff + R(anis) * fqf);
}

// Force of a spatially varying H field at site i (planes hx, hy, hz of n):
// minus the reference's fieldForce, as computeSpatiallyVaryingFieldCPU
// subtracts it (landauDeGennesLCOtherForces.cpp:254-283); P = deltaChi mu0.
template <typename R>
__device__ __forceinline__ void varying_field_force(const double* hvar, long long n, long long i, double P, R* out) {
    const double hx = hvar[i], hy = hvar[n + i], hz = hvar[2 * n + i];
    out[0] = R(-(-0.5 * P * (hx * hx - hz * hz)));
    out[1] = R(-(-P * hx * hy));
    out[2] = R(-(-P * hx * hz));
    out[3] = R(-(-0.5 * P * (hy * hy - hz * hz)));
    out[4] = R(-(-P * hy * hz));
}

// QT inner product (inc/functions.h:140-147): sum a_i b_i + a_0 b_3
template <typename R>
__device__ __forceinline__ double qt_dot(const R* a, const R* b) {
    double s = 0;
#pragma unroll
    for (int i = 0; i < 5; ++i) s += double(a[i]) * double(b[i]);
    return s + double(a[0]) * double(b[3]);
}

}  // namespace qmin
