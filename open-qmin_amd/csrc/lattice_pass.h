// Device code shared by the lattice passes of the qmin HIP library (gfx950):
// the fused pass over halo-padded tiles (march_body) and the boundary pass
// (boundary_body), each compiled per pass mode (force only, FIRE, or the
// force with the Adam or Nesterov update) by the translation units that
// launch them.
#pragma once
#include "qmin_device.h"
#include "ldg_math.h"

namespace qmin {

#include "gen/ldg_stencils.inc"

__device__ __forceinline__ int wrapi(int v, int m) {
    v %= m;
    return v < 0 ? v + m : v;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Block-wide sum of three doubles in a fixed order; result valid in thread 0.
template <int NT>
__device__ __forceinline__ void block_sum3(double& a, double& b, double& c, double (*red)[3]) {
    a = wave_sum(a);
    b = wave_sum(b);
    c = wave_sum(c);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) {
        red[w][0] = a;
        red[w][1] = b;
        red[w][2] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
        for (int i = 0; i < NT / 64; ++i) {
            s0 += red[i][0];
            s1 += red[i][1];
            s2 += red[i][2];
        }
        a = s0;
        b = s1;
        c = s2;
    }
}

// Derivative at neighbour n along one transverse axis, with the reference's
// rules (landauDeGennesLCOtherForces.cpp:15-145): zero on objects, central on
// t == 0 or when neither side is an object, otherwise one-sided with an object
// above taking precedence.
template <typename R>
__device__ __forceinline__ void nbr_derivative(int tn, int tu, int td, const R* nq, const R* qu, const R* qd, R* out) {
    if (tn > 0) {
#pragma unroll
        for (int c = 0; c < 5; ++c) out[c] = R(0);
    } else if (tn == 0 || (td <= 0 && tu <= 0)) {
#pragma unroll
        for (int c = 0; c < 5; ++c) out[c] = R(0.5) * (qu[c] - qd[c]);
    } else if (tu > 0) {
#pragma unroll
        for (int c = 0; c < 5; ++c) out[c] = nq[c] - qd[c];
    } else {
#pragma unroll
        for (int c = 0; c < 5; ++c) out[c] = qu[c] - nq[c];
    }
}

// Raw force (phase - distortion) of a site whose neighbours all have central
// derivatives: the bulk L1..L4 forms composed with those differences into one
// linear 19-point operator (weights W), plus the quadratic L6 form.
// Uniform read-only data seen through the constant address space, so that the
// compiler fetches it with scalar loads.
template <typename R>
using ConstPtr = const R __attribute__((address_space(4)))*;

template <typename R, int TERMS, int HX, int HP>
__device__ __forceinline__ void fast_raw_force(const ForceConst& fc, ConstPtr<R> W, const R (*q)[5][HP],
                                               int sm, int sc, int sp, int h, const R* qc, R* F) {
#pragma unroll
    for (int c = 0; c < 5; ++c) F[c] = R(0);
    phase_force(R(fc.a), R(fc.b), R(fc.c), qc, F);
    R spatial[5] = {R(0), R(0), R(0), R(0), R(0)};
    if constexpr (TERMS == kOneConstant) {
#pragma unroll
        for (int c = 0; c < 5; ++c)
            spatial[c] = R(fc.L1) * (R(6) * qc[c] - q[sc][c][h - 1] - q[sc][c][h + 1] - q[sc][c][h - HX] -
                                     q[sc][c][h + HX] - q[sm][c][h] - q[sp][c][h]);
        l1_shape(spatial);
    } else if constexpr (TERMS == kMultiLinear) {
        ldg_fast_lin_a<HX, HP>(W, q, sm, sc, sp, h, spatial);
    } else if constexpr (TERMS == kMultiL4) {
        ldg_fast_lin_b<HX, HP>(W, q, sm, sc, sp, h, spatial);
    } else {
        R o6[5] = {R(0), R(0), R(0), R(0), R(0)};
        ldg_fast_lin_b_l6<HX, HP>(W, q, sm, sc, sp, h, qc, spatial, o6);
#pragma unroll
        for (int c = 0; c < 5; ++c) spatial[c] += R(fc.L6) * o6[c];
    }
#pragma unroll
    for (int c = 0; c < 5; ++c) F[c] -= spatial[c];
}

// fv: the site's spatially varying field force, or null
template <typename R>
__device__ __forceinline__ void finish_bulk(const ForceConst& fc, R* F, const R* fv) {
    if (fc.semantics == 0) metric(F);
    if (fc.hasField) {
#pragma unroll
        for (int c = 0; c < 5; ++c) F[c] += R(fc.field[c]);
    }
    if (fv) {
#pragma unroll
        for (int c = 0; c < 5; ++c) F[c] += fv[c];
    }
    metric(F);
}

// Stores of the next state: not read again in this pass, so they are streamed
// past L2 (non-temporal), which keeps L2 for the halo rows that neighbouring
// tiles read.
template <typename R>
__device__ __forceinline__ void st_out(R* p, R v) {
    __builtin_nontemporal_store(v, p);
}

// Coefficients of one Adam or Nesterov update, from the minimizer's device
// scalars as the previous iteration's finalize kernel left them.
template <typename R>
struct StepCoef {
    R b1, b1m1, b2, omb2, inv1, inv2, eps, mdt;  // Adam
    R dt, mu;                                     // Nesterov
};

template <typename R, int MODE>
__device__ __forceinline__ StepCoef<R> step_coef(const StepArgs& S) {
    StepCoef<R> k{};
    if constexpr (MODE == kAdamStep) {
        const double b1t = S.scal[0], b2t = S.scal[1];
        k.b1 = R(S.b1);
        k.b2 = R(S.b2);
        k.b1m1 = R(S.b1 - 1);
        k.omb2 = R(1 - S.b2);
        k.inv1 = R(1.0 / (1.0 - b1t));
        k.inv2 = R(1.0 / (1.0 - b2t));
        k.eps = R(S.eps);
        k.mdt = R(-S.dt);
    } else if constexpr (MODE == kNesterovStep) {
        // lambda and the scheduled momentum of this iteration
        // (energyMinimizerNesterovAG.cpp:101-108)
        const double oldLambda = S.scal[0];
        const double lambda = 0.5 * (1.0 + sqrt(1.0 + 4.0 * oldLambda * oldLambda));
        const double gamma = (1 - oldLambda) / lambda;
        k.mu = R(S.scheduled ? 1. + gamma : S.scal[1]);
        k.dt = R(S.dt);
    }
    return k;
}

// The minimizer's own entries of site i (Adam: m and s; Nesterov: the
// alternate sequence), loaded before the site's force is computed so that the
// force computation covers their latency.
template <typename R, int MODE>
__device__ __forceinline__ void step_load(const StepArgs& S, long long n, long long i, R (&pre)[2][5]) {
    const R* m = static_cast<const R*>(S.m);
#pragma unroll
    for (int c = 0; c < 5; ++c) pre[0][c] = m[c * n + i];
    if constexpr (MODE == kAdamStep) {
        const R* v = static_cast<const R*>(S.v);
#pragma unroll
        for (int c = 0; c < 5; ++c) pre[1][c] = v[c * n + i];
    }
}

// The update of site i from its force F and its entries pre (step_load):
// energyMinimizerAdam::adamStepCPU (energyMinimizerAdam.cpp:25-54; epsilon
// replaces sqrt(vc) where that is 0) or energyMinimizerNesterovAG::
// nesterovStepCPU (energyMinimizerNesterovAG.cpp:75-92).  Stores the site's
// advanced entries (no other work-item reads them in the pass) and returns the
// new Q in qn.
template <typename R, int MODE>
__device__ __forceinline__ void step_site(const StepArgs& S, const StepCoef<R>& k, long long n, long long i,
                                          const R* q, const R* F, const R (&pre)[2][5], R* qn) {
    R* m = static_cast<R*>(S.m);
#pragma unroll
    for (int c = 0; c < 5; ++c) {
        const long long j = c * n + i;
        if constexpr (MODE == kAdamStep) {
            R* v = static_cast<R*>(S.v);
            const R mm = k.b1 * pre[0][c] + k.b1m1 * F[c];
            const R vv = k.b2 * pre[1][c] + k.omb2 * (F[c] * F[c]);
            const R mc = mm * k.inv1;
            const R vc = vv * k.inv2;
            R root = sqrt(vc);
            if (root == R(0)) root = k.eps;
            st_out(&m[j], mm);
            st_out(&v[j], vv);
            qn[c] = q[c] + k.mdt * mc / root;
        } else {
            const R old = pre[0][c];
            const R alt = q[c] + k.dt * F[c];
            st_out(&m[j], alt);
            qn[c] = alt + k.mu * (alt - old);
        }
    }
}

// Sum of squares of the five stored components (the minimizers' |F|^2).
template <typename R>
__device__ __forceinline__ double sq5(const R* F) {
    double s = 0;
#pragma unroll
    for (int c = 0; c < 5; ++c) s += double(F[c]) * double(F[c]);
    return s;
}

constexpr bool is_step_mode(int mode) { return mode == kAdamStep || mode == kNesterovStep; }

// Whether the fused pass loads the minimizer's entries before the site's force
// computation (their latency hidden behind it) or right before the update.
// Early loads keep 5 or 10 more values live across the force computation: in
// FP64 with L6 that costs a wave per SIMD (Adam 202 VGPRs against 146), and
// the late load is faster there (per iteration at 384^3: Adam 4.68 against
// 5.29 ms, Nesterov 3.51 against 4.33 ms); everywhere else the early load is
// (FP32 with L6: 2.30 against 2.51 and 1.85 against 2.02 ms; 256^3 with
// L1..L3: FP64 0.74 against 0.78 ms for Nesterov, FP32 0.56 against 0.59 ms
// for Adam).
template <typename R, int TERMS>
constexpr bool step_early_load() { return !(sizeof(R) == 8 && TERMS == kMultiL6); }

// FP32 force-only passes on the 512-thread tiles, without L6 (55-70 of 128
// registers): the next plane but one is loaded into registers before the force
// of the current one, and written to LDS a step later; each step waits for
// those loads before it stores its forces, so that the stores drain behind the
// next step's barrier instead of being waited for with the next plane's loads
// (vector memory operations complete in order on one counter).  At 256^3 a
// force computation (pass, boundary pass, synchronisation) takes 0.223 against
// 0.281 ms so; in FP64 (107 registers against 76) 0.369 against 0.365 ms, and
// FP64 keeps loading each plane at the top of its step.
template <typename R, int MODE, int TERMS, int TILE>
constexpr bool march_prefetch() {
    return sizeof(R) == 4 && MODE == kForce && TERMS != kMultiL6 && kTileThreads[TILE] == 512;
}

template <typename R>
__device__ __forceinline__ void wait_plane(const R (&a)[5], const R (&b)[5], int c) {
    asm volatile("" ::"v"(a[0]), "v"(a[1]), "v"(a[2]), "v"(a[3]), "v"(a[4]), "v"(c) : "memory");
    asm volatile("" ::"v"(b[0]), "v"(b[1]), "v"(b[2]), "v"(b[3]), "v"(b[4]) : "memory");
}

// One work-group's share of the fused pass: the column of tile `logical` (x
// tile fastest, then y tile, then z chunk) of the lattice described by A; its
// partial sums go to slot `slot`.
template <typename R, int MODE, int TERMS, int TILE>
__device__ __forceinline__ void march_body(const MarchArgs<R>& A, const int logical, const int slot) {
    using T = Tile<kTileTX[TILE], kTileTY[TILE]>;
    constexpr int kTX = T::TX, kTY = T::TY, kHX = T::HX, kHY = T::HY, kHP = T::HP, kHalo = T::Halo;
    __shared__ R sQ[3][5][kHP];
    __shared__ double sRed[T::Threads / 64][3];

    Ctrl* ctrl = A.ctrl;
    if ((MODE != kForce || A.gate) && ctrl->done) return;
    const int cur = ctrl->cur;
    const StateSet<R> in = A.set[cur];
    const StateSet<R> out = A.set[MODE == kForce ? cur : cur ^ 1];
    // pending FIRE velocity mix v = ck u + cf F: (1, 0) when none is pending,
    // (0, 0) after a reset (energyMinimizerFIRE.cpp:117-144)
    R dt = 0, hdt = 0, hdt2 = 0, ck = 0, cf = 0;
    if (MODE == kFire) {
        const double dtd = ctrl->dt;
        dt = R(dtd);
        hdt = R(0.5 * dtd);
        hdt2 = R(0.5 * dtd * dtd);
        const int reset = ctrl->reset, identity = ctrl->identity;
        ck = reset ? R(0) : (identity ? R(1) : R(ctrl->cKeep));
        cf = (reset || identity) ? R(0) : R(ctrl->cForce);
    }
    StepCoef<R> sk{};
    if constexpr (is_step_mode(MODE)) sk = step_coef<R, MODE>(A.step);

    const int txi = logical % A.nbx;
    const int rest = logical / A.nbx;
    const int tyi = rest % A.nby;
    const int zc = rest / A.nby;
    const int z0 = zc * A.zchunk;
    const int z1 = min(z0 + A.zchunk, A.lz);

    const int lx = A.lx, ly = A.ly, lz = A.lz;
    const long long n = A.n;
    const long long plane = (long long)lx * ly;
    const int tid = threadIdx.x;
    const int ox = tid % kTX, oy = tid / kTX;
    const int x = txi * kTX + ox, y = tyi * kTY + oy;
    const bool active = x < lx && y < ly;
    const int hOwn = (oy + 1) * kHX + (ox + 1);
    const long long colOwn = wrapi(x, lx) + (long long)lx * wrapi(y, ly);
    int hHalo = -1;
    long long colHalo = 0;
    if (tid < kHalo) {
        const int k = tid;
        if (k < kHX)
            hHalo = k;
        else if (k < 2 * kHX)
            hHalo = (kHY - 1) * kHX + (k - kHX);
        else if (k < 2 * kHX + (kHY - 2))
            hHalo = (1 + k - 2 * kHX) * kHX;
        else
            hHalo = (1 + k - 2 * kHX - (kHY - 2)) * kHX + (kHX - 1);
        const int hx = hHalo % kHX, hy = hHalo / kHX;
        colHalo = wrapi(txi * kTX - 1 + hx, lx) + (long long)lx * wrapi(tyi * kTY - 1 + hy, ly);
    }

    // Stage plane z into slot s (FIRE: pending mix, drift and first half kick);
    // returns the owner's half-kicked velocity and class.
    auto stage = [&](int z, int s, R* vh, int& cOwn) {
        const long long zoff = plane * wrapi(z, lz);
        {
            const long long i = colOwn + zoff;
            cOwn = A.cls[i];
            R q[5];
#pragma unroll
            for (int c = 0; c < 5; ++c) q[c] = in.q[c * n + i];
            if (MODE == kFire) {
                R u[5], f[5];
#pragma unroll
                for (int c = 0; c < 5; ++c) {
                    u[c] = in.u[c * n + i];
                    f[c] = in.f[c * n + i];
                }
#pragma unroll
                for (int c = 0; c < 5; ++c) {
                    const R v = ck * u[c] + cf * f[c];
                    q[c] = q[c] + (dt * v + hdt2 * f[c]);
                    vh[c] = v + hdt * f[c];
                }
            }
#pragma unroll
            for (int c = 0; c < 5; ++c) sQ[s][c][hOwn] = q[c];
        }
        if (hHalo >= 0) {
            const long long i = colHalo + zoff;
            R q[5];
#pragma unroll
            for (int c = 0; c < 5; ++c) q[c] = in.q[c * n + i];
            if (MODE == kFire) {
                R u[5], f[5];
#pragma unroll
                for (int c = 0; c < 5; ++c) {
                    u[c] = in.u[c * n + i];
                    f[c] = in.f[c * n + i];
                }
#pragma unroll
                for (int c = 0; c < 5; ++c) {
                    const R v = ck * u[c] + cf * f[c];
                    q[c] = q[c] + (dt * v + hdt2 * f[c]);
                }
            }
#pragma unroll
            for (int c = 0; c < 5; ++c) sQ[s][c][hHalo] = q[c];
        }
    };

    constexpr bool kPF = march_prefetch<R, MODE, TERMS, TILE>();
    R pq[5], ph[5];   // kPF: Q of the next plane but one (own site, halo site) and the own class
    int pc = 0;
    auto fetch = [&](int z) {
        const long long zoff = plane * wrapi(z, lz);
        const long long i = colOwn + zoff;
        pc = A.cls[i];
#pragma unroll
        for (int c = 0; c < 5; ++c) pq[c] = in.q[c * n + i];
        if (hHalo >= 0) {
            const long long j = colHalo + zoff;
#pragma unroll
            for (int c = 0; c < 5; ++c) ph[c] = in.q[c * n + j];
        }
    };

    double pF = 0, pP = 0, pV = 0;
    R vhCur[5], vhNext[5];
    int cCur = 0, cNext = 0;
    stage(z0 - 1, 0, vhNext, cNext);
    stage(z0, 1, vhCur, cCur);
    if constexpr (kPF) {
        // waited for here, so that the loop's head waits for nothing
        fetch(z0 + 1);
        wait_plane(pq, ph, pc);
    }
    for (int z = z0; z < z1; ++z) {
        const int sm = (z - z0) % 3, sc = (z - z0 + 1) % 3, sp = (z - z0 + 2) % 3;
        if constexpr (kPF) {
#pragma unroll
            for (int c = 0; c < 5; ++c) sQ[sp][c][hOwn] = pq[c];
            if (hHalo >= 0) {
#pragma unroll
                for (int c = 0; c < 5; ++c) sQ[sp][c][hHalo] = ph[c];
            }
            cNext = pc;
            if (z + 1 < z1) fetch(z + 2);
        } else {
            stage(z + 1, sp, vhNext, cNext);
        }
        __syncthreads();
        R Fst[5];
        bool putF = false;
        if (active) {
            const long long i = colOwn + plane * z;
            R pre[2][5];
            if constexpr (is_step_mode(MODE) && step_early_load<R, TERMS>())
                if (cCur >= 0) step_load<R, MODE>(A.step, n, i, pre);
            R q[5];
#pragma unroll
            for (int c = 0; c < 5; ++c) q[c] = sQ[sc][c][hOwn];
            if (cCur >= 0) {
                R F[5];
                if (cCur == 0) {
                    // re-opaque the weight pointer every step: the loads stay
                    // inside the loop instead of being hoisted and spilled
                    uint64_t wa = reinterpret_cast<uint64_t>(A.W);
                    asm volatile("" : "+s"(wa));
                    fast_raw_force<R, TERMS, kHX, kHP>(A.fc, reinterpret_cast<ConstPtr<R>>(wa), sQ, sm, sc, sp,
                                                       hOwn, q, F);
                    if (A.hvar) {
                        R fv[5];
                        varying_field_force(A.hvar, n, i, A.hvarP, fv);
                        finish_bulk(A.fc, F, fv);
                    } else {
                        finish_bulk<R>(A.fc, F, nullptr);
                    }
                } else {
#pragma unroll
                    for (int c = 0; c < 5; ++c) F[c] = R(0);
                }
                if (MODE == kFire) {
                    R u[5];
#pragma unroll
                    for (int c = 0; c < 5; ++c) u[c] = vhCur[c] + hdt * F[c];
#pragma unroll
                    for (int c = 0; c < 5; ++c) {
                        st_out(&out.q[c * n + i], q[c]);
                        st_out(&out.u[c * n + i], u[c]);
                        st_out(&out.f[c * n + i], F[c]);
                    }
                    pF += qt_dot(F, F);
                    pP += qt_dot(F, u);
                    pV += qt_dot(u, u);
                } else if constexpr (is_step_mode(MODE)) {
                    R qn[5];
                    if constexpr (!step_early_load<R, TERMS>()) step_load<R, MODE>(A.step, n, i, pre);
                    step_site<R, MODE>(A.step, sk, n, i, q, F, pre, qn);
#pragma unroll
                    for (int c = 0; c < 5; ++c) {
                        st_out(&out.q[c * n + i], qn[c]);
                        st_out(&out.f[c * n + i], F[c]);
                    }
                    pF += sq5(F);
                } else if constexpr (kPF) {
#pragma unroll
                    for (int c = 0; c < 5; ++c) Fst[c] = F[c];
                    putF = true;
                } else {
#pragma unroll
                    for (int c = 0; c < 5; ++c) out.f[c * n + i] = F[c];
                }
            } else if (MODE == kFire) {
                // the boundary pass finishes U and F of this site
#pragma unroll
                for (int c = 0; c < 5; ++c) {
                    st_out(&out.q[c * n + i], q[c]);
                    st_out(&out.u[c * n + i], vhCur[c]);
                }
            }
        }
        if constexpr (kPF) {
            wait_plane(pq, ph, pc);
            if (putF) {
                const long long i = colOwn + plane * z;
#pragma unroll
                for (int c = 0; c < 5; ++c) out.f[c * n + i] = Fst[c];
            }
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < 5; ++c) vhCur[c] = vhNext[c];
        cCur = cNext;
    }
    if (MODE != kForce) {
        block_sum3<T::Threads>(pF, pP, pV, sRed);
        if (tid == 0) {
            A.partials[3 * (long long)slot + 0] = pF;
            A.partials[3 * (long long)slot + 1] = pP;
            A.partials[3 * (long long)slot + 2] = pV;
        }
    }
}

// XCD-aware work-group order: consecutive hardware ids go round-robin over the
// 8 XCDs; each XCD gets a contiguous run of logical tiles, so that tiles sharing
// halos share an L2.  With PAIR, within each run of 64 tiles of an XCD its
// first 32 work-groups take the even tiles and the next 32 the odd ones, so
// that, if the XCD's 32 CUs receive its work-groups in order (two each),
// x-neighbouring tiles share a CU.  Measured at 256^3: 2.6% faster for the FP32
// 64x8 pass; 27% slower, with 64% more HBM reads, for the FP64 32x16 pass,
// which keeps the plain order.
template <bool PAIR>
__device__ __forceinline__ int xcd_logical(int b, int nb) {
    const int qb = nb >> 3, rb = nb & 7, xcd = b & 7;
    int j = b >> 3;
    if constexpr (PAIR) {
        const int per = qb + (xcd < rb ? 1 : 0);
        if (j < (per & ~63)) j = (j & ~63) + ((j & 31) << 1) + ((j >> 5) & 1);
    }
    return xcd * qb + min(xcd, rb) + j;
}

// Waves per SIMD asked of the compiler for the 256-thread tiles: four (at most
// 128 VGPRs) for the FP64 force-only pass with L6 on 32 x 8 (its tile by
// default), which needs 132 otherwise and
// spills 4 at 128, and is 2% faster so at 384^3 (2.68-2.69 against 2.74-2.77 ms);
// the Adam and Nesterov passes with L6 spill 18 and 11 registers at 128 and are
// slower (5.31 against 4.69 ms, 4.00 against 3.55 ms), so they and the rest keep
// what their registers give.
template <typename R, int MODE, int TERMS, int TILE>
constexpr int march_min_waves() {
    return (sizeof(R) == 8 && TERMS == kMultiL6 && MODE == kForce && TILE == kTile32x8) ? 4 : 1;
}

template <typename R, int MODE, int TERMS, int TILE>
__global__ void __launch_bounds__(kTileThreads[TILE], (march_min_waves<R, MODE, TERMS, TILE>())) march_kernel(MarchArgs<R> A) {
    march_body<R, MODE, TERMS, TILE>(A, xcd_logical<false>(blockIdx.x, A.nblocks), blockIdx.x);
}

// The 512-thread tiles (32 x 16, 64 x 8; 8 waves): two work-groups per CU need
// at least four waves per SIMD, i.e. at most 128 VGPRs.
template <typename R, int MODE, int TERMS, int TILE>
__global__ void __launch_bounds__(kTileThreads[TILE], 4) march_wide_kernel(MarchArgs<R> A) {
    static_assert(kTileThreads[TILE] == 512, "512-thread tiles only");
    constexpr bool kPair = sizeof(R) == 4 && TILE == kTile64x8;
    march_body<R, MODE, TERMS, TILE>(A, xcd_logical<kPair>(blockIdx.x, A.nblocks), blockIdx.x);
}

// The fused pass over many lattices in one launch: map[b] = {lattice, logical
// tile} for hardware work-group b, laid out by the host in XCD-aware order.
template <typename R, int MODE, int TERMS, int TILE>
__global__ void __launch_bounds__(kTileThreads[TILE]) march_batch_kernel(const MarchArgs<R>* __restrict__ descs,
                                                               const int2* __restrict__ map) {
    // uniform over the work-group: read once, keep in scalar registers
    const int lat = __builtin_amdgcn_readfirstlane(map[blockIdx.x].x);
    const int logical = __builtin_amdgcn_readfirstlane(map[blockIdx.x].y);
    march_body<R, MODE, TERMS, TILE>(descs[lat], logical, logical);
}

template <typename R, int MODE, int TERMS, int TILE>
__global__ void __launch_bounds__(kTileThreads[TILE], 4) march_wide_batch_kernel(
    const MarchArgs<R>* __restrict__ descs, const int2* __restrict__ map) {
    static_assert(kTileThreads[TILE] == 512, "512-thread tiles only");
    const int lat = __builtin_amdgcn_readfirstlane(map[blockIdx.x].x);
    const int logical = __builtin_amdgcn_readfirstlane(map[blockIdx.x].y);
    march_body<R, MODE, TERMS, TILE>(descs[lat], logical, logical);
}

// ---------------------------------------------------------------- boundary pass

template <typename R>
__device__ __forceinline__ void gload(const R* base, long long n, long long i, R* v) {
#pragma unroll
    for (int c = 0; c < 5; ++c) v[c] = base[c * n + i];
}

// Work-group `blk` of the boundary pass of the lattice described by A.
template <typename R, int MODE, int TERMS>
__device__ __forceinline__ void boundary_body(const BoundaryArgs<R>& A, const int blk) {
    __shared__ double sRed[kBoundaryThreads / 64][3];
    Ctrl* ctrl = A.ctrl;
    if ((MODE != kForce || A.gate) && ctrl->done) return;
    const int cur = ctrl->cur;
    // FIRE: the drifted state the fused pass has written; otherwise the current one
    const StateSet<R> st = A.set[MODE == kFire ? cur ^ 1 : cur];
    const R hdt = MODE == kFire ? R(0.5 * ctrl->dt) : R(0);
    StepCoef<R> sk{};
    if constexpr (is_step_mode(MODE)) sk = step_coef<R, MODE>(A.step);
    const long long n = A.n;
    const int lx = A.lx, ly = A.ly, lz = A.lz;
    double pF = 0, pP = 0, pV = 0;

    const long long k = (long long)blk * kBoundaryThreads + threadIdx.x;
    if (k < A.nsites) {
        const long long i = A.sites[k];
        R pre[2][5];
        if constexpr (is_step_mode(MODE)) step_load<R, MODE>(A.step, n, i, pre);
        const int self = A.types[i];
        const int x = int(i % lx), y = int((i / lx) % ly), z = int(i / ((long long)lx * ly));
        // neighbour sites in the reference order xd, xu, yd, yu, zd, zu
        long long nbi[6];
        nbi[0] = wrapi(x - 1, lx) + (long long)lx * (y + (long long)ly * z);
        nbi[1] = wrapi(x + 1, lx) + (long long)lx * (y + (long long)ly * z);
        nbi[2] = x + (long long)lx * (wrapi(y - 1, ly) + (long long)ly * z);
        nbi[3] = x + (long long)lx * (wrapi(y + 1, ly) + (long long)ly * z);
        nbi[4] = x + (long long)lx * (y + (long long)ly * wrapi(z - 1, lz));
        nbi[5] = x + (long long)lx * (y + (long long)ly * wrapi(z + 1, lz));
        R xin[125];
        int tn[6];
        gload(st.q, n, i, xin);
#pragma unroll
        for (int d = 0; d < 6; ++d) {
            tn[d] = A.types[nbi[d]];
            gload(st.q, n, nbi[d], xin + 5 + 5 * d);
        }
        // transverse derivatives of each neighbour (along-axis entries are unused)
#pragma unroll
        for (int d = 0; d < 6; ++d) {
            const int axis = d / 2;
            const int nx = d == 0 ? wrapi(x - 1, lx) : (d == 1 ? wrapi(x + 1, lx) : x);
            const int ny = d == 2 ? wrapi(y - 1, ly) : (d == 3 ? wrapi(y + 1, ly) : y);
            const int nz = d == 4 ? wrapi(z - 1, lz) : (d == 5 ? wrapi(z + 1, lz) : z);
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) {
                R* out = xin + 35 + 15 * d + 5 * ax;
                if (ax == axis) {
#pragma unroll
                    for (int c = 0; c < 5; ++c) out[c] = R(0);
                    continue;
                }
                const long long iu = ax == 0 ? wrapi(nx + 1, lx) + (long long)lx * (ny + (long long)ly * nz)
                                   : ax == 1 ? nx + (long long)lx * (wrapi(ny + 1, ly) + (long long)ly * nz)
                                             : nx + (long long)lx * (ny + (long long)ly * wrapi(nz + 1, lz));
                const long long idn = ax == 0 ? wrapi(nx - 1, lx) + (long long)lx * (ny + (long long)ly * nz)
                                    : ax == 1 ? nx + (long long)lx * (wrapi(ny - 1, ly) + (long long)ly * nz)
                                              : nx + (long long)lx * (ny + (long long)ly * wrapi(nz - 1, lz));
                R qu[5], qd[5];
                gload(st.q, n, iu, qu);
                gload(st.q, n, idn, qd);
                nbr_derivative(tn[d], int(A.types[iu]), int(A.types[idn]), xin + 5 + 5 * d, qu, qd, out);
            }
        }
        const R* q = xin;
        R F[5];
#pragma unroll
        for (int c = 0; c < 5; ++c) F[c] = R(0);
        phase_force(R(A.fc.a), R(A.fc.b), R(A.fc.c), q, F);
        R spatial[5];
        if (self == -1) {
            // boundaryL1Force: only object-facing directions (lcForces.h:39-63)
#pragma unroll
            for (int c = 0; c < 5; ++c) spatial[c] = R(0);
#pragma unroll
            for (int d = 0; d < 6; ++d)
                if (tn[d] > 0) {
                    const R* opp = xin + 5 + 5 * (d ^ 1);
#pragma unroll
                    for (int c = 0; c < 5; ++c) spatial[c] -= (opp[c] - q[c]);
                }
#pragma unroll
            for (int c = 0; c < 5; ++c) spatial[c] *= R(A.fc.L1);
        } else {
#pragma unroll
            for (int c = 0; c < 5; ++c) {
                R lap = R(6) * q[c];
#pragma unroll
                for (int d = 0; d < 6; ++d) lap -= xin[5 + 5 * d + c];
                spatial[c] = R(A.fc.L1) * lap;
            }
        }
        l1_shape(spatial);
        if constexpr (TERMS != kOneConstant) {
            R acc[3][5];
#pragma unroll
            for (int t = 0; t < 3; ++t)
#pragma unroll
                for (int c = 0; c < 5; ++c) acc[t][c] = R(0);
            int bcase = 0;
            if (self == -1) {
                int xx = 0, yy = 0, zz = 0;
                if (tn[1] > 0) xx = 1;
                if (tn[0] > 0) xx = 2;
                if (tn[3] > 0) yy = 1;
                if (tn[2] > 0) yy = 2;
                if (tn[5] > 0) zz = 1;
                if (tn[4] > 0) zz = 2;
                bcase = xx + 4 * yy + 16 * zz;
            }
            ldg_boundary_lin(bcase, xin, acc);
#pragma unroll
            for (int c = 0; c < 5; ++c) spatial[c] += R(A.fc.L2) * acc[0][c];
#pragma unroll
            for (int c = 0; c < 5; ++c) spatial[c] += R(A.fc.L3) * acc[1][c];
#pragma unroll
            for (int c = 0; c < 5; ++c) spatial[c] += R(A.fc.L4) * acc[2][c];
            if constexpr (TERMS == kMultiL6) {
                R o6[5];
                ldg_boundary_l6(bcase, xin, o6);
#pragma unroll
                for (int c = 0; c < 5; ++c) spatial[c] += R(A.fc.L6) * o6[c];
            }
        }
#pragma unroll
        for (int c = 0; c < 5; ++c) F[c] -= spatial[c];
        R fv[5];
        if (A.hvar) varying_field_force(A.hvar, n, i, A.hvarP, fv);
        if (self == 0) {
            // a bulk site with a one-sided neighbour derivative: bulk semantics
            finish_bulk(A.fc, F, A.hvar ? fv : nullptr);
        } else {
#pragma unroll
            for (int d = 0; d < 6; ++d)
                if (tn[d] > 0) add_anchoring_force(A.objects[tn[d] - 1], q, xin + 5 + 5 * d, F);
            if (A.fc.hasField) {
#pragma unroll
                for (int c = 0; c < 5; ++c) F[c] += R(A.fc.field[c]);
            }
            if (A.hvar) {
#pragma unroll
                for (int c = 0; c < 5; ++c) F[c] += fv[c];
            }
            metric(F);
        }
        if (MODE == kFire) {
            R u[5];
            gload(st.u, n, i, u);
#pragma unroll
            for (int c = 0; c < 5; ++c) {
                u[c] = u[c] + hdt * F[c];
                st.u[c * n + i] = u[c];
                st.f[c * n + i] = F[c];
            }
            pF = qt_dot(F, F);
            pP = qt_dot(F, u);
            pV = qt_dot(u, u);
        } else if constexpr (is_step_mode(MODE)) {
            const StateSet<R> out = A.set[cur ^ 1];
            R qn[5];
            step_site<R, MODE>(A.step, sk, n, i, q, F, pre, qn);
#pragma unroll
            for (int c = 0; c < 5; ++c) {
                out.q[c * n + i] = qn[c];
                out.f[c * n + i] = F[c];
            }
            pF = sq5(F);
        } else {
#pragma unroll
            for (int c = 0; c < 5; ++c) st.f[c * n + i] = F[c];
        }
    }
    if (MODE == kForce) return;
    block_sum3<kBoundaryThreads>(pF, pP, pV, sRed);
    if (threadIdx.x == 0) {
        const long long slot = A.marchBlocks + blk;
        A.partials[3 * slot + 0] = pF;
        A.partials[3 * slot + 1] = pP;
        A.partials[3 * slot + 2] = pV;
    }
}

template <typename R, int MODE, int TERMS>
__global__ void __launch_bounds__(kBoundaryThreads) boundary_kernel(BoundaryArgs<R> A) {
    boundary_body<R, MODE, TERMS>(A, blockIdx.x);
}

template <typename R, int MODE, int TERMS>
__global__ void __launch_bounds__(kBoundaryThreads) boundary_batch_kernel(
    const BoundaryArgs<R>* __restrict__ descs, const int2* __restrict__ map) {
    const int lat = __builtin_amdgcn_readfirstlane(map[blockIdx.x].x);
    const int blk = __builtin_amdgcn_readfirstlane(map[blockIdx.x].y);
    boundary_body<R, MODE, TERMS>(descs[lat], blk);
}

}  // namespace qmin
