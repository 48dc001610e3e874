// gfx950 FIRE pass that keeps only Q and v between iterations.
//
// The fused pass of ldg_kernels.hip reads Q, v and F of every site and writes
// all three (240 B per FP64 site).  F of a bulk site is a function of Q alone,
// so this pass does not store it: it recomputes F_k from Q_k on the tile and a
// one-site halo, applies the pending FIRE velocity mix, the drift and the first
// half kick there (Q_{k+1} = Q_k + dt v' + dt^2/2 F_k, v' = cKeep v_k + cForce F_k,
// energyMinimizerFIRE.cpp:113-116, velocityVerlet.cpp:8-40), computes F_{k+1}
// on the tile from Q_{k+1}, and writes Q_{k+1} and v_{k+1} (160 B per FP64 site).
// Sites that the boundary pass finishes (next to objects, class < 0) keep their
// F in memory: the boundary pass writes it and this pass reads it on the halo.
// With storeF (the last pass of a run of known length) F_{k+1} of the other
// sites is stored too, so that the lattice holds the final state's force.
//
// Streaming: a work-group owns a TX x TY column and marches along z.  Each step
// stages one plane of Q_k over the tile plus two sites on each side into a ring
// of three LDS planes (sA); the plane below it is then complete on the tile
// plus one site on each side: its F_k (composed 19-point operator), mix, drift
// and first half kick give Q_{k+1} there, into a second ring of three planes
// (sB), from which F_{k+1} of the tile's sites one plane further down follows;
// those sites are finished and stored.  All global loads of a step (the next
// plane's Q and class, v and, next to objects, F of the plane drifted next) are
// issued before its second half, so they are in flight while it computes; the
// step then waits for them and only after that issues its stores, which stay in
// flight through the next step's barrier and first half (waiting for a load
// issued after a store would wait for the store as well).  Measured at 256^3 in
// FP64: 0.786 against 0.898 ms per pass with the stores issued before that wait
// (same box).  The composed operator is applied over its few distinct weights
// (ldg_ring_lin_*), read with scalar loads.
#include "qmin_device.h"
#include "ldg_math.h"

namespace qmin {

#include "gen/ldg_stencils.inc"

namespace {

// Points of the composed operator between scheduling barriers.  At 256^3 FP64
// a barrier every 2 points gave 0.802 ms per pass against 0.823-0.825 ms for
// 4, 6 or 8 (and 168 registers with spills for none): more LDS reads in flight
// do not pay, as the pass is bound by latency between its barriers.
constexpr int kSchedPoints = 2;

__device__ __forceinline__ int swrap(int v, int m) {
    v %= m;
    return v < 0 ? v + m : v;
}

__device__ __forceinline__ double swave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <typename R>
using SConstPtr = const R __attribute__((address_space(4)))*;

// one thread per site of the tile plus a two-site margin
template <int TX_, int TY_>
struct StreamTile {
    static constexpr int TX = TX_, TY = TY_;
    static constexpr int HX2 = TX + 4, HY2 = TY + 4, N2 = HX2 * HY2;  // Q_k plane with margin 2
    static constexpr int HX1 = TX + 2, HY1 = TY + 2, N1 = HX1 * HY1;  // Q_{k+1} plane with margin 1
    static constexpr int NT = (N2 + 63) / 64 * 64;
    static_assert(NT <= 1024, "work-group too large");
};

// the two rings of three planes fit in a CU's 160 KB of LDS
template <typename R, int TX, int TY>
constexpr bool stream_fits() {
    using T = StreamTile<TX, TY>;
    return sizeof(R) * 15 * (T::N2 + T::N1) + 24 * (T::NT / 64) <= 160 * 1024;
}

// minimum waves per SIMD asked of the compiler: for the 32 x 16 tile (12
// waves) one work-group per CU.  In FP64 its rings fill the LDS; in FP32 two
// work-groups would fit the LDS but not 80 registers, which spill (0.654 against
// 0.540 ms per pass at 256^3 with one).  The other tiles take what their
// registers give.
template <typename R, int TX, int TY>
constexpr int stream_waves() {
    return (TX == 32 && TY == 16) ? 3 : 1;
}

}  // namespace

// Distortion term of a class-0 site read from a ring of three planes (slots sm,
// sc, sp below, at and above it): the unit L1 Laplacian (scaled and shaped by
// stream_force), or the composed operator over its distinct weights W.
// qc: the site's own value (held in registers by the caller).
template <typename R, int TERMS, int HXS, int NP>
__device__ __forceinline__ void direct_spatial(SConstPtr<R> W, const R (*q)[5][NP], int sm, int sc, int sp, int h,
                                               const R* qc, R* s) {
#pragma unroll
    for (int k = 0; k < 5; ++k) s[k] = R(0);
    R cm[5], cp[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        cm[k] = q[sm][k][h];
        cp[k] = q[sp][k][h];
    }
    if constexpr (TERMS == kOneConstant) {
        // bulkL1Force (lcForces.h:26-37) before its scale and shape: 6 q - faces
#pragma unroll
        for (int k = 0; k < 5; ++k)
            s[k] = R(6) * qc[k] - q[sc][k][h - 1] - q[sc][k][h + 1] - q[sc][k][h - HXS] - q[sc][k][h + HXS] - cm[k] -
                   cp[k];
    } else if constexpr (TERMS == kMultiLinear) {
        ldg_ring_lin_a<HXS, NP, kSchedPoints>(W, q, sm, sc, sp, h, cm, qc, cp, s);
    } else {
        ldg_ring_lin_b<HXS, NP, kSchedPoints>(W, q, sm, sc, sp, h, cm, qc, cp, s);
    }
}

// Force of a class-0 site from its distortion sum s (unit L1 sum in the
// one-constant case): phase - distortion, fields, metric (as finish_bulk of
// the fused pass).
template <typename R, int TERMS>
__device__ __forceinline__ void stream_force(const ForceConst& fc, const R* q, R* s, const double* hvar, long long n,
                                             long long i, double hvarP, R* F) {
#pragma unroll
    for (int k = 0; k < 5; ++k) F[k] = R(0);
    phase_force(R(fc.a), R(fc.b), R(fc.c), q, F);
    if constexpr (TERMS == kOneConstant) {
#pragma unroll
        for (int k = 0; k < 5; ++k) s[k] = R(fc.L1) * s[k];
        l1_shape(s);
    }
#pragma unroll
    for (int k = 0; k < 5; ++k) F[k] -= s[k];
    if (fc.semantics == 0) metric(F);
    if (fc.hasField) {
#pragma unroll
        for (int k = 0; k < 5; ++k) F[k] += R(fc.field[k]);
    }
    if (hvar) {
        R fv[5];
        varying_field_force(hvar, n, i, hvarP, fv);
#pragma unroll
        for (int k = 0; k < 5; ++k) F[k] += fv[k];
    }
    metric(F);
}

// Waits for the step's global loads before its stores are issued.  Vector
// memory loads and stores share one in-order completion counter, so a wait on a
// load issued before a store also waits for that store; with the loads
// complete first, the stores stay in flight through the next step's barrier
// and first half, where no wait on a load follows them.
template <typename R>
__device__ __forceinline__ void wait_loads(const R (&qn)[5], const R (&un)[5], const R (&fb)[5], int cn) {
    asm volatile("" ::"v"(qn[0]), "v"(qn[1]), "v"(qn[2]), "v"(qn[3]), "v"(qn[4]), "v"(cn) : "memory");
    asm volatile("" ::"v"(un[0]), "v"(un[1]), "v"(un[2]), "v"(un[3]), "v"(un[4]) : "memory");
    asm volatile("" ::"v"(fb[0]), "v"(fb[1]), "v"(fb[2]), "v"(fb[3]), "v"(fb[4]) : "memory");
}

template <typename R, int TERMS, int TX, int TY>
__device__ __forceinline__ void stream_body(const MarchArgs<R>& A, const int logical, const int slot) {
    using T = StreamTile<TX, TY>;
    constexpr int HX2 = T::HX2, HY2 = T::HY2, N2 = T::N2, HX1 = T::HX1, N1 = T::N1, NT = T::NT;
    __shared__ R sA[3][5][N2];   // Q_k: planes s-2, s-1, s (rotating slots)
    __shared__ R sB[3][5][N1];   // Q_{k+1}: planes s-3, s-2, s-1
    __shared__ double sRed[NT / 64][3];

    Ctrl* ctrl = A.ctrl;
    if (ctrl->done) return;
    const int cur = ctrl->cur;
    const StateSet<R> in = A.set[cur];
    const StateSet<R> out = A.set[cur ^ 1];
    const double dtd = ctrl->dt;
    const R dt = R(dtd), hdt = R(0.5 * dtd), hdt2 = R(0.5 * dtd * dtd);
    const int reset = ctrl->reset, identity = ctrl->identity;
    const R ck = reset ? R(0) : (identity ? R(1) : R(ctrl->cKeep));
    const R cf = (reset || identity) ? R(0) : R(ctrl->cForce);

    const int txi = logical % A.nbx;
    const int rest = logical / A.nbx;
    const int tyi = rest % A.nby;
    const int zc = rest / A.nby;
    const int z0 = zc * A.zchunk;
    const int z1 = min(z0 + A.zchunk, A.lz);
    const int lx = A.lx, ly = A.ly, lz = A.lz;
    const long long n = A.n;
    const long long plane = (long long)lx * ly;

    const int t = threadIdx.x;
    const int hx = t % HX2, hy = t / HX2;
    const bool valid = t < N2;
    const bool in1 = valid && hx >= 1 && hx <= HX1 && hy >= 1 && hy <= HY2 - 2;
    const bool own = valid && hx >= 2 && hx <= TX + 1 && hy >= 2 && hy <= TY + 1;
    const int gx = txi * TX - 2 + hx, gy = tyi * TY - 2 + hy;
    const bool active = own && gx < lx && gy < ly;
    const long long col = valid ? swrap(gx, lx) + (long long)lx * swrap(gy, ly) : 0;
    const int h1 = (hy - 1) * HX1 + (hx - 1);

    R qn[5];              // Q_k of this column, plane staged next
    int cn = 1;           // its class
    R un[5];              // v_k of this column, plane drifted next
    int cprev = 1;        // class of the plane staged last
    R fb[5];              // F of the plane drifted next if the boundary pass finished its site (class < 0)
    R va[5], vb[5];       // half-kicked v of planes s-2 and s-1 (own sites)
    int cla = 1, clb = 1; // their classes
#pragma unroll
    for (int k = 0; k < 5; ++k) un[k] = fb[k] = va[k] = vb[k] = R(0);
    double pF = 0, pP = 0, pV = 0;
    int a0 = 0, a1 = 1, a2 = 2;   // slots of Q_k planes s-2, s-1, s
    int b0 = 0, b1 = 1, b2 = 2;   // slots of Q_{k+1} planes s-3, s-2, s-1

    {
        const long long i = col + plane * swrap(z0 - 2, lz);
#pragma unroll
        for (int k = 0; k < 5; ++k) qn[k] = valid ? in.q[k * n + i] : R(0);
        cn = valid ? A.cls[i] : 1;
    }
    for (int s = z0 - 2; s <= z1 + 1; ++s) {
        if (valid) {
#pragma unroll
            for (int k = 0; k < 5; ++k) sA[a2][k][t] = qn[k];
        }
        const int cs = cn;
        __syncthreads();
        if (in1 && s >= z0) {
            // plane s-1: F_k, then the FIRE mix, drift and first half kick
            const long long i = col + plane * swrap(s - 1, lz);
            R qc[5], F[5];
#pragma unroll
            for (int k = 0; k < 5; ++k) qc[k] = sA[a1][k][t];
            if (cprev == 0) {
                uint64_t wa = reinterpret_cast<uint64_t>(A.W);
                asm volatile("" : "+s"(wa));
                R sp[5];
                direct_spatial<R, TERMS, HX2, N2>(reinterpret_cast<SConstPtr<R>>(wa) + kFastWeights, sA, a0, a1, a2,
                                                  t, qc, sp);
                stream_force<R, TERMS>(A.fc, qc, sp, A.hvar, n, i, A.hvarP, F);
            } else if (cprev < 0) {
#pragma unroll
                for (int k = 0; k < 5; ++k) F[k] = fb[k];
            } else {
#pragma unroll
                for (int k = 0; k < 5; ++k) F[k] = R(0);
            }
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                const R v = ck * un[k] + cf * F[k];
                sB[b2][k][h1] = qc[k] + (dt * v + hdt2 * F[k]);
                va[k] = vb[k];
                vb[k] = v + hdt * F[k];
            }
            cla = clb;
            clb = cprev;
        }
        cprev = cs;
        // Every global load of the step: in flight during the second half.
        if (s < z1 + 1 && valid) {
            const long long i = col + plane * swrap(s + 1, lz);
#pragma unroll
            for (int k = 0; k < 5; ++k) qn[k] = in.q[k * n + i];
            cn = A.cls[i];
            if (in1) {
                const long long iu = col + plane * swrap(s, lz);
#pragma unroll
                for (int k = 0; k < 5; ++k) un[k] = in.u[k * n + iu];
                if (cprev < 0) {
#pragma unroll
                    for (int k = 0; k < 5; ++k) fb[k] = in.f[k * n + iu];
                }
            }
        }
        __syncthreads();
        // plane s-2: F_{k+1}, second half kick and sums; stored after the wait below
        const long long iw = col + plane * (s - 2);
        R q1[5], u1[5], F1[5];
        int put = 0;   // 1: Q and v; 2: Q, v and F
        if (own && s >= z0 + 2) {
            const long long i = iw;
#pragma unroll
            for (int k = 0; k < 5; ++k) q1[k] = sB[b1][k][h1];
            if (cla >= 0) {
                R F[5];
                if (cla == 0) {
                    uint64_t wa = reinterpret_cast<uint64_t>(A.W);
                    asm volatile("" : "+s"(wa));
                    R sp[5];
                    direct_spatial<R, TERMS, HX1, N1>(reinterpret_cast<SConstPtr<R>>(wa) + kFastWeights, sB, b0, b1,
                                                      b2, h1, q1, sp);
                    stream_force<R, TERMS>(A.fc, q1, sp, A.hvar, n, i, A.hvarP, F);
                } else {
#pragma unroll
                    for (int k = 0; k < 5; ++k) F[k] = R(0);
                }
#pragma unroll
                for (int k = 0; k < 5; ++k) {
                    u1[k] = va[k] + hdt * F[k];
                    F1[k] = F[k];
                }
                if (active) {
                    pF += qt_dot(F, F);
                    pP += qt_dot(F, u1);
                    pV += qt_dot(u1, u1);
                    put = A.storeF ? 2 : 1;
                }
            } else if (active) {
                // the boundary pass finishes v and F of this site
#pragma unroll
                for (int k = 0; k < 5; ++k) u1[k] = va[k];
                put = 1;
            }
        }
        wait_loads(qn, un, fb, cn);
        if (put) {
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                __builtin_nontemporal_store(q1[k], &out.q[k * n + iw]);
                __builtin_nontemporal_store(u1[k], &out.u[k * n + iw]);
            }
            if (put == 2) {
#pragma unroll
                for (int k = 0; k < 5; ++k) __builtin_nontemporal_store(F1[k], &out.f[k * n + iw]);
            }
        }
        const int ta = a0;
        a0 = a1;
        a1 = a2;
        a2 = ta;
        if (s >= z0) {
            const int tb = b0;
            b0 = b1;
            b1 = b2;
            b2 = tb;
        }
    }
    // work-group sums of F.F, F.v, v.v in a fixed order
    pF = swave_sum(pF);
    pP = swave_sum(pP);
    pV = swave_sum(pV);
    const int w = t >> 6, l = t & 63;
    if (l == 0) {
        sRed[w][0] = pF;
        sRed[w][1] = pP;
        sRed[w][2] = pV;
    }
    __syncthreads();
    if (t == 0) {
        double s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
        for (int j = 0; j < NT / 64; ++j) {
            s0 += sRed[j][0];
            s1 += sRed[j][1];
            s2 += sRed[j][2];
        }
        A.partials[3 * (long long)slot + 0] = s0;
        A.partials[3 * (long long)slot + 1] = s1;
        A.partials[3 * (long long)slot + 2] = s2;
    }
}

// work-group order as the fused pass: each XCD a contiguous run of tiles
__device__ __forceinline__ int stream_logical(int b, int nb) {
    const int qb = nb >> 3, rb = nb & 7, xcd = b & 7;
    return xcd * qb + min(xcd, rb) + (b >> 3);
}

template <typename R, int TERMS, int TX, int TY>
__global__ void __launch_bounds__((StreamTile<TX, TY>::NT), (stream_waves<R, TX, TY>())) fire_stream_kernel(MarchArgs<R> A) {
    stream_body<R, TERMS, TX, TY>(A, stream_logical(blockIdx.x, A.nblocks), blockIdx.x);
}

// ---------------------------------------------------------------- host side

template <typename R, int TERMS, int TX, int TY>
static hipError_t launch_stream_t(const MarchArgs<R>& a, hipStream_t s) {
    hipLaunchKernelGGL((fire_stream_kernel<R, TERMS, TX, TY>), dim3(a.nblocks), dim3(StreamTile<TX, TY>::NT), 0, s,
                       a);
    return hipGetLastError();
}

template <typename R, int TX, int TY>
static hipError_t launch_stream_terms(const MarchArgs<R>& a, int terms, hipStream_t s) {
    if constexpr (!stream_fits<R, TX, TY>()) {
        return hipErrorInvalidValue;   // the two LDS rings would not fit a CU
    } else {
        switch (terms) {
            case kOneConstant: return launch_stream_t<R, kOneConstant, TX, TY>(a, s);
            case kMultiLinear: return launch_stream_t<R, kMultiLinear, TX, TY>(a, s);
            case kMultiL4: return launch_stream_t<R, kMultiL4, TX, TY>(a, s);
            default: return hipErrorInvalidValue;
        }
    }
}

// tile: kStream32x16, kStream32x8 or kStream64x8
template <typename R>
hipError_t launch_fire_stream(const MarchArgs<R>& a, int terms, int tile, hipStream_t s) {
    switch (tile) {
        case kStream32x16: return launch_stream_terms<R, 32, 16>(a, terms, s);
        case kStream32x8: return launch_stream_terms<R, 32, 8>(a, terms, s);
        case kStream64x8: return launch_stream_terms<R, 64, 8>(a, terms, s);
        default: return hipErrorInvalidValue;
    }
}

template <typename R, int TX, int TY>
static int stream_occupancy(int terms) {
    if constexpr (!stream_fits<R, TX, TY>()) {
        return 1;
    } else {
        int blocks = 0;
        constexpr int nt = StreamTile<TX, TY>::NT;
        hipError_t e = hipErrorInvalidValue;
        switch (terms) {
            case kOneConstant:
                e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, fire_stream_kernel<R, kOneConstant, TX, TY>,
                                                                 nt, 0);
                break;
            case kMultiLinear:
                e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, fire_stream_kernel<R, kMultiLinear, TX, TY>,
                                                                 nt, 0);
                break;
            case kMultiL4:
                e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, fire_stream_kernel<R, kMultiL4, TX, TY>, nt,
                                                                 0);
                break;
        }
        return (e == hipSuccess && blocks > 0) ? blocks : 1;
    }
}

template <typename R>
int stream_blocks_per_cu(int terms, int tile) {
    switch (tile) {
        case kStream32x16: return stream_occupancy<R, 32, 16>(terms);
        case kStream32x8: return stream_occupancy<R, 32, 8>(terms);
        default: return stream_occupancy<R, 64, 8>(terms);
    }
}

template hipError_t launch_fire_stream<double>(const MarchArgs<double>&, int, int, hipStream_t);
template hipError_t launch_fire_stream<float>(const MarchArgs<float>&, int, int, hipStream_t);
template int stream_blocks_per_cu<double>(int, int);
template int stream_blocks_per_cu<float>(int, int);

}  // namespace qmin
