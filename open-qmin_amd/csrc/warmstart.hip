// Learned warm start: a small fully-convolutional 3-D network maps a lattice's
// geometry (site types, anchoring tensors and strengths, a chosen far-field
// director) to an initial Q field, from which FIRE then minimizes.  Random-start
// FIRE stays the fallback and decides the final state.
//
// Kernels (BF16 operands on the matrix cores, FP32 accumulation):
//  * conv_fwd_kernel: 3x3x3 dilated periodic convolution as an implicit GEMM,
//    D[site][cout] = sum_{tap, cin} X[site + dil*tap][cin] Wt[tap][cout][cin],
//    on the matrix cores: a wave owns 64 consecutive sites and all output
//    channels; with 64 input channels each lane feeds 8 consecutive input
//    channels of one neighbour site to v_mfma_f32_16x16x32_bf16 (activations
//    are channels-last, so that is one 16-byte load), with 16 or 32 input
//    channels 4 of them to v_mfma_f32_16x16x16_bf16; the tap's weights come
//    from LDS.
//    Optional ReLU on the input as it is read, bias, and an output mask (the
//    ReLU derivative of the previous layer when the kernel runs backwards).
//  * The input gradient is the same kernel with the taps mirrored and the
//    weights transposed (Wbd[26 - tap][cin][cout] = Wt[tap][cout][cin]).
//  * conv_wgrad_kernel: dWt[tap][cout][cin] = sum_site G[site][cout]
//    X[site + dil*tap][cin], a GEMM over sites per (tap, chunk of sites); the
//    chunks' partial sums are added in a fixed order by wgrad_reduce_kernel.
//  * Elementwise kernels: features, loss gradient, bias gradient, Adam with the
//    BF16 (and mirrored) weight copies, writing the prediction into a lattice.
#include <cstdint>

#include "qmin_internal.h"

namespace qmin {
namespace ws {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef unsigned v4u __attribute__((ext_vector_type(4)));

// BF16 pairs with the sign bit set (negative values and -0) -> +0
__device__ __forceinline__ unsigned relu2(unsigned w) {
    const unsigned lo = (w & 0x8000u) ? 0xffffu : 0u, hi = (w & 0x80000000u) ? 0xffff0000u : 0u;
    return w & ~(lo | hi);
}

constexpr int kTaps = 27;
constexpr int kWsThreads = 256;

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }

// round to nearest even
__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}

__device__ __forceinline__ int wrapd(int v, int m) {
    v %= m;
    return v < 0 ? v + m : v;
}

__device__ __forceinline__ long long tap_site(int x, int y, int z, int tap, int dil, int lx, int ly, int lz) {
    const int dx = tap % 3 - 1, dy = (tap / 3) % 3 - 1, dz = tap / 9 - 1;
    return wrapd(x + dil * dx, lx) + (long long)lx * (wrapd(y + dil * dy, ly) + (long long)ly * wrapd(z + dil * dz, lz));
}

struct ConvArgs {
    const uint16_t* x;     // [n][cin] bf16
    const uint16_t* wt;    // [27][cout][cin] bf16
    const float* bias;     // [cout] or null
    const uint16_t* mask;  // [n][cout] bf16: output zeroed where mask <= 0 (null: none)
    uint16_t* ybf;         // [n][cout] bf16 output, or
    float* yf;             // [n][cout] fp32 output
    long long n;
    int cin, cout, lx, ly, lz, dil, reluIn;
};

constexpr int kFwdMT = 4;   // 16-site row tiles per wave in the forward kernel

// Each wave owns 64 consecutive sites (kFwdMT tiles of 16) and all output
// channels: per tap and 16-channel slice it gathers one 8-byte A fragment per
// tile and reads each B fragment once from LDS for all four tiles.  The
// work-group stages a tap's weights [cout][cin] in LDS (rows padded by 4 BF16
// against bank conflicts), double-buffered across taps.
template <int NT, int KT>
__global__ void __launch_bounds__(kWsThreads) conv_fwd_kernel(ConvArgs a) {
    // K32: 32 input channels per matrix instruction (for 64 input channels; with
    // 32, one such step per tap ran 9% slower at 128^3 than two 16-channel
    // ones); LDS rows padded by 8 BF16 so that every lane's 16-byte weight
    // fragment stays 16-byte aligned
    constexpr bool K32 = KT == 4;
    constexpr int CO = NT * 16, CI = KT * 16, RS = CI + (K32 ? 8 : 4);   // LDS row stride in BF16
    __shared__ __align__(16) uint2 sW[2][CO * RS / 4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int kq = lane >> 4, l16 = lane & 15;
    const long long wbase = ((long long)blockIdx.x * (kWsThreads / 64) + wave) * (16 * kFwdMT);
    const bool waveActive = wbase < a.n;
    int xs[kFwdMT], ys[kFwdMT], zs[kFwdMT];
    bool valid[kFwdMT];
#pragma unroll
    for (int mt = 0; mt < kFwdMT; ++mt) {
        const long long s = wbase + 16 * mt + l16;
        valid[mt] = s < a.n;
        const long long sc = valid[mt] ? s : 0;
        xs[mt] = int(sc % a.lx);
        ys[mt] = int((sc / a.lx) % a.ly);
        zs[mt] = int(sc / ((long long)a.lx * a.ly));
    }
    v4f acc[kFwdMT][NT];
#pragma unroll
    for (int mt = 0; mt < kFwdMT; ++mt)
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[mt][t] = v4f{0.f, 0.f, 0.f, 0.f};

    auto stage = [&](int tap, int buf) {
        const uint2* src = reinterpret_cast<const uint2*>(a.wt + (long long)tap * CO * CI);
        for (int e = threadIdx.x; e < CO * CI / 4; e += kWsThreads) {
            const int co = e / (CI / 4), c4 = e % (CI / 4);
            sW[buf][co * (RS / 4) + c4] = src[e];
        }
    };
    stage(0, 0);
    __syncthreads();
    for (int tap = 0; tap < kTaps; ++tap) {
        const int buf = tap & 1;
        // the other buffer was last read in tap - 1, before the barrier below
        if (tap + 1 < kTaps) stage(tap + 1, buf ^ 1);
        if (K32 && waveActive) {
            const uint16_t* xr[kFwdMT];
#pragma unroll
            for (int mt = 0; mt < kFwdMT; ++mt)
                xr[mt] = a.x + tap_site(xs[mt], ys[mt], zs[mt], tap, a.dil, a.lx, a.ly, a.lz) * CI + kq * 8;
#pragma unroll
            for (int k2 = 0; k2 < KT / 2; ++k2) {
                v8bf av[kFwdMT];
#pragma unroll
                for (int mt = 0; mt < kFwdMT; ++mt) {
                    v4u raw = *reinterpret_cast<const v4u*>(xr[mt] + k2 * 32);
                    if (a.reluIn) {
#pragma unroll
                        for (int j = 0; j < 4; ++j) raw[j] = relu2(raw[j]);
                    }
                    if (!valid[mt]) raw = v4u{0u, 0u, 0u, 0u};
                    av[mt] = __builtin_bit_cast(v8bf, raw);
                }
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                    const v4u wraw =
                        *reinterpret_cast<const v4u*>(&sW[buf][((t * 16 + l16) * RS + k2 * 32 + kq * 8) / 4]);
                    const v8bf bv = __builtin_bit_cast(v8bf, wraw);
#pragma unroll
                    for (int mt = 0; mt < kFwdMT; ++mt)
                        acc[mt][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[mt], bv, acc[mt][t], 0, 0, 0);
                }
            }
        } else if (waveActive) {
            const uint16_t* xr[kFwdMT];
#pragma unroll
            for (int mt = 0; mt < kFwdMT; ++mt)
                xr[mt] = a.x + tap_site(xs[mt], ys[mt], zs[mt], tap, a.dil, a.lx, a.ly, a.lz) * CI + kq * 4;
#pragma unroll
            for (int kt = 0; kt < KT; ++kt) {
                v4s av[kFwdMT];
#pragma unroll
                for (int mt = 0; mt < kFwdMT; ++mt) {
                    const uint2 raw = *reinterpret_cast<const uint2*>(xr[mt] + kt * 16);
                    av[mt][0] = short(raw.x & 0xffffu);
                    av[mt][1] = short(raw.x >> 16);
                    av[mt][2] = short(raw.y & 0xffffu);
                    av[mt][3] = short(raw.y >> 16);
                    if (a.reluIn) {
#pragma unroll
                        for (int j = 0; j < 4; ++j) av[mt][j] = av[mt][j] < 0 ? short(0) : av[mt][j];
                    }
                    if (!valid[mt]) av[mt] = v4s{0, 0, 0, 0};
                }
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                    const uint2 wraw = sW[buf][((t * 16 + l16) * RS + kt * 16 + kq * 4) / 4];
                    v4s bv;
                    bv[0] = short(wraw.x & 0xffffu);
                    bv[1] = short(wraw.x >> 16);
                    bv[2] = short(wraw.y & 0xffffu);
                    bv[3] = short(wraw.y >> 16);
#pragma unroll
                    for (int mt = 0; mt < kFwdMT; ++mt)
                        acc[mt][t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av[mt], bv, acc[mt][t], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
    if (!waveActive) return;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int co = t * 16 + l16;
        const float b = a.bias ? a.bias[co] : 0.f;
#pragma unroll
        for (int mt = 0; mt < kFwdMT; ++mt) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const long long site = wbase + 16 * mt + 4 * kq + j;
                if (site >= a.n) continue;
                float v = acc[mt][t][j] + b;
                const long long o = site * CO + co;
                if (a.mask && !(bf2f(a.mask[o]) > 0.f)) v = 0.f;
                if (a.yf)
                    a.yf[o] = v;
                else
                    a.ybf[o] = f2bf(v);
            }
        }
    }
}

struct WgradArgs {
    const uint16_t* g;  // [n][cout] bf16 output gradient
    const uint16_t* x;  // [n][cin] bf16 layer input (before the layer's input ReLU)
    float* part;        // [chunks][27][cout][cin]
    long long n, chunk;
    int cin, cout, lx, ly, lz, dil, reluIn;
};

constexpr int kWgSites = 64;             // sites staged per step of the weight gradient
constexpr int kWgLd = kWgSites + 8;      // LDS row stride (BF16) of the transposed tiles

// One work-group per (chunk of sites, tap).  Per step it stages 64 sites of the
// output gradient G and of the tap's neighbours of X (ReLU applied) in LDS,
// transposed to [channel][site], so that both operands of
// v_mfma_f32_16x16x32_bf16 read 8 consecutive sites (the contraction axis) as
// one 16-byte LDS load; the four waves split the 16 x 16 output tiles of
// dW[tap][cout][cin].
template <int NT, int KT>
__global__ void __launch_bounds__(kWsThreads) conv_wgrad_kernel(WgradArgs a) {
    constexpr int CO = NT * 16, CI = KT * 16, TILES = NT * KT;
    constexpr int TPW = (TILES + 3) / 4;    // tiles per wave
    __shared__ __align__(16) uint16_t gT[CO * kWgLd];
    __shared__ __align__(16) uint16_t xT[CI * kWgLd];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int kq = lane >> 4, l16 = lane & 15;
    const int tap = blockIdx.y;
    const long long c = blockIdx.x;
    const long long s0 = c * a.chunk, s1 = min(a.n, s0 + a.chunk);
    v4f acc[TPW];
#pragma unroll
    for (int k = 0; k < TPW; ++k) acc[k] = v4f{0.f, 0.f, 0.f, 0.f};
    for (long long sb = s0; sb < s1; sb += kWgSites) {
        // stage: 8-byte loads (4 channels of one site), 2-byte transposed LDS stores
        for (int e = threadIdx.x; e < kWgSites * CO / 4; e += kWsThreads) {
            const int site = e / (CO / 4), c4 = (e % (CO / 4)) * 4;
            const long long sg = sb + site;
            uint2 raw = {0u, 0u};
            if (sg < s1) raw = *reinterpret_cast<const uint2*>(a.g + sg * CO + c4);
            gT[(c4 + 0) * kWgLd + site] = uint16_t(raw.x & 0xffffu);
            gT[(c4 + 1) * kWgLd + site] = uint16_t(raw.x >> 16);
            gT[(c4 + 2) * kWgLd + site] = uint16_t(raw.y & 0xffffu);
            gT[(c4 + 3) * kWgLd + site] = uint16_t(raw.y >> 16);
        }
        for (int e = threadIdx.x; e < kWgSites * CI / 4; e += kWsThreads) {
            const int site = e / (CI / 4), c4 = (e % (CI / 4)) * 4;
            const long long sg = sb + site;
            uint2 raw = {0u, 0u};
            if (sg < s1) {
                const int x = int(sg % a.lx), y = int((sg / a.lx) % a.ly), z = int(sg / ((long long)a.lx * a.ly));
                const long long nb = tap_site(x, y, z, tap, a.dil, a.lx, a.ly, a.lz);
                raw = *reinterpret_cast<const uint2*>(a.x + nb * CI + c4);
            }
            uint16_t v[4] = {uint16_t(raw.x & 0xffffu), uint16_t(raw.x >> 16), uint16_t(raw.y & 0xffffu),
                             uint16_t(raw.y >> 16)};
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (a.reluIn && (v[j] & 0x8000u)) v[j] = 0;   // negative BF16 (and -0) -> +0
                xT[(c4 + j) * kWgLd + site] = v[j];
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < TPW; ++k) {
            const int tile = wave * TPW + k;
            if (tile >= TILES) break;
            const int mt = tile / KT, nt = tile % KT;
#pragma unroll
            for (int ks = 0; ks < kWgSites / 32; ++ks) {
                const int site = ks * 32 + 8 * kq;
                const v4u ar = *reinterpret_cast<const v4u*>(&gT[(mt * 16 + l16) * kWgLd + site]);
                const v4u br = *reinterpret_cast<const v4u*>(&xT[(nt * 16 + l16) * kWgLd + site]);
                acc[k] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, ar),
                                                                 __builtin_bit_cast(v8bf, br), acc[k], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    // D[m = cout][n = cin]: lane l, element j = D[4 (l / 16) + j][l % 16]
#pragma unroll
    for (int k = 0; k < TPW; ++k) {
        const int tile = wave * TPW + k;
        if (tile >= TILES) break;
        const int mt = tile / KT, nt = tile % KT;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int row = mt * 16 + 4 * kq + j;
            a.part[((c * kTaps + tap) * CO + row) * CI + nt * 16 + l16] = acc[k][j];
        }
    }
}

// fixed-order sum over chunks
__global__ void wgrad_reduce_kernel(const float* part, long long chunks, long long per, float* out) {
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < per; e += (long long)gridDim.x * blockDim.x) {
        float s = 0.f;
        for (long long c = 0; c < chunks; ++c) s += part[c * per + e];
        out[e] = s;
    }
}

// bias gradient, first stage: work-group b sums its chunk of sites for every
// channel (consecutive threads read consecutive channels of a site: coalesced);
// wgrad_reduce_kernel then adds the work-groups' rows in a fixed order
__global__ void __launch_bounds__(kWsThreads) bias_grad_kernel(const uint16_t* g, long long n, int cout,
                                                              long long chunk, float* part) {
    __shared__ float red[kWsThreads];
    const int rows = kWsThreads / cout, c = threadIdx.x % cout, r = threadIdx.x / cout;
    const long long s0 = (long long)blockIdx.x * chunk, s1 = min(n, s0 + chunk);
    float s = 0.f;
    if (r < rows)
        for (long long i = s0 + r; i < s1; i += rows) s += bf2f(g[i * cout + c]);
    red[threadIdx.x] = s;
    __syncthreads();
    if (r == 0) {
        float t = 0.f;
        for (int k = 0; k < rows; ++k) t += red[k * cout + c];
        part[(long long)blockIdx.x * cout + c] = t;
    }
}

// Adam on FP32 master weights; refreshes the BF16 copy and, for [27][cout][cin]
// weights (cin > 0), the mirrored transpose used by the input gradient.
__global__ void adam_kernel(float* w, const float* g, float* m, float* v, long long count, float lr, float b1, float b2,
                            float eps, float c1, float c2, uint16_t* wbf, uint16_t* wbd, int cout, int cin) {
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < count; e += (long long)gridDim.x * blockDim.x) {
        const float gi = g[e];
        const float mi = b1 * m[e] + (1.f - b1) * gi;
        const float vi = b2 * v[e] + (1.f - b2) * gi * gi;
        m[e] = mi;
        v[e] = vi;
        const float wi = w[e] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
        w[e] = wi;
        if (wbf) wbf[e] = f2bf(wi);
        if (wbd) {
            const long long per = (long long)cout * cin;
            const int tap = int(e / per), co = int((e / cin) % cout), ci = int(e % cin);
            wbd[((long long)(kTaps - 1 - tap) * cin + ci) * cout + co] = f2bf(wi);
        }
    }
}

// BF16 copies of FP32 weights (after initialisation or loading)
__global__ void weights_to_bf16_kernel(const float* w, long long count, uint16_t* wbf, uint16_t* wbd, int cout, int cin) {
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < count; e += (long long)gridDim.x * blockDim.x) {
        const uint16_t b = f2bf(w[e]);
        wbf[e] = b;
        if (wbd) {
            const long long per = (long long)cout * cin;
            const int tap = int(e / per), co = int((e / cin) % cout), ci = int(e % cin);
            wbd[((long long)(kTaps - 1 - tap) * cin + ci) * cout + co] = b;
        }
    }
}

// Offset added to the first five output channels: the network predicts Q as a
// correction to the uniform far-field state.
struct Off5 {
    float v[5];
};

// Gradient of the mean squared error over liquid-crystal sites (types <= 0) on
// the first five output channels (plus the offset); per-block partial losses in FP64.
__global__ void __launch_bounds__(kWsThreads) mse_grad_kernel(const float* y, int cy, const float* target,
                                                             const int8_t* types, long long n, float scale,
                                                             uint16_t* g, double* lossPart, Off5 off) {
    __shared__ double red[kWsThreads];
    double l = 0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const bool lc = types[i] <= 0;
        for (int c = 0; c < cy; ++c) {
            float gv = 0.f;
            if (lc && c < 5) {
                const float d = (y[i * cy + c] + off.v[c]) - target[i * 5 + c];
                l += double(d) * double(d);
                gv = 2.f * d * scale;
            }
            g[i * cy + c] = f2bf(gv);
        }
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int w = kWsThreads / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) lossPart[blockIdx.x] = red[0];
}

// Energy term of the training loss: g[s][c] -= scale F_c(s) for the first five
// channels of liquid-crystal sites, F being the force computer's force on the
// predicted state (minus the energy gradient, shaped by the metric).
template <typename R>
__global__ void __launch_bounds__(kWsThreads) energy_grad_kernel(const R* f, const SiteType* types, long long n,
                                                               int cy, float scale, uint16_t* g) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        if (types[i] > 0) continue;
#pragma unroll
        for (int c = 0; c < 5; ++c) g[i * cy + c] = f2bf(bf2f(g[i * cy + c]) - scale * float(f[c * n + i]));
    }
}

// Input features, channels-last [n][16] bf16:
//  0-2 one-hot of bulk / next to an object / object,
//  3-7 anchoring tensor of object sites: Q^B / S0 for homeotropic objects,
//      -(nu nu - I/3) for degenerate planar (nu = the first three entries of Q^B),
//  8   planar flag, 9 log(W) / 3 (object sites),
//  10-14 far-field tensor d d - I/3 (every site), 15 zero.
template <typename R>
__global__ void features_kernel(const R* q, const SiteType* types, const DevObject* objects, long long n, double dx,
                                double dy, double dz, uint16_t* out) {
    const float far[5] = {float(dx * dx - 1.0 / 3.0), float(dx * dy), float(dx * dz), float(dy * dy - 1.0 / 3.0),
                          float(dy * dz)};
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        float f[16];
#pragma unroll
        for (int c = 0; c < 16; ++c) f[c] = 0.f;
        const int t = types[i];
        f[t == 0 ? 0 : (t < 0 ? 1 : 2)] = 1.f;
        if (t > 0) {
            const DevObject o = objects[t - 1];
            float qb[5];
#pragma unroll
            for (int c = 0; c < 5; ++c) qb[c] = float(q[c * n + i]);
            if (o.type == 0) {
                const float s = o.P2 != 0 ? float(1.0 / o.P2) : 1.f;
#pragma unroll
                for (int c = 0; c < 5; ++c) f[3 + c] = qb[c] * s;
            } else {
                const float nx = qb[0], ny = qb[1], nz = qb[2];
                f[3] = -(nx * nx - 1.f / 3.f);
                f[4] = -nx * ny;
                f[5] = -nx * nz;
                f[6] = -(ny * ny - 1.f / 3.f);
                f[7] = -ny * nz;
                f[8] = 1.f;
            }
            f[9] = o.P1 > 0 ? float(log(o.P1) / 3.0) : 0.f;
        }
#pragma unroll
        for (int c = 0; c < 5; ++c) f[10 + c] = far[c];
#pragma unroll
        for (int c = 0; c < 16; ++c) out[i * 16 + c] = f2bf(f[c]);
    }
}

// Q of liquid-crystal sites from the first five channels of y; velocities zero.
__global__ void types_to_i8_kernel(const SiteType* types, long long n, int8_t* out) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        out[i] = int8_t(types[i] > 127 ? 127 : types[i]);
}

template <typename R>
__global__ void apply_kernel(const float* y, int cy, const SiteType* types, long long n, R* q, R* u, Off5 off) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        if (types[i] <= 0) {
#pragma unroll
            for (int c = 0; c < 5; ++c) q[c * n + i] = R(y[i * cy + c] + off.v[c]);
        }
#pragma unroll
        for (int c = 0; c < 5; ++c) u[c * n + i] = R(0);
    }
}

template <typename R>
__global__ void extract_q_kernel(const R* q, long long n, float* out) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
#pragma unroll
        for (int c = 0; c < 5; ++c) out[i * 5 + c] = float(q[c * n + i]);
    }
}

inline int grid_for(long long work, int per = kWsThreads) {
    long long b = (work + per - 1) / per;
    return int(b < 1 ? 1 : (b > 65536 ? 65536 : b));
}

template <int NT, int KT>
hipError_t launch_fwd_t(const ConvArgs& a, hipStream_t s) {
    const long long waves = (a.n + 16 * kFwdMT - 1) / (16 * kFwdMT);
    const int blocks = int((waves + 3) / 4);
    hipLaunchKernelGGL((conv_fwd_kernel<NT, KT>), dim3(blocks), dim3(kWsThreads), 0, s, a);
    return hipGetLastError();
}

template <int NT, int KT>
hipError_t launch_wgrad_t(const WgradArgs& a, long long chunks, hipStream_t s) {
    hipLaunchKernelGGL((conv_wgrad_kernel<NT, KT>), dim3(unsigned(chunks), kTaps), dim3(kWsThreads), 0, s, a);
    return hipGetLastError();
}

// channel counts supported: multiples of 16 up to 64
#define QMIN_WS_DISPATCH(FN, NT_, KT_, ...)                                        \
    switch ((NT_) * 8 + (KT_)) {                                                  \
        case 1 * 8 + 1: return FN<1, 1>(__VA_ARGS__);                             \
        case 1 * 8 + 2: return FN<1, 2>(__VA_ARGS__);                             \
        case 1 * 8 + 4: return FN<1, 4>(__VA_ARGS__);                             \
        case 2 * 8 + 1: return FN<2, 1>(__VA_ARGS__);                             \
        case 2 * 8 + 2: return FN<2, 2>(__VA_ARGS__);                             \
        case 2 * 8 + 4: return FN<2, 4>(__VA_ARGS__);                             \
        case 4 * 8 + 1: return FN<4, 1>(__VA_ARGS__);                             \
        case 4 * 8 + 2: return FN<4, 2>(__VA_ARGS__);                             \
        case 4 * 8 + 4: return FN<4, 4>(__VA_ARGS__);                             \
        default: return hipErrorInvalidValue;                                     \
    }

hipError_t launch_fwd(const ConvArgs& a, hipStream_t s) { QMIN_WS_DISPATCH(launch_fwd_t, a.cout / 16, a.cin / 16, a, s) }

hipError_t launch_wgrad(const WgradArgs& a, long long chunks, hipStream_t s) {
    QMIN_WS_DISPATCH(launch_wgrad_t, a.cout / 16, a.cin / 16, a, chunks, s)
}

bool channels_ok(int c) { return c == 16 || c == 32 || c == 64; }

}  // namespace ws
}  // namespace qmin

using namespace qmin;
using namespace qmin::ws;

extern "C" {

int qm_ws_conv_forward(const void* x, int cin, const void* wt, const float* bias, int cout, const void* mask,
                       void* ybf16, float* yf32, int lx, int ly, int lz, int dil, int reluIn, void* stream) {
    if (!channels_ok(cin) || !channels_ok(cout)) return fail(QM_EINVAL, "channels must be 16, 32 or 64");
    if (!x || !wt || (!ybf16 && !yf32) || lx < 1 || ly < 1 || lz < 1 || dil < 1)
        return fail(QM_EINVAL, "bad convolution argument");
    ConvArgs a{};
    a.x = static_cast<const uint16_t*>(x);
    a.wt = static_cast<const uint16_t*>(wt);
    a.bias = bias;
    a.mask = static_cast<const uint16_t*>(mask);
    a.ybf = static_cast<uint16_t*>(ybf16);
    a.yf = yf32;
    a.n = (long long)lx * ly * lz;
    a.cin = cin;
    a.cout = cout;
    a.lx = lx;
    a.ly = ly;
    a.lz = lz;
    a.dil = dil;
    a.reluIn = reluIn;
    HIP_TRY(launch_fwd(a, static_cast<hipStream_t>(stream)));
    return QM_OK;
}

int qm_ws_conv_wgrad(const void* g, int cout, const void* x, int cin, int reluIn, int lx, int ly, int lz, int dil,
                     float* part, long long chunk, float* dw, void* stream) {
    if (!channels_ok(cin) || !channels_ok(cout)) return fail(QM_EINVAL, "channels must be 16, 32 or 64");
    if (!g || !x || !part || !dw || chunk < 16 || chunk % 16) return fail(QM_EINVAL, "bad weight-gradient argument");
    WgradArgs a{};
    a.g = static_cast<const uint16_t*>(g);
    a.x = static_cast<const uint16_t*>(x);
    a.part = part;
    a.n = (long long)lx * ly * lz;
    a.chunk = chunk;
    a.cin = cin;
    a.cout = cout;
    a.lx = lx;
    a.ly = ly;
    a.lz = lz;
    a.dil = dil;
    a.reluIn = reluIn;
    const long long chunks = (a.n + chunk - 1) / chunk;
    hipStream_t s = static_cast<hipStream_t>(stream);
    HIP_TRY(launch_wgrad(a, chunks, s));
    const long long per = (long long)kTaps * cout * cin;
    hipLaunchKernelGGL(wgrad_reduce_kernel, dim3(grid_for(per)), dim3(kWsThreads), 0, s, part, chunks, per, dw);
    HIP_TRY(hipGetLastError());
    return QM_OK;
}

int qm_ws_bias_grad(const void* g, long long n, int cout, float* part, int blocks, float* db, void* stream) {
    if (!g || !db || !part || !channels_ok(cout) || blocks < 1 || n < 1) return fail(QM_EINVAL, "bad argument");
    const long long chunk = (n + blocks - 1) / blocks;
    const auto st = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(bias_grad_kernel, dim3(blocks), dim3(kWsThreads), 0, st, static_cast<const uint16_t*>(g), n, cout,
                       chunk, part);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(wgrad_reduce_kernel, dim3(1), dim3(64), 0, st, part, (long long)blocks, (long long)cout, db);
    HIP_TRY(hipGetLastError());
    return QM_OK;
}

int qm_ws_adam(float* w, const float* g, float* m, float* v, long long count, float lr, float b1, float b2, float eps,
               int step, void* wbf16, void* wbd16, int cout, int cin, void* stream) {
    if (!w || !g || !m || !v || step < 1) return fail(QM_EINVAL, "bad argument");
    const float c1 = 1.f / (1.f - powf(b1, float(step))), c2 = 1.f / (1.f - powf(b2, float(step)));
    hipLaunchKernelGGL(adam_kernel, dim3(grid_for(count)), dim3(kWsThreads), 0, static_cast<hipStream_t>(stream), w, g,
                       m, v, count, lr, b1, b2, eps, c1, c2, static_cast<uint16_t*>(wbf16),
                       static_cast<uint16_t*>(wbd16), cout, cin);
    HIP_TRY(hipGetLastError());
    return QM_OK;
}

int qm_ws_weights_to_bf16(const float* w, long long count, void* wbf16, void* wbd16, int cout, int cin, void* stream) {
    if (!w || !wbf16) return fail(QM_EINVAL, "bad argument");
    hipLaunchKernelGGL(weights_to_bf16_kernel, dim3(grid_for(count)), dim3(kWsThreads), 0,
                       static_cast<hipStream_t>(stream), w, count, static_cast<uint16_t*>(wbf16),
                       static_cast<uint16_t*>(wbd16), cout, cin);
    HIP_TRY(hipGetLastError());
    return QM_OK;
}

static Off5 off5(const double* offset5) {
    Off5 o{};
    for (int c = 0; c < 5; ++c) o.v[c] = offset5 ? float(offset5[c]) : 0.f;
    return o;
}

int qm_ws_mse_grad(const float* y, int cy, const float* target, const void* types, long long n, float scale,
                   const double* offset5, void* g, double* lossPart, int blocks, void* stream) {
    if (!y || !target || !types || !g || !lossPart || blocks < 1 || cy < 5) return fail(QM_EINVAL, "bad argument");
    hipLaunchKernelGGL(mse_grad_kernel, dim3(blocks), dim3(kWsThreads), 0, static_cast<hipStream_t>(stream), y, cy,
                       target, static_cast<const int8_t*>(types), n, scale, static_cast<uint16_t*>(g), lossPart,
                       off5(offset5));
    HIP_TRY(hipGetLastError());
    return QM_OK;
}

// The energy term of the warm-start loss: the lattice holds the prediction
// (qm_ws_apply); its forces are computed and -scale F added to the first five
// channels of g at liquid-crystal sites.  The caller's stream is drained first.
int qm_ws_energy_grad(qm_ldg* f, qm_lattice* L, float scale, void* g, int cy, void* stream) {
    if (!f || !L || !g || cy < 5) return fail(QM_EINVAL, "bad argument");
    HIP_TRY(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    if (int rc = qm_ldg_compute_forces(f, L)) return rc;
    const int blocks = grid_for(L->n);
    if (L->precision == QM_FP64)
        hipLaunchKernelGGL(energy_grad_kernel<double>, dim3(blocks), dim3(kWsThreads), 0, L->stream,
                           state_set<double>(L, L->hCtrl.cur).f, L->dTypes, L->n, cy, scale,
                           static_cast<uint16_t*>(g));
    else
        hipLaunchKernelGGL(energy_grad_kernel<float>, dim3(blocks), dim3(kWsThreads), 0, L->stream,
                           state_set<float>(L, L->hCtrl.cur).f, L->dTypes, L->n, cy, scale,
                           static_cast<uint16_t*>(g));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(L->stream));
    return QM_OK;
}

// features of a lattice's current geometry; out is [n][16] bf16 on the device
int qm_ws_features(qm_lattice* L, double dx, double dy, double dz, void* out) {
    if (!L || !out) return fail(QM_EINVAL, "bad argument");
    if (int rc = sync_ctrl_to_host(L)) return rc;
    const int blocks = grid_for(L->n);
    if (L->precision == QM_FP64)
        hipLaunchKernelGGL(features_kernel<double>, dim3(blocks), dim3(kWsThreads), 0, L->stream,
                           state_set<double>(L, L->hCtrl.cur).q, L->dTypes, L->dObjects, L->n, dx, dy, dz,
                           static_cast<uint16_t*>(out));
    else
        hipLaunchKernelGGL(features_kernel<float>, dim3(blocks), dim3(kWsThreads), 0, L->stream,
                           state_set<float>(L, L->hCtrl.cur).q, L->dTypes, L->dObjects, L->n, dx, dy, dz,
                           static_cast<uint16_t*>(out));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(L->stream));
    return QM_OK;
}

// the device site types as int8 (reference encoding, object ids above 127 clamped
// to 127: the network's inputs and loss only tell objects from the rest)
int qm_ws_types(qm_lattice* L, void* out) {
    if (!L || !out) return fail(QM_EINVAL, "bad argument");
    hipLaunchKernelGGL(types_to_i8_kernel, dim3(grid_for(L->n)), dim3(kWsThreads), 0, L->stream, L->dTypes, L->n,
                       static_cast<int8_t*>(out));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(L->stream));
    return QM_OK;
}

// Q of liquid-crystal sites from y (first five of cy channels, FP32, device)
// plus offset5 (host, may be NULL); velocities are cleared and any pending FIRE
// velocity mix dropped
int qm_ws_apply(qm_lattice* L, const float* y, int cy, const double* offset5) {
    if (!L || !y || cy < 5) return fail(QM_EINVAL, "bad argument");
    if (int rc = sync_ctrl_to_host(L)) return rc;
    const int blocks = grid_for(L->n);
    if (L->precision == QM_FP64) {
        const StateSet<double> st = state_set<double>(L, L->hCtrl.cur);
        hipLaunchKernelGGL(apply_kernel<double>, dim3(blocks), dim3(kWsThreads), 0, L->stream, y, cy, L->dTypes, L->n,
                           st.q, st.u, off5(offset5));
    } else {
        const StateSet<float> st = state_set<float>(L, L->hCtrl.cur);
        hipLaunchKernelGGL(apply_kernel<float>, dim3(blocks), dim3(kWsThreads), 0, L->stream, y, cy, L->dTypes, L->n,
                           st.q, st.u, off5(offset5));
    }
    HIP_TRY(hipGetLastError());
    L->hCtrl.identity = 1;
    L->hCtrl.reset = 0;
    if (int rc = push_ctrl(L)) return rc;
    HIP_TRY(hipStreamSynchronize(L->stream));
    return QM_OK;
}

// current Q as FP32 [n][5] on the device
int qm_ws_extract_q(qm_lattice* L, float* out) {
    if (!L || !out) return fail(QM_EINVAL, "bad argument");
    if (int rc = sync_ctrl_to_host(L)) return rc;
    const int blocks = grid_for(L->n);
    if (L->precision == QM_FP64)
        hipLaunchKernelGGL(extract_q_kernel<double>, dim3(blocks), dim3(kWsThreads), 0, L->stream,
                           state_set<double>(L, L->hCtrl.cur).q, L->n, out);
    else
        hipLaunchKernelGGL(extract_q_kernel<float>, dim3(blocks), dim3(kWsThreads), 0, L->stream,
                           state_set<float>(L, L->hCtrl.cur).q, L->n, out);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(L->stream));
    return QM_OK;
}

}  // extern "C"
