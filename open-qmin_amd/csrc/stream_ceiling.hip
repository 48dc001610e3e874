// Streaming ceiling of the fused FIRE pass's access pattern on one GPU: read 15
// planes and write 15 planes of n values (structure of arrays, as Q, v, F),
// with no stencil and no halo.  Reports the best time over launch shapes; the
// fused pass cannot beat it, so it is the practical bound that its achieved
// bandwidth is compared with.  Blocked layouts (fewer concurrent DRAM streams)
// are timed beside it, for what a change of the state layout could give.
//   qmin-stream-ceiling [n = 16777216]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define CHECK(x)                                                                          \
    do {                                                                                  \
        hipError_t e_ = (x);                                                              \
        if (e_ != hipSuccess) {                                                           \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                  \
            std::exit(1);                                                                 \
        }                                                                                 \
    } while (0)

// Variant 0: one value per thread and plane, ordinary stores.  Variant 1: the
// same with non-temporal stores (as the fused pass).  Variant 2: V consecutive
// values per thread and plane through one vector load and one non-temporal
// vector store (16-byte accesses).
template <typename R, bool NT>
__device__ __forceinline__ void put(R* p, R v) {
    if constexpr (NT)
        __builtin_nontemporal_store(v, p);
    else
        *p = v;
}

template <typename R, bool NT>
__global__ void __launch_bounds__(256) read15_write15(const R* __restrict__ in, R* __restrict__ out, long long n,
                                                      R a) {
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        R q[5], u[5], f[5];
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            q[c] = in[c * n + i];
            u[c] = in[(5 + c) * n + i];
            f[c] = in[(10 + c) * n + i];
        }
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            put<R, NT>(&out[c * n + i], q[c] + a * u[c]);
            put<R, NT>(&out[(5 + c) * n + i], u[c] + a * f[c]);
            put<R, NT>(&out[(10 + c) * n + i], a * f[c]);
        }
    }
}

typedef double vd2 __attribute__((ext_vector_type(2)));
typedef float vf4 __attribute__((ext_vector_type(4)));

// VT: a 16-byte vector of R (vd2 or vf4); n is a multiple of its width
template <typename VT, typename R>
__global__ void __launch_bounds__(256) read15_write15_vec(const VT* __restrict__ in, VT* __restrict__ out,
                                                          long long nv, R a) {
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < nv; i += (long long)gridDim.x * 256) {
        VT q[5], u[5], f[5];
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            q[c] = in[c * nv + i];
            u[c] = in[(5 + c) * nv + i];
            f[c] = in[(10 + c) * nv + i];
        }
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            __builtin_nontemporal_store(q[c] + a * u[c], &out[c * nv + i]);
            __builtin_nontemporal_store(u[c] + a * f[c], &out[(5 + c) * nv + i]);
            __builtin_nontemporal_store(a * f[c], &out[(10 + c) * nv + i]);
        }
    }
}

// Blocked layouts with fewer concurrent DRAM streams: the values of each site
// grouped in blocks of B sites, [block][plane][B].  FIELDS = 1: one array of
// all 15 planes (one read and one write stream per wave); FIELDS = 3: Q, v and
// F each an array of 5 planes (three and three).  Non-temporal stores.
template <typename R, int B, int FIELDS>
__global__ void __launch_bounds__(256) read15_write15_blocked(const R* __restrict__ in, R* __restrict__ out,
                                                              long long n, R a) {
    constexpr int P = 15 / FIELDS;  // planes per array
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        const long long blk = i / B;
        const int l = int(i % B);
        R v[15];
#pragma unroll
        for (int f = 0; f < FIELDS; ++f)
#pragma unroll
            for (int c = 0; c < P; ++c) v[f * P + c] = in[f * P * n + (blk * P + c) * B + l];
        R w[15];
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            w[c] = v[c] + a * v[5 + c];
            w[5 + c] = v[5 + c] + a * v[10 + c];
            w[10 + c] = a * v[10 + c];
        }
#pragma unroll
        for (int f = 0; f < FIELDS; ++f)
#pragma unroll
            for (int c = 0; c < P; ++c) __builtin_nontemporal_store(w[f * P + c], &out[f * P * n + (blk * P + c) * B + l]);
    }
}

template <typename R, typename VT>
static void run(const char* name, long long n) {
    R *a = nullptr, *b = nullptr;
    CHECK(hipMalloc(&a, 15 * n * sizeof(R)));
    CHECK(hipMalloc(&b, 15 * n * sizeof(R)));
    CHECK(hipMemset(a, 0, 15 * n * sizeof(R)));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    const int width = int(sizeof(VT) / sizeof(R));
    const long long nv = n / width;
    const char* names[7] = {"scalar", "scalar, non-temporal stores", "16-byte vectors, non-temporal stores",
                            "blocked [block][15][64] (2 streams)", "blocked Q, v, F each [block][5][64] (6 streams)",
                            "blocked [block][15][256] (2 streams)", "blocked Q, v, F each [block][5][256] (6 streams)"};
    double best = 1e30, bestBlocked = 1e30;
    for (int variant = 0; variant < 7; ++variant) {
        double vbest = 1e30;
        int bestGrid = 0;
        for (int grid : {1024, 2048, 4096, 8192, 16384, 65536}) {
            auto launch = [&] {
                if (variant == 0) read15_write15<R, false><<<grid, 256>>>(a, b, n, R(0.5));
                else if (variant == 1) read15_write15<R, true><<<grid, 256>>>(a, b, n, R(0.5));
                else if (variant == 2)
                    read15_write15_vec<VT, R><<<grid, 256>>>(reinterpret_cast<const VT*>(a), reinterpret_cast<VT*>(b),
                                                             nv, R(0.5));
                else if (variant == 3) read15_write15_blocked<R, 64, 1><<<grid, 256>>>(a, b, n, R(0.5));
                else if (variant == 4) read15_write15_blocked<R, 64, 3><<<grid, 256>>>(a, b, n, R(0.5));
                else if (variant == 5) read15_write15_blocked<R, 256, 1><<<grid, 256>>>(a, b, n, R(0.5));
                else read15_write15_blocked<R, 256, 3><<<grid, 256>>>(a, b, n, R(0.5));
            };
            for (int w = 0; w < 3; ++w) launch();
            const int reps = 20;
            CHECK(hipEventRecord(e0));
            for (int r = 0; r < reps; ++r) launch();
            CHECK(hipEventRecord(e1));
            CHECK(hipEventSynchronize(e1));
            CHECK(hipGetLastError());
            float ms = 0;
            CHECK(hipEventElapsedTime(&ms, e0, e1));
            if (ms / reps < vbest) {
                vbest = ms / reps;
                bestGrid = grid;
            }
        }
        const double bytes = 30.0 * n * sizeof(R);
        std::printf("%s read 15 + write 15 planes of %lld, %s: %.4f ms (grid %d), %.3f TB/s\n", name, n,
                    names[variant], vbest, bestGrid, bytes / vbest / 1e9);
        if (variant < 3 && vbest < best) best = vbest;
        if (variant >= 3 && vbest < bestBlocked) bestBlocked = vbest;
    }
    std::printf("%s ceiling (planes, as the state is laid out): %.4f ms, %.3f TB/s\n", name, best,
                30.0 * n * sizeof(R) / best / 1e9);
    std::printf("%s blocked layouts, best: %.4f ms, %.3f TB/s\n", name, bestBlocked,
                30.0 * n * sizeof(R) / bestBlocked / 1e9);
    CHECK(hipFree(a));
    CHECK(hipFree(b));
}

int main(int argc, char** argv) {
    const long long n = argc > 1 ? std::atoll(argv[1]) : 16777216LL;
    if (n % 256 != 0) {
        std::fprintf(stderr, "n must be a multiple of 256\n");
        return 1;
    }
    run<double, vd2>("fp64", n);
    run<float, vf4>("fp32", n);
    return 0;
}
