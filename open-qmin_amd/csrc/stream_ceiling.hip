// Streaming ceiling of the fused FIRE pass's access pattern on one GPU: read 15
// planes and write 15 planes of n values (structure of arrays, as Q, v, F),
// with no stencil and no halo.  Reports the best time over launch shapes; the
// fused pass cannot beat it, so it is the practical bound that its achieved
// bandwidth is compared with.
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
    const char* names[3] = {"scalar", "scalar, non-temporal stores", "16-byte vectors, non-temporal stores"};
    double best = 1e30;
    for (int variant = 0; variant < 3; ++variant) {
        double vbest = 1e30;
        int bestGrid = 0;
        for (int grid : {1024, 2048, 4096, 8192, 16384, 65536}) {
            auto launch = [&] {
                if (variant == 0) read15_write15<R, false><<<grid, 256>>>(a, b, n, R(0.5));
                else if (variant == 1) read15_write15<R, true><<<grid, 256>>>(a, b, n, R(0.5));
                else read15_write15_vec<VT, R><<<grid, 256>>>(reinterpret_cast<const VT*>(a),
                                                            reinterpret_cast<VT*>(b), nv, R(0.5));
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
        if (vbest < best) best = vbest;
    }
    std::printf("%s ceiling: %.4f ms, %.3f TB/s\n", name, best, 30.0 * n * sizeof(R) / best / 1e9);
    CHECK(hipFree(a));
    CHECK(hipFree(b));
}

int main(int argc, char** argv) {
    const long long n = argc > 1 ? std::atoll(argv[1]) : 16777216LL;
    if (n % 4 != 0) {
        std::fprintf(stderr, "n must be a multiple of 4\n");
        return 1;
    }
    run<double, vd2>("fp64", n);
    run<float, vf4>("fp32", n);
    return 0;
}
