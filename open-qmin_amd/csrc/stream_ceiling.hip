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

template <typename R>
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
            out[c * n + i] = q[c] + a * u[c];
            out[(5 + c) * n + i] = u[c] + a * f[c];
            out[(10 + c) * n + i] = a * f[c];
        }
    }
}

template <typename R>
static void run(const char* name, long long n) {
    R *a = nullptr, *b = nullptr;
    CHECK(hipMalloc(&a, 15 * n * sizeof(R)));
    CHECK(hipMalloc(&b, 15 * n * sizeof(R)));
    CHECK(hipMemset(a, 0, 15 * n * sizeof(R)));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    double best = 1e30;
    int bestGrid = 0;
    for (int grid : {1024, 2048, 4096, 8192, 16384, 65536}) {
        for (int w = 0; w < 3; ++w) read15_write15<R><<<grid, 256>>>(a, b, n, R(0.5));
        const int reps = 20;
        CHECK(hipEventRecord(e0));
        for (int r = 0; r < reps; ++r) read15_write15<R><<<grid, 256>>>(a, b, n, R(0.5));
        CHECK(hipEventRecord(e1));
        CHECK(hipEventSynchronize(e1));
        float ms = 0;
        CHECK(hipEventElapsedTime(&ms, e0, e1));
        if (ms / reps < best) {
            best = ms / reps;
            bestGrid = grid;
        }
    }
    const double bytes = 30.0 * n * sizeof(R);
    std::printf("%s read 15 + write 15 planes of %lld: %.4f ms (grid %d), %.3f TB/s\n", name, n, best, bestGrid,
                bytes / best / 1e9);
    CHECK(hipFree(a));
    CHECK(hipFree(b));
}

int main(int argc, char** argv) {
    const long long n = argc > 1 ? std::atoll(argv[1]) : 16777216LL;
    run<double>("fp64", n);
    run<float>("fp32", n);
    return 0;
}
