// Adam, Nesterov accelerated gradient and online L-BFGS on the device.
//
// An Adam or Nesterov iteration is one lattice pass (the fused and boundary
// passes in mode kAdamStep / kNesterovStep, step_kernels.hip): each site's
// force, then the minimizer's update of that site, written to the other state
// set.  An L-BFGS iteration is the shared force pass (force mode) and
// streaming kernels over the 5 n components of Q, F and the minimizer's own
// vectors (same plane layout as Q).  Every pass reduces the dot products it
// needs into per-work-group partials; a one-work-group finalize kernel sums
// them in a fixed order and finishes the minimizer's scalar logic in device
// memory (step coefficients, forceMax, the stopping test).  (A separate kernel, not a last-work-group
// pattern: the device-wide fences that pattern needs write back each XCD's L2,
// which the streaming kernel has just filled.)  A batch of iterations is
// therefore queued without host round trips, every kernel after the stopping
// test is a no-op, and the host reads the control block once per batch.
//
// Arithmetic follows the reference CPU steps: energyMinimizerAdam.cpp:25-54,
// energyMinimizerNesterovAG.cpp:70-92 and energyMinimizerLoLBFGS.cpp:104-193.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "qmin_internal.h"

namespace qmin {
namespace {

constexpr int kVecThreads = 256;
constexpr int kVecUnroll = 4;
constexpr int kVecMaxBlocks = 2048;
constexpr int kMaxDots = 4;

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Fixed-order work-group sum of K doubles; the result is valid in thread 0.
template <int K>
__device__ __forceinline__ void block_sum(double (&v)[K], double (*red)[kMaxDots]) {
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = wsum(v[k]);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) red[w][k] = v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            double s = 0;
#pragma unroll
            for (int i = 0; i < kVecThreads / 64; ++i) s += red[i][k];
            v[k] = s;
        }
    }
    __syncthreads();
}

// Writes this work-group's K partial sums (stride kMaxDots).
template <int K>
__device__ __forceinline__ void write_partials(double (&v)[K], double* partials) {
    __shared__ double red[kVecThreads / 64][kMaxDots];
    block_sum<K>(v, red);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) partials[kMaxDots * blockIdx.x + k] = v[k];
    }
}

// Iteration bookkeeping shared by the three minimizers (thread 0 of the last
// work-group): count, forceMax, trace, and the loop test of minimize().
__device__ void finish_iteration(Ctrl* ctrl, double forceMax, double fc, double* trace) {
    const int it = ctrl->iterations + 1;
    ctrl->iterations = it;
    ctrl->forceMax = forceMax;
    if (trace && ctrl->traceLen < ctrl->traceCap) {
        trace[ctrl->traceLen] = forceMax;
        ctrl->traceLen += 1;
    }
    if (isnan(forceMax)) ctrl->nanFlag = 1;
    ctrl->done = !(it < ctrl->maxIter && forceMax > fc);
}

struct Common {
    Ctrl* ctrl;
    double* partials;   // [blocks][kMaxDots]
    double* trace;      // [traceCap] forceMax per iteration
    double fc, ndof;
    long long len;      // 5 n
};

// ---------------------------------------------------------------- online L-BFGS
// Device scalars of the two-loop recursion, per memory slot t and loop index ii:
//   SY[t] = s_t . y_t and YY[t] = y_t . y_t, cached when y_t is completed
//   (s_t and y_t do not change until slot t is reused),
//   A[ii] = alpha_ii, NA[ii] = -alpha_ii, CS[ii] = alpha_ii - beta_ii (0 when
//   s.y is 0, where the reference skips the update), K3 = the step-3 factor.
struct LbLayout {
    int m;
    __host__ __device__ int SY(int t) const { return t; }
    __host__ __device__ int YY(int t) const { return m + t; }
    __host__ __device__ int A(int i) const { return 2 * m + i; }
    __host__ __device__ int NA(int i) const { return 3 * m + i; }
    __host__ __device__ int CS(int i) const { return 4 * m + i; }
    __host__ __device__ int K3() const { return 5 * m; }
    __host__ __device__ int size() const { return 5 * m + 1; }
};


template <typename R>
struct LbStageArgs {
    R* p;
    const R* x1;
    const R* x2;      // may be null
    const R* d;
    int k1, k2;       // scalar slots of the coefficients of x1, x2
    double* scal;
    Common c;
};

// p = (p + k1 x1) [+ k2 x2]; then partials of d.p, finished into alpha_ii
// (step 2, energyMinimizerLoLBFGS.cpp:117-131) or alpha_ii - beta_ii (step 4,
// :145-163) by the finalize kernel.
template <typename R>
__global__ void __launch_bounds__(kVecThreads) lb_stage_kernel(LbStageArgs<R> A) {
    Ctrl* ctrl = A.c.ctrl;
    if (ctrl->done) return;
    const R k1 = R(A.scal[A.k1]);
    const R k2 = A.x2 ? R(A.scal[A.k2]) : R(0);
    double dp = 0;
    const long long stride = (long long)gridDim.x * kVecThreads * kVecUnroll;
    for (long long base = (long long)blockIdx.x * kVecThreads * kVecUnroll + threadIdx.x; base < A.c.len;
         base += stride) {
#pragma unroll
        for (int u = 0; u < kVecUnroll; ++u) {
            const long long i = base + u * kVecThreads;
            if (i < A.c.len) {
                R p = A.p[i] + k1 * A.x1[i];
                if (A.x2) p = p + k2 * A.x2[i];
                A.p[i] = p;
                dp += double(A.d[i]) * double(p);
            }
        }
    }
    double part[1] = {dp};
    write_partials<1>(part, A.c.partials);
}

template <typename R>
struct LbMoveArgs {
    const R* p;
    const R* s0;      // s of the newest slot, added with CS[0]
    R* sc;            // s of the slot being replaced
    R* yc;            // y of the slot being replaced: receives F
    R* q;
    const R* f;
    double eta;
    int k;
    double* scal;
    Common c;
};

// Last step-4 update, then s_c = eta p, y_c = F (old forces), Q += s_c
// (energyMinimizerLoLBFGS.cpp:158-176); p itself is replaced by the new forces.
template <typename R>
__global__ void __launch_bounds__(kVecThreads) lb_move_kernel(LbMoveArgs<R> A) {
    if (A.c.ctrl->done) return;
    const R k = R(A.scal[A.k]), eta = R(A.eta);
    const long long stride = (long long)gridDim.x * kVecThreads * kVecUnroll;
    for (long long base = (long long)blockIdx.x * kVecThreads * kVecUnroll + threadIdx.x; base < A.c.len;
         base += stride) {
#pragma unroll
        for (int u = 0; u < kVecUnroll; ++u) {
            const long long i = base + u * kVecThreads;
            if (i < A.c.len) {
                const R p = A.p[i] + k * A.s0[i];
                const R s = eta * p;
                A.sc[i] = s;
                A.yc[i] = A.f[i];
                A.q[i] = A.q[i] + s;
            }
        }
    }
}

template <typename R>
struct LbFinishArgs {
    R* p;
    const R* f;
    const R* s;       // s of the newest slot
    R* y;             // y of the newest slot (null on the first call: no update)
    Common c;
};

// p = F; y_t = y_t - F (:178-188); partials of F.F, s_t.y_t, y_t.y_t and s_t.p,
// from which the finalize kernel takes forceMax = sqrt(F.F)/N, caches s_t.y_t
// and y_t.y_t, and starts the next two-loop pass: alpha_0 = s_t.p / s_t.y_t and
// the step-3 factor (s.y)/(y.y) of the same slot (:117-142).
template <typename R>
__global__ void __launch_bounds__(kVecThreads) lb_finish_kernel(LbFinishArgs<R> A) {
    Ctrl* ctrl = A.c.ctrl;
    if (ctrl->done) return;
    double pp = 0, sy = 0, yy = 0, sp = 0;
    const long long stride = (long long)gridDim.x * kVecThreads * kVecUnroll;
    for (long long base = (long long)blockIdx.x * kVecThreads * kVecUnroll + threadIdx.x; base < A.c.len;
         base += stride) {
#pragma unroll
        for (int u = 0; u < kVecUnroll; ++u) {
            const long long i = base + u * kVecThreads;
            if (i < A.c.len) {
                const R f = A.f[i];
                const double s = double(A.s[i]);
                A.p[i] = f;
                pp += double(f) * double(f);
                sp += s * double(f);
                if (A.y) {
                    const R y = A.y[i] - f;
                    A.y[i] = y;
                    sy += s * double(y);
                    yy += double(y) * double(y);
                }
            }
        }
    }
    double part[4] = {pp, sy, yy, sp};
    write_partials<4>(part, A.c.partials);
}

// ---------------------------------------------------------------- finalize
enum FinOp { kFinAdam, kFinNesterov, kFinAlpha, kFinBeta, kFinLbFinish, kFinLbStart };

struct FinArgs {
    int op;
    int blocks;       // partial-sum slots to sum
    int stride;       // doubles per slot: kMaxDots for the streaming kernels, 3 for the lattice passes
    int flip;         // the pass wrote the other state set: make it current
    int ii, t;        // two-loop index and memory slot
    LbLayout lay;
    double b1, b2;    // Adam
    int scheduled;    // Nesterov
    double* scal;
    Common c;
};

// One work-group: fixed-order sums of the streaming kernel's partials, then the
// minimizer's scalar step in thread 0.
__global__ void __launch_bounds__(kVecThreads) vec_finalize_kernel(FinArgs A) {
    __shared__ double red[kVecThreads / 64][kMaxDots];
    Ctrl* ctrl = A.c.ctrl;
    if (ctrl->done) return;
    double tot[kMaxDots] = {0, 0, 0, 0};
    const int stride = A.stride;
    for (int j = threadIdx.x; j < A.blocks; j += kVecThreads) {
#pragma unroll
        for (int k = 0; k < kMaxDots; ++k)
            if (k < stride) tot[k] += A.c.partials[(long long)stride * j + k];
    }
    block_sum<kMaxDots>(tot, red);
    if (threadIdx.x != 0) return;
    double* scal = A.scal;
    const LbLayout& L = A.lay;
    switch (A.op) {
        case kFinAdam:
            // beta^t advances once per executed step (energyMinimizerAdam.cpp:52-53)
            scal[0] = scal[0] * A.b1;
            scal[1] = scal[1] * A.b2;
            finish_iteration(ctrl, sqrt(tot[0]) / A.c.ndof, A.c.fc, A.c.trace);
            if (A.flip) ctrl->cur ^= 1;
            break;
        case kFinNesterov: {
            // the same lambda and momentum the step kernel computed from scal[0]
            const double oldLambda = scal[0];
            const double lambda = 0.5 * (1.0 + sqrt(1.0 + 4.0 * oldLambda * oldLambda));
            const double gamma = (1 - oldLambda) / lambda;
            if (A.scheduled) scal[1] = 1. + gamma;
            scal[0] = lambda;
            finish_iteration(ctrl, sqrt(tot[0]) / A.c.ndof, A.c.fc, A.c.trace);
            if (A.flip) ctrl->cur ^= 1;
            break;
        }
        case kFinAlpha: {
            const double sy = scal[L.SY(A.t)];
            double a = 0;
            if (sy != 0) a = tot[0] / sy;
            scal[L.A(A.ii)] = a;
            scal[L.NA(A.ii)] = -a;
            break;
        }
        case kFinBeta: {
            const double sy = scal[L.SY(A.t)];
            double cs = 0;
            if (sy != 0) {
                const double beta = tot[0] / sy;
                cs = scal[L.A(A.ii)] - beta;
            }
            scal[L.CS(A.ii)] = cs;
            break;
        }
        case kFinLbFinish:
        case kFinLbStart: {
            if (A.op == kFinLbFinish) {
                scal[L.SY(A.t)] = tot[1];
                scal[L.YY(A.t)] = tot[2];
            }
            const double syt = scal[L.SY(A.t)], yyt = scal[L.YY(A.t)];
            double a = 0;
            if (syt != 0) a = tot[3] / syt;
            scal[L.A(0)] = a;
            scal[L.NA(0)] = -a;
            scal[L.K3()] = yyt != 0 ? syt / yyt : 0.0;
            if (A.op == kFinLbFinish) finish_iteration(ctrl, sqrt(tot[0]) / A.c.ndof, A.c.fc, A.c.trace);
            break;
        }
    }
}

int vec_blocks(long long len) {
    const long long per = (long long)kVecThreads * kVecUnroll;
    return (int)std::max<long long>(1, std::min<long long>((len + per - 1) / per, kVecMaxBlocks));
}

}  // namespace
}  // namespace qmin

using namespace qmin;

// ------------------------------------------------------------------ host side

// Device state common to the three minimizers: vectors of 5 n reals bound to
// one lattice, scalars, partial sums and the forceMax trace.
struct VecMinimizer {
    uint64_t bound = 0;   // serial of the lattice the vectors belong to (0: none)
    long long boundN = 0;
    int boundPrecision = -1;
    std::vector<void*> vecs;
    double* dScal = nullptr;
    int scalCount = 0;
    double* dPartials = nullptr;
    double* dTrace = nullptr;
    int dTraceCap = 0;
    int traceCap = 0;
    std::vector<double> trace;
    int iterations = 0, maxIter = 1000;
    double forceMax = 100.0, fc = 1e-12;
    int converged = 0, nan = 0;

    ~VecMinimizer() {
        release();
        if (dTrace) (void)hipFree(dTrace);
    }
    // drops the device history (the trace buffer stays)
    void release() {
        for (void* p : vecs) (void)hipFree(p);
        vecs.clear();
        if (dScal) (void)hipFree(dScal);
        if (dPartials) (void)hipFree(dPartials);
        dScal = nullptr;
        dPartials = nullptr;
        bound = 0;
        boundN = 0;
        boundPrecision = -1;
    }
    bool needs_bind(const qm_lattice* L) const {
        return bound != L->serial || boundN != L->n || boundPrecision != L->precision;
    }
    // (re)allocates nv zeroed vectors and nscal zeroed scalars for lattice L
    int bind(qm_lattice* L, int nv, int nscal) {
        release();
        const size_t bytes = 5 * (size_t)L->n * L->realSize;
        for (int k = 0; k < nv; ++k) {
            void* p = nullptr;
            if (hipMalloc(&p, bytes) != hipSuccess) {
                release();
                return fail(QM_ENOMEM, "minimizer vectors: out of device memory");
            }
            vecs.push_back(p);
            HIP_TRY(hipMemsetAsync(p, 0, bytes, L->stream));
        }
        HIP_TRY(hipMalloc(&dScal, std::max(nscal, 1) * sizeof(double)));
        HIP_TRY(hipMemsetAsync(dScal, 0, std::max(nscal, 1) * sizeof(double), L->stream));
        scalCount = nscal;
        HIP_TRY(hipMalloc(&dPartials, (size_t)kVecMaxBlocks * kMaxDots * sizeof(double)));
        bound = L->serial;
        boundN = L->n;
        boundPrecision = L->precision;
        return QM_OK;
    }
    int ensure_trace() {
        if (traceCap > dTraceCap) {
            if (dTrace) HIP_TRY(hipFree(dTrace));
            HIP_TRY(hipMalloc(&dTrace, (size_t)traceCap * sizeof(double)));
            dTraceCap = traceCap;
        }
        return QM_OK;
    }
    Common common(const qm_lattice* L) const {
        Common c{};
        c.ctrl = L->dCtrl;
        c.partials = dPartials;
        c.trace = traceCap > 0 ? dTrace : nullptr;
        c.fc = fc;
        c.ndof = (double)L->n;
        c.len = 5 * L->n;
        return c;
    }
    int set_scalars(const qm_lattice* L, const double* v, int count, int offset = 0) {
        HIP_TRY(hipMemcpyAsync(dScal + offset, v, count * sizeof(double), hipMemcpyHostToDevice, L->stream));
        return QM_OK;
    }
    // the finalize kernel for a streaming kernel of `blocks` work-groups, or
    // (partials given) for a lattice pass of `blocks` slots of 3 doubles
    int finalize(const qm_lattice* L, FinArgs a, long long blocks, double* partials = nullptr) const {
        a.blocks = (int)blocks;
        a.stride = partials ? 3 : kMaxDots;
        a.flip = partials ? 1 : 0;
        a.scal = dScal;
        a.c = common(L);
        if (partials) a.c.partials = partials;
        vec_finalize_kernel<<<1, kVecThreads, 0, L->stream>>>(a);
        HIP_TRY(hipGetLastError());
        return QM_OK;
    }
    int get_scalars(const qm_lattice* L, double* v, int count, int offset = 0) {
        HIP_TRY(hipMemcpyAsync(v, dScal + offset, count * sizeof(double), hipMemcpyDeviceToHost, L->stream));
        HIP_TRY(hipStreamSynchronize(L->stream));
        return QM_OK;
    }
};

// Starts a minimize call: FIRE's pending velocity is made concrete (it is kept
// in the same state set), the control block is set for this minimizer.
static int begin_run(VecMinimizer& s, qm_lattice* L, const qm_ldg* f, bool done) {
    if (int rc = prepare(L, f)) return rc;
    if (int rc = materialize(L)) return rc;
    if (int rc = s.ensure_trace()) return rc;
    if (int rc = sync_ctrl_to_host(L)) return rc;
    Ctrl& c = L->hCtrl;
    c.done = done ? 1 : 0;
    c.iterations = s.iterations;
    c.maxIter = s.maxIter;
    c.nanFlag = 0;
    c.traceLen = 0;
    c.traceCap = s.traceCap;
    c.forceMax = s.forceMax;
    return push_ctrl(L);
}

// Queues batches of iterations until the control block reports done.
template <typename Enqueue>
static int drive(VecMinimizer& s, qm_lattice* L, Enqueue&& enqueue) {
    const int batch = L->n >= (1 << 22) ? 16 : 64;
    int queued = 0;
    while (!L->hCtrl.done) {
        int todo = std::min(batch, s.maxIter - (s.iterations + queued));
        if (todo <= 0) todo = 1;
        for (int k = 0; k < todo; ++k) {
            if (int rc = enqueue(s.iterations + queued)) return rc;
            ++queued;
        }
        if (int rc = sync_ctrl_to_host(L)) return rc;
        // iterations the device actually ran: the rest were skipped
        queued = L->hCtrl.iterations - s.iterations;
    }
    const Ctrl& h = L->hCtrl;
    s.iterations = h.iterations;
    s.forceMax = h.forceMax;
    s.nan = h.nanFlag;
    s.converged = !h.nanFlag && h.forceMax <= s.fc;
    if (s.traceCap > 0 && h.traceLen > 0) {
        const size_t old = s.trace.size();
        s.trace.resize(old + (size_t)h.traceLen);
        HIP_TRY(copy_on(L, s.trace.data() + old, s.dTrace, (size_t)h.traceLen * sizeof(double),
                          hipMemcpyDeviceToHost));
    }
    if (h.nanFlag) return fail(QM_ENAN, "minimizer produced a NaN force norm");
    return QM_OK;
}

template <typename R>
static R* rp(void* p) {
    return static_cast<R*>(p);
}

// ------------------------------------------------------------------ Adam
struct qm_adam : VecMinimizer {
    // setAdamParameters defaults (energyMinimizerAdam.h:16-25)
    double b1 = 0.9, b2 = 0.99, eps = 1e-8, dt = 1e-4;
    double b1t = 0.9, b2t = 0.99;
};

static int adam_enqueue(qm_adam* m, qm_lattice* L, const qm_ldg* f) {
    StepArgs st{};
    st.m = m->vecs[0];
    st.v = m->vecs[1];
    st.scal = m->dScal;
    st.b1 = m->b1;
    st.b2 = m->b2;
    st.eps = m->eps;
    st.dt = m->dt;
    long long slots = 0;
    if (int rc = step_pass(L, f, kAdamStep, st, &slots)) return rc;
    FinArgs fa{};
    fa.op = kFinAdam;
    fa.b1 = m->b1;
    fa.b2 = m->b2;
    return m->finalize(L, fa, slots, L->dPartials);
}

extern "C" {

int qm_adam_create(qm_adam** out) {
    if (!out) return fail(QM_EINVAL, "null output handle");
    *out = new qm_adam;
    (*out)->forceMax = 110.0;
    return QM_OK;
}

int qm_adam_destroy(qm_adam* m) {
    delete m;
    return QM_OK;
}

int qm_adam_set_parameters(qm_adam* m, double beta1, double beta2, double eps, double dt, double forceCutoff) {
    if (!m) return fail(QM_EINVAL, "null minimizer");
    m->b1 = beta1;
    m->b1t = beta1;
    m->b2 = beta2;
    m->b2t = beta2;
    m->eps = eps;
    m->dt = dt;
    m->fc = forceCutoff;
    return QM_OK;
}

int qm_adam_set_max_iterations(qm_adam* m, int maxIterations) {
    if (!m) return fail(QM_EINVAL, "null minimizer");
    m->maxIter = maxIterations;
    return QM_OK;
}

int qm_adam_minimize(qm_adam* m, qm_ldg* f, qm_lattice* L) {
    if (!m || !f || !L) return fail(QM_EINVAL, "null handle");
    if (m->needs_bind(L)) {
        // initializeFromModel (energyMinimizerAdam.cpp:9-23): zero moments, iteration 0
        if (int rc = m->bind(L, 2, 2)) return rc;
        m->iterations = 0;
        m->b1t = m->b1;
        m->b2t = m->b2;
    }
    m->forceMax = 110.0;
    if (int rc = begin_run(*m, L, f, m->iterations >= m->maxIter)) return rc;
    // the iterations alternate the state sets and leave U alone
    if (int rc = share_velocity(L)) return rc;
    const double sc[2] = {m->b1t, m->b2t};
    if (int rc = m->set_scalars(L, sc, 2)) return rc;
    auto enq = [&](int) { return adam_enqueue(m, L, f); };
    int rc = drive(*m, L, enq);
    double back[2];
    if (int r2 = m->get_scalars(L, back, 2)) return r2;
    m->b1t = back[0];
    m->b2t = back[1];
    return rc;
}

int qm_adam_get_state(const qm_adam* m, int* iterations, double* forceMax) {
    if (!m) return fail(QM_EINVAL, "null minimizer");
    if (iterations) *iterations = m->iterations;
    if (forceMax) *forceMax = m->forceMax;
    return QM_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ Nesterov
struct qm_nesterov : VecMinimizer {
    // setNesterovAGParameters defaults (energyMinimizerNesterovAG.h:22-27)
    double dt = 1e-4, mu = 0.01;
    double mu0 = 0.01;   // mu as set: the scheduled momentum restarts from it
    double lambda = 1.0;
    int scheduled = 0;
};

static int nesterov_enqueue(qm_nesterov* m, qm_lattice* L, const qm_ldg* f) {
    StepArgs st{};
    st.m = m->vecs[0];
    st.scal = m->dScal;
    st.dt = m->dt;
    st.scheduled = m->scheduled;
    long long slots = 0;
    if (int rc = step_pass(L, f, kNesterovStep, st, &slots)) return rc;
    FinArgs fa{};
    fa.op = kFinNesterov;
    fa.scheduled = m->scheduled;
    return m->finalize(L, fa, slots, L->dPartials);
}

extern "C" {

int qm_nesterov_create(qm_nesterov** out) {
    if (!out) return fail(QM_EINVAL, "null output handle");
    *out = new qm_nesterov;
    return QM_OK;
}

int qm_nesterov_destroy(qm_nesterov* m) {
    delete m;
    return QM_OK;
}

int qm_nesterov_set_parameters(qm_nesterov* m, double dt, double mu, double forceCutoff, int scheduledMomentum) {
    if (!m) return fail(QM_EINVAL, "null minimizer");
    m->dt = dt;
    m->mu = mu;
    m->mu0 = mu;
    m->fc = forceCutoff;
    m->scheduled = scheduledMomentum != 0;
    return QM_OK;
}

int qm_nesterov_set_max_iterations(qm_nesterov* m, int maxIterations) {
    if (!m) return fail(QM_EINVAL, "null minimizer");
    m->maxIter = maxIterations;
    return QM_OK;
}

int qm_nesterov_minimize(qm_nesterov* m, qm_ldg* f, qm_lattice* L) {
    if (!m || !f || !L) return fail(QM_EINVAL, "null handle");
    m->forceMax = 110.0;
    // a new lattice (or precision) starts a new run: iteration 0, momentum as set
    const bool rebind = m->needs_bind(L);
    if (rebind) {
        if (int rc = m->bind(L, 1, 2)) return rc;
        m->iterations = 0;
        m->lambda = 1.0;
        m->mu = m->mu0;
    }
    if (int rc = begin_run(*m, L, f, m->iterations >= m->maxIter)) return rc;
    // the iterations alternate the state sets and leave U alone
    if (int rc = share_velocity(L)) return rc;
    if (rebind) {
        // initializeFromModel (energyMinimizerNesterovAG.cpp:31-37): the
        // alternate sequence starts as a copy of the positions
        HIP_TRY(hipMemcpyAsync(m->vecs[0], L->dSet[L->hCtrl.cur], 5 * (size_t)L->n * L->realSize,
                               hipMemcpyDeviceToDevice, L->stream));
    }
    const double sc[2] = {m->lambda, m->mu};
    if (int rc = m->set_scalars(L, sc, 2)) return rc;
    auto enq = [&](int) { return nesterov_enqueue(m, L, f); };
    int rc = drive(*m, L, enq);
    double back[2];
    if (int r2 = m->get_scalars(L, back, 2)) return r2;
    m->lambda = back[0];
    m->mu = back[1];
    return rc;
}

int qm_nesterov_get_state(const qm_nesterov* m, int* iterations, double* forceMax) {
    if (!m) return fail(QM_EINVAL, "null minimizer");
    if (iterations) *iterations = m->iterations;
    if (forceMax) *forceMax = m->forceMax;
    return QM_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ online L-BFGS
struct qm_lolbfgs : VecMinimizer {
    // setLoLBFGSParameters defaults (energyMinimizerLoLBFGS.h:29-45); the memory
    // only grows, starting from the constructor's 5
    int m = 5;
    double dt = 1e-4, c = 1.0, tau = 10.0;
    int scheduled = 0;
    int cur = 0;   // currentIterationInMLoop
    bool fresh = true;   // p must be (re)started from the forces: new or re-initialised history
    // vecs: [0] p, [1 .. m] s, [m+1 .. 2m] y
    void* p() const { return vecs[0]; }
    void* s(int t) const { return vecs[1 + t]; }
    void* y(int t) const { return vecs[1 + m + t]; }
};

template <typename R>
static int lb_enqueue(qm_lolbfgs* M, qm_lattice* L, const qm_ldg* f, int iteration) {
    const int m = M->m, c = M->cur;
    const LbLayout lay{m};
    const StateSet<R> st = state_set<R>(L, L->hCtrl.cur);
    const Common com = M->common(L);
    const int blocks = vec_blocks(com.len);
    auto slot = [&](int ii) { return (c - ii - 1 + m) % m; };
    if (iteration == 0 || M->fresh) {
        // step 1 (energyMinimizerLoLBFGS.cpp:107-114): p = F(Q); then alpha_0 and
        // the step-3 factor of the (empty) newest slot.  A history re-initialised
        // after iterations also restarts here: its direction vector is new.
        M->fresh = false;
        if (int rc = force_pass(L, f, true)) return rc;
        LbFinishArgs<R> a{};
        a.p = rp<R>(M->p());
        a.f = st.f;
        a.s = rp<R>(M->s(slot(0)));
        a.y = nullptr;
        a.c = com;
        lb_finish_kernel<R><<<blocks, kVecThreads, 0, L->stream>>>(a);
        HIP_TRY(hipGetLastError());
        FinArgs fa{};
        fa.op = kFinLbStart;
        fa.t = slot(0);
        fa.lay = lay;
        if (int rc = M->finalize(L, fa, blocks)) return rc;
    }
    auto stage = [&](const R* x1, int k1, const R* x2, int k2, const R* d, int op, int ii, int t) {
        LbStageArgs<R> a{};
        a.p = rp<R>(M->p());
        a.x1 = x1;
        a.x2 = x2;
        a.d = d;
        a.k1 = k1;
        a.k2 = k2;
        a.scal = M->dScal;
        a.c = com;
        lb_stage_kernel<R><<<blocks, kVecThreads, 0, L->stream>>>(a);
        HIP_TRY(hipGetLastError());
        FinArgs fa{};
        fa.op = op;
        fa.ii = ii;
        fa.t = t;
        fa.lay = lay;
        return M->finalize(L, fa, blocks);
    };
    // step 2: p -= alpha_ii y; alpha_{ii+1} = s.p / s.y (newest first)
    for (int ii = 0; ii + 1 < m; ++ii)
        if (int rc = stage(rp<R>(M->y(slot(ii))), lay.NA(ii), nullptr, 0, rp<R>(M->s(slot(ii + 1))), kFinAlpha,
                           ii + 1, slot(ii + 1)))
            return rc;
    // last step-2 update fused with step 3 (p += (s.y / y.y) y of the newest
    // slot), then beta of the oldest slot
    if (int rc = stage(rp<R>(M->y(slot(m - 1))), lay.NA(m - 1), rp<R>(M->y(slot(0))), lay.K3(),
                       rp<R>(M->y(slot(m - 1))), kFinBeta, m - 1, slot(m - 1)))
        return rc;
    // step 4: p += (alpha_ii - beta_ii) s; beta_{ii-1} = y.p / s.y (oldest first)
    for (int ii = m - 1; ii >= 1; --ii)
        if (int rc = stage(rp<R>(M->s(slot(ii))), lay.CS(ii), nullptr, 0, rp<R>(M->y(slot(ii - 1))), kFinBeta,
                           ii - 1, slot(ii - 1)))
            return rc;
    // update step: eta as minimize() sets it (energyMinimizerLoLBFGS.cpp:203-209)
    const double eta = M->scheduled ? M->tau / (M->tau + iteration) * M->dt : M->dt;
    {
        LbMoveArgs<R> a{};
        a.p = rp<R>(M->p());
        a.s0 = rp<R>(M->s(slot(0)));
        a.sc = rp<R>(M->s(c));
        a.yc = rp<R>(M->y(c));
        a.q = st.q;
        a.f = st.f;
        a.eta = eta;
        a.k = lay.CS(0);
        a.scal = M->dScal;
        a.c = com;
        lb_move_kernel<R><<<blocks, kVecThreads, 0, L->stream>>>(a);
        HIP_TRY(hipGetLastError());
    }
    if (int rc = force_pass(L, f, true)) return rc;
    {
        LbFinishArgs<R> a{};
        a.p = rp<R>(M->p());
        a.f = st.f;
        a.s = rp<R>(M->s(c));
        a.y = rp<R>(M->y(c));
        a.c = com;
        lb_finish_kernel<R><<<blocks, kVecThreads, 0, L->stream>>>(a);
        HIP_TRY(hipGetLastError());
        FinArgs fa{};
        fa.op = kFinLbFinish;
        fa.t = c;
        fa.lay = lay;
        if (int rc = M->finalize(L, fa, blocks)) return rc;
    }
    M->cur = (c + 1) % m;
    return QM_OK;
}

extern "C" {

int qm_lolbfgs_create(qm_lolbfgs** out) {
    if (!out) return fail(QM_EINVAL, "null output handle");
    *out = new qm_lolbfgs;
    return QM_OK;
}

int qm_lolbfgs_destroy(qm_lolbfgs* m) {
    delete m;
    return QM_OK;
}

int qm_lolbfgs_set_parameters(qm_lolbfgs* m, int memory, double dt, double c, double forceCutoff, double tau,
                              int scheduledMomentum) {
    if (!m) return fail(QM_EINVAL, "null minimizer");
    m->c = c;
    m->dt = dt;
    m->tau = tau;
    m->fc = forceCutoff;
    m->scheduled = scheduledMomentum != 0;
    if (memory > m->m) {
        // a larger memory re-initialises the history (energyMinimizerLoLBFGS.h:36-43)
        m->m = memory;
        m->release();
    }
    return QM_OK;
}

int qm_lolbfgs_set_max_iterations(qm_lolbfgs* m, int maxIterations) {
    if (!m) return fail(QM_EINVAL, "null minimizer");
    m->maxIter = maxIterations;
    return QM_OK;
}

int qm_lolbfgs_minimize(qm_lolbfgs* M, qm_ldg* f, qm_lattice* L) {
    if (!M || !f || !L) return fail(QM_EINVAL, "null handle");
    if (M->needs_bind(L)) {
        // initializeFromModel (energyMinimizerLoLBFGS.cpp:32-52): zero history
        if (int rc = M->bind(L, 1 + 2 * M->m, LbLayout{M->m}.size())) return rc;
        M->fresh = true;
        M->iterations = 0;
        M->cur = 0;
    }
    // at least one iteration per call (energyMinimizerLoLBFGS.cpp:199-201)
    if (int rc = begin_run(*M, L, f, false)) return rc;
    const int cur0 = M->cur, it0 = M->iterations;
    auto enq = [&](int iteration) {
        return L->precision == QM_FP64 ? lb_enqueue<double>(M, L, f, iteration)
                                       : lb_enqueue<float>(M, L, f, iteration);
    };
    int rc = drive(*M, L, enq);
    M->cur = (cur0 + (M->iterations - it0)) % M->m;
    return rc;
}

int qm_lolbfgs_get_state(const qm_lolbfgs* m, int* iterations, double* forceMax) {
    if (!m) return fail(QM_EINVAL, "null minimizer");
    if (iterations) *iterations = m->iterations;
    if (forceMax) *forceMax = m->forceMax;
    return QM_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ traces
template <typename M>
static int set_trace(M* m, int capacity) {
    if (!m || capacity < 0) return fail(QM_EINVAL, "bad argument");
    m->traceCap = capacity;
    m->trace.clear();
    return QM_OK;
}

template <typename M>
static int get_trace(const M* m, double* out, int cap, int* n) {
    if (!m || !n) return fail(QM_EINVAL, "bad argument");
    const int k = std::min<int>(cap, (int)m->trace.size());
    if (out && k > 0) std::memcpy(out, m->trace.data(), (size_t)k * sizeof(double));
    *n = k;
    return QM_OK;
}

extern "C" {
int qm_adam_set_trace(qm_adam* m, int capacity) { return set_trace(m, capacity); }
int qm_adam_get_trace(const qm_adam* m, double* forceMax, int cap, int* n) { return get_trace(m, forceMax, cap, n); }
int qm_nesterov_set_trace(qm_nesterov* m, int capacity) { return set_trace(m, capacity); }
int qm_nesterov_get_trace(const qm_nesterov* m, double* forceMax, int cap, int* n) {
    return get_trace(m, forceMax, cap, n);
}
int qm_lolbfgs_set_trace(qm_lolbfgs* m, int capacity) { return set_trace(m, capacity); }
int qm_lolbfgs_get_trace(const qm_lolbfgs* m, double* forceMax, int cap, int* n) {
    return get_trace(m, forceMax, cap, n);
}
}
