// gfx950 kernels of the lattice Landau-de Gennes minimizer.
//
// Fused lattice pass (march_kernel): one work-group owns a 32 x 8 column of
// sites and marches it along a z chunk.  Each step stages one halo-padded
// plane of the new state into a 3-plane LDS ring: in FIRE mode the pending
// FIRE velocity mix, the velocity-Verlet drift and the first half kick are
// applied while loading (Q' = Q + dt v + dt^2/2 F, v' = v + dt/2 F), so the
// force at Q' is computed from LDS without another pass over HBM.  Every
// site's force needs only its 3x3x3 neighbourhood minus the corners: the
// reference's stencils use neighbour derivatives transverse to the neighbour
// direction only (src/forces/lcForces.h:66-1043).  The second half kick and
// the FP64 partial sums of F.F, F.v, v.v (QT metric, inc/functions.h:140-147)
// close the step.  Sites next to an object (t < 0) are left to the boundary
// pass, which keeps the rarely-needed 27-case boundary stencils out of this
// kernel's register budget.
//
// Boundary pass (boundary_kernel): one thread per site of the compacted t < 0
// list; 27-case stencils, anchoring, fields and metric.  A one-work-group
// kernel (fire_finalize_kernel) then sums all partials in a fixed order and
// advances the FIRE scalars on the device (energyMinimizerFIRE.cpp:77-145): no
// host round trip.
#include "lattice_pass.h"

namespace qmin {

// FIRE scalar step of energyMinimizerFIRE::fireStepCPU (energyMinimizerFIRE.cpp:77-145),
// run by one thread once all partial sums are in.
__device__ void fire_scalar_step(Ctrl* ctrl, const FireConst& fp, double fnorm, double power, double vnorm,
                                 double* trace) {
    const int it = ctrl->iterations + 1;
    const double forceMax = sqrt(fnorm) / fp.nTotal;
    const double scaling = fnorm > 0. ? sqrt(vnorm / fnorm) : 0.0;
    double alpha = ctrl->alpha, dt = ctrl->dt;
    int nPos = ctrl->nPositive;
    ctrl->cKeep = 1.0 - alpha;
    ctrl->cForce = alpha * scaling;
    ctrl->identity = 0;
    if (power > 0 && it % 500 != 0) {
        if (nPos > fp.nMin) {
            dt = fmin(dt * fp.dtInc, fp.dtMax);
            alpha = alpha * fp.alphaDec;
            alpha = fmax(alpha, fp.alphaMin);
        }
        nPos += 1;
        ctrl->reset = 0;
    } else {
        nPos = 0;
        dt = dt * fp.dtDec;
        dt = fmax(dt, fp.dtMin);
        alpha = fp.alphaStart;
        ctrl->reset = 1;
    }
    ctrl->alpha = alpha;
    ctrl->dt = dt;
    ctrl->nPositive = nPos;
    ctrl->iterations = it;
    ctrl->forceMax = forceMax;
    ctrl->power = power;
    ctrl->fnorm = fnorm;
    ctrl->vnorm = vnorm;
    ctrl->scaling = scaling;
    if (trace && ctrl->traceLen < ctrl->traceCap) {
        double* t = trace + 6 * (long long)ctrl->traceLen;
        t[0] = forceMax;
        t[1] = power;
        t[2] = alpha;
        t[3] = dt;
        t[4] = scaling;
        t[5] = nPos;
        ctrl->traceLen += 1;
    }
    ctrl->cur ^= 1;
    if (isnan(forceMax)) ctrl->nanFlag = 1;
    ctrl->done = !(it < ctrl->maxIter && forceMax > fp.fc);
}

// FIRE scalar step after both lattice passes of an iteration: one work-group
// sums every partial in a fixed order (fused-pass blocks, then boundary-pass
// blocks) and advances the FIRE scalars.  A kernel of its own: the launch
// boundary makes the partials visible without device-wide fences, which cost
// a write-back of each XCD's L2 right after the lattice pass has filled it.
__device__ __forceinline__ void fire_finalize_body(Ctrl* ctrl, const FireConst& fp, const double* part,
                                                   long long total, double* trace) {
    __shared__ double sRed[kReduceThreads / 64][3];
    if (ctrl->done) return;
    double s0 = 0, s1 = 0, s2 = 0;
    for (long long j = threadIdx.x; j < total; j += kReduceThreads) {
        s0 += part[3 * j + 0];
        s1 += part[3 * j + 1];
        s2 += part[3 * j + 2];
    }
    block_sum3<kReduceThreads>(s0, s1, s2, sRed);
    if (threadIdx.x == 0) fire_scalar_step(ctrl, fp, s0, s1, s2, trace);
}

__global__ void __launch_bounds__(kReduceThreads) fire_finalize_kernel(Ctrl* ctrl, FireConst fp, const double* part,
                                                                      long long total, double* trace) {
    fire_finalize_body(ctrl, fp, part, total, trace);
}

// One work-group per lattice of an ensemble; each also copies its lattice's
// progress into a compact status row for the host.
__global__ void __launch_bounds__(kReduceThreads) fire_finalize_batch_kernel(const FinalizeDesc* descs,
                                                                            EnsembleStatus* status) {
    const FinalizeDesc& d = descs[blockIdx.x];
    fire_finalize_body(d.ctrl, d.fire, d.partials, d.total, nullptr);
    if (threadIdx.x == 0) {
        EnsembleStatus st;
        st.done = d.ctrl->done;
        st.iterations = d.ctrl->iterations;
        st.forceMax = d.ctrl->forceMax;
        status[blockIdx.x] = st;
    }
}

// ---------------------------------------------------------------- other passes

// Materialize the pending FIRE velocity mix into U of the current set.
template <typename R>
__global__ void materialize_velocity_kernel(StateSet<R> s0, StateSet<R> s1, long long n, Ctrl* ctrl) {
    const StateSet<R> st = ctrl->cur ? s1 : s0;
    const int mix = ctrl->reset ? 2 : (ctrl->identity ? 0 : 1);
    const R ck = R(ctrl->cKeep), cf = R(ctrl->cForce);
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < 5 * n; i += (long long)gridDim.x * blockDim.x) {
        const R u = st.u[i];
        st.u[i] = mix == 2 ? R(0) : (mix == 0 ? u : ck * u + cf * st.f[i]);
    }
}

// Energy density of liquid-crystal site i split in its components [phase,
// distortion, anchoring, E, H] (landauDeGennesLC.cpp:204-353), and the first
// derivatives there (central; one-sided towards the liquid crystal on an axis
// with an object neighbour, the upper neighbour taking precedence).
template <typename R>
__device__ __forceinline__ void site_energy(const R* q, const SiteType* types, const DevObject* objects, int lx, int ly,
                                            int lz, const EnergyConst& ec, long long i, double* parts,
                                            R (*dv)[5]) {
    const long long n = (long long)lx * ly * lz;
    const int t = types[i];
    const int x = int(i % lx), y = int((i / lx) % ly), z = int(i / ((long long)lx * ly));
    long long nbi[6];
    nbi[0] = wrapi(x - 1, lx) + (long long)lx * (y + (long long)ly * z);
    nbi[1] = wrapi(x + 1, lx) + (long long)lx * (y + (long long)ly * z);
    nbi[2] = x + (long long)lx * (wrapi(y - 1, ly) + (long long)ly * z);
    nbi[3] = x + (long long)lx * (wrapi(y + 1, ly) + (long long)ly * z);
    nbi[4] = x + (long long)lx * (y + (long long)ly * wrapi(z - 1, lz));
    nbi[5] = x + (long long)lx * (y + (long long)ly * wrapi(z + 1, lz));
    R qc[5], nq[6][5];
    gload(q, n, i, qc);
#pragma unroll
    for (int d = 0; d < 6; ++d) gload(q, n, nbi[d], nq[d]);
    const R t2 = trQ2(qc);
    parts[0] = double(R(ec.a) * t2 + R(ec.b) * trQ3(qc) + R(ec.c) * t2 * t2);
    parts[2] = 0;
    parts[3] = ec.eOn ? double(field_energy(ec.E, ec.eps, ec.deps, ec.eps0, qc)) : 0.0;
    parts[4] = ec.hOn ? double(field_energy(ec.H, ec.chi, ec.dchi, ec.mu0, qc)) : 0.0;
    if (ec.hvar) {
        const double h3[3] = {ec.hvar[i], ec.hvar[n + i], ec.hvar[2 * n + i]};
        parts[4] += double(field_energy(h3, ec.chi, ec.dchi, ec.mu0, qc));
    }
#pragma unroll
    for (int ax = 0; ax < 3; ++ax)
#pragma unroll
        for (int c = 0; c < 5; ++c) dv[ax][c] = R(0.5) * (nq[2 * ax + 1][c] - nq[2 * ax][c]);
    if (t < 0) {
        R anch = 0;
#pragma unroll
        for (int d = 0; d < 6; ++d) {
            const int tn = types[nbi[d]];
            if (tn <= 0) continue;
            anch += anchoring_energy(objects[tn - 1], qc, nq[d]);
            const int ax = d / 2;
#pragma unroll
            for (int c = 0; c < 5; ++c)
                dv[ax][c] = (d & 1) ? (qc[c] - nq[d ^ 1][c]) : (nq[d ^ 1][c] - qc[c]);
        }
        parts[2] = double(anch);
    }
    R e[5];
    distortion_densities(qc, dv[0], dv[1], dv[2], e);
    double dist = 0;
#pragma unroll
    for (int k = 0; k < 5; ++k)
        if (ec.L[k] != 0) dist += ec.L[k] * double(e[k]);
    parts[1] = dist;
}

// Energy: per-block partial sums of the five components over t <= 0 sites.
template <typename R>
__global__ void __launch_bounds__(kReduceThreads) energy_kernel(const R* q, const SiteType* types,
                                                                const DevObject* objects, int lx, int ly, int lz,
                                                                EnergyConst ec, double* partials) {
    __shared__ double sRed[kReduceThreads / 64][5];
    const long long n = (long long)lx * ly * lz;
    double acc[5] = {0, 0, 0, 0, 0};
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        if (types[i] > 0) continue;
        double parts[5];
        R dv[3][5];
        site_energy(q, types, objects, lx, ly, lz, ec, i, parts, dv);
        acc[0] += parts[0];
        acc[3] += parts[3];
        acc[4] += parts[4];
        acc[2] += parts[2];
        acc[1] += parts[1];
    }
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        const double v = wave_sum(acc[k]);
        if (l == 0) sRed[w][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < 5) {
        double s = 0;
        for (int i = 0; i < kReduceThreads / 64; ++i) s += sRed[i][threadIdx.x];
        partials[5 * (long long)blockIdx.x + threadIdx.x] = s;
    }
}

// Per-site energy density (landauDeGennesLC::energyDensity): 0 on object
// sites, else phase, fields, anchoring and distortion, added in the order of
// computeEnergyCPU (landauDeGennesLC.cpp:204-353).
template <typename R>
__global__ void __launch_bounds__(kReduceThreads) energy_density_kernel(const R* q, const SiteType* types,
                                                                        const DevObject* objects, int lx, int ly,
                                                                        int lz, EnergyConst ec, double* out) {
    const long long n = (long long)lx * ly * lz;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        if (types[i] > 0) {
            out[i] = 0.0;
            continue;
        }
        double parts[5];
        R dv[3][5];
        site_energy(q, types, objects, lx, ly, lz, ec, i, parts, dv);
        out[i] = parts[0] + parts[3] + parts[4] + parts[2] + parts[1];
    }
}

// Force on an object from its surface sites (computeObjectForces with
// computeStressTensors, landauDeGennesLCOtherForces.cpp:147-251; one-constant):
// per surface site, the stress tensor of the first derivatives and L1 with the
// site's energy density on the diagonal, applied to the surface-area vector
// (+-1 per axis towards an object neighbour, the upper one taking precedence).
// out[3 k + 0..2] is site k's contribution; the host adds them in site order.
template <typename R>
__global__ void object_force_kernel(const R* q, const SiteType* types, const DevObject* objects, int lx, int ly, int lz,
                                    EnergyConst ec, double L1, const int64_t* sites, long long m, double* out) {
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < m; k += (long long)gridDim.x * blockDim.x) {
        const long long i = sites[k];
        double* o = out + 3 * k;
        if (types[i] > 0) {   // a surface site taken by a later object
            o[0] = o[1] = o[2] = 0.0;
            continue;
        }
        double parts[5];
        R dvr[3][5];
        site_energy(q, types, objects, lx, ly, lz, ec, i, parts, dvr);
        // the energy density in the order computeEnergyCPU adds it
        const double e = (((parts[0] + parts[3]) + parts[4]) + parts[2]) + parts[1];
        double X[5], Y[5], Z[5];
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            X[c] = double(dvr[0][c]);
            Y[c] = double(dvr[1][c]);
            Z[c] = double(dvr[2][c]);
        }
        auto diag = [&](const double* D) {
            return -2 * L1 * (D[0] * D[0] + 2 * (D[1] * D[1]) + 2 * (D[2] * D[2]) + D[3] * D[3] + 2 * (D[4] * D[4]) +
                              D[0] * D[3]);
        };
        auto off = [&](const double* P, const double* S) {
            return L1 * (-(S[0] * (2 * P[0] + P[3])) - S[3] * (P[0] + 2 * P[3]) -
                         4 * (S[1] * P[1] + S[2] * P[2] + S[4] * P[4]));
        };
        const double m11 = diag(X) + e, m22 = diag(Y) + e, m33 = diag(Z) + e;
        const double mxy = off(X, Y), mxz = off(X, Z), myz = off(Y, Z);
        const int x = int(i % lx), y = int((i / lx) % ly), z = int(i / ((long long)lx * ly));
        double vx = 0, vy = 0, vz = 0;
        if (types[wrapi(x - 1, lx) + (long long)lx * (y + (long long)ly * z)] > 0) vx = -1.0;
        if (types[wrapi(x + 1, lx) + (long long)lx * (y + (long long)ly * z)] > 0) vx = 1.0;
        if (types[x + (long long)lx * (wrapi(y - 1, ly) + (long long)ly * z)] > 0) vy = -1.0;
        if (types[x + (long long)lx * (wrapi(y + 1, ly) + (long long)ly * z)] > 0) vy = 1.0;
        if (types[x + (long long)lx * (y + (long long)ly * wrapi(z - 1, lz))] > 0) vz = -1.0;
        if (types[x + (long long)lx * (y + (long long)ly * wrapi(z + 1, lz))] > 0) vz = 1.0;
        // the stress is symmetric: x12 = x21 etc. (scalar3 * Matrix3x3, inc/matrix.h:281-288)
        o[0] = m11 * vx + mxy * vy + mxz * vz;
        o[1] = mxy * vx + m22 * vy + myz * vz;
        o[2] = mxz * vx + myz * vy + m33 * vz;
    }
}

// Moves the Q of sites src[k] to dst[k] (k < m) in two passes through tmp, so
// that every source is read before any destination is written.
template <typename R>
__global__ void gather_q_kernel(const R* q, long long n, const int64_t* src, long long m, R* tmp) {
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < m; k += (long long)gridDim.x * blockDim.x)
#pragma unroll
        for (int c = 0; c < 5; ++c) tmp[c * m + k] = q[c * n + src[k]];
}

template <typename R>
__global__ void scatter_q_kernel(R* q, long long n, const int64_t* dst, long long m, const R* tmp) {
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < m; k += (long long)gridDim.x * blockDim.x)
#pragma unroll
        for (int c = 0; c < 5; ++c) q[c * n + dst[k]] = tmp[c * m + k];
}

// qTensorLatticeModel::computeDefectMeasures (qTensorLatticeModel.cpp:95-133) with
// eigenvaluesOfQ (qTensorFunctions.h:215-241)
template <typename R>
__global__ void defect_kernel(const R* q, const SiteType* types, long long n, int kind, double* out) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        if (types[i] > 0) {
            out[i] = 0.0;
            continue;
        }
        double v[5];
#pragma unroll
        for (int c = 0; c < 5; ++c) v[c] = double(q[c * n + i]);
        const double w = -v[0] - v[3];
        const double det = v[0] * (v[3] * w - v[4] * v[4]) - v[1] * (v[1] * w - v[4] * v[2]) +
                           v[2] * (v[1] * v[4] - v[3] * v[2]);
        if (kind == 1) {
            out[i] = det;
            continue;
        }
        if (kind == 2) {
            const double t = trQ2(v);
            out[i] = t * t * t - 54.0 * det * det;
            continue;
        }
        const double p1 = v[1] * v[1] + v[2] * v[2] + v[4] * v[4];
        if (p1 == 0) {
            out[i] = fmax(fmax(v[0], v[3]), w);
            continue;
        }
        const double p2 = 2 * (v[0] * v[0] + v[3] * v[3] + v[0] * v[3]) + 2 * p1;
        const double p = sqrt(p2 / 6.0);
        const double r = 0.5 * det / (p * p * p);
        const double pi = 3.14159265358979323846;
        double phi;
        if (r <= -1)
            phi = pi / 3.0;
        else if (r >= 1)
            phi = 0.0;
        else
            phi = acos(r) / 3.0;
        const double a = 2.0 * p * cos(phi);
        const double c = 2.0 * p * cos(phi + 2.0 * pi / 3.0);
        out[i] = fmax(fmax(a, -a - c), c);
    }
}

// ---------------------------------------------------------------- host launchers

template <typename R, int MODE, int TERMS>
static const void* march_fn_tile(int tile) {
    switch (tile) {
        case kTile64x4: return reinterpret_cast<const void*>(&march_kernel<R, MODE, TERMS, kTile64x4>);
        case kTile32x16: return reinterpret_cast<const void*>(&march_wide_kernel<R, MODE, TERMS, kTile32x16>);
        case kTile64x8: return reinterpret_cast<const void*>(&march_wide_kernel<R, MODE, TERMS, kTile64x8>);
        default: return reinterpret_cast<const void*>(&march_kernel<R, MODE, TERMS, kTile32x8>);
    }
}

template <typename R, int MODE>
static const void* march_fn_terms(int terms, int tile) {
    switch (terms) {
        case kOneConstant: return march_fn_tile<R, MODE, kOneConstant>(tile);
        case kMultiLinear: return march_fn_tile<R, MODE, kMultiLinear>(tile);
        case kMultiL4: return march_fn_tile<R, MODE, kMultiL4>(tile);
        default: return march_fn_tile<R, MODE, kMultiL6>(tile);
    }
}

template <typename R>
static const void* march_fn(int mode, int terms, int tile) {
    return mode == kFire ? march_fn_terms<R, kFire>(terms, tile) : march_fn_terms<R, kForce>(terms, tile);
}

template <typename R, int MODE, int TERMS>
static void launch_boundary_t(const BoundaryArgs<R>& a, int blocks, hipStream_t s) {
    hipLaunchKernelGGL((boundary_kernel<R, MODE, TERMS>), dim3(blocks), dim3(kBoundaryThreads), 0, s, a);
}

template <typename R, int MODE, int TERMS>
static const void* march_batch_fn_tile(int tile) {
    switch (tile) {
        case kTile64x4: return reinterpret_cast<const void*>(&march_batch_kernel<R, MODE, TERMS, kTile64x4>);
        case kTile32x16: return reinterpret_cast<const void*>(&march_wide_batch_kernel<R, MODE, TERMS, kTile32x16>);
        case kTile64x8: return reinterpret_cast<const void*>(&march_wide_batch_kernel<R, MODE, TERMS, kTile64x8>);
        default: return reinterpret_cast<const void*>(&march_batch_kernel<R, MODE, TERMS, kTile32x8>);
    }
}

template <typename R, int MODE>
static const void* march_batch_fn_terms(int terms, int tile) {
    switch (terms) {
        case kOneConstant: return march_batch_fn_tile<R, MODE, kOneConstant>(tile);
        case kMultiLinear: return march_batch_fn_tile<R, MODE, kMultiLinear>(tile);
        case kMultiL4: return march_batch_fn_tile<R, MODE, kMultiL4>(tile);
        default: return march_batch_fn_tile<R, MODE, kMultiL6>(tile);
    }
}

template <typename R>
hipError_t launch_march_batch(const MarchArgs<R>* descs, const int2* map, int blocks, int mode, int terms, int tile,
                              hipStream_t s) {
    const void* fn = mode == kFire ? march_batch_fn_terms<R, kFire>(terms, tile)
                                   : march_batch_fn_terms<R, kForce>(terms, tile);
    void* args[] = {&descs, &map};
    return hipLaunchKernel(fn, dim3(blocks), dim3(kTileThreads[tile]), args, 0, s);
}

template <typename R>
hipError_t launch_boundary_batch(const BoundaryArgs<R>* descs, const int2* map, int blocks, int mode, int terms,
                                 hipStream_t s) {
    const void* fn;
    if (mode == kFire)
        fn = terms == kOneConstant ? reinterpret_cast<const void*>(&boundary_batch_kernel<R, kFire, kOneConstant>)
           : terms == kMultiL6     ? reinterpret_cast<const void*>(&boundary_batch_kernel<R, kFire, kMultiL6>)
                                   : reinterpret_cast<const void*>(&boundary_batch_kernel<R, kFire, kMultiLinear>);
    else
        fn = terms == kOneConstant ? reinterpret_cast<const void*>(&boundary_batch_kernel<R, kForce, kOneConstant>)
           : terms == kMultiL6     ? reinterpret_cast<const void*>(&boundary_batch_kernel<R, kForce, kMultiL6>)
                                   : reinterpret_cast<const void*>(&boundary_batch_kernel<R, kForce, kMultiLinear>);
    void* args[] = {&descs, &map};
    return hipLaunchKernel(fn, dim3(blocks), dim3(kBoundaryThreads), args, 0, s);
}

hipError_t launch_fire_finalize_batch(const FinalizeDesc* descs, EnsembleStatus* status, int count, hipStream_t s) {
    hipLaunchKernelGGL(fire_finalize_batch_kernel, dim3(count), dim3(kReduceThreads), 0, s, descs, status);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_march(const MarchArgs<R>& a, int mode, int terms, int tile, hipStream_t s) {
    void* args[] = {const_cast<MarchArgs<R>*>(&a)};
    return hipLaunchKernel(march_fn<R>(mode, terms, tile), dim3(a.nblocks), dim3(kTileThreads[tile]), args, 0, s);
}

template <typename R>
void launch_boundary(const BoundaryArgs<R>& a, int mode, int terms, hipStream_t s) {
    long long blocks = (a.nsites + kBoundaryThreads - 1) / kBoundaryThreads;
    if (blocks < 1) blocks = 1;
    // the boundary pass does not distinguish L1..L3 from L1..L4
    if (mode == kFire) {
        if (terms == kOneConstant) launch_boundary_t<R, kFire, kOneConstant>(a, int(blocks), s);
        else if (terms == kMultiL6) launch_boundary_t<R, kFire, kMultiL6>(a, int(blocks), s);
        else launch_boundary_t<R, kFire, kMultiLinear>(a, int(blocks), s);
    } else {
        if (a.nsites == 0) return;
        if (terms == kOneConstant) launch_boundary_t<R, kForce, kOneConstant>(a, int(blocks), s);
        else if (terms == kMultiL6) launch_boundary_t<R, kForce, kMultiL6>(a, int(blocks), s);
        else launch_boundary_t<R, kForce, kMultiLinear>(a, int(blocks), s);
    }
}

template <typename R>
void launch_object_force(const R* q, const SiteType* types, const DevObject* objects, int lx, int ly, int lz,
                         const EnergyConst& ec, double L1, const int64_t* sites, long long m, double* out,
                         hipStream_t s) {
    const long long blocks = std::max<long long>(1, std::min<long long>((m + 127) / 128, 4096));
    hipLaunchKernelGGL(object_force_kernel<R>, dim3(unsigned(blocks)), dim3(128), 0, s, q, types, objects, lx, ly, lz,
                       ec, L1, sites, m, out);
}

template <typename R>
void launch_move_q(R* q, long long n, const int64_t* src, const int64_t* dst, long long m, R* tmp, hipStream_t s) {
    const long long blocks = std::max<long long>(1, std::min<long long>((m + 255) / 256, 4096));
    hipLaunchKernelGGL(gather_q_kernel<R>, dim3(unsigned(blocks)), dim3(256), 0, s, q, n, src, m, tmp);
    hipLaunchKernelGGL(scatter_q_kernel<R>, dim3(unsigned(blocks)), dim3(256), 0, s, q, n, dst, m, tmp);
}

// Work-groups of the fused pass that fit on one CU at once.
template <typename R>
int march_blocks_per_cu(int mode, int terms, int tile) {
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, march_fn<R>(mode, terms, tile), kTileThreads[tile], 0) !=
        hipSuccess)
        return 1;
    return blocks < 1 ? 1 : blocks;
}

void launch_fire_finalize(Ctrl* ctrl, const FireConst& fp, const double* partials, long long total, double* trace,
                          hipStream_t s) {
    hipLaunchKernelGGL(fire_finalize_kernel, dim3(1), dim3(kReduceThreads), 0, s, ctrl, fp, partials, total, trace);
}

int boundary_blocks(long long nsites) {
    long long b = (nsites + kBoundaryThreads - 1) / kBoundaryThreads;
    return b < 1 ? 1 : int(b);
}

template <typename R>
void launch_materialize(StateSet<R> s0, StateSet<R> s1, long long n, Ctrl* ctrl, hipStream_t s) {
    long long blocks = (5 * n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL((materialize_velocity_kernel<R>), dim3(blocks), dim3(256), 0, s, s0, s1, n, ctrl);
}

template <typename R>
int launch_energy(const R* q, const SiteType* types, const DevObject* objects, int lx, int ly, int lz,
                  const EnergyConst& ec, double* partials, hipStream_t s) {
    long long n = (long long)lx * ly * lz;
    long long blocks = (n + kReduceThreads - 1) / kReduceThreads;
    if (blocks > kMaxEnergyBlocks) blocks = kMaxEnergyBlocks;
    hipLaunchKernelGGL((energy_kernel<R>), dim3(blocks), dim3(kReduceThreads), 0, s, q, types, objects, lx, ly, lz,
                       ec, partials);
    return int(blocks);
}

template <typename R>
void launch_energy_density(const R* q, const SiteType* types, const DevObject* objects, int lx, int ly, int lz,
                           const EnergyConst& ec, double* out, hipStream_t s) {
    const long long n = (long long)lx * ly * lz;
    long long blocks = (n + kReduceThreads - 1) / kReduceThreads;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL((energy_density_kernel<R>), dim3(blocks), dim3(kReduceThreads), 0, s, q, types, objects, lx,
                       ly, lz, ec, out);
}

template <typename R>
void launch_defects(const R* q, const SiteType* types, long long n, int kind, double* out, hipStream_t s) {
    long long blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL((defect_kernel<R>), dim3(blocks), dim3(256), 0, s, q, types, n, kind, out);
}

#define QMIN_INSTANTIATE(R)                                                                                      \
    template hipError_t launch_march<R>(const MarchArgs<R>&, int, int, int, hipStream_t);                                   \
    template void launch_boundary<R>(const BoundaryArgs<R>&, int, int, hipStream_t);                             \
    template void launch_materialize<R>(StateSet<R>, StateSet<R>, long long, Ctrl*, hipStream_t);                \
    template int launch_energy<R>(const R*, const SiteType*, const DevObject*, int, int, int, const EnergyConst&,   \
                                  double*, hipStream_t);                                                        \
    template void launch_defects<R>(const R*, const SiteType*, long long, int, double*, hipStream_t);             \
    template void launch_energy_density<R>(const R*, const SiteType*, const DevObject*, int, int, int,              \
                                           const EnergyConst&, double*, hipStream_t);                            \
    template int march_blocks_per_cu<R>(int, int, int);                                                          \
    template hipError_t launch_march_batch<R>(const MarchArgs<R>*, const int2*, int, int, int, int, hipStream_t);  \
    template hipError_t launch_boundary_batch<R>(const BoundaryArgs<R>*, const int2*, int, int, int, hipStream_t);  \
    template void launch_object_force<R>(const R*, const SiteType*, const DevObject*, int, int, int, const EnergyConst&, \
                                         double, const int64_t*, long long, double*, hipStream_t);                 \
    template void launch_move_q<R>(R*, long long, const int64_t*, const int64_t*, long long, R*, hipStream_t);

QMIN_INSTANTIATE(double)
QMIN_INSTANTIATE(float)

}  // namespace qmin
