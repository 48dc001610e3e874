// Ensembles: many independent lattices (different sizes, seeds, objects and
// anchoring) minimized with FIRE together.  Each kernel of an iteration is one
// launch over every lattice: a work-group looks up {lattice, tile} in a map
// built here (laid out in XCD-aware order over the whole ensemble) and reads
// that lattice's arguments from a device array.  Each lattice keeps its own
// control block, so converged lattices drop out of later launches on their own.
// A batch of iterations is captured once in a hipGraph and replayed; the host
// reads one compact status row per lattice between replays.
#include <algorithm>
#include <memory>
#include <numeric>
#include <vector>

#include "qmin_internal.h"

namespace qmin {
template <typename R>
hipError_t launch_march_batch(const MarchArgs<R>* descs, const int2* map, int blocks, int mode, int terms, int tile,
                              hipStream_t s);
template <typename R>
hipError_t launch_boundary_batch(const BoundaryArgs<R>* descs, const int2* map, int blocks, int mode, int terms,
                                 hipStream_t s);
hipError_t launch_fire_finalize_batch(const FinalizeDesc* descs, EnsembleStatus* status, int count, hipStream_t s);
}  // namespace qmin

using namespace qmin;

constexpr int kTilePerLattice = QM_TILE_PER_LATTICE;

struct qm_ensemble {
    std::vector<qm_lattice*> lats;
    int precision = QM_FP64;
    hipStream_t stream = nullptr;
    void* dMarchDescs = nullptr;
    void* dBoundaryDescs = nullptr;
    FinalizeDesc* dFinal = nullptr;
    int2* dMarchMap = nullptr;
    int2* dBoundaryMap = nullptr;
    EnsembleStatus* dStatus = nullptr;
    int boundaryBlocks = 0;
    // work-groups of the batched fused pass per tile shape, in one map: entries
    // [segOff[t], segOff[t] + segCount[t]) run with tile t (one launch each)
    int segOff[kNumMarchTiles] = {0, 0, 0, 0}, segCount[kNumMarchTiles] = {0, 0, 0, 0};
    // tile shape of the batched pass: -1 the first lattice's default, kTilePerLattice
    // each lattice's own (below), else that tile for all
    int tile = -1;
    int zchunk = 0;    // z chunk (0: each lattice's own setting, else whole columns up to 32 planes)
    hipGraphExec_t graph = nullptr;
    int graphIters = 0;

    void release_device() {
        if (graph) (void)hipGraphExecDestroy(graph);
        graph = nullptr;
        graphIters = 0;
        for (void* p : {dMarchDescs, dBoundaryDescs, static_cast<void*>(dFinal), static_cast<void*>(dMarchMap),
                        static_cast<void*>(dBoundaryMap), static_cast<void*>(dStatus)})
            if (p) (void)hipFree(p);
        dMarchDescs = dBoundaryDescs = nullptr;
        dFinal = nullptr;
        dMarchMap = dBoundaryMap = nullptr;
        dStatus = nullptr;
    }
};

// Work-group order of one batched launch: hardware id b goes to the logical
// entry that gives each XCD a contiguous run (as in the single-lattice pass).
static std::vector<int2> xcd_order(const std::vector<int2>& logical) {
    const int nb = (int)logical.size();
    std::vector<int2> map(nb);
    const int qb = nb >> 3, rb = nb & 7;
    for (int b = 0; b < nb; ++b) {
        const int xcd = b & 7;
        map[b] = logical[xcd * qb + std::min(xcd, rb) + (b >> 3)];
    }
    return map;
}

// Tile of lattice L in the batched pass.  Per lattice: the 8-row tile, 32 or
// 64 sites wide, that pads the lattice's x edge less (64 on a tie), so that a
// 24- or 32-wide lattice does not run half-empty 64-wide tiles.
static int ensemble_tile(const qm_ensemble* E, const qm_lattice* L) {
    if (E->tile >= 0) return E->tile;
    if (E->tile == kTilePerLattice) {
        const int w32 = (L->lx + 31) / 32 * 32, w64 = (L->lx + 63) / 64 * 64;
        return w32 < w64 ? kTile32x8 : kTile64x8;
    }
    // batched passes have the tiles of the fused pass only
    return active_tile(E->lats[0], kForce);
}

template <typename R>
static int build_descs(qm_ensemble* E, const qm_ldg* f, const qm_fire* params) {
    const int count = (int)E->lats.size();
    std::vector<MarchArgs<R>> md(count);
    std::vector<BoundaryArgs<R>> bd(count);
    std::vector<FinalizeDesc> fd(count);
    std::vector<int2> mlog[kNumMarchTiles], blog;
    for (int k = 0; k < count; ++k) {
        qm_lattice* L = E->lats[k];
        const int tile = ensemble_tile(E, L);
        // the ensemble's tile and z chunk live in its descriptors only: the
        // lattice's own settings are restored once its arguments are filled
        const int ownTile = L->tileOverride, ownZ = L->zchunk;
        L->tileOverride = tile;
        // whole columns (or chunks of 32 planes): an ensemble has work-groups enough
        L->zchunk = E->zchunk > 0 ? E->zchunk : L->zchunkOverride > 0 ? L->zchunkOverride : std::min(L->lz, 32);
        FireConst fc = fire_const(params, L);
        fill_args<R>(L, f, md[k], bd[k], fc, tile);
        L->tileOverride = ownTile;
        L->zchunk = ownZ;
        md[k].gate = 0;
        bd[k].gate = 0;
        const int bb = boundary_blocks(L->nBoundary);
        for (int j = 0; j < md[k].nblocks; ++j) mlog[tile].push_back(make_int2(k, j));
        for (int j = 0; j < bb; ++j) blog.push_back(make_int2(k, j));
        fd[k].ctrl = L->dCtrl;
        fd[k].fire = fc;
        fd[k].partials = L->dPartials;
        fd[k].total = (long long)md[k].nblocks + bb;
    }
    std::vector<int2> mmap;
    for (int t = 0; t < kNumMarchTiles; ++t) {
        const std::vector<int2> seg = xcd_order(mlog[t]);
        E->segOff[t] = (int)mmap.size();
        E->segCount[t] = (int)seg.size();
        mmap.insert(mmap.end(), seg.begin(), seg.end());
    }
    // the old descriptors may still be read by work queued on the ensemble's stream
    HIP_TRY(hipStreamSynchronize(E->stream));
    E->release_device();
    E->boundaryBlocks = (int)blog.size();
    HIP_TRY(hipMalloc(&E->dMarchDescs, count * sizeof(MarchArgs<R>)));
    HIP_TRY(hipMalloc(&E->dBoundaryDescs, count * sizeof(BoundaryArgs<R>)));
    HIP_TRY(hipMalloc(&E->dFinal, count * sizeof(FinalizeDesc)));
    HIP_TRY(hipMalloc(&E->dMarchMap, mmap.size() * sizeof(int2)));
    HIP_TRY(hipMalloc(&E->dBoundaryMap, blog.size() * sizeof(int2)));
    HIP_TRY(hipMalloc(&E->dStatus, count * sizeof(EnsembleStatus)));
    // ordered on the ensemble's stream, before the launches that read them
    HIP_TRY(hipMemcpyAsync(E->dMarchDescs, md.data(), count * sizeof(MarchArgs<R>), hipMemcpyHostToDevice, E->stream));
    HIP_TRY(hipMemcpyAsync(E->dBoundaryDescs, bd.data(), count * sizeof(BoundaryArgs<R>), hipMemcpyHostToDevice,
                           E->stream));
    HIP_TRY(hipMemcpyAsync(E->dFinal, fd.data(), count * sizeof(FinalizeDesc), hipMemcpyHostToDevice, E->stream));
    HIP_TRY(hipMemcpyAsync(E->dMarchMap, mmap.data(), mmap.size() * sizeof(int2), hipMemcpyHostToDevice, E->stream));
    HIP_TRY(hipMemcpyAsync(E->dBoundaryMap, blog.data(), blog.size() * sizeof(int2), hipMemcpyHostToDevice,
                           E->stream));
    // the host vectors end with this function
    HIP_TRY(hipStreamSynchronize(E->stream));
    return QM_OK;
}

// the batched fused pass: one launch per tile shape in use
template <typename R>
static hipError_t launch_march_segments(qm_ensemble* E, int mode, int terms) {
    for (int t = 0; t < kNumMarchTiles; ++t) {
        if (E->segCount[t] == 0) continue;
        hipError_t err = launch_march_batch<R>(static_cast<const MarchArgs<R>*>(E->dMarchDescs),
                                               E->dMarchMap + E->segOff[t], E->segCount[t], mode, terms, t, E->stream);
        if (err != hipSuccess) return err;
    }
    return hipSuccess;
}

template <typename R>
static int enqueue_pass(qm_ensemble* E, int mode, int terms) {
    HIP_TRY(launch_march_segments<R>(E, mode, terms));
    HIP_TRY(launch_boundary_batch<R>(static_cast<const BoundaryArgs<R>*>(E->dBoundaryDescs), E->dBoundaryMap,
                                     E->boundaryBlocks, mode, terms, E->stream));
    if (mode == kFire)
        HIP_TRY(launch_fire_finalize_batch(E->dFinal, E->dStatus, (int)E->lats.size(), E->stream));
    return QM_OK;
}

template <typename R>
static int ensemble_fire(qm_ensemble* E, const qm_fire* params, qm_ldg* f, int graphIters, int* itOut,
                         double* fmOut) {
    const int count = (int)E->lats.size();
    for (qm_lattice* L : E->lats) {
        if (int rc = prepare(L, f)) return rc;
        if (int rc = sync_ctrl_to_host(L)) return rc;
        // a fresh FIRE minimizer on every lattice (energyMinimizerFIRE.cpp:24-43, 151-170)
        Ctrl& c = L->hCtrl;
        c.done = 0;
        c.iterations = 0;
        c.maxIter = params->maxIter;
        c.nPositive = 0;
        c.dt = params->dt;
        c.alpha = params->alphaStart;
        c.nanFlag = 0;
        c.traceLen = 0;
        c.traceCap = 0;
        c.forceMax = 100.0;
        if (int rc = push_ctrl(L)) return rc;
    }
    if (int rc = build_descs<R>(E, f, params)) return rc;
    const int terms = terms_of(f);
    // sim->computeForces() at the start of minimize, on every lattice at once
    if (int rc = enqueue_pass<R>(E, kForce, terms)) return rc;
    // capture graphIters FIRE iterations once, replay until every lattice is done
    hipGraph_t g = nullptr;
    HIP_TRY(hipStreamBeginCapture(E->stream, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < graphIters; ++k) {
        if (int rc = enqueue_pass<R>(E, kFire, terms)) {
            (void)hipStreamEndCapture(E->stream, &g);
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
    }
    HIP_TRY(hipStreamEndCapture(E->stream, &g));
    HIP_TRY(hipGraphInstantiate(&E->graph, g, nullptr, nullptr, 0));
    HIP_TRY(hipGraphDestroy(g));
    E->graphIters = graphIters;
    std::vector<EnsembleStatus> st(count);
    int maxIt = params->maxIter;
    for (long long replays = 0;; ++replays) {
        HIP_TRY(hipGraphLaunch(E->graph, E->stream));
        HIP_TRY(hipMemcpyAsync(st.data(), E->dStatus, count * sizeof(EnsembleStatus), hipMemcpyDeviceToHost,
                               E->stream));
        HIP_TRY(hipStreamSynchronize(E->stream));
        bool all = true;
        for (const auto& s : st) all = all && s.done;
        // every lattice stops at maxIter on its own; the bound guards the loop
        if (all || (replays + 1) * (long long)graphIters >= (long long)maxIt + graphIters) break;
    }
    int nan = 0;
    for (int k = 0; k < count; ++k) {
        if (itOut) itOut[k] = st[k].iterations;
        if (fmOut) fmOut[k] = st[k].forceMax;
        if (int rc = sync_ctrl_to_host(E->lats[k])) return rc;
        nan |= E->lats[k]->hCtrl.nanFlag;
        // the deferred FIRE velocity mix, made concrete as after qm_fire_minimize
        if (int rc = materialize(E->lats[k])) return rc;
    }
    if (nan) return fail(QM_ENAN, "FIRE produced a NaN force norm on an ensemble lattice");
    return QM_OK;
}

// The tile autotuner for ensembles (the role of kernelTuner,
// src/utilities/kernelTuner.cpp:5-125, for the batched pass): times the batched
// fused FIRE pass for every tile shape and z chunks {8, 16, 32, 64} with HIP
// events and keeps the fastest for later fire_minimize calls.  The timed passes
// read each lattice's current state set and write only the other set and the
// partial sums, which the next real pass overwrites; the control blocks are
// restored, so the lattices are left as they were.
template <typename R>
static int ensemble_autotune(qm_ensemble* E, qm_ldg* f, int* tileOut, int* zchunkOut, double* msOut) {
    const int count = (int)E->lats.size();
    std::vector<Ctrl> saved(count);
    for (int k = 0; k < count; ++k) {
        qm_lattice* L = E->lats[k];
        if (int rc = prepare(L, f)) return rc;
        if (int rc = sync_ctrl_to_host(L)) return rc;
        saved[k] = L->hCtrl;
        L->hCtrl.done = 0;  // a done lattice would skip its work-groups
        if (int rc = push_ctrl(L)) return rc;
    }
    const qm_fire defaults{};  // FIRE constants of the descriptors; the timed kernel does not read them
    const int terms = terms_of(f);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    const int savedTile = E->tile, savedZ = E->zchunk;
    double best = 1e300;
    int bestTile = 0, bestZ = 0, rc = QM_OK;
    const int reps = 3;
    for (int tile = kTilePerLattice; tile < kNumMarchTiles && rc == QM_OK; ++tile) {
        if (tile == -1) continue;
        for (int zc : {8, 16, 32, 64}) {
            E->tile = tile;
            E->zchunk = zc;
            if ((rc = build_descs<R>(E, f, &defaults))) break;
            hipError_t err = launch_march_segments<R>(E, kFire, terms);
            if (err == hipSuccess) err = hipEventRecord(e0, E->stream);
            for (int r = 0; r < reps && err == hipSuccess; ++r) err = launch_march_segments<R>(E, kFire, terms);
            if (err == hipSuccess) err = hipEventRecord(e1, E->stream);
            if (err == hipSuccess) err = hipEventSynchronize(e1);
            float t = 0;
            if (err == hipSuccess) err = hipEventElapsedTime(&t, e0, e1);
            if (err != hipSuccess) {
                rc = fail(QM_EHIP, std::string("HIP error: ") + hipGetErrorString(err));
                break;
            }
            if (t / reps < best) {
                best = t / reps;
                bestTile = tile;
                bestZ = zc;
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    E->tile = rc == QM_OK ? bestTile : savedTile;
    E->zchunk = rc == QM_OK ? bestZ : savedZ;
    E->release_device();
    for (int k = 0; k < count; ++k) {
        E->lats[k]->hCtrl = saved[k];
        if (int r2 = push_ctrl(E->lats[k])) return r2;
    }
    if (rc) return rc;
    if (tileOut) *tileOut = bestTile;
    if (zchunkOut) *zchunkOut = bestZ;
    if (msOut) *msOut = best;
    return QM_OK;
}

extern "C" {

int qm_ensemble_create(qm_lattice* const* lattices, int count, qm_ensemble** out) {
    if (!out || !lattices || count < 1) return fail(QM_EINVAL, "bad argument");
    std::unique_ptr<qm_ensemble> E(new qm_ensemble);
    for (int k = 0; k < count; ++k) {
        if (!lattices[k]) return fail(QM_EINVAL, "null lattice in the ensemble");
        if (lattices[k]->precision != lattices[0]->precision)
            return fail(QM_EINVAL, "ensemble lattices must share one precision");
        for (int j = 0; j < k; ++j)
            if (lattices[j] == lattices[k]) return fail(QM_EINVAL, "a lattice appears twice in the ensemble");
        E->lats.push_back(lattices[k]);
    }
    E->precision = lattices[0]->precision;
    HIP_TRY(hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking));
    *out = E.release();
    return QM_OK;
}

int qm_ensemble_destroy(qm_ensemble* E) {
    if (!E) return QM_OK;
    (void)hipStreamSynchronize(E->stream);
    E->release_device();
    (void)hipStreamDestroy(E->stream);
    delete E;
    return QM_OK;
}

int qm_ensemble_autotune(qm_ensemble* E, qm_ldg* f, int* tile, int* zchunk, double* ms) {
    if (!E || !f) return fail(QM_EINVAL, "bad argument");
    for (qm_lattice* L : E->lats)
        if (L->precision != E->precision) return fail(QM_ESTATE, "a lattice changed precision after joining");
    return E->precision == QM_FP64 ? ensemble_autotune<double>(E, f, tile, zchunk, ms)
                                   : ensemble_autotune<float>(E, f, tile, zchunk, ms);
}

int qm_ensemble_set_tile(qm_ensemble* E, int tile, int zchunk) {
    if (!E) return fail(QM_EINVAL, "null ensemble");
    if (!(tile == -1 || tile == kTilePerLattice || (tile >= 0 && tile < kNumMarchTiles)))
        return fail(QM_EINVAL, "ensemble tile must be -1, QM_TILE_PER_LATTICE or 0..3");
    if (zchunk < 0 || (zchunk > 0 && zchunk < 4)) return fail(QM_EINVAL, "z chunk must be 0 (automatic) or at least 4");
    E->tile = tile;
    E->zchunk = zchunk;
    return QM_OK;
}

int qm_ensemble_fire_minimize(qm_ensemble* E, const qm_fire* params, qm_ldg* f, int graphIterations, int* iterations,
                              double* forceMax) {
    if (!E || !params || !f || graphIterations < 1) return fail(QM_EINVAL, "bad argument");
    for (qm_lattice* L : E->lats)
        if (L->precision != E->precision) return fail(QM_ESTATE, "a lattice changed precision after joining");
    return E->precision == QM_FP64 ? ensemble_fire<double>(E, params, f, graphIterations, iterations, forceMax)
                                   : ensemble_fire<float>(E, params, f, graphIterations, iterations, forceMax);
}

}  // extern "C"
