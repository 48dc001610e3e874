// Host side of the qmin C ABI (include/qmin.h): lattice state in HBM,
// objects, the seeded initial state, and the minimizer drivers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <mutex>
#include <random>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "qmin_internal.h"
#include "host_init.h"
#include "gen/ldg_fast_host.inc"

namespace qmin {
template <typename R>
hipError_t launch_march(const MarchArgs<R>& a, int mode, int terms, int tile, hipStream_t s);
template <typename R>
void launch_boundary(const BoundaryArgs<R>& a, int mode, int terms, hipStream_t s);
int boundary_blocks(long long nsites);
void launch_fire_finalize(Ctrl* ctrl, const FireConst& fp, const double* partials, long long total, double* trace,
                          hipStream_t s);
template <typename R>
int march_blocks_per_cu(int mode, int terms, int tile);
template <typename R>
hipError_t launch_fire_stream(const MarchArgs<R>& a, int terms, int tile, hipStream_t s);
template <typename R>
int stream_blocks_per_cu(int terms, int tile);
template <typename R>
hipError_t launch_step(const MarchArgs<R>& ma, const BoundaryArgs<R>& ba, int mode, int terms, int tile,
                       hipStream_t s);
template <typename R>
void launch_materialize(StateSet<R> s0, StateSet<R> s1, long long n, Ctrl* ctrl, hipStream_t s);
template <typename R>
int launch_energy(const R* q, const SiteType* types, const DevObject* objects, int lx, int ly, int lz,
                  const EnergyConst& ec, double* partials, hipStream_t s);
template <typename R>
void launch_defects(const R* q, const SiteType* types, long long n, int kind, double* out, hipStream_t s);
template <typename R>
void launch_energy_density(const R* q, const SiteType* types, const DevObject* objects, int lx, int ly, int lz,
                           const EnergyConst& ec, double* out, hipStream_t s);
template <typename R>
void launch_object_force(const R* q, const SiteType* types, const DevObject* objects, int lx, int ly, int lz,
                         const EnergyConst& ec, double L1, const int64_t* sites, long long m, double* out,
                         hipStream_t s);
template <typename R>
void launch_move_q(R* q, long long n, const int64_t* src, const int64_t* dst, long long m, R* tmp, hipStream_t s);
}  // namespace qmin

using namespace qmin;

// ------------------------------------------------------------------ errors
static thread_local std::string g_error;

int qmin::fail(int code, const std::string& msg) {
    g_error = msg;
    return code;
}

// ------------------------------------------------------------------ timing
namespace {
struct Timing {
    bool on = false;
    int64_t launches[2] = {0, 0};
    double ms[2] = {0, 0};
};
Timing g_timing;
}  // namespace

const char* qm_last_error(void) { return g_error.c_str(); }

int qm_device_count(int* count) {
    if (!count) return fail(QM_EINVAL, "null pointer");
    int c = 0;
    HIP_TRY(hipGetDeviceCount(&c));
    *count = c;
    return QM_OK;
}

// ------------------------------------------------------------------ lattice helpers
static int64_t site_of(const qm_lattice* L, int x, int y, int z) {
    return (int64_t)x + (int64_t)L->lx * ((int64_t)y + (int64_t)L->ly * z);
}

static int wrapc(int v, int m) {
    // wrap (inc/functions.h:299-307)
    int a = v;
    if (v >= m) a = v % m;
    while (a < 0) a += m;
    return a;
}

int qmin::sync_ctrl_to_host(qm_lattice* L) {
    HIP_TRY(hipMemcpyAsync(&L->hCtrl, L->dCtrl, sizeof(Ctrl), hipMemcpyDeviceToHost, L->stream));
    HIP_TRY(hipStreamSynchronize(L->stream));
    return QM_OK;
}

// Host <-> device copy ordered on the lattice's stream (which does not
// synchronize with the null stream), completed before returning.
hipError_t qmin::copy_on(qm_lattice* L, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, L->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(L->stream);
    return e;
}

int qmin::push_ctrl(qm_lattice* L) {
    HIP_TRY(hipMemcpyAsync(L->dCtrl, &L->hCtrl, sizeof(Ctrl), hipMemcpyHostToDevice, L->stream));
    HIP_TRY(hipStreamSynchronize(L->stream));
    return QM_OK;
}

// Upload types/objects, classify sites for the fused pass and rebuild the
// compacted list of sites finished by the boundary pass: t < 0 sites, and t == 0
// sites with a neighbour whose derivative is one-sided (an object two steps away
// along an edge diagonal, landauDeGennesLCOtherForces.cpp:77-129).  Every other
// t == 0 site has central derivatives at all six neighbours, where the composed
// bulk operator is exact.
static int refresh_types(qm_lattice* L) {
    if (!L->typesDirty) return QM_OK;
    const long long n = L->n;
    const int lx = L->lx, ly = L->ly, lz = L->lz;
    std::vector<SiteType> t8(n);
    std::vector<int8_t> cls(n);
    std::vector<int32_t> sites;
    int64_t active = 0;
    auto T = [&](int x, int y, int z) { return L->types[site_of(L, wrapc(x, lx), wrapc(y, ly), wrapc(z, lz))]; };
    for (int z = 0; z < lz; ++z)
        for (int y = 0; y < ly; ++y)
            for (int x = 0; x < lx; ++x) {
                const int64_t i = site_of(L, x, y, z);
                const int t = L->types[i];
                t8[i] = (SiteType)t;
                int c = t > 0 ? 1 : (t < 0 ? -1 : 0);
                if (t == 0) {
                    static const int dir[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
                    for (int d = 0; d < 6 && c == 0; ++d) {
                        const int nx = x + dir[d][0], ny = y + dir[d][1], nz = z + dir[d][2];
                        const int tn = T(nx, ny, nz);
                        if (tn > 0) c = -1;
                        if (tn >= 0) continue;
                        for (int e = 0; e < 6 && c == 0; ++e) {
                            if (e / 2 == d / 2) continue;
                            if (T(nx + dir[e][0], ny + dir[e][1], nz + dir[e][2]) > 0) c = -1;
                        }
                    }
                }
                cls[i] = (int8_t)c;
                if (c < 0) sites.push_back((int32_t)i);
                active += t <= 0;
            }
    HIP_TRY(copy_on(L, L->dTypes, t8.data(), n * sizeof(SiteType), hipMemcpyHostToDevice));
    HIP_TRY(copy_on(L, L->dClass, cls.data(), n, hipMemcpyHostToDevice));
    std::vector<DevObject> objs(kMaxObjects);
    for (size_t j = 0; j < L->objects.size(); ++j) objs[j] = {L->objects[j].type, 0, L->objects[j].W, L->objects[j].S0};
    HIP_TRY(copy_on(L, L->dObjects, objs.data(), sizeof(DevObject) * kMaxObjects, hipMemcpyHostToDevice));
    if (L->dSites) HIP_TRY(hipFree(L->dSites));
    L->dSites = nullptr;
    L->nBoundary = (long long)sites.size();
    if (!sites.empty()) {
        HIP_TRY(hipMalloc(&L->dSites, sites.size() * sizeof(int32_t)));
        HIP_TRY(copy_on(L, L->dSites, sites.data(), sites.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    }
    L->nActive = active;
    L->typesDirty = false;
    return QM_OK;
}

int qmin::ensure_partials(qm_lattice* L, size_t count) {
    if (count <= L->partialsCap) return QM_OK;
    if (L->dPartials) HIP_TRY(hipFree(L->dPartials));
    HIP_TRY(hipMalloc(&L->dPartials, count * sizeof(double)));
    L->partialsCap = count;
    return QM_OK;
}

// AoS host field -> SoA device plane set (current set)
static int upload_field(qm_lattice* L, int which, const double* aos) {
    const long long n = L->n;
    if (which == QM_FIELD_Q) {
        // The objects' FP64 Q, restored when an FP32 state converts back to
        // FP64.  On an FP32 lattice a value equal to the kept one in FP32 is the
        // same value come back through a download (random_init and the like
        // rewrite only liquid-crystal sites): the FP64 value is kept.
        const bool fp32 = L->precision == QM_FP32;
        for (HostObject& ob : L->objects) {
            ob.qb.resize(5 * ob.sites.size(), 0.0);
            for (size_t k = 0; k < ob.sites.size(); ++k) {
                const double* v = aos + 5 * ob.sites[k];
                double* kept = &ob.qb[5 * k];
                for (int c = 0; c < 5; ++c)
                    if (!fp32 || (float)v[c] != (float)kept[c]) kept[c] = v[c];
            }
        }
    }
    if (L->precision == QM_FP64) {
        std::vector<double> soa(5 * n);
        for (long long i = 0; i < n; ++i)
            for (int c = 0; c < 5; ++c) soa[c * n + i] = aos[5 * i + c];
        double* base = static_cast<double*>(L->dSet[L->hCtrl.cur]) + which * 5 * n;
        HIP_TRY(copy_on(L, base, soa.data(), 5 * n * sizeof(double), hipMemcpyHostToDevice));
    } else {
        std::vector<float> soa(5 * n);
        for (long long i = 0; i < n; ++i)
            for (int c = 0; c < 5; ++c) soa[c * n + i] = (float)aos[5 * i + c];
        float* base = static_cast<float*>(L->dSet[L->hCtrl.cur]) + which * 5 * n;
        HIP_TRY(copy_on(L, base, soa.data(), 5 * n * sizeof(float), hipMemcpyHostToDevice));
    }
    return QM_OK;
}

static int download_field(qm_lattice* L, int which, double* aos) {
    const long long n = L->n;
    HIP_TRY(hipStreamSynchronize(L->stream));
    if (L->precision == QM_FP64) {
        std::vector<double> soa(5 * n);
        const double* base = static_cast<double*>(L->dSet[L->hCtrl.cur]) + which * 5 * n;
        HIP_TRY(copy_on(L, soa.data(), base, 5 * n * sizeof(double), hipMemcpyDeviceToHost));
        for (long long i = 0; i < n; ++i)
            for (int c = 0; c < 5; ++c) aos[5 * i + c] = soa[c * n + i];
    } else {
        std::vector<float> soa(5 * n);
        const float* base = static_cast<float*>(L->dSet[L->hCtrl.cur]) + which * 5 * n;
        HIP_TRY(copy_on(L, soa.data(), base, 5 * n * sizeof(float), hipMemcpyDeviceToHost));
        for (long long i = 0; i < n; ++i)
            for (int c = 0; c < 5; ++c) aos[5 * i + c] = soa[c * n + i];
    }
    return QM_OK;
}

// ------------------------------------------------------------------ lattice API
int qm_lattice_create(int lx, int ly, int lz, int precision, qm_lattice** out) {
    if (!out) return fail(QM_EINVAL, "null output handle");
    if (lx < 3 || ly < 3 || lz < 3) return fail(QM_EINVAL, "each lattice edge must be at least 3 sites");
    if (precision != QM_FP64 && precision != QM_FP32) return fail(QM_EINVAL, "precision must be QM_FP64 or QM_FP32");
    const long long n = (long long)lx * ly * lz;
    if (n > (1LL << 31) - 1) return fail(QM_EINVAL, "lattice exceeds 2^31-1 sites");
    static std::atomic<uint64_t> serials{0};
    std::unique_ptr<qm_lattice> L(new qm_lattice);
    L->serial = ++serials;
    L->lx = lx;
    L->ly = ly;
    L->lz = lz;
    L->n = n;
    L->precision = precision;
    L->realSize = precision == QM_FP64 ? 8 : 4;
    L->types.assign(n, 0);
    HIP_TRY(hipGetDevice(&L->device));
    HIP_TRY(hipStreamCreateWithFlags(&L->ownStream, hipStreamNonBlocking));
    L->stream = L->ownStream;
    HIP_TRY(hipMalloc(&L->dTypes, n * sizeof(SiteType)));
    HIP_TRY(hipMalloc(&L->dClass, n));
    HIP_TRY(hipMalloc(&L->dW, sizeof(double) * 2 * kFastWeights));
    HIP_TRY(hipMalloc(&L->dObjects, sizeof(DevObject) * kMaxObjects));
    for (int s = 0; s < 2; ++s) {
        HIP_TRY(hipMalloc(&L->dSet[s], 15 * n * L->realSize));
        HIP_TRY(hipMemset(L->dSet[s], 0, 15 * n * L->realSize));
    }
    HIP_TRY(hipMalloc(&L->dCtrl, sizeof(Ctrl)));
    std::memset(&L->hCtrl, 0, sizeof(Ctrl));
    L->hCtrl.identity = 1;
    if (int rc = push_ctrl(L.get())) return rc;
    if (int rc = refresh_types(L.get())) return rc;
    *out = L.release();
    return QM_OK;
}

int qm_lattice_destroy(qm_lattice* L) {
    if (!L) return QM_OK;
    (void)hipStreamSynchronize(L->stream);
    (void)hipFree(L->dTypes);
    (void)hipFree(L->dClass);
    (void)hipFree(L->dW);
    (void)hipFree(L->dObjects);
    (void)hipFree(L->dSet[0]);
    (void)hipFree(L->dSet[1]);
    if (L->dSites) (void)hipFree(L->dSites);
    (void)hipFree(L->dCtrl);
    if (L->dPartials) (void)hipFree(L->dPartials);
    if (L->dTrace) (void)hipFree(L->dTrace);
    if (L->dHField) (void)hipFree(L->dHField);
    (void)hipStreamDestroy(L->ownStream);
    delete L;
    return QM_OK;
}

template <typename S, typename D>
__global__ void convert_kernel(const S* src, D* dst, long long count) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (long long)gridDim.x * blockDim.x)
        dst[i] = D(src[i]);
}

// q[c n + sites[k]] = vals[5 k + c] (each site listed once)
__global__ void scatter_object_q_kernel(double* q, long long n, const int64_t* sites, long long m,
                                        const double* vals) {
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < m; k += (long long)gridDim.x * blockDim.x) {
#pragma unroll
        for (int c = 0; c < 5; ++c) q[c * n + sites[k]] = vals[5 * k + c];
    }
}

// Object sites of an FP64 state converted from FP32 get back their FP64 Q (the
// fixed boundary values the FP32 state rounded), in both state sets.  A site
// listed by several objects takes the value of the last, as when they were
// added.
static int restore_object_q(qm_lattice* L) {
    std::vector<int64_t> sites;
    std::vector<double> vals;
    std::unordered_map<int64_t, size_t> at;
    for (const HostObject& ob : L->objects) {
        if (ob.qb.size() != 5 * ob.sites.size()) continue;
        for (size_t k = 0; k < ob.sites.size(); ++k) {
            auto it = at.find(ob.sites[k]);
            size_t slot;
            if (it == at.end()) {
                slot = sites.size();
                at.emplace(ob.sites[k], slot);
                sites.push_back(ob.sites[k]);
                vals.resize(vals.size() + 5);
            } else {
                slot = it->second;
            }
            std::memcpy(&vals[5 * slot], &ob.qb[5 * k], 5 * sizeof(double));
        }
    }
    if (sites.empty()) return QM_OK;
    const long long m = (long long)sites.size();
    int64_t* dSites = nullptr;
    double* dVals = nullptr;
    if (hipMalloc(&dSites, m * sizeof(int64_t)) != hipSuccess) return fail(QM_ENOMEM, "object Q: out of device memory");
    if (hipMalloc(&dVals, 5 * m * sizeof(double)) != hipSuccess) {
        (void)hipFree(dSites);
        return fail(QM_ENOMEM, "object Q: out of device memory");
    }
    hipError_t err = hipMemcpyAsync(dSites, sites.data(), m * sizeof(int64_t), hipMemcpyHostToDevice, L->stream);
    if (err == hipSuccess)
        err = hipMemcpyAsync(dVals, vals.data(), 5 * m * sizeof(double), hipMemcpyHostToDevice, L->stream);
    for (int s = 0; s < 2 && err == hipSuccess; ++s) {
        const int blocks = (int)std::min<long long>((m + 255) / 256, 4096);
        scatter_object_q_kernel<<<blocks, 256, 0, L->stream>>>(static_cast<double*>(L->dSet[s]), L->n, dSites, m,
                                                               dVals);
        err = hipGetLastError();
    }
    if (err == hipSuccess) err = hipStreamSynchronize(L->stream);
    (void)hipFree(dSites);
    (void)hipFree(dVals);
    if (err != hipSuccess) return fail(QM_EHIP, std::string("restoring object Q: ") + hipGetErrorString(err));
    return QM_OK;
}

// Converts the device state (Q, velocity with any pending FIRE mix applied,
// forces) to the other storage precision in place of the old buffers.
int qm_lattice_set_precision(qm_lattice* L, int precision) {
    if (!L) return fail(QM_EINVAL, "null lattice");
    if (precision != QM_FP64 && precision != QM_FP32) return fail(QM_EINVAL, "precision must be QM_FP64 or QM_FP32");
    if (precision == L->precision) return QM_OK;
    if (int rc = materialize(L)) return rc;
    const size_t realSize = precision == QM_FP64 ? 8 : 4;
    const long long count = 15 * L->n;
    const int blocks = (int)std::min<long long>((count + 255) / 256, 8192);
    void* fresh[2] = {nullptr, nullptr};
    for (int s = 0; s < 2; ++s) {
        if (hipMalloc(&fresh[s], count * realSize) != hipSuccess) {
            if (fresh[0]) (void)hipFree(fresh[0]);
            return fail(QM_ENOMEM, "precision change: out of device memory");
        }
        if (precision == QM_FP64)
            convert_kernel<float, double><<<blocks, 256, 0, L->stream>>>(static_cast<const float*>(L->dSet[s]),
                                                                        static_cast<double*>(fresh[s]), count);
        else
            convert_kernel<double, float><<<blocks, 256, 0, L->stream>>>(static_cast<const double*>(L->dSet[s]),
                                                                        static_cast<float*>(fresh[s]), count);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipStreamSynchronize(L->stream));
    for (int s = 0; s < 2; ++s) {
        HIP_TRY(hipFree(L->dSet[s]));
        L->dSet[s] = fresh[s];
    }
    L->precision = precision;
    L->realSize = realSize;
    if (precision == QM_FP64) return restore_object_q(L);
    return QM_OK;
}

// Spatially varying H field of the lattice, AoS [N][3] (NULL removes it).
int qm_lattice_set_hfield(qm_lattice* L, const double* h3) {
    if (!L) return fail(QM_EINVAL, "null lattice");
    if (!h3) {
        if (L->dHField) HIP_TRY(hipFree(L->dHField));
        L->dHField = nullptr;
        return QM_OK;
    }
    const long long n = L->n;
    std::vector<double> planes(3 * (size_t)n);
    for (long long i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) planes[(size_t)a * n + i] = h3[3 * i + a];
    if (!L->dHField) HIP_TRY(hipMalloc(&L->dHField, 3 * (size_t)n * sizeof(double)));
    HIP_TRY(copy_on(L, L->dHField, planes.data(), 3 * (size_t)n * sizeof(double), hipMemcpyHostToDevice));
    return QM_OK;
}

// landauDeGennesLC::setSpatiallyVaryingField's file reading
// (landauDeGennesLC.cpp:81-123): lines "x y z Hx Hy Hz", unlisted sites zero;
// path is the full file name.
int qm_lattice_load_hfield(qm_lattice* L, const char* path) {
    if (!L || !path) return fail(QM_EINVAL, "bad argument");
    FILE* fp = std::fopen(path, "r");
    if (!fp) return fail(QM_EIO, std::string("cannot read ") + path);
    std::vector<double> h3(3 * (size_t)L->n, 0.0);
    double fx, fy, fz, hx, hy, hz;
    while (std::fscanf(fp, "%lf %lf %lf %lf %lf %lf", &fx, &fy, &fz, &hx, &hy, &hz) == 6) {
        const long long i = site_of(L, wrapc((int)fx, L->lx), wrapc((int)fy, L->ly), wrapc((int)fz, L->lz));
        h3[3 * i + 0] = hx;
        h3[3 * i + 1] = hy;
        h3[3 * i + 2] = hz;
    }
    std::fclose(fp);
    return qm_lattice_set_hfield(L, h3.data());
}

int qm_lattice_set_tile(qm_lattice* L, int tile) {
    if (!L || tile < -1 || tile >= kNumTiles)
        return fail(QM_EINVAL, "tile is -1 (automatic), 0 (32x8), 1 (64x4), 2 (32x16), 3 (64x8), or for FIRE "
                               "passes that keep only Q and v 4 (32x16), 5 (32x8), 6 (64x8)");
    L->tileOverride = tile;
    return QM_OK;
}

int qm_lattice_set_stream(qm_lattice* L, void* stream) {
    if (!L) return fail(QM_EINVAL, "null lattice");
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : L->ownStream;
    if (s != L->ownStream) {
        // the lattice's own stream was created with the default priority
        int prio = 0, own = 0;
        HIP_TRY(hipStreamGetPriority(s, &prio));
        HIP_TRY(hipStreamGetPriority(L->ownStream, &own));
        if (prio != own) return fail(QM_EINVAL, "the stream must have the default priority");
    }
    // work queued so far finishes on the stream it was queued on
    HIP_TRY(hipStreamSynchronize(L->stream));
    L->stream = s;
    return QM_OK;
}

int qm_lattice_set_zchunk(qm_lattice* L, int zchunk) {
    if (!L || zchunk < 0) return fail(QM_EINVAL, "bad argument");
    if (zchunk > 0 && zchunk < 4) return fail(QM_EINVAL, "z chunk must be 0 (automatic) or at least 4");
    L->zchunkOverride = zchunk;
    return QM_OK;
}

int qm_lattice_dims(const qm_lattice* L, int* lx, int* ly, int* lz, int* precision) {
    if (!L) return fail(QM_EINVAL, "null lattice");
    if (lx) *lx = L->lx;
    if (ly) *ly = L->ly;
    if (lz) *lz = L->lz;
    if (precision) *precision = L->precision;
    return QM_OK;
}

int qm_lattice_set_field(qm_lattice* L, int which, const double* aos) {
    if (!L || !aos || which < 0 || which > 2) return fail(QM_EINVAL, "bad argument");
    if (which == QM_FIELD_V) {
        // an explicit velocity replaces any pending FIRE mix
        L->hCtrl.identity = 1;
        L->hCtrl.reset = 0;
        if (int rc = push_ctrl(L)) return rc;
    }
    return upload_field(L, which, aos);
}


int qm_lattice_get_field(qm_lattice* L, int which, double* aos) {
    if (!L || !aos || which < 0 || which > 2) return fail(QM_EINVAL, "bad argument");
    if (which == QM_FIELD_V)
        if (int rc = materialize(L)) return rc;
    return download_field(L, which, aos);
}

// setNematicQTensorRandomly with a reproducible seed (host_init.cpp)
int qm_lattice_random_init(qm_lattice* L, int seed, double S0) {
    if (!L) return fail(QM_EINVAL, "null lattice");
    std::vector<double> q(5 * L->n);
    if (int rc = download_field(L, QM_FIELD_Q, q.data())) return rc;
    random_nematic(L->n, L->types.data(), seed, S0, q.data());
    return upload_field(L, QM_FIELD_Q, q.data());
}

int qm_lattice_uniform_init(qm_lattice* L, double nx, double ny, double nz, double S0) {
    if (!L) return fail(QM_EINVAL, "null lattice");
    std::vector<double> q(5 * L->n);
    if (int rc = download_field(L, QM_FIELD_Q, q.data())) return rc;
    double d[5];
    director_q(nx, ny, nz, S0, d);
    for (long long i = 0; i < L->n; ++i)
        if (L->types[i] <= 0) std::memcpy(&q[5 * i], d, sizeof(d));
    return upload_field(L, QM_FIELD_Q, q.data());
}

// setNematicQTensorRandomly / setRandomDirectors with globallyAligned = true
int qm_lattice_random_uniform_init(qm_lattice* L, int seed, double S0) {
    if (!L) return fail(QM_EINVAL, "null lattice");
    double n[3];
    random_director(seed, n);
    return qm_lattice_uniform_init(L, n[0], n[1], n[2], S0);
}

// multirankQTensorLatticeModel::setDirectorFromFunction (multirankQTensorLatticeModel.cpp:90-110):
// directors holds (nx, ny, nz, S0) per site in site order, the function's values
int qm_lattice_director_init(qm_lattice* L, const double* directors) {
    if (!L || !directors) return fail(QM_EINVAL, "bad argument");
    std::vector<double> q(5 * L->n);
    if (int rc = download_field(L, QM_FIELD_Q, q.data())) return rc;
    for (long long i = 0; i < L->n; ++i)
        if (L->types[i] <= 0) {
            const double* d = directors + 4 * i;
            director_q(d[0], d[1], d[2], d[3], &q[5 * i]);
        }
    return upload_field(L, QM_FIELD_Q, q.data());
}

// createMultirankBoundaryObject (multirankSimulationBoundaries.cpp:512-543) then
// cubicLattice::createBoundaryObject (cubicLattice.cpp:267-318): write Q^B,
// mark the sites j, mark their non-object neighbours -1.
int qm_lattice_add_object(qm_lattice* L, const int32_t* xyz, int64_t n, const double* qb, int btype, double W,
                          double S0, int* obj) {
    if (!L || (n > 0 && (!xyz || !qb))) return fail(QM_EINVAL, "bad argument");
    if (btype != QM_HOMEOTROPIC && btype != QM_DEGENERATE_PLANAR) return fail(QM_EINVAL, "unknown boundary type");
    if ((int)L->objects.size() >= kMaxObjects) return fail(QM_EINVAL, "at most 32767 objects per lattice");
    std::vector<double> q(5 * L->n);
    if (int rc = download_field(L, QM_FIELD_Q, q.data())) return rc;
    std::vector<int64_t> sites;
    sites.reserve(n);
    for (int64_t k = 0; k < n; ++k) {
        const int x = wrapc(xyz[3 * k], L->lx), y = wrapc(xyz[3 * k + 1], L->ly), z = wrapc(xyz[3 * k + 2], L->lz);
        const int64_t i = site_of(L, x, y, z);
        sites.push_back(i);
        std::memcpy(&q[5 * i], qb + 5 * k, 5 * sizeof(double));
    }
    L->objects.push_back({btype, W, S0, {}, {}, {}});
    const int j = (int)L->objects.size();
    for (int64_t i : sites) L->types[i] = j;
    // surface sites in order of discovery, first occurrence kept
    // (removeDuplicateVectorElements, inc/functions.h:23-35)
    std::vector<int64_t> surface;
    std::vector<char> seen(L->n, 0);
    for (int64_t i : sites) {
        const int x = int(i % L->lx), y = int((i / L->lx) % L->ly), z = int(i / ((int64_t)L->lx * L->ly));
        const int64_t nb[6] = {site_of(L, wrapc(x - 1, L->lx), y, z), site_of(L, wrapc(x + 1, L->lx), y, z),
                               site_of(L, x, wrapc(y - 1, L->ly), z), site_of(L, x, wrapc(y + 1, L->ly), z),
                               site_of(L, x, y, wrapc(z - 1, L->lz)), site_of(L, x, y, wrapc(z + 1, L->lz))};
        for (int d = 0; d < 6; ++d)
            if (L->types[nb[d]] < 1) {
                L->types[nb[d]] = -1;
                if (!seen[nb[d]]) {
                    seen[nb[d]] = 1;
                    surface.push_back(nb[d]);
                }
            }
    }
    L->objects.back().sites = std::move(sites);
    L->objects.back().surface = std::move(surface);
    L->typesDirty = true;
    if (int rc = refresh_types(L)) return rc;
    if (obj) *obj = j;
    return upload_field(L, QM_FIELD_Q, q.data());
}

int qm_lattice_wall(qm_lattice* L, int xyz, int plane, int btype, double W, double S0, int* obj) {
    if (!L) return fail(QM_EINVAL, "null lattice");
    if (xyz < 0 || xyz > 2) return fail(QM_EINVAL, "wall normal must be 0 (x), 1 (y) or 2 (z)");
    double Qt[5] = {0, 0, 0, 0, 0};
    if (btype == QM_HOMEOTROPIC) {
        if (xyz == 0) { Qt[0] = S0; Qt[3] = -0.5 * S0; }
        else if (xyz == 1) { Qt[0] = -0.5 * S0; Qt[3] = S0; }
        else { Qt[0] = -0.5 * S0; Qt[3] = -0.5 * S0; }
    } else {
        Qt[xyz] = 1.0;
    }
    const int s1 = xyz == 0 ? L->ly : L->lx, s2 = xyz == 2 ? L->ly : L->lz;
    std::vector<int32_t> pts;
    std::vector<double> qs;
    pts.reserve(3 * (size_t)s1 * s2);
    qs.reserve(5 * (size_t)s1 * s2);
    for (int a = 0; a < s1; ++a)
        for (int b = 0; b < s2; ++b) {
            if (xyz == 0) { pts.push_back(plane); pts.push_back(a); pts.push_back(b); }
            else if (xyz == 1) { pts.push_back(a); pts.push_back(plane); pts.push_back(b); }
            else { pts.push_back(a); pts.push_back(b); pts.push_back(plane); }
            qs.insert(qs.end(), Qt, Qt + 5);
        }
    return qm_lattice_add_object(L, pts.data(), (int64_t)(pts.size() / 3), qs.data(), btype, W, S0, obj);
}

int qm_lattice_sphere(qm_lattice* L, double cx, double cy, double cz, double radius, int btype, double W, double S0,
                      int* obj) {
    if (!L) return fail(QM_EINVAL, "null lattice");
    std::vector<int32_t> pts;
    std::vector<double> qs;
    sphere_sites(cx, cy, cz, radius, btype, S0, pts, qs);
    return qm_lattice_add_object(L, pts.data(), (int64_t)(pts.size() / 3), qs.data(), btype, W, S0, obj);
}

// createBoundaryFromFile (multirankSimulationBoundaries.cpp:41-96): first line the
// object count; per object a line "a W S0 nSites" (a = 0 oriented, 1 degenerate
// planar) then nSites lines "x y z Q0 Q1 Q2 Q3 Q4".
int qm_lattice_boundary_file(qm_lattice* L, const char* path) {
    if (!L || !path) return fail(QM_EINVAL, "bad argument");
    std::ifstream in(path);
    if (!in) return fail(QM_EIO, std::string("cannot open boundary file ") + path);
    std::string line;
    if (!std::getline(in, line)) return fail(QM_EIO, "empty boundary file");
    int nObjects = 0;
    std::istringstream(line) >> nObjects;
    for (int o = 0; o < nObjects; ++o) {
        if (!std::getline(in, line)) return fail(QM_EIO, "boundary file ends before object header");
        int bt = 0, nEntries = 0;
        double W = 0, S0 = 0;
        std::istringstream hs(line);
        hs >> bt >> W >> S0 >> nEntries;
        if (!hs) return fail(QM_EIO, "bad object header: " + line);
        std::vector<int32_t> pts;
        std::vector<double> qs;
        int read = 0;
        while (read < nEntries && std::getline(in, line)) {
            std::istringstream ls(line);
            int x, y, z;
            double q[5];
            ls >> x >> y >> z >> q[0] >> q[1] >> q[2] >> q[3] >> q[4];
            pts.push_back(x);
            pts.push_back(y);
            pts.push_back(z);
            qs.insert(qs.end(), q, q + 5);
            ++read;
        }
        if (read != nEntries) return fail(QM_EIO, "boundary file ends inside an object");
        int rc = qm_lattice_add_object(L, pts.data(), read, qs.data(), bt == 0 ? QM_HOMEOTROPIC : QM_DEGENERATE_PLANAR,
                                       W, S0, nullptr);
        if (rc) return rc;
    }
    return QM_OK;
}

int qm_lattice_get_types(const qm_lattice* L, int32_t* types) {
    if (!L || !types) return fail(QM_EINVAL, "bad argument");
    std::memcpy(types, L->types.data(), L->n * sizeof(int32_t));
    return QM_OK;
}

int qm_lattice_get_neighbors(const qm_lattice* L, int32_t* nb) {
    if (!L || !nb) return fail(QM_EINVAL, "bad argument");
    for (int z = 0; z < L->lz; ++z)
        for (int y = 0; y < L->ly; ++y)
            for (int x = 0; x < L->lx; ++x) {
                const int64_t i = site_of(L, x, y, z);
                int32_t* o = nb + 6 * i;
                o[0] = (int32_t)site_of(L, wrapc(x - 1, L->lx), y, z);
                o[1] = (int32_t)site_of(L, wrapc(x + 1, L->lx), y, z);
                o[2] = (int32_t)site_of(L, x, wrapc(y - 1, L->ly), z);
                o[3] = (int32_t)site_of(L, x, wrapc(y + 1, L->ly), z);
                o[4] = (int32_t)site_of(L, x, y, wrapc(z - 1, L->lz));
                o[5] = (int32_t)site_of(L, x, y, wrapc(z + 1, L->lz));
            }
    return QM_OK;
}

int qm_lattice_n_active(const qm_lattice* L, int64_t* n) {
    if (!L || !n) return fail(QM_EINVAL, "bad argument");
    *n = qmin::n_active(L);
    return QM_OK;
}

int qm_lattice_n_boundary(const qm_lattice* L, int64_t* n) {
    if (!L || !n) return fail(QM_EINVAL, "bad argument");
    int64_t c = 0;
    for (long long i = 0; i < L->n; ++i) c += L->types[i] < 0;
    *n = c;
    return QM_OK;
}

int qm_lattice_defects(qm_lattice* L, int kind, double* out) {
    if (!L || !out || kind < 0 || kind > 2) return fail(QM_EINVAL, "bad argument");
    double* d = nullptr;
    HIP_TRY(hipMalloc(&d, L->n * sizeof(double)));
    if (L->precision == QM_FP64)
        launch_defects<double>(state_set<double>(L, L->hCtrl.cur).q, L->dTypes, L->n, kind, d, L->stream);
    else
        launch_defects<float>(state_set<float>(L, L->hCtrl.cur).q, L->dTypes, L->n, kind, d, L->stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, d, L->n * sizeof(double), hipMemcpyDeviceToHost, L->stream));
    HIP_TRY(hipStreamSynchronize(L->stream));
    HIP_TRY(hipFree(d));
    return QM_OK;
}

// cubicLattice::boundarySites / surfaceSites of object `object` (0-based)
int qm_lattice_object_sites(const qm_lattice* L, int object, int surface, int64_t* sites, int64_t* count) {
    if (!L || !count || object < 0 || object >= (int)L->objects.size()) return fail(QM_EINVAL, "bad argument");
    const std::vector<int64_t>& v = surface ? L->objects[object].surface : L->objects[object].sites;
    *count = (int64_t)v.size();
    if (sites) std::memcpy(sites, v.data(), v.size() * sizeof(int64_t));
    return QM_OK;
}

namespace {
template <typename R>
int move_q(qm_lattice* L, const std::vector<int64_t>& src, const std::vector<int64_t>& dst) {
    const long long m = (long long)src.size();
    if (m == 0) return QM_OK;
    int64_t* d = nullptr;
    R* tmp = nullptr;
    HIP_TRY(hipMalloc(&d, 2 * m * sizeof(int64_t)));
    hipError_t err = hipMalloc(&tmp, 5 * m * sizeof(R));
    if (err == hipSuccess) err = hipMemcpyAsync(d, src.data(), m * sizeof(int64_t), hipMemcpyHostToDevice, L->stream);
    if (err == hipSuccess)
        err = hipMemcpyAsync(d + m, dst.data(), m * sizeof(int64_t), hipMemcpyHostToDevice, L->stream);
    if (err == hipSuccess) {
        launch_move_q<R>(state_set<R>(L, L->hCtrl.cur).q, L->n, d, d + m, m, tmp, L->stream);
        err = hipGetLastError();
    }
    if (err == hipSuccess) err = hipStreamSynchronize(L->stream);
    (void)hipFree(d);
    if (tmp) (void)hipFree(tmp);
    if (err != hipSuccess) return fail(QM_EHIP, std::string("moving an object: ") + hipGetErrorString(err));
    return QM_OK;
}
}  // namespace

// cubicLattice::displaceBoundaryObject (cubicLattice.cpp:329-372): per unit
// step the object's sites and its surface sites, with their Q, move one site
// along direction (0..5: x-, x+, y-, y+, z-, z+); the old surface sites become
// bulk first, then the moved sites take the object's type and -1.
int qm_lattice_displace_object(qm_lattice* L, int object, int direction, int magnitude) {
    if (!L || object < 0 || object >= (int)L->objects.size() || direction < 0 || direction > 5 || magnitude < 0)
        return fail(QM_EINVAL, "bad argument");
    HostObject& ob = L->objects[object];
    static const int dir[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
    auto moved = [&](int64_t i) {
        const int x = int(i % L->lx), y = int((i / L->lx) % L->ly), z = int(i / ((int64_t)L->lx * L->ly));
        return site_of(L, wrapc(x + dir[direction][0], L->lx), wrapc(y + dir[direction][1], L->ly),
                       wrapc(z + dir[direction][2], L->lz));
    };
    for (int mm = 0; mm < magnitude; ++mm) {
        std::vector<int64_t> src(ob.sites), dst;
        src.insert(src.end(), ob.surface.begin(), ob.surface.end());
        dst.reserve(src.size());
        for (int64_t i : src) dst.push_back(moved(i));
        for (int64_t i : ob.surface) L->types[i] = 0;
        const size_t nb = ob.sites.size();
        for (size_t k = 0; k < nb; ++k) {
            ob.sites[k] = dst[k];
            L->types[dst[k]] = object + 1;
        }
        for (size_t k = 0; k < ob.surface.size(); ++k) {
            ob.surface[k] = dst[nb + k];
            L->types[dst[nb + k]] = -1;
        }
        const int rc = L->precision == QM_FP64 ? move_q<double>(L, src, dst) : move_q<float>(L, src, dst);
        if (rc) return rc;
    }
    L->typesDirty = true;
    return refresh_types(L);
}

// saveState (multirankSimulation.cpp:630-665): "x y z Q0 Q1 Q2 Q3 Q4 type defect",
// tab separated, one line per saved site in index order, numbers as an
// ostream writes them (6 significant digits, %g); sites with x, y and z all
// multiples of latticeSkip; defect = computeDefectMeasures(defectType) on
// liquid-crystal sites, 0 on object sites (never computed there)
int qm_lattice_save_state_ex(qm_lattice* L, const char* path, int latticeSkip, int defectType) {
    if (!L || !path) return fail(QM_EINVAL, "bad argument");
    if (defectType < 0 || defectType > 2) return fail(QM_EINVAL, "defect type is 0, 1 or 2");
    const int stride = latticeSkip < 1 ? 1 : latticeSkip;
    std::vector<double> q(5 * L->n), defects(L->n);
    if (int rc = qm_lattice_get_field(L, QM_FIELD_Q, q.data())) return rc;
    if (int rc = qm_lattice_defects(L, defectType, defects.data())) return rc;
    FILE* f = std::fopen(path, "w");
    if (!f) return fail(QM_EIO, std::string("cannot write ") + path);
    for (long long i = 0; i < L->n; ++i) {
        const int x = int(i % L->lx), y = int((i / L->lx) % L->ly), z = int(i / ((long long)L->lx * L->ly));
        if (x % stride != 0 || y % stride != 0 || z % stride != 0) continue;
        std::fprintf(f, "%i\t%i\t%i\t%g\t%g\t%g\t%g\t%g\t%i\t%g\n", x, y, z, q[5 * i], q[5 * i + 1], q[5 * i + 2],
                     q[5 * i + 3], q[5 * i + 4], L->types[i], L->types[i] > 0 ? 0.0 : defects[i]);
    }
    if (std::fclose(f) != 0) return fail(QM_EIO, std::string("cannot write ") + path);
    return QM_OK;
}

int qm_lattice_save_state(qm_lattice* L, const char* path) { return qm_lattice_save_state_ex(L, path, 1, 0); }

// loadState (multirankSimulation.cpp:573-612): reads Q only; types are ignored
int qm_lattice_load_state(qm_lattice* L, const char* path) {
    if (!L || !path) return fail(QM_EINVAL, "bad argument");
    std::ifstream in(path);
    if (!in) return fail(QM_EIO, std::string("cannot open ") + path);
    std::vector<double> q(5 * L->n);
    if (int rc = download_field(L, QM_FIELD_Q, q.data())) return rc;
    int x, y, z, t;
    double v[5], s;
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        if (!(ls >> x >> y >> z >> v[0] >> v[1] >> v[2] >> v[3] >> v[4])) continue;
        ls >> t >> s;
        if (x < 0 || y < 0 || z < 0 || x >= L->lx || y >= L->ly || z >= L->lz)
            return fail(QM_EIO, "state file site outside the lattice");
        std::memcpy(&q[5 * site_of(L, x, y, z)], v, sizeof(v));
    }
    return upload_field(L, QM_FIELD_Q, q.data());
}

int qm_lattice_copy_q(qm_lattice* dst, const qm_lattice* src) {
    if (!dst || !src || dst->n != src->n) return fail(QM_EINVAL, "lattices differ in size");
    std::vector<double> q(5 * src->n);
    if (int rc = download_field(const_cast<qm_lattice*>(src), QM_FIELD_Q, q.data())) return rc;
    return upload_field(dst, QM_FIELD_Q, q.data());
}

// ------------------------------------------------------------------ force computer
int qm_ldg_create(qm_ldg** out) {
    if (!out) return fail(QM_EINVAL, "null output handle");
    *out = new qm_ldg;
    return QM_OK;
}
int qm_ldg_destroy(qm_ldg* f) {
    delete f;
    return QM_OK;
}
int qm_ldg_set_phase(qm_ldg* f, double A, double B, double C) {
    if (!f) return fail(QM_EINVAL, "null force");
    f->A = A;
    f->B = B;
    f->C = C;
    return QM_OK;
}
int qm_ldg_set_elastic(qm_ldg* f, double L1, double L2, double L3, double L4, double L6, int multi) {
    if (!f) return fail(QM_EINVAL, "null force");
    if (!multi && (L2 != 0 || L3 != 0 || L4 != 0 || L6 != 0))
        return fail(QM_EINVAL, "the one-constant path takes L1 only");
    f->L1 = L1;
    f->L2 = L2;
    f->L3 = L3;
    f->L4 = L4;
    f->L6 = L6;
    f->multi = multi;
    return QM_OK;
}
int qm_ldg_set_efield(qm_ldg* f, int on, double Ex, double Ey, double Ez, double eps, double eps0, double deps) {
    if (!f) return fail(QM_EINVAL, "null force");
    f->eOn = on;
    f->E[0] = Ex; f->E[1] = Ey; f->E[2] = Ez;
    f->eps = eps; f->eps0 = eps0; f->deps = deps;
    return QM_OK;
}
int qm_ldg_set_hfield(qm_ldg* f, int on, double Hx, double Hy, double Hz, double chi, double mu0, double dchi) {
    if (!f) return fail(QM_EINVAL, "null force");
    f->hOn = on;
    f->H[0] = Hx; f->H[1] = Hy; f->H[2] = Hz;
    f->chi = chi; f->mu0 = mu0; f->dchi = dchi;
    return QM_OK;
}
int qm_ldg_set_varying_hfield(qm_ldg* f, int on, double chi, double mu0, double dchi) {
    if (!f) return fail(QM_EINVAL, "null force");
    f->hvarOn = on != 0;
    // shared with the uniform H field, as in the reference
    f->chi = chi;
    f->mu0 = mu0;
    f->dchi = dchi;
    return QM_OK;
}

int qm_ldg_set_semantics(qm_ldg* f, int s) {
    if (!f || (s != QM_SEMANTICS_REF_CPU && s != QM_SEMANTICS_REF_GPU)) return fail(QM_EINVAL, "bad semantics");
    f->semantics = s;
    return QM_OK;
}

static void field_force(const double* F, double anis, double perm, double* out) {
    // fieldForce of computeEorHFieldForcesCPU; the site force is F -= fieldForce
    const double p = anis * perm;
    out[0] = -0.5 * p * (F[0] * F[0] - F[2] * F[2]);
    out[1] = -p * F[0] * F[1];
    out[2] = -p * F[0] * F[2];
    out[3] = -0.5 * p * (F[1] * F[1] - F[2] * F[2]);
    out[4] = -p * F[1] * F[2];
}

static ForceConst force_const(const qm_ldg* f) {
    ForceConst fc{};
    fc.a = 0.5 * f->A;
    fc.b = f->B / 3.0;
    fc.c = 0.25 * f->C;
    fc.L1 = f->L1;
    fc.L2 = f->L2;
    fc.L3 = f->L3;
    fc.L4 = f->L4;
    fc.L6 = f->L6;
    fc.semantics = f->semantics;
    double tmp[5];
    for (int c = 0; c < 5; ++c) fc.field[c] = 0;
    if (f->eOn) {
        field_force(f->E, f->deps, f->eps0, tmp);
        for (int c = 0; c < 5; ++c) fc.field[c] -= tmp[c];
    }
    if (f->hOn) {
        field_force(f->H, f->dchi, f->mu0, tmp);
        for (int c = 0; c < 5; ++c) fc.field[c] -= tmp[c];
    }
    fc.hasField = f->eOn || f->hOn;
    return fc;
}

int qmin::terms_of(const qm_ldg* f) {
    if (!f->multi) return kOneConstant;
    if (f->L6 != 0) return kMultiL6;
    return f->L4 != 0 ? kMultiL4 : kMultiLinear;
}

// Weights of the composed bulk operator for the force's constants.
template <typename R>
static void composed_weights(const qm_ldg* f, R* w) {
    for (int k = 0; k < kFastWeights; ++k) w[k] = R(0);
    const int terms = terms_of(f);
    if (terms == kOneConstant) return;
    const double Ls[4] = {f->L1, f->L2, f->L3, f->L4};
    const bool b = terms != kMultiLinear;
    const int cnt = b ? kFastLinBCount : kFastLinACount;
    for (int k = 0; k < cnt; ++k) {
        const double* c = b ? kFastLinB[k] : kFastLinA[k];
        double acc = 0;
        for (int t = 0; t < 4; ++t) acc += Ls[t] * c[t];
        w[k] = R(acc);
    }
}
static_assert(kFastLinBCount <= kFastWeights, "composed operator larger than the weight buffer");

// The device buffer holds the weights of the 19-point operator, then the few
// distinct weights of its plane-by-plane form (each computed as the entries'
// own weights are, so that both forms multiply by the same values).
template <typename R>
static void weight_buffer(const qm_ldg* f, R* w) {
    composed_weights<R>(f, w);
    R* plane = w + kFastWeights;
    for (int k = 0; k < kFastWeights; ++k) plane[k] = R(0);
    const int terms = terms_of(f);
    if (terms == kOneConstant) return;
    const double Ls[4] = {f->L1, f->L2, f->L3, f->L4};
    const bool b = terms != kMultiLinear;
    const int cnt = b ? kPlaneLinBCount : kPlaneLinACount;
    for (int j = 0; j < cnt; ++j) {
        const double* c = b ? kPlaneLinB[j] : kPlaneLinA[j];
        double acc = 0;
        for (int t = 0; t < 4; ++t) acc += Ls[t] * c[t];
        plane[j] = R(acc);
    }
}

static int upload_weights(qm_lattice* L, const qm_ldg* f) {
    if (L->precision == QM_FP64) {
        double w[2 * kFastWeights];
        weight_buffer<double>(f, w);
        HIP_TRY(copy_on(L, L->dW, w, sizeof(w), hipMemcpyHostToDevice));
    } else {
        float w[2 * kFastWeights];
        weight_buffer<float>(f, w);
        HIP_TRY(copy_on(L, L->dW, w, sizeof(w), hipMemcpyHostToDevice));
    }
    return QM_OK;
}

// Tile shape of the lattice pass: the one set on the lattice, else the measured
// better default.  On lattices at least 128 sites wide in x and y: for FP64 FIRE
// passes, the 32 x 16 pass that keeps only Q and v (256^3: 0.87-0.92 ms per
// iteration against 0.93-0.96 for the fused 32 x 16 pass, which reads and
// writes F as well); for FP64 force-only passes 32 x 16 (2% faster than 32 x 8,
// 3% faster than 64 x 8); in FP32 64 x 8 (6% faster than 32 x 16, 11% faster
// than 64 x 4, and 20% faster than the FP32 passes that keep only Q and v).
// On smaller lattices, and with the L6 term (whose pass needs more than the 128
// registers the 512-thread tiles allow), 32 x 8 for FP64 and 64 x 4 for FP32.
// The tiles of the FIRE pass that keeps only Q and v (4..6) serve FIRE passes
// without the L6 term; a force-only pass, or one with L6, then takes the
// default of its kind.
int qmin::active_tile(const qm_lattice* L, int mode) {
    // the 64 x 8 FIRE tile that keeps only Q and v needs more LDS than a CU has in FP64
    const int t = (L->tileOverride == kStream64x8 && L->precision == QM_FP64) ? kStream32x16 : L->tileOverride;
    if (t >= 0 && !(is_stream_tile(t) && (mode != kFire || L->terms == kMultiL6))) return t;
    if (L->terms != kMultiL6 && L->lx >= 128 && L->ly >= 128) {
        if (L->precision == QM_FP64) return mode == kFire ? kStream32x16 : kTile32x16;
        return kTile64x8;
    }
    return L->precision == QM_FP64 ? kTile32x8 : kTile64x4;
}

long long qmin::tile_columns(const qm_lattice* L, int tile) {
    return (long long)((L->lx + kTileTX[tile] - 1) / kTileTX[tile]) * ((L->ly + kTileTY[tile] - 1) / kTileTY[tile]);
}

template <typename R>
void qmin::fill_args(qm_lattice* L, const qm_ldg* f, MarchArgs<R>& ma, BoundaryArgs<R>& ba, const FireConst& fire,
                     int tile) {
    ma.set[0] = state_set<R>(L, 0);
    ma.set[1] = state_set<R>(L, 1);
    ma.cls = L->dClass;
    ma.W = static_cast<const R*>(L->dW);
    ma.lx = L->lx;
    ma.ly = L->ly;
    ma.lz = L->lz;
    ma.zchunk = std::min(L->zchunk, L->lz);
    ma.n = L->n;
    ma.nbx = (L->lx + kTileTX[tile] - 1) / kTileTX[tile];
    ma.nby = (L->ly + kTileTY[tile] - 1) / kTileTY[tile];
    const int nbz = (L->lz + ma.zchunk - 1) / ma.zchunk;
    ma.nblocks = ma.nbx * ma.nby * nbz;
    ma.fc = force_const(f);
    ma.ctrl = L->dCtrl;
    ma.partials = L->dPartials;
    ba.set[0] = ma.set[0];
    ba.set[1] = ma.set[1];
    ba.types = L->dTypes;
    ba.objects = L->dObjects;
    ba.sites = L->dSites;
    ba.nsites = L->nBoundary;
    ba.lx = L->lx;
    ba.ly = L->ly;
    ba.lz = L->lz;
    ba.n = L->n;
    ba.marchBlocks = ma.nblocks;
    ba.fc = ma.fc;
    ba.fire = fire;
    ba.ctrl = L->dCtrl;
    ba.partials = L->dPartials;
    ba.trace = L->dTrace;
    ma.hvar = f->hvarOn ? L->dHField : nullptr;
    ma.hvarP = f->dchi * f->mu0;
    ba.hvar = ma.hvar;
    ba.hvarP = ma.hvarP;
}

template void qmin::fill_args<double>(qm_lattice*, const qm_ldg*, MarchArgs<double>&, BoundaryArgs<double>&,
                                      const FireConst&, int);
template void qmin::fill_args<float>(qm_lattice*, const qm_ldg*, MarchArgs<float>&, BoundaryArgs<float>&,
                                     const FireConst&, int);

// z chunk of the fused pass: as long as possible while the grid still fills
// every CU to its occupancy, so that the whole grid is resident at once, all
// columns march through z in step and neighbouring columns find each other's
// halos in L2; minimum 4 planes.
static void choose_zchunk(qm_lattice* L, int mode, int terms) {
    if (L->zchunkOverride > 0) {
        L->zchunk = L->zchunkOverride;
        return;
    }
    if (L->numCUs == 0) {
        int dev = 0;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
            L->numCUs = prop.multiProcessorCount;
        if (L->numCUs <= 0) L->numCUs = 256;
    }
    const int tile = active_tile(L, mode);
    int perCU;
    if (is_stream_tile(tile))
        perCU = L->precision == QM_FP64 ? stream_blocks_per_cu<double>(terms, tile)
                                        : stream_blocks_per_cu<float>(terms, tile);
    else
        perCU = L->precision == QM_FP64 ? march_blocks_per_cu<double>(mode, terms, tile)
                                        : march_blocks_per_cu<float>(mode, terms, tile);
    const long long columns = tile_columns(L, tile);
    long long chunks = (long long)perCU * L->numCUs / columns;
    if (chunks < 1) chunks = 1;
    int zc = (int)((L->lz + chunks - 1) / chunks);
    L->zchunk = std::max(4, std::min(zc, L->lz));
}

static int march_blocks(const qm_lattice* L, int mode) {
    const int zc = std::min(L->zchunk, L->lz);
    return int(tile_columns(L, active_tile(L, mode)) * ((L->lz + zc - 1) / zc));
}

// One pass (force only or one FIRE iteration); kernels only, no host sync.
// partials: the work-group partial sums (null: the lattice's own buffer)
template <typename R>
// boundaryOnly: only the boundary pass (a force pass ahead of the FIRE pass
// that keeps only Q and v, which computes the other sites' force itself)
static int run_pass(qm_lattice* L, const qm_ldg* f, int mode, const FireConst& fire, bool gated = false,
                    double* partials = nullptr, bool storeF = false, bool boundaryOnly = false) {
    choose_zchunk(L, mode, terms_of(f));
    MarchArgs<R> ma{};
    BoundaryArgs<R> ba{};
    const int tile = active_tile(L, mode);
    fill_args<R>(L, f, ma, ba, fire, tile);
    if (partials) {
        ma.partials = partials;
        ba.partials = partials;
    }
    ma.gate = gated ? 1 : 0;
    ba.gate = ma.gate;
    ma.storeF = storeF ? 1 : 0;
    const int terms = terms_of(f);
    hipEvent_t ev[3];
    if (g_timing.on && mode == kFire) {
        for (auto& e : ev) HIP_TRY(hipEventCreate(&e));
        HIP_TRY(hipEventRecord(ev[0], L->stream));
    }
    if (boundaryOnly)
        ;
    else if (is_stream_tile(tile))
        HIP_TRY(launch_fire_stream<R>(ma, terms, tile, L->stream));
    else
        HIP_TRY(launch_march<R>(ma, mode, terms, tile, L->stream));
    if (g_timing.on && mode == kFire) HIP_TRY(hipEventRecord(ev[1], L->stream));
    launch_boundary<R>(ba, mode, terms, L->stream);
    HIP_TRY(hipGetLastError());
    if (mode == kFire) {
        launch_fire_finalize(L->dCtrl, fire, ma.partials, (long long)ma.nblocks + boundary_blocks(L->nBoundary),
                             L->dTrace, L->stream);
        HIP_TRY(hipGetLastError());
    }
    if (g_timing.on && mode == kFire) {
        HIP_TRY(hipEventRecord(ev[2], L->stream));
        HIP_TRY(hipEventSynchronize(ev[2]));
        float a = 0, b = 0;
        HIP_TRY(hipEventElapsedTime(&a, ev[0], ev[1]));
        HIP_TRY(hipEventElapsedTime(&b, ev[1], ev[2]));
        g_timing.launches[0]++;
        g_timing.launches[1]++;
        g_timing.ms[0] += a;
        g_timing.ms[1] += b;
        for (auto& e : ev) (void)hipEventDestroy(e);
    }
    return QM_OK;
}

// Average duration of the fused FIRE pass alone, on the current state: it
// reads the current state set and writes only the other one and the partial
// sums, which the next real pass overwrites, so timing it changes nothing.
template <typename R>
static int time_march(qm_lattice* L, const qm_ldg* f, int reps, double* ms) {
    choose_zchunk(L, kFire, terms_of(f));
    MarchArgs<R> ma{};
    BoundaryArgs<R> ba{};
    FireConst fire{};
    const int terms = terms_of(f), tile = active_tile(L, kFire);
    fill_args<R>(L, f, ma, ba, fire, tile);
    auto launch = [&]() {
        return is_stream_tile(tile) ? launch_fire_stream<R>(ma, terms, tile, L->stream)
                                    : launch_march<R>(ma, kFire, terms, tile, L->stream);
    };
    for (int w = 0; w < 2; ++w) HIP_TRY(launch());
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, L->stream));
    for (int r = 0; r < reps; ++r) HIP_TRY(launch());
    HIP_TRY(hipEventRecord(e1, L->stream));
    HIP_TRY(hipEventSynchronize(e1));
    float t = 0;
    HIP_TRY(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = t / reps;
    return QM_OK;
}

// Tile autotuner (the role of the reference's kernelTuner, kernelTuner.cpp:5-125,
// measured with HIP events on the library's stream instead of host clocks):
// every compiled tile shape x z chunks {automatic, 8, 16, 32, 64}.
int qm_lattice_autotune(qm_lattice* L, qm_ldg* f, int* tileOut, int* zchunkOut, double* msOut) {
    if (int rc = prepare(L, f)) return rc;
    if (int rc = sync_ctrl_to_host(L)) return rc;
    const int savedDone = L->hCtrl.done;
    if (savedDone) {
        // the pass returns at once on a finished control block
        L->hCtrl.done = 0;
        if (int rc = push_ctrl(L)) return rc;
    }
    const int reps = L->n >= (1 << 22) ? 5 : 20;
    double best = 1e30;
    int bestTile = 0, bestZ = 0;
    for (int tile = 0; tile < kNumTiles; ++tile) {
        for (int zc : {0, 8, 16, 32, 64}) {
            if (zc > L->lz) continue;
            L->tileOverride = tile;
            L->zchunkOverride = zc;
            double ms = 0;
            int rc = L->precision == QM_FP64 ? time_march<double>(L, f, reps, &ms)
                                             : time_march<float>(L, f, reps, &ms);
            if (rc) return rc;
            if (ms < best) {
                best = ms;
                bestTile = tile;
                bestZ = zc;
            }
        }
    }
    L->tileOverride = bestTile;
    L->zchunkOverride = bestZ;
    if (savedDone) {
        L->hCtrl.done = savedDone;
        if (int rc = push_ctrl(L)) return rc;
    }
    if (tileOut) *tileOut = bestTile;
    if (zchunkOut) *zchunkOut = bestZ;
    if (msOut) *msOut = best;
    return QM_OK;
}

int qmin::prepare(qm_lattice* L, const qm_ldg* f) {
    if (!L || !f) return fail(QM_EINVAL, "null handle");
    L->terms = terms_of(f);
    if (f->hvarOn && !L->dHField)
        return fail(QM_ESTATE, "the force uses a spatially varying field but the lattice has none");
    if (int rc = refresh_types(L)) return rc;
    if (int rc = upload_weights(L, f)) return rc;
    // room for the smallest z chunk (4 planes) the fused pass may pick
    long long worst = 0;
    for (int t = 0; t < kNumTiles; ++t) worst = std::max(worst, tile_columns(L, t) * ((L->lz + 3) / 4));
    const size_t need = 3 * (size_t)(worst + boundary_blocks(L->nBoundary));
    return ensure_partials(L, need);
}

int qmin::force_pass(qm_lattice* L, const qm_ldg* f, bool gated) {
    FireConst fire{};
    return L->precision == QM_FP64 ? run_pass<double>(L, f, kForce, fire, gated)
                                   : run_pass<float>(L, f, kForce, fire, gated);
}

template <typename R>
static int step_pass_t(qm_lattice* L, const qm_ldg* f, int mode, const StepArgs& step, long long* slots) {
    // the force pass's tile and z chunk: the update adds no staging
    choose_zchunk(L, kForce, terms_of(f));
    MarchArgs<R> ma{};
    BoundaryArgs<R> ba{};
    FireConst fire{};
    const int tile = active_tile(L, kForce);
    fill_args<R>(L, f, ma, ba, fire, tile);
    ma.step = step;
    ba.step = step;
    HIP_TRY(launch_step<R>(ma, ba, mode, terms_of(f), tile, L->stream));
    *slots = (long long)ma.nblocks + boundary_blocks(L->nBoundary);
    return QM_OK;
}

int qmin::step_pass(qm_lattice* L, const qm_ldg* f, int mode, const StepArgs& step, long long* slots) {
    return L->precision == QM_FP64 ? step_pass_t<double>(L, f, mode, step, slots)
                                   : step_pass_t<float>(L, f, mode, step, slots);
}

int qmin::share_velocity(qm_lattice* L) {
    const int cur = L->hCtrl.cur;
    const size_t bytes = 5 * (size_t)L->n * L->realSize;
    char* src = static_cast<char*>(L->dSet[cur]) + bytes;
    char* dst = static_cast<char*>(L->dSet[cur ^ 1]) + bytes;
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, L->stream));
    return QM_OK;
}

// every change of the types ends in refresh_types, which counts them
int64_t qmin::n_active(const qm_lattice* L) { return L->nActive; }

int qm_ldg_compute_forces(qm_ldg* f, qm_lattice* L) {
    if (int rc = prepare(L, f)) return rc;
    FireConst fire{};
    int rc = L->precision == QM_FP64 ? run_pass<double>(L, f, kForce, fire) : run_pass<float>(L, f, kForce, fire);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(L->stream));
    return QM_OK;
}

static EnergyConst energy_const(const qm_ldg* f, const qm_lattice* L) {
    EnergyConst ec{};
    ec.a = 0.5 * f->A;
    ec.b = f->B / 3.0;
    ec.c = 0.25 * f->C;
    ec.L[0] = f->L1;
    ec.L[1] = f->multi ? f->L2 : 0;
    ec.L[2] = f->multi ? f->L3 : 0;
    ec.L[3] = f->multi ? f->L4 : 0;
    ec.L[4] = f->multi ? f->L6 : 0;
    ec.eOn = f->eOn;
    ec.hOn = f->hOn;
    for (int i = 0; i < 3; ++i) {
        ec.E[i] = f->E[i];
        ec.H[i] = f->H[i];
    }
    ec.eps = f->eps;
    ec.eps0 = f->eps0;
    ec.deps = f->deps;
    ec.chi = f->chi;
    ec.mu0 = f->mu0;
    ec.dchi = f->dchi;
    ec.hvar = f->hvarOn ? L->dHField : nullptr;
    return ec;
}

// landauDeGennesLC::computeObjectForces (landauDeGennesLCOtherForces.cpp:147-208)
// with computeStressTensors (:214-251): the force on object `object` (0-based)
// from the stress tensor on its surface sites, summed in surface-site order.
// One-constant models only, as in the reference.
int qm_ldg_object_force(qm_ldg* f, qm_lattice* L, int object, double* force) {
    if (!f || !L || !force || object < 0 || object >= (int)L->objects.size()) return fail(QM_EINVAL, "bad argument");
    if (f->multi)
        return fail(QM_EINVAL, "object forces need the one-constant model: the reference computes the stress "
                               "tensor only there (landauDeGennesLCOtherForces.cpp:214-251)");
    if (int rc = prepare(L, f)) return rc;
    const std::vector<int64_t>& sur = L->objects[object].surface;
    const long long m = (long long)sur.size();
    force[0] = force[1] = force[2] = 0.0;
    if (m == 0) return QM_OK;
    const EnergyConst ec = energy_const(f, L);
    int64_t* dSites = nullptr;
    double* dOut = nullptr;
    HIP_TRY(hipMalloc(&dSites, m * sizeof(int64_t)));
    hipError_t err = hipMalloc(&dOut, 3 * m * sizeof(double));
    std::vector<double> out(3 * m);
    if (err == hipSuccess)
        err = hipMemcpyAsync(dSites, sur.data(), m * sizeof(int64_t), hipMemcpyHostToDevice, L->stream);
    if (err == hipSuccess) {
        if (L->precision == QM_FP64)
            launch_object_force<double>(state_set<double>(L, L->hCtrl.cur).q, L->dTypes, L->dObjects, L->lx, L->ly,
                                        L->lz, ec, f->L1, dSites, m, dOut, L->stream);
        else
            launch_object_force<float>(state_set<float>(L, L->hCtrl.cur).q, L->dTypes, L->dObjects, L->lx, L->ly,
                                       L->lz, ec, f->L1, dSites, m, dOut, L->stream);
        err = hipGetLastError();
    }
    if (err == hipSuccess)
        err = hipMemcpyAsync(out.data(), dOut, 3 * m * sizeof(double), hipMemcpyDeviceToHost, L->stream);
    if (err == hipSuccess) err = hipStreamSynchronize(L->stream);
    (void)hipFree(dSites);
    if (dOut) (void)hipFree(dOut);
    if (err != hipSuccess) return fail(QM_EHIP, std::string("object force: ") + hipGetErrorString(err));
    double fx = 0, fy = 0, fz = 0;
    for (long long k = 0; k < m; ++k) {
        fx = fx + out[3 * k];
        fy = fy + out[3 * k + 1];
        fz = fz + out[3 * k + 2];
    }
    force[0] = fx;
    force[1] = fy;
    force[2] = fz;
    return QM_OK;
}

int qm_ldg_compute_energy(qm_ldg* f, qm_lattice* L, double* total, double* comps) {
    if (int rc = prepare(L, f)) return rc;
    const EnergyConst ec = energy_const(f, L);
    if (int rc = ensure_partials(L, 5 * (size_t)kMaxEnergyBlocks)) return rc;
    int blocks;
    if (L->precision == QM_FP64)
        blocks = launch_energy<double>(state_set<double>(L, L->hCtrl.cur).q, L->dTypes, L->dObjects, L->lx, L->ly,
                                       L->lz, ec, L->dPartials, L->stream);
    else
        blocks = launch_energy<float>(state_set<float>(L, L->hCtrl.cur).q, L->dTypes, L->dObjects, L->lx, L->ly,
                                      L->lz, ec, L->dPartials, L->stream);
    HIP_TRY(hipGetLastError());
    std::vector<double> part(5 * (size_t)blocks);
    HIP_TRY(hipMemcpyAsync(part.data(), L->dPartials, part.size() * sizeof(double), hipMemcpyDeviceToHost, L->stream));
    HIP_TRY(hipStreamSynchronize(L->stream));
    double c5[5] = {0, 0, 0, 0, 0};
    for (int b = 0; b < blocks; ++b)
        for (int k = 0; k < 5; ++k) c5[k] += part[5 * (size_t)b + k];
    if (comps)
        for (int k = 0; k < 5; ++k) comps[k] = c5[k];
    if (total) *total = c5[0] + c5[1] + c5[2] + c5[3] + c5[4];
    return QM_OK;
}

int qm_ldg_energy_density(qm_ldg* f, qm_lattice* L, double* density) {
    if (!density) return fail(QM_EINVAL, "null output");
    if (int rc = prepare(L, f)) return rc;
    const EnergyConst ec = energy_const(f, L);
    double* dOut = nullptr;
    HIP_TRY(hipMalloc(&dOut, sizeof(double) * (size_t)L->n));
    if (L->precision == QM_FP64)
        launch_energy_density<double>(state_set<double>(L, L->hCtrl.cur).q, L->dTypes, L->dObjects, L->lx, L->ly,
                                      L->lz, ec, dOut, L->stream);
    else
        launch_energy_density<float>(state_set<float>(L, L->hCtrl.cur).q, L->dTypes, L->dObjects, L->lx, L->ly,
                                     L->lz, ec, dOut, L->stream);
    hipError_t err = hipGetLastError();
    if (err == hipSuccess)
        err = hipMemcpyAsync(density, dOut, sizeof(double) * (size_t)L->n, hipMemcpyDeviceToHost, L->stream);
    if (err == hipSuccess) err = hipStreamSynchronize(L->stream);
    (void)hipFree(dOut);
    if (err != hipSuccess) return fail(QM_EHIP, std::string("energy density: ") + hipGetErrorString(err));
    return QM_OK;
}

// ------------------------------------------------------------------ FIRE
int qmin::materialize(qm_lattice* L) {
    if (int rc = sync_ctrl_to_host(L)) return rc;
    if (L->hCtrl.identity) return QM_OK;
    if (L->precision == QM_FP64)
        launch_materialize<double>(state_set<double>(L, 0), state_set<double>(L, 1), L->n, L->dCtrl, L->stream);
    else
        launch_materialize<float>(state_set<float>(L, 0), state_set<float>(L, 1), L->n, L->dCtrl, L->stream);
    HIP_TRY(hipGetLastError());
    L->hCtrl.identity = 1;
    L->hCtrl.reset = 0;
    HIP_TRY(hipMemcpyAsync(&L->dCtrl->identity, &L->hCtrl.identity, sizeof(int), hipMemcpyHostToDevice, L->stream));
    HIP_TRY(hipMemcpyAsync(&L->dCtrl->reset, &L->hCtrl.reset, sizeof(int), hipMemcpyHostToDevice, L->stream));
    HIP_TRY(hipStreamSynchronize(L->stream));
    return QM_OK;
}

int qm_fire_create(qm_fire** out) {
    if (!out) return fail(QM_EINVAL, "null output handle");
    *out = new qm_fire;
    return QM_OK;
}
int qm_fire_destroy(qm_fire* m) {
    delete m;
    return QM_OK;
}

int qm_fire_set_parameters(qm_fire* m, double dt, double alphaStart, double dtMax, double dtInc, double dtDec,
                           double alphaDec, int nMin, double fc, double alphaMin) {
    if (!m) return fail(QM_EINVAL, "null minimizer");
    m->dt = dt;
    m->dtMin = dt * .01;
    m->alphaStart = alphaStart;
    m->alpha = alphaStart;
    m->dtMax = dtMax;
    m->dtInc = dtInc;
    m->dtDec = dtDec;
    m->alphaDec = alphaDec;
    m->nMin = nMin;
    m->fc = fc;
    m->alphaMin = alphaMin;
    return QM_OK;
}

int qm_fire_set_force_cutoff(qm_fire* m, double fc) {
    if (!m) return fail(QM_EINVAL, "null minimizer");
    m->fc = fc;
    return QM_OK;
}

int qm_fire_set_max_iterations(qm_fire* m, int it) {
    if (!m) return fail(QM_EINVAL, "null minimizer");
    m->maxIter = it;
    return QM_OK;
}

int qm_fire_set_trace(qm_fire* m, int capacity) {
    if (!m || capacity < 0) return fail(QM_EINVAL, "bad argument");
    m->traceCap = capacity;
    m->trace.clear();
    return QM_OK;
}

int qm_fire_get_trace(qm_fire* m, double* rows, int cap, int* n) {
    if (!m || !n) return fail(QM_EINVAL, "bad argument");
    const int k = std::min<int>(cap, (int)(m->trace.size() / 6));
    if (rows && k > 0) std::memcpy(rows, m->trace.data(), 6 * (size_t)k * sizeof(double));
    *n = k;
    return QM_OK;
}

int qm_fire_get_state(const qm_fire* m, qm_fire_state* s) {
    if (!m || !s) return fail(QM_EINVAL, "bad argument");
    s->iterations = m->iterations;
    s->nPositive = m->nPositive;
    s->converged = m->converged;
    s->nanDetected = m->nan;
    s->forceMax = m->forceMax;
    s->power = m->power;
    s->alpha = m->alpha;
    s->deltaT = m->dt;
    s->scaling = m->scaling;
    return QM_OK;
}

FireConst qmin::fire_const(const qm_fire* m, const qm_lattice* L) {
    FireConst fp{};
    fp.alphaStart = m->alphaStart;
    fp.dtMax = m->dtMax;
    fp.dtInc = m->dtInc;
    fp.dtDec = m->dtDec;
    fp.alphaDec = m->alphaDec;
    fp.alphaMin = m->alphaMin;
    fp.dtMin = m->dtMin;
    fp.fc = m->fc;
    int64_t na = 0;
    qm_lattice_n_active(L, &na);
    fp.nTotal = (double)na;
    fp.nMin = m->nMin;
    return fp;
}

// ------------------------------------------------------------------ allocation-free FIRE
// Work-group partial sums of qm_fire_run_iterations_noalloc: a fixed array in
// the library's device image, so that the call needs no hipMalloc.  Each GPU the
// library runs on holds its own copy until the process ends.  Every slot a pass
// reads is written by that pass first, so nothing has to start at zero; calls
// that overlap on one GPU share the array and race.
constexpr long long kScratchSlots = 65536;   // partial-sum slots of 3 doubles: 1.5 MiB
constexpr int kScratchDevices = 64;
static __device__ __align__(256) double scratch[3 * kScratchSlots];
static std::mutex g_scratchMutex;
static double* g_scratchPtr[kScratchDevices] = {};

// Device address of `scratch` on the lattice's GPU, looked up once per device
// with that device current; host code never uses the array's own name.
static int scratch_for(const qm_lattice* L, double** out) {
    if (L->device < 0 || L->device >= kScratchDevices) return fail(QM_EINVAL, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    if (!g_scratchPtr[L->device]) {
        int prev = 0;
        HIP_TRY(hipGetDevice(&prev));
        HIP_TRY(hipSetDevice(L->device));
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        const hipError_t r = hipSetDevice(prev);
        if (e != hipSuccess) return fail(QM_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        if (r != hipSuccess) return fail(QM_EHIP, std::string("hipSetDevice: ") + hipGetErrorString(r));
        g_scratchPtr[L->device] = static_cast<double*>(p);
    }
    *out = g_scratchPtr[L->device];
    return QM_OK;
}

// What prepare() does, minus every allocation (the site types are on the
// device already: qm_lattice_create and every object call upload them at
// once), provided the passes' partial sums fit the scratch array.
static int prepare_noalloc(qm_lattice* L, const qm_ldg* f, double** partials) {
    if (!L || !f) return fail(QM_EINVAL, "null handle");
    if (f->hvarOn && !L->dHField)
        return fail(QM_ESTATE, "the force uses a spatially varying field but the lattice has none");
    L->terms = terms_of(f);
    long long need = 0;
    for (int mode : {kForce, kFire}) {
        choose_zchunk(L, mode, L->terms);
        need = std::max(need, (long long)march_blocks(L, mode) + boundary_blocks(L->nBoundary));
    }
    if (need > kScratchSlots)
        return fail(QM_EINVAL, "the lattice needs " + std::to_string(need) + " partial-sum slots; the "
                               "allocation-free call holds at most " + std::to_string(kScratchSlots));
    if (int rc = upload_weights(L, f)) return rc;
    return scratch_for(L, partials);
}

// Runs iterations on the device until the control block reports done, syncing
// with the host once per batch.  With noalloc the partial sums go to the scratch
// array and no trace is kept.
static int fire_drive(qm_fire* m, qm_ldg* f, qm_lattice* L, int maxIter, bool tolerance, bool noalloc = false) {
    double* partials = nullptr;
    if (noalloc) {
        if (int rc = prepare_noalloc(L, f, &partials)) return rc;
    } else {
        if (int rc = prepare(L, f)) return rc;
        if (m->traceCap > 0 && L->traceCap < m->traceCap) {
            if (L->dTrace) HIP_TRY(hipFree(L->dTrace));
            HIP_TRY(hipMalloc(&L->dTrace, 6 * (size_t)m->traceCap * sizeof(double)));
            L->traceCap = m->traceCap;
        }
    }
    const int traceCap = noalloc ? 0 : m->traceCap;
    if (int rc = sync_ctrl_to_host(L)) return rc;
    Ctrl& c = L->hCtrl;
    c.done = 0;
    c.iterations = m->iterations;
    c.maxIter = maxIter;
    c.nPositive = m->nPositive;
    c.dt = m->dt;
    c.alpha = m->alpha;
    c.arrive = 0;
    c.nanFlag = 0;
    c.traceLen = 0;
    c.traceCap = traceCap;
    c.forceMax = m->forceMax;
    if (int rc = push_ctrl(L)) return rc;
    FireConst fp = fire_const(m, L);
    if (!tolerance) fp.fc = -1.0;  // never converged: exactly maxIter - iterations steps
    // sim->computeForces() at the start of minimize (energyMinimizerFIRE.cpp:157).
    // The FIRE pass that keeps only Q and v computes the force of the sites it
    // finishes itself and reads only that of the boundary pass's sites.
    const bool streamFire = is_stream_tile(active_tile(L, kFire));
    int rc = L->precision == QM_FP64 ? run_pass<double>(L, f, kForce, fp, false, partials, false, streamFire)
                                     : run_pass<float>(L, f, kForce, fp, false, partials, false, streamFire);
    if (rc) return rc;
    // With a tolerance the host looks at the control block every batch (the
    // device skips the passes after the stop anyway); a run of known length is
    // queued in batches of up to 256 iterations.
    const int batch = !tolerance ? 256 : (L->n >= (1 << 22) ? 16 : 64);
    // without a tolerance the run's last pass is known: it stores F as well
    bool lastStoresF = false;
    while (true) {
        int todo = std::min(batch, maxIter - L->hCtrl.iterations);
        if (todo <= 0) todo = 1;  // at least one iteration per minimize call
        for (int k = 0; k < todo; ++k) {
            const bool last = !tolerance && L->hCtrl.iterations + k + 1 == maxIter;
            rc = L->precision == QM_FP64 ? run_pass<double>(L, f, kFire, fp, false, partials, last)
                                         : run_pass<float>(L, f, kFire, fp, false, partials, last);
            if (rc) return rc;
            lastStoresF = last;
        }
        if ((rc = sync_ctrl_to_host(L))) return rc;
        if (L->hCtrl.done) break;
    }
    const Ctrl& h = L->hCtrl;
    m->iterations = h.iterations;
    m->nPositive = h.nPositive;
    m->dt = h.dt;
    m->alpha = h.alpha;
    m->forceMax = h.forceMax;
    m->power = h.power;
    m->scaling = h.scaling;
    m->nan = h.nanFlag;
    m->converged = !h.nanFlag && h.forceMax <= m->fc;
    if (traceCap > 0 && h.traceLen > 0) {
        const size_t old = m->trace.size();
        m->trace.resize(old + 6 * (size_t)h.traceLen);
        HIP_TRY(copy_on(L, m->trace.data() + old, L->dTrace, 6 * (size_t)h.traceLen * sizeof(double),
                          hipMemcpyDeviceToHost));
    }
    // The pass that keeps only Q and v leaves F stale on sites it finishes
    // itself, unless it was the run's known last pass: recompute F of the final
    // state, which the caller may read.
    if (streamFire && !(lastStoresF && h.iterations == maxIter && !h.nanFlag)) {
        rc = L->precision == QM_FP64 ? run_pass<double>(L, f, kForce, fp, false, partials)
                                     : run_pass<float>(L, f, kForce, fp, false, partials);
        if (rc) return rc;
    }
    // energyMinimizerFIRE::fireStep mixes the velocity with the force of the
    // step's end at once (energyMinimizerFIRE.cpp:113-116); the fused pass defers
    // that mix to the next pass.  Make it concrete before returning, so that
    // whatever changes Q or F between minimize calls (object moves, a new Q, a
    // force computation) cannot change the mix.
    if (int rc2 = materialize(L)) return rc2;
    if (h.nanFlag) return fail(QM_ENAN, "FIRE produced a NaN force norm");
    return QM_OK;
}

int qm_fire_minimize(qm_fire* m, qm_ldg* f, qm_lattice* L) {
    if (!m) return fail(QM_EINVAL, "null minimizer");
    return fire_drive(m, f, L, m->maxIter, true);
}

int qm_fire_run_iterations(qm_fire* m, qm_ldg* f, qm_lattice* L, int n) {
    if (!m || n < 1) return fail(QM_EINVAL, "bad argument");
    return fire_drive(m, f, L, m->iterations + n, false);
}

int qm_fire_run_iterations_noalloc(qm_fire* m, qm_ldg* f, qm_lattice* L, int n) {
    if (!m || n < 1) return fail(QM_EINVAL, "bad argument");
    return fire_drive(m, f, L, m->iterations + n, false, true);
}

// ------------------------------------------------------------------ timing
int qm_timing_enable(int on) {
    g_timing.on = on != 0;
    return QM_OK;
}
int qm_timing_reset(void) {
    g_timing.launches[0] = g_timing.launches[1] = 0;
    g_timing.ms[0] = g_timing.ms[1] = 0;
    return QM_OK;
}
int qm_timing_get(int k, int64_t* launches, double* ms) {
    if (k < 0 || k > 1) return fail(QM_EINVAL, "kernel index is 0 or 1");
    if (launches) *launches = g_timing.launches[k];
    if (ms) *ms = g_timing.ms[k];
    return QM_OK;
}

