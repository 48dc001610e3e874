// C++ classes of open-Qmin's lattice Landau-de Gennes path, backed by libqmin.
//
// The classes keep the reference's names, constructors and method signatures
// for the single-rank path (src/model/multirankQTensorLatticeModel.h,
// src/forces/landauDeGennesLC.h, src/updaters/energyMinimizer*.h,
// src/simulation/multirankSimulation.h), so a driver written against the
// reference (openQmin.cpp) compiles against this header with its include lines
// changed.  All state lives on the GPU; every call goes through the C ABI in
// include/qmin.h.  Errors throw qminError (the reference throws
// std::exception from HANDLE_ERROR, src/utilities/std_include.h:170-183).
//
// Differences a caller can see:
//  * one rank only: halos, rank topologies > 1x1x1 and GPU selection by MPI
//    rank throw;
//  * GPUArray<dVec> / ArrayHandle keep the reference's host view (AoS dVec, one
//    per site, {Qxx,Qxy,Qxz,Qyy,Qyz}); the arrays a model returns are views of
//    its device state (structure of arrays), copied on a host handle's
//    acquisition and written back on its release; device handles throw, as the
//    device layout is not an array of dVec;
//  * force::computeForces(GPUArray<dVec>&, zeroOutForce, type) computes every
//    site's force at type 0; type 1 (the CPU path's second pass over boundary
//    sites) adds nothing, as the device passes cover all sites together;
//  * setCPUOperation(true) selects the reference CPU path's metric semantics,
//    (false) the GPU path's; the arithmetic always runs on the GPU.
#pragma once

#include <chrono>
#include <cmath>
#include <cstdio>
#include <iostream>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "qmin.h"

typedef double scalar;

struct scalar3 {
    double x, y, z;
};
struct scalar4 {
    double x, y, z, w;
};
struct int3 {
    int x, y, z;
};
inline scalar3 make_scalar3(double x, double y, double z) { return scalar3{x, y, z}; }

// normalizeDirector (inc/functions.h:193-205): unit vector, x for a zero vector
inline void normalizeDirector(scalar3& p) {
    const scalar a = std::sqrt(p.x * p.x + p.y * p.y + p.z * p.z);
    if (a == 0) {
        p.x = 1;
        p.y = 0;
        p.z = 0;
    } else {
        p.x = p.x / a;
        p.y = p.y / a;
        p.z = p.z / a;
    }
}

// src/utilities/std_include.h (distortionEnergyType) and src/model/latticeBoundaries.h:8-40
enum class distortionEnergyType { oneConstant, multiConstant };
enum class boundaryType { homeotropic, degeneratePlanar };

class boundaryObject {
   public:
    boundaryObject(boundaryType b, double p1, double p2) : boundary(b), P1(p1), P2(p2) {}
    boundaryType boundary;
    double P1, P2;   // W and S0
};

class qminError : public std::runtime_error {
   public:
    qminError(int code, const std::string& what) : std::runtime_error(what), code(code) {}
    int code;
};

namespace qmin_host {
inline void check(int rc, const char* where) {
    if (rc != QM_OK) throw qminError(rc, std::string(where) + ": " + qm_last_error());
}
inline int btype(boundaryType b) { return b == boundaryType::homeotropic ? QM_HOMEOTROPIC : QM_DEGENERATE_PLANAR; }
}  // namespace qmin_host

// src/utilities/noiseSource.h:20-90: a reproducible source is seeded by
// setReproducibleSeed; the lattice initialiser consumes it.
class noiseSource {
   public:
    noiseSource(bool rep = false) : Reproducible(rep) {}
    void setReproducibleSeed(int seed) {
        Reproducible = true;
        seed_ = seed;
    }
    // the seed the next lattice initialisation uses (a fresh random one when
    // the source is not reproducible)
    int seedForDraw() {
        if (!Reproducible) return int(std::random_device{}() & 0x7fffffff);
        return seed_;
    }
    bool Reproducible;

   private:
    int seed_ = 13377;
};

// logSpacedIntegers (inc/logSpacedIntegers.h): the iterations of logarithmically
// spaced saves, round(10^(exponent k)) for k = 1, 2, ..., repeats skipped
class logSpacedIntegers {
   public:
    logSpacedIntegers(int firstSave = 0, scalar _exp = 0.05) {
        nextSave = firstSave;
        exponent = _exp;
        base = std::pow(10.0, exponent);
        logSaveIdx = 0;
        if (nextSave != 0) {
            int tempCur = 0;
            while (tempCur < nextSave) {
                logSaveIdx += 1;
                tempCur = (int)std::round(std::pow(base, logSaveIdx));
            }
        }
    }
    void update() {
        logSaveIdx += 1;
        int curSave = (int)std::round(std::pow(base, logSaveIdx));
        while (curSave == nextSave) {
            logSaveIdx += 1;
            curSave = (int)std::round(std::pow(base, logSaveIdx));
        }
        nextSave = curSave;
    }
    int nextSave;
    int logSaveIdx;
    scalar exponent;
    scalar base;
};

// profiler (src/utilities/profiler.h): wall-clock time of start()/end()
// brackets, accumulated over calls
class profiler {
   public:
    explicit profiler(std::string profilerName) : name(std::move(profilerName)) {}
    void start() { startTime = std::chrono::high_resolution_clock::now(); }
    void end() {
        timeTaken += std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - startTime).count();
        functionCalls += 1;
    }
    double timing() { return functionCalls > 0 ? timeTaken / functionCalls : 0.0; }
    void print() {
        std::cout << "profiler \"" << name << "\": " << functionCalls << " calls, " << timing() << " s per call, "
                  << timeTaken << " s in all" << std::endl;
    }
    std::string name;
    int functionCalls = 0;
    double timeTaken = 0.0;
    std::chrono::time_point<std::chrono::high_resolution_clock> startTime;
};

// dVec (src/utilities/dDimensionalVectorTypes.h) with DIMENSION = 5: one
// Q tensor {Qxx, Qxy, Qxz, Qyy, Qyz}
struct dVec {
    double x[5];
    dVec() : x{0, 0, 0, 0, 0} {}
    explicit dVec(double v) : x{v, v, v, v, v} {}
    double& operator[](int i) { return x[i]; }
    const double& operator[](int i) const { return x[i]; }
    dVec& operator+=(const dVec& o) {
        for (int k = 0; k < 5; ++k) x[k] += o.x[k];
        return *this;
    }
    dVec& operator-=(const dVec& o) {
        for (int k = 0; k < 5; ++k) x[k] -= o.x[k];
        return *this;
    }
};
inline dVec make_dVec(double v) { return dVec(v); }
static_assert(sizeof(dVec) == 5 * sizeof(double), "dVec must be five packed doubles (AoS host layout)");

// GPUArray / ArrayHandle (src/utilities/gpuarray.h:7-124)
struct access_location {
    enum Enum { host, device };
};
struct access_mode {
    enum Enum { read, readwrite, overwrite };
};

template <class T>
class ArrayHandle;

// A host array of T.  Arrays returned by a model are views of one of its
// device fields (Q, v, F: T = dVec; site types: T = int, read only): a host
// ArrayHandle copies the field in (unless it overwrites) and, unless it only
// reads, writes it back when it is destroyed.
template <class T>
class GPUArray {
   public:
    GPUArray() = default;
    explicit GPUArray(unsigned int num_elements, bool = false) : host_(num_elements) {}
    virtual ~GPUArray() = default;
    unsigned int getNumElements() { return size(); }
    unsigned int size() {
        return bound() ? (unsigned int)nSites() : (unsigned int)host_.size();
    }
    virtual void resize(unsigned int num_elements) {
        if (bound()) throw qminError(QM_EINVAL, "GPUArray: a model's array has the model's size");
        host_.resize(num_elements);
    }
    void swap(GPUArray& other) {
        if (bound() || other.bound()) throw qminError(QM_EINVAL, "GPUArray: a model's array cannot be swapped");
        host_.swap(other.host_);
    }
    // binding to a lattice field (QM_FIELD_Q/V/F, or kTypes)
    static constexpr int kTypes = 3;
    void bind(qm_lattice* lat, int field) {
        lat_ = lat;
        field_ = field;
    }
    bool bound() const { return lat_ != nullptr; }
    qm_lattice* lattice() const { return lat_; }
    int field() const { return field_; }

   private:
    friend class ArrayHandle<T>;
    long long nSites() const {
        int lx = 0, ly = 0, lz = 0, prec = 0;
        qm_lattice_dims(lat_, &lx, &ly, &lz, &prec);
        return (long long)lx * ly * lz;
    }
    T* acquire(access_mode::Enum mode);
    void release(access_mode::Enum mode);
    std::vector<T> host_;
    qm_lattice* lat_ = nullptr;
    int field_ = -1;
};

template <class T>
class ArrayHandle {
   public:
    ArrayHandle(GPUArray<T>& gpu_array, access_location::Enum location = access_location::host,
                access_mode::Enum mode = access_mode::readwrite)
        : arr_(gpu_array), mode_(mode) {
        if (location == access_location::device)
            throw qminError(QM_EINVAL, "ArrayHandle: device access is not offered (device storage is structure of "
                                       "arrays, not an array of dVec)");
        data = arr_.acquire(mode);
    }
    ~ArrayHandle() {
        try {
            arr_.release(mode_);
        } catch (...) {
        }
    }
    ArrayHandle(const ArrayHandle&) = delete;
    ArrayHandle& operator=(const ArrayHandle&) = delete;
    T* data;

   private:
    GPUArray<T>& arr_;
    access_mode::Enum mode_;
};

// ---------------------------------------------------------------- models
class simpleModel {
   public:
    virtual ~simpleModel() = default;
    virtual int getNumberOfParticles() = 0;
    virtual qm_lattice* qmLattice() = 0;
    // simpleModel::returnPositions/returnForces/returnTypes/returnVelocities
    // (simpleModel.h:63-73): views of the device state
    GPUArray<dVec>& returnPositions() { return bound(positions, QM_FIELD_Q); }
    GPUArray<dVec>& returnVelocities() { return bound(velocities, QM_FIELD_V); }
    GPUArray<dVec>& returnForces() { return bound(forces, QM_FIELD_F); }
    GPUArray<int>& returnTypes() { return bound(types, GPUArray<int>::kTypes); }
    // simpleModel::getClassSize (simpleModel.h:95): GB of data the model holds
    // (here its device state: two sets of Q, v, F, the site types and classes)
    virtual scalar getClassSize() {
        int lx = 0, ly = 0, lz = 0, prec = 0;
        qm_lattice_dims(qmLattice(), &lx, &ly, &lz, &prec);
        const double n = double(lx) * ly * lz;
        return 1e-9 * (30.0 * n * (prec == QM_FP64 ? 8 : 4) + 2.0 * n);
    }

   private:
    template <class T>
    GPUArray<T>& bound(GPUArray<T>& a, int field) {
        if (!a.bound()) a.bind(qmLattice(), field);
        return a;
    }
    GPUArray<dVec> positions, velocities, forces;
    GPUArray<int> types;
};

template <class T>
T* GPUArray<T>::acquire(access_mode::Enum mode) {
    if (!bound()) return host_.data();
    host_.resize((size_t)nSites());
    if (mode == access_mode::overwrite) return host_.data();
    if constexpr (std::is_same<T, dVec>::value) {
        qmin_host::check(qm_lattice_get_field(lat_, field_, reinterpret_cast<double*>(host_.data())), "ArrayHandle");
    } else {
        static_assert(std::is_same<T, int>::value, "model arrays are dVec or int");
        qmin_host::check(qm_lattice_get_types(lat_, reinterpret_cast<int32_t*>(host_.data())), "ArrayHandle");
    }
    return host_.data();
}

template <class T>
void GPUArray<T>::release(access_mode::Enum mode) {
    if (!bound() || mode == access_mode::read) return;
    // site types are read only through a handle: they change through objects
    // (createWall, createSphericalColloid, createBoundaryFromFile)
    if constexpr (std::is_same<T, dVec>::value)
        qmin_host::check(qm_lattice_set_field(lat_, field_, reinterpret_cast<const double*>(host_.data())),
                         "ArrayHandle");
}

class qTensorLatticeModel : public simpleModel {
   public:
    // cubicLattice(lx, ly, lz, ...) on one rank; precision is QM_FP64 (the
    // reference's arithmetic) or QM_FP32 (mixed form)
    qTensorLatticeModel(int lx, int ly, int lz, int precision = QM_FP64) {
        qmin_host::check(qm_lattice_create(lx, ly, lz, precision, &lat_), "qm_lattice_create");
        latticeSites = {lx, ly, lz};
        N = lx * ly * lz;
    }
    ~qTensorLatticeModel() override { qm_lattice_destroy(lat_); }
    qTensorLatticeModel(const qTensorLatticeModel&) = delete;
    qTensorLatticeModel& operator=(const qTensorLatticeModel&) = delete;

    int getNumberOfParticles() override { return N; }
    qm_lattice* qmLattice() override { return lat_; }

    // qTensorLatticeModel::setNematicQTensorRandomly (qTensorLatticeModel.cpp:134-177)
    void setNematicQTensorRandomly(noiseSource& noise, scalar s0, bool globallyAligned = false) {
        if (globallyAligned)
            qmin_host::check(qm_lattice_random_uniform_init(lat_, noise.seedForDraw(), s0), "setNematicQTensorRandomly");
        else
            qmin_host::check(qm_lattice_random_init(lat_, noise.seedForDraw(), s0), "setNematicQTensorRandomly");
    }
    // multirankQTensorLatticeModel::setRandomDirectors (multirankQTensorLatticeModel.cpp:52-77): the same
    // draws as setNematicQTensorRandomly
    void setRandomDirectors(noiseSource& noise, scalar s0, bool globallyAligned = false) {
        setNematicQTensorRandomly(noise, s0, globallyAligned);
    }
    // multirankQTensorLatticeModel::setUniformDirectors (multirankQTensorLatticeModel.cpp:79-88)
    void setUniformDirectors(scalar3 d, scalar s0) {
        qmin_host::check(qm_lattice_uniform_init(lat_, d.x, d.y, d.z, s0), "setUniformDirectors");
    }
    // multirankQTensorLatticeModel::setDirectorFromFunction (multirankQTensorLatticeModel.cpp:90-110):
    // func(x, y, z) gives the director and S0 at a site; it is evaluated at every site (the values of
    // object sites are not used)
    template <typename F>
    void setDirectorFromFunction(F&& func) {
        std::vector<double> d(4 * (size_t)N);
        size_t k = 0;
        for (int z = 0; z < latticeSites.z; ++z)
            for (int y = 0; y < latticeSites.y; ++y)
                for (int x = 0; x < latticeSites.x; ++x, k += 4) {
                    const scalar4 r = func(scalar(x), scalar(y), scalar(z));
                    d[k] = r.x;
                    d[k + 1] = r.y;
                    d[k + 2] = r.z;
                    d[k + 3] = r.w;
                }
        qmin_host::check(qm_lattice_director_init(lat_, d.data()), "setDirectorFromFunction");
    }
    // qTensorLatticeModel::computeDefectMeasures (qTensorLatticeModel.cpp:95-133)
    void computeDefectMeasures(int defectType) {
        defectMeasures.resize(N);
        qmin_host::check(qm_lattice_defects(lat_, defectType, defectMeasures.data()), "computeDefectMeasures");
    }
    void getPositions(std::vector<double>& q) { get(QM_FIELD_Q, q); }
    void getVelocities(std::vector<double>& v) { get(QM_FIELD_V, v); }
    void getForces(std::vector<double>& f) { get(QM_FIELD_F, f); }
    void setPositions(const std::vector<double>& q) { set(QM_FIELD_Q, q); }
    void setVelocities(const std::vector<double>& v) { set(QM_FIELD_V, v); }
    void getTypes(std::vector<int>& t) {
        t.resize(N);
        qmin_host::check(qm_lattice_get_types(lat_, t.data()), "getTypes");
    }
    // FP32 <-> FP64 storage (mixed form and FP64 polish); no reference counterpart
    void setPrecision(int precision) { qmin_host::check(qm_lattice_set_precision(lat_, precision), "setPrecision"); }
    // cubicLattice::displaceBoundaryObject (cubicLattice.cpp:329-372); motionDirection 0..5 = x-, x+, y-, y+, z-, z+
    void displaceBoundaryObject(int objectIndex, int motionDirection, int magnitude) {
        qmin_host::check(qm_lattice_displace_object(lat_, objectIndex, motionDirection, magnitude),
                         "displaceBoundaryObject");
    }
    // cubicLattice::boundarySites / surfaceSites of one object
    std::vector<int64_t> objectSites(int objectIndex, bool surface) {
        int64_t count = 0;
        qmin_host::check(qm_lattice_object_sites(lat_, objectIndex, surface ? 1 : 0, nullptr, &count), "objectSites");
        std::vector<int64_t> v((size_t)count);
        qmin_host::check(qm_lattice_object_sites(lat_, objectIndex, surface ? 1 : 0, v.data(), &count), "objectSites");
        return v;
    }

    struct {
        int x, y, z;
    } latticeSites;
    std::vector<scalar> defectMeasures;
    // cubicLattice::boundaryForce (cubicLattice.h:85-86): per object, the forces
    // landauDeGennesLC::computeObjectForces has added
    std::vector<scalar3> boundaryForce;

   protected:
    void get(int which, std::vector<double>& a) {
        a.resize(5 * (size_t)N);
        qmin_host::check(qm_lattice_get_field(lat_, which, a.data()), "qm_lattice_get_field");
    }
    void set(int which, const std::vector<double>& a) {
        if (a.size() != 5 * (size_t)N) throw qminError(QM_EINVAL, "field size must be 5 N");
        qmin_host::check(qm_lattice_set_field(lat_, which, a.data()), "qm_lattice_set_field");
    }
    qm_lattice* lat_ = nullptr;
    int N = 0;
};
using cubicLattice = qTensorLatticeModel;

// multirankQTensorLatticeModel(lx, ly, lz, xHalo, yHalo, zHalo, useGPU, neverGPU)
// (multirankQTensorLatticeModel.cpp:8-50) on one rank: no halos.
class multirankQTensorLatticeModel : public qTensorLatticeModel {
   public:
    multirankQTensorLatticeModel(int lx, int ly, int lz, bool xHalo = false, bool yHalo = false, bool zHalo = false,
                                 bool useGPU = false, bool neverGPU = false, int precision = QM_FP64)
        : qTensorLatticeModel(lx, ly, lz, precision) {
        (void)useGPU;
        (void)neverGPU;
        if (xHalo || yHalo || zHalo) throw qminError(QM_EINVAL, "halo sites (more than one rank) are not provided");
    }
};

// ---------------------------------------------------------------- force
class landauDeGennesLC {
   public:
    landauDeGennesLC(bool neverGPU = false) {
        (void)neverGPU;
        qmin_host::check(qm_ldg_create(&f_), "qm_ldg_create");
        setPhaseConstants();
        setElasticConstants();
        setNumberOfConstants(distortionEnergyType::oneConstant);
    }
    landauDeGennesLC(scalar A, scalar B, scalar C, scalar L1) : landauDeGennesLC() {
        setPhaseConstants(A, B, C);
        setElasticConstants(L1);
    }
    landauDeGennesLC(scalar A, scalar B, scalar C, scalar L1, scalar L2, scalar L3, scalar L4, scalar L6)
        : landauDeGennesLC() {
        setPhaseConstants(A, B, C);
        setElasticConstants(L1, L2, L3, L4, L6);
        setNumberOfConstants(distortionEnergyType::multiConstant);
    }
    ~landauDeGennesLC() { qm_ldg_destroy(f_); }
    landauDeGennesLC(const landauDeGennesLC&) = delete;
    landauDeGennesLC& operator=(const landauDeGennesLC&) = delete;

    // landauDeGennesLC.h:34-36 (same defaults)
    void setPhaseConstants(scalar a = -1, scalar b = -12.325581395, scalar c = 10.058139535) {
        A = a;
        B = b;
        C = c;
        qmin_host::check(qm_ldg_set_phase(f_, A, B, C), "setPhaseConstants");
    }
    void setElasticConstants(scalar l1 = 2.32, scalar l2 = 0, scalar l3 = 0, scalar l4 = 0, scalar l6 = 0) {
        L1 = l1;
        L2 = l2;
        L3 = l3;
        L4 = l4;
        L6 = l6;
        push_elastic();
    }
    void setNumberOfConstants(distortionEnergyType t) {
        numberOfConstants = t;
        push_elastic();
    }
    // landauDeGennesLC.h:76-91
    void setEField(scalar3 E, scalar eps, scalar eps0, scalar deltaEps) {
        qmin_host::check(qm_ldg_set_efield(f_, 1, E.x, E.y, E.z, eps, eps0, deltaEps), "setEField");
    }
    void setHField(scalar3 H, scalar chi, scalar mu0, scalar deltaChi) {
        qmin_host::check(qm_ldg_set_hfield(f_, 1, H.x, H.y, H.z, chi, mu0, deltaChi), "setHField");
    }
    // landauDeGennesLC::setSpatiallyVaryingField (landauDeGennesLC.cpp:81-123):
    // reads fname_x0y0z0.txt (one rank) into the model's field
    void setSpatiallyVaryingField(std::string fname, scalar chi, scalar mu0, scalar deltaChi) {
        const std::string fn = fname + "_x0y0z0.txt";
        qmin_host::check(qm_lattice_load_hfield(lattice(), fn.c_str()), "setSpatiallyVaryingField");
        qmin_host::check(qm_ldg_set_varying_hfield(f_, 1, chi, mu0, deltaChi), "setSpatiallyVaryingField");
    }
    // the reference's signature (landauDeGennesLC.h:93); rank (0,0,0) on one rank
    void setSpatiallyVaryingField(std::string fname, scalar chi, scalar mu0, scalar deltaChi, int3 rankParity) {
        (void)rankParity;
        setSpatiallyVaryingField(fname, chi, mu0, deltaChi);
    }
    // QM_SEMANTICS_REF_CPU or QM_SEMANTICS_REF_GPU
    void setSemantics(int s) { qmin_host::check(qm_ldg_set_semantics(f_, s), "setSemantics"); }
    void setModel(std::shared_ptr<cubicLattice> m) { model = m; }

    // force::computeForces into the model's force field
    void computeForces() {
        qmin_host::check(qm_ldg_compute_forces(f_, lattice()), "computeForces");
    }
    // force::computeForces(GPUArray<dVec>&, zeroOutForce, type) (baseForce.h:23,
    // landauDeGennesLC.cpp:125-135): the forces of every site into `forces`
    // (replacing its contents, or added to them), metric-corrected.  type 1 (the
    // CPU path's boundary pass, which adds the boundary sites' terms) adds
    // nothing here: the type-0 call has computed every site.
    void computeForces(GPUArray<dVec>& forces, bool zeroOutForce = true, int type = 0) {
        if (type != 0) return;
        const bool ownField = forces.bound() && forces.lattice() == lattice() && forces.field() == QM_FIELD_F;
        const size_t n = (size_t)model->getNumberOfParticles();
        std::vector<dVec> before;
        if (!zeroOutForce) {
            ArrayHandle<dVec> h(forces, access_location::host, access_mode::read);
            before.assign(h.data, h.data + n);
        }
        computeForces();
        if (ownField && zeroOutForce) return;
        std::vector<double> f(5 * n);
        qmin_host::check(qm_lattice_get_field(lattice(), QM_FIELD_F, f.data()), "computeForces");
        if (!forces.bound() && forces.getNumElements() < n) forces.resize((unsigned int)n);
        ArrayHandle<dVec> h(forces, access_location::host, access_mode::overwrite);
        for (size_t i = 0; i < n; ++i)
            for (int k = 0; k < 5; ++k) h.data[i][k] = (zeroOutForce ? 0.0 : before[i][k]) + f[5 * i + k];
    }
    // landauDeGennesLC::getClassSize (landauDeGennesLC.h:102-106): GB of data the
    // force holds (its energy components and density on the host)
    scalar getClassSize() { return 1e-9 * sizeof(scalar) * (energyComponents.size() + energyDensity.size() + 20); }
    // landauDeGennesLC::computeEnergy: the total (not per site); fills
    // energyComponents {phase, distortion, anchoring, E, H} and, as the
    // reference does, the per-site energyDensity (n doubles copied to the host
    // on every call)
    scalar computeEnergy(bool verbose = false) {
        double total = 0;
        energyComponents.assign(5, 0.0);
        qmin_host::check(qm_ldg_compute_energy(f_, lattice(), &total, energyComponents.data()), "computeEnergy");
        energyDensity.assign((size_t)model->getNumberOfParticles(), 0.0);
        qmin_host::check(qm_ldg_energy_density(f_, lattice(), energyDensity.data()), "computeEnergy");
        if (verbose)
            std::printf("energy: phase %g distortion %g anchoring %g E %g H %g\n", energyComponents[0],
                        energyComponents[1], energyComponents[2], energyComponents[3], energyComponents[4]);
        return total;
    }

    // landauDeGennesLC::computeObjectForces (landauDeGennesLCOtherForces.cpp:147-208):
    // adds the stress-tensor force on the object to the model's boundaryForce and
    // prints the running total, as the reference does; one-constant only
    void computeObjectForces(int objectIdx) {
        double f[3];
        qmin_host::check(qm_ldg_object_force(f_, lattice(), objectIdx, f), "computeObjectForces");
        if ((int)model->boundaryForce.size() <= objectIdx) model->boundaryForce.resize(objectIdx + 1, scalar3{0, 0, 0});
        scalar3& b = model->boundaryForce[objectIdx];
        b = scalar3{b.x + f[0], b.y + f[1], b.z + f[2]};
        std::printf("%f\t%f\t%f\n", b.x, b.y, b.z);
    }

    qm_ldg* handle() { return f_; }
    qm_lattice* lattice() {
        if (!model) throw qminError(QM_ESTATE, "landauDeGennesLC: setModel was not called");
        return model->qmLattice();
    }

    std::vector<scalar> energyComponents;
    //! the free energy density at each lattice site (landauDeGennesLC.h:96), site order
    std::vector<scalar> energyDensity;
    scalar A = -1, B = 0, C = 0, L1 = 2.32, L2 = 0, L3 = 0, L4 = 0, L6 = 0;
    distortionEnergyType numberOfConstants = distortionEnergyType::oneConstant;

   private:
    void push_elastic() {
        // the reference stores all constants and lets setNumberOfConstants pick the path
        if (numberOfConstants == distortionEnergyType::multiConstant)
            qmin_host::check(qm_ldg_set_elastic(f_, L1, L2, L3, L4, L6, 1), "setElasticConstants");
        else
            qmin_host::check(qm_ldg_set_elastic(f_, L1, 0, 0, 0, 0, 0), "setElasticConstants");
    }
    qm_ldg* f_ = nullptr;
    std::shared_ptr<cubicLattice> model;
};

class multirankSimulation;

// ---------------------------------------------------------------- updaters
// updater (src/updaters/baseUpdater.h:18-115)
class updater {
   public:
    virtual ~updater() = default;
    void setSimulation(multirankSimulation* s) { sim = s; }
    virtual void setModel(std::shared_ptr<simpleModel> m) { model = m; }
    void setMaximumIterations(int maxIt) { maxIterations = maxIt; }
    virtual scalar getMaxForce() { return 0.0; }
    virtual void performUpdate() = 0;
    // updater::getClassSize (baseUpdater.h:82-85): GB of data the updater holds
    // on the host (its device history belongs to the lattice's allocation)
    virtual scalar getClassSize() { return 1e-9 * (6 * sizeof(int) + 2 * sizeof(bool) + sizeof(scalar)); }
    void Update(int) { performUpdate(); }
    int iterations = 0;
    int maxIterations = 1000;

   protected:
    qm_ldg* force();
    qm_lattice* lattice() {
        if (!model) throw qminError(QM_ESTATE, "updater: no model");
        return model->qmLattice();
    }
    multirankSimulation* sim = nullptr;
    std::shared_ptr<simpleModel> model;
};

// energyMinimizerFIRE (src/updaters/energyMinimizerFIRE.h:20-110)
class energyMinimizerFIRE : public updater {
   public:
    energyMinimizerFIRE() { qmin_host::check(qm_fire_create(&m_), "qm_fire_create"); }
    explicit energyMinimizerFIRE(std::shared_ptr<simpleModel> system) : energyMinimizerFIRE() { setModel(system); }
    ~energyMinimizerFIRE() override { qm_fire_destroy(m_); }

    void setFIREParameters(scalar deltaT, scalar alphaStart, scalar deltaTMax, scalar deltaTInc, scalar deltaTDec,
                           scalar alphaDec, int nMin, scalar forceCutoff, scalar alphaMin = 0.75) {
        qmin_host::check(qm_fire_set_parameters(m_, deltaT, alphaStart, deltaTMax, deltaTInc, deltaTDec, alphaDec,
                                                nMin, forceCutoff, alphaMin),
                         "setFIREParameters");
    }
    // energyMinimizerFIRE::setForceCutoff (energyMinimizerFIRE.h:29)
    void setForceCutoff(scalar fc) { qmin_host::check(qm_fire_set_force_cutoff(m_, fc), "setForceCutoff"); }
    void minimize() {
        qmin_host::check(qm_fire_set_max_iterations(m_, maxIterations), "setMaximumIterations");
        qmin_host::check(qm_fire_minimize(m_, force(), lattice()), "energyMinimizerFIRE::minimize");
        qm_fire_state s;
        qm_fire_get_state(m_, &s);
        iterations = s.iterations;
        forceMax = s.forceMax;
    }
    void performUpdate() override { minimize(); }
    scalar getMaxForce() override { return forceMax; }
    qm_fire_state state() const {
        qm_fire_state s;
        qm_fire_get_state(m_, &s);
        return s;
    }

   private:
    qm_fire* m_ = nullptr;
    scalar forceMax = 100.0;
};

// energyMinimizerAdam (src/updaters/energyMinimizerAdam.h:7-50)
class energyMinimizerAdam : public updater {
   public:
    energyMinimizerAdam() {
        qmin_host::check(qm_adam_create(&m_), "qm_adam_create");
        setAdamParameters();
    }
    explicit energyMinimizerAdam(std::shared_ptr<simpleModel> system) : energyMinimizerAdam() { setModel(system); }
    ~energyMinimizerAdam() override { qm_adam_destroy(m_); }
    void setAdamParameters(scalar b1 = 0.9, scalar b2 = 0.99, scalar eps = 0.00000001, scalar dt = 0.0001,
                           scalar fc = 1e-12) {
        qmin_host::check(qm_adam_set_parameters(m_, b1, b2, eps, dt, fc), "setAdamParameters");
    }
    void minimize() {
        qmin_host::check(qm_adam_set_max_iterations(m_, maxIterations), "setMaximumIterations");
        qmin_host::check(qm_adam_minimize(m_, force(), lattice()), "energyMinimizerAdam::minimize");
        qm_adam_get_state(m_, &iterations, &forceMax);
    }
    void performUpdate() override { minimize(); }
    scalar getMaxForce() override { return forceMax; }

   private:
    qm_adam* m_ = nullptr;
    scalar forceMax = 110.0;
};

// energyMinimizerNesterovAG (src/updaters/energyMinimizerNesterovAG.h:8-61)
class energyMinimizerNesterovAG : public updater {
   public:
    energyMinimizerNesterovAG() {
        qmin_host::check(qm_nesterov_create(&m_), "qm_nesterov_create");
        setNesterovAGParameters();
    }
    explicit energyMinimizerNesterovAG(std::shared_ptr<simpleModel> system) : energyMinimizerNesterovAG() {
        setModel(system);
    }
    ~energyMinimizerNesterovAG() override { qm_nesterov_destroy(m_); }
    void setNesterovAGParameters(scalar dt = 0.0001, scalar mu = 0.01, scalar fc = 1e-12) {
        dt_ = dt;
        mu_ = mu;
        fc_ = fc;
        push();
    }
    void minimize() {
        push();
        qmin_host::check(qm_nesterov_set_max_iterations(m_, maxIterations), "setMaximumIterations");
        qmin_host::check(qm_nesterov_minimize(m_, force(), lattice()), "energyMinimizerNesterovAG::minimize");
        qm_nesterov_get_state(m_, &iterations, &forceMax);
    }
    void performUpdate() override { minimize(); }
    scalar getMaxForce() override { return forceMax; }
    bool scheduledMomentum = false;

   private:
    void push() {
        qmin_host::check(qm_nesterov_set_parameters(m_, dt_, mu_, fc_, scheduledMomentum ? 1 : 0),
                         "setNesterovAGParameters");
    }
    qm_nesterov* m_ = nullptr;
    scalar dt_ = 1e-4, mu_ = 0.01, fc_ = 1e-12, forceMax = 110.0;
};

// energyMinimizerLoLBFGS (src/updaters/energyMinimizerLoLBFGS.h:13-110)
class energyMinimizerLoLBFGS : public updater {
   public:
    energyMinimizerLoLBFGS() {
        qmin_host::check(qm_lolbfgs_create(&m_), "qm_lolbfgs_create");
        setLoLBFGSParameters();
    }
    explicit energyMinimizerLoLBFGS(std::shared_ptr<simpleModel> system) : energyMinimizerLoLBFGS() {
        setModel(system);
    }
    ~energyMinimizerLoLBFGS() override { qm_lolbfgs_destroy(m_); }
    void setLoLBFGSParameters(int m = 5, scalar dt = 0.0001, scalar c = 1.00, scalar fc = 1e-12, scalar tau = 10) {
        mem_ = m;
        dt_ = dt;
        c_ = c;
        fc_ = fc;
        tau_ = tau;
        push();
    }
    void minimize() {
        push();
        qmin_host::check(qm_lolbfgs_set_max_iterations(m_, maxIterations), "setMaximumIterations");
        qmin_host::check(qm_lolbfgs_minimize(m_, force(), lattice()), "energyMinimizerLoLBFGS::minimize");
        qm_lolbfgs_get_state(m_, &iterations, &forceMax);
    }
    void performUpdate() override { minimize(); }
    scalar getMaxForce() override { return forceMax; }
    bool scheduledMomentum = false;

   private:
    void push() {
        qmin_host::check(qm_lolbfgs_set_parameters(m_, mem_, dt_, c_, fc_, tau_, scheduledMomentum ? 1 : 0),
                         "setLoLBFGSParameters");
    }
    qm_lolbfgs* m_ = nullptr;
    int mem_ = 5;
    scalar dt_ = 1e-4, c_ = 1.0, fc_ = 1e-12, tau_ = 10, forceMax = 100.0;
};

// ---------------------------------------------------------------- simulation
// multirankSimulation (src/simulation/multirankSimulation.h:13-160) on one rank.
class multirankSimulation {
   public:
    multirankSimulation(int myRank = 0, int xDiv = 1, int yDiv = 1, int zDiv = 1, bool edges = false,
                        bool corners = false) {
        (void)edges;
        (void)corners;
        if (myRank != 0 || xDiv != 1 || yDiv != 1 || zDiv != 1)
            throw qminError(QM_EINVAL, "one rank only: the lattice is one domain on one GPU");
    }
    void setConfiguration(std::shared_ptr<multirankQTensorLatticeModel> c) { config = c; }
    void addForce(std::shared_ptr<landauDeGennesLC> f) {
        force = f;
        force->setSemantics(cpuSemantics ? QM_SEMANTICS_REF_CPU : QM_SEMANTICS_REF_GPU);
    }
    void addForce(std::shared_ptr<landauDeGennesLC> f, std::shared_ptr<multirankQTensorLatticeModel> c) {
        f->setModel(c);
        addForce(f);
    }
    void addUpdater(std::shared_ptr<updater> u) {
        u->setSimulation(this);
        updaters.push_back(u);
    }
    void addUpdater(std::shared_ptr<updater> u, std::shared_ptr<multirankQTensorLatticeModel> c) {
        u->setModel(c);
        addUpdater(u);
    }
    void clearUpdaters() { updaters.clear(); }
    // the reference CPU path corrects bulk forces with the metric twice; the
    // GPU path once (multirankSimulation.cpp:452-473)
    void setCPUOperation(bool cpu) {
        cpuSemantics = cpu;
        if (force) force->setSemantics(cpu ? QM_SEMANTICS_REF_CPU : QM_SEMANTICS_REF_GPU);
    }
    void setReproducible(bool) {}

    // multirankSimulationBoundaries.cpp:101-215
    void createWall(int xyz, int plane, boundaryObject& b) {
        qmin_host::check(qm_lattice_wall(lat(), xyz, plane, qmin_host::btype(b.boundary), b.P1, b.P2, nullptr),
                         "createWall");
    }
    void createSphericalColloid(scalar3 center, scalar radius, boundaryObject& b) {
        qmin_host::check(qm_lattice_sphere(lat(), center.x, center.y, center.z, radius, qmin_host::btype(b.boundary),
                                           b.P1, b.P2, nullptr),
                         "createSphericalColloid");
    }
    void createBoundaryFromFile(std::string fname, bool verbose = false) {
        (void)verbose;
        qmin_host::check(qm_lattice_boundary_file(lat(), fname.c_str()), "createBoundaryFromFile");
    }
    void finalizeObjects() {
        int64_t n = 0;
        qmin_host::check(qm_lattice_n_active(lat(), &n), "finalizeObjects");
        NActive = n;
    }

    void performTimestep() {
        integerTimestep += 1;
        for (auto& u : updaters) u->Update(integerTimestep);
    }
    void computeForces() {
        if (!force) throw qminError(QM_ESTATE, "multirankSimulation: no force computer");
        force->computeForces();
    }
    // multirankSimulation.cpp:475-490: energy per active site
    scalar computePotentialEnergy(bool verbose = false) {
        if (!force) throw qminError(QM_ESTATE, "multirankSimulation: no force computer");
        finalizeObjects();
        const scalar e = force->computeEnergy(verbose);
        return NActive != 0 ? e / (1.0 * NActive) : e;
    }
    scalar getMaxForce() { return updaters.empty() ? 0.0 : updaters.back()->getMaxForce(); }

    // multirankSimulation.cpp:573-665, rank (0,0,0) file name
    void saveState(std::string fname, int latticeSkip = 1, int defectType = 0) {
        const std::string fn = fname + "_x0y0z0.txt";
        qmin_host::check(qm_lattice_save_state_ex(lat(), fn.c_str(), latticeSkip, defectType), "saveState");
    }
    void loadState(std::string fname) {
        const std::string fn = fname + "_x0y0z0.txt";
        qmin_host::check(qm_lattice_load_state(lat(), fn.c_str()), "loadState");
    }

    qm_ldg* forceHandle() {
        if (!force) throw qminError(QM_ESTATE, "multirankSimulation: no force computer");
        return force->handle();
    }

    int integerTimestep = 0;
    long long NActive = 0;
    // multirankSimulation.h:57,131: communication time (none on one rank) and
    // the rank's parity in the rank grid
    profiler p1 = profiler("total communication time");
    int3 rankParity{0, 0, 0};

   private:
    qm_lattice* lat() {
        if (!config) throw qminError(QM_ESTATE, "multirankSimulation: setConfiguration was not called");
        return config->qmLattice();
    }
    std::shared_ptr<multirankQTensorLatticeModel> config;
    std::shared_ptr<landauDeGennesLC> force;
    std::vector<std::shared_ptr<updater>> updaters;
    bool cpuSemantics = false;
};

inline qm_ldg* updater::force() {
    if (!sim) throw qminError(QM_ESTATE, "updater: add it to a simulation first");
    return sim->forceHandle();
}
