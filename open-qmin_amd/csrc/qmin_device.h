// Shared host/device definitions of the qmin HIP library (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qmin {

// Work-group of the fused pass: one thread per site of a TX x TY tile with a
// one-site halo in x and y; the work-group marches along z through a chunk.
// Four tile shapes are compiled: 32 x 8 and 64 x 4 (256 threads; wider rows use
// whole cache lines for the side halo but have more halo sites per tile), and
// 32 x 16 and 64 x 8 (512 threads, 73-79 KB of LDS in FP64: fewer halo sites per
// site; with 64 x 8 each wave covers one contiguous 64-site row).
// Tiles 4..6 are those of the FIRE pass that keeps only Q and v between
// iterations (fire_stream.hip): TX x TY owned sites, one thread per site of the
// tile plus a two-site margin.
constexpr int kThreads = 256;   // work-group size of the streaming kernels
enum TileId {
    kTile32x8 = 0, kTile64x4 = 1, kTile32x16 = 2, kTile64x8 = 3, kNumMarchTiles = 4,
    kStream32x16 = 4, kStream32x8 = 5, kStream64x8 = 6, kNumTiles = 7
};
constexpr int kTileTX[kNumTiles] = {32, 64, 32, 64, 32, 32, 64};
constexpr int kTileTY[kNumTiles] = {8, 4, 16, 8, 16, 8, 8};
constexpr int kTileThreads[kNumMarchTiles] = {32 * 8, 64 * 4, 32 * 16, 64 * 8};
__host__ __device__ constexpr bool is_stream_tile(int tile) { return tile >= kNumMarchTiles; }

template <int TX_, int TY_>
struct Tile {
    static constexpr int TX = TX_, TY = TY_;
    static constexpr int Threads = TX * TY;     // one site per thread
    static constexpr int HX = TX + 2, HY = TY + 2;
    static constexpr int HP = HX * HY;          // sites of one halo-padded plane
    static constexpr int Halo = HP - TX * TY;   // perimeter sites of one plane
    static_assert(Threads % 64 == 0, "whole wave64s");
    static_assert(Halo <= Threads, "one halo site per thread at most");
};

using SiteType = int16_t;                // site types on the device (reference encoding)
constexpr int kMaxObjects = 32767;      // object ids a SiteType holds
constexpr int kBoundaryThreads = 128;
constexpr int kReduceThreads = 256;
constexpr int kMaxEnergyBlocks = 4096;
constexpr int kFastWeights = 160;  // >= entries of the composed L1..L4 operator

// kAdamStep, kNesterovStep: the force of every site, then that minimizer's
// update of the site, written to the other state set (one pass per iteration)
enum Mode { kForce = 0, kFire = 1, kAdamStep = 2, kNesterovStep = 3 };
// kMultiLinear: L1..L3, kMultiL4: L1..L4, kMultiL6: L1..L4 and L6
enum Terms { kOneConstant = 0, kMultiLinear = 1, kMultiL4 = 2, kMultiL6 = 3 };

struct DevObject {
    int type;  // 0 homeotropic, 1 degenerate planar
    int pad;
    double P1, P2;
};

struct ForceConst {
    double a, b, c;        // A/2, B/3, C/4 of the phase energy
    double L1, L2, L3, L4, L6;
    double field[5];       // uniform E and H field force, added to the raw force
    int hasField;
    int semantics;         // 0 reference CPU two-pass, 1 reference GPU single pass
};

// Device-resident control block of one lattice: which state set is current and
// the FIRE scalars, advanced by the last work-group of each iteration.
struct Ctrl {
    int cur;
    int done;
    int iterations;
    int maxIter;
    int nPositive;
    int reset;      // pending velocity reset (v = 0) from the last FIRE step
    int identity;   // no pending velocity mix (v = u)
    int nanFlag;
    unsigned int arrive;
    int traceLen;
    int traceCap;
    int pad;
    double dt, alpha;
    double cKeep, cForce;   // pending mix v = cKeep u + cForce F
    double slot[40];        // minimizer scalars finished on the device (L-BFGS s.y, a_i, coefficients)
    double forceMax, power, fnorm, vnorm, scaling;
};

struct FireConst {
    double alphaStart, dtMax, dtInc, dtDec, alphaDec, alphaMin, dtMin, fc, nTotal;
    int nMin;
    int pad;
};

struct EnergyConst {
    double a, b, c;
    double L[5];  // L1, L2, L3, L4, L6
    int eOn, hOn;
    double E[3], eps, eps0, deps;
    double H[3], chi, mu0, dchi;
    const double* hvar;   // spatially varying H, planes hx, hy, hz of n (null: none)
};

// The Adam or Nesterov update that a pass in mode kAdamStep / kNesterovStep
// applies to each site once its force is known.
struct StepArgs {
    void* m;             // Adam: first moment; Nesterov: the alternate sequence (5 planes of n, lattice precision)
    void* v;             // Adam: second moment
    const double* scal;  // device scalars: Adam beta1^t, beta2^t; Nesterov lambda, mu
    double dt, b1, b2, eps;
    int scheduled;       // Nesterov: momentum 1 + gamma from lambda
    int pad;
};

template <typename R>
struct StateSet {
    R* q;  // 5 planes of n
    R* u;
    R* f;
};

template <typename R>
struct MarchArgs {
    StateSet<R> set[2];
    const int8_t* cls;      // per site: 0 composed bulk formula, > 0 object, < 0 finished by the boundary pass
    int gate;               // force-only pass: skip once the control block reports done
    int lx, ly, lz, zchunk;
    long long n;
    int nbx, nby, nblocks;
    ForceConst fc;
    Ctrl* ctrl;
    double* partials;   // [nblocks + boundary blocks][3]
    const R* W;         // composed bulk operator weights (device buffer, read with scalar loads)
    const double* hvar; // spatially varying H, planes hx, hy, hz of n (null: none)
    double hvarP;       // deltaChi mu0
    int storeF;         // FIRE pass that keeps only Q and v: store F of the sites it finishes as well
    StepArgs step;      // modes kAdamStep, kNesterovStep
};

template <typename R>
struct BoundaryArgs {
    StateSet<R> set[2];
    const SiteType* types;
    const DevObject* objects;
    const int32_t* sites;   // compacted list of sites finished here
    int gate;
    long long nsites;
    int lx, ly, lz;
    long long n;
    int marchBlocks;        // partial slots used by the fused pass
    ForceConst fc;
    FireConst fire;
    Ctrl* ctrl;
    double* partials;
    double* trace;          // [traceCap][6]
    const double* hvar;
    double hvarP;
    StepArgs step;          // modes kAdamStep, kNesterovStep
};

// Per-lattice arguments of the batched FIRE scalar step.
struct FinalizeDesc {
    Ctrl* ctrl;
    FireConst fire;
    const double* partials;
    long long total;
};

struct EnsembleStatus {
    int done;
    int iterations;
    double forceMax;
};

}  // namespace qmin
