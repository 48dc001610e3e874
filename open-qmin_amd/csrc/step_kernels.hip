// Lattice passes of the Adam and Nesterov iterations: the fused and boundary
// passes in modes kAdamStep / kNesterovStep compute each site's force and then
// apply the minimizer's update to that site (energyMinimizerAdam.cpp:25-54,
// energyMinimizerNesterovAG.cpp:75-92), writing the new Q and the force to the
// other state set and the work-group partial sums of |F|^2 to the lattice's
// partials.  One pass per iteration: the update needs only the site's own
// force and its own entries of the minimizer's vectors, so it rides on the
// store of the force instead of a second pass over Q, F and those vectors.
#include "lattice_pass.h"

namespace qmin {
namespace {

template <typename R, int MODE, int TERMS>
const void* step_fn_tile(int tile) {
    switch (tile) {
        case kTile64x4: return reinterpret_cast<const void*>(&march_kernel<R, MODE, TERMS, kTile64x4>);
        case kTile32x16: return reinterpret_cast<const void*>(&march_wide_kernel<R, MODE, TERMS, kTile32x16>);
        case kTile64x8: return reinterpret_cast<const void*>(&march_wide_kernel<R, MODE, TERMS, kTile64x8>);
        default: return reinterpret_cast<const void*>(&march_kernel<R, MODE, TERMS, kTile32x8>);
    }
}

template <typename R, int MODE>
const void* step_fn_terms(int terms, int tile) {
    switch (terms) {
        case kOneConstant: return step_fn_tile<R, MODE, kOneConstant>(tile);
        case kMultiLinear: return step_fn_tile<R, MODE, kMultiLinear>(tile);
        case kMultiL4: return step_fn_tile<R, MODE, kMultiL4>(tile);
        default: return step_fn_tile<R, MODE, kMultiL6>(tile);
    }
}

// the boundary pass does not distinguish L1..L3 from L1..L4
template <typename R, int MODE>
const void* step_boundary_fn(int terms) {
    if (terms == kOneConstant) return reinterpret_cast<const void*>(&boundary_kernel<R, MODE, kOneConstant>);
    if (terms == kMultiL6) return reinterpret_cast<const void*>(&boundary_kernel<R, MODE, kMultiL6>);
    return reinterpret_cast<const void*>(&boundary_kernel<R, MODE, kMultiLinear>);
}

}  // namespace

// Queues the fused pass (tile `tile`, one of the four halo-padded tiles) and
// the boundary pass of one Adam or Nesterov iteration.
template <typename R>
hipError_t launch_step(const MarchArgs<R>& ma, const BoundaryArgs<R>& ba, int mode, int terms, int tile,
                       hipStream_t s) {
    if (tile < 0 || tile >= kNumMarchTiles || !is_step_mode(mode)) return hipErrorInvalidValue;
    const void* fn = mode == kAdamStep ? step_fn_terms<R, kAdamStep>(terms, tile)
                                       : step_fn_terms<R, kNesterovStep>(terms, tile);
    void* margs[] = {const_cast<MarchArgs<R>*>(&ma)};
    hipError_t e = hipLaunchKernel(fn, dim3(ma.nblocks), dim3(kTileThreads[tile]), margs, 0, s);
    if (e != hipSuccess) return e;
    long long blocks = (ba.nsites + kBoundaryThreads - 1) / kBoundaryThreads;
    if (blocks < 1) blocks = 1;
    const void* bfn = mode == kAdamStep ? step_boundary_fn<R, kAdamStep>(terms) : step_boundary_fn<R, kNesterovStep>(terms);
    void* bargs[] = {const_cast<BoundaryArgs<R>*>(&ba)};
    e = hipLaunchKernel(bfn, dim3((unsigned)blocks), dim3(kBoundaryThreads), bargs, 0, s);
    if (e != hipSuccess) return e;
    return hipGetLastError();
}

template hipError_t launch_step<double>(const MarchArgs<double>&, const BoundaryArgs<double>&, int, int, int,
                                        hipStream_t);
template hipError_t launch_step<float>(const MarchArgs<float>&, const BoundaryArgs<float>&, int, int, int,
                                       hipStream_t);

}  // namespace qmin
