// htx_strict_fvh.hip -- the FP64 drop-in build of the Hermite-TENO stage
// kernels (htx_fvh2d.hip): no FMA contraction, every division of the
// reference kept as a division, as the TENO5 kernels of htx_strict.hip.  A
// translation unit of its own so that it can be compiled with machine LICM
// off (constants rematerialised at their uses instead of hoisted into scalar
// registers that spill to vector lanes across the long Gauss-point loop),
// which the TENO5 kernels do not want.
#define HTX_NS htx_strict_fvh
#define HTX_EXACT_DIVISION 1
#include "htx_fvh2d.hip"

// entry point for htx_api.hip and htx_batch.hip (the argument struct has the
// same layout in every build)
hipError_t htx_strict_launch_fvh2d_stage(const void* a, int flux, hipStream_t s) {
    return htx_strict_fvh::launch_fvh2d_stage(*static_cast<const htx_strict_fvh::FvhArgs*>(a), flux, s);
}
