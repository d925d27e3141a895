// mrm_batched_lu_block_s96.hip -- the eight-wave small solve for up to 96
// unknowns, FP32 factors with FP64 refinement.
#include "mrm_batched_lu_block.h"

template <>
const void* mrm::batched_lu_block_kernel<1, 96>() {
    return reinterpret_cast<const void*>(&blk::bl_block_kernel<1, 96>);
}
