// mrm_batched_lu_block_s128.hip -- the eight-wave small solve for up to 128
// unknowns, FP32 factors with FP64 refinement.
#include "mrm_batched_lu_block.h"

template <>
const void* mrm::batched_lu_block_kernel<1, 128>() {
    return reinterpret_cast<const void*>(&blk::bl_block_kernel<1, 128>);
}
