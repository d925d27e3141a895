// mrm_batched_lu_block_d96.hip -- the eight-wave small solve for up to 96
// unknowns, FP64 factors.
#include "mrm_batched_lu_block.h"

template <>
const void* mrm::batched_lu_block_kernel<0, 96>() {
    return reinterpret_cast<const void*>(&blk::bl_block_kernel<0, 96>);
}
