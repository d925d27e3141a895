// mrm_batched_lu_block_d128.hip -- the eight-wave small solve for up to 128
// unknowns, FP64 factors.
#include "mrm_batched_lu_block.h"

template <>
const void* mrm::batched_lu_block_kernel<0, 128>() {
    return reinterpret_cast<const void*>(&blk::bl_block_kernel<0, 128>);
}
