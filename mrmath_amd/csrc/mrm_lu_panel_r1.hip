// mrm_lu_panel_r1.hip -- the LU panel kernel with 1 row per thread.
#include "mrm_lu_panel_kernel.h"

namespace mrm {
template <>
const void* lu_panel_fn<1>() {
    return reinterpret_cast<const void*>(&lu_panel_kernel<double, LU_W, 1>);
}
}  // namespace mrm
