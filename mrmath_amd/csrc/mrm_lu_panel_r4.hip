// mrm_lu_panel_r4.hip -- the LU panel kernel with 4 rows per thread.
#include "mrm_lu_panel_kernel.h"

namespace mrm {
template <>
const void* lu_panel_fn<4>() {
    return reinterpret_cast<const void*>(&lu_panel_kernel<double, LU_W, 4>);
}
}  // namespace mrm
