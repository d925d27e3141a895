// mrm_lu_panel_r2.hip -- the LU panel kernel with 2 rows per thread.
#include "mrm_lu_panel_kernel.h"

namespace mrm {
template <>
const void* lu_panel_fn<2>() {
    return reinterpret_cast<const void*>(&lu_panel_kernel<double, LU_W, 2>);
}
}  // namespace mrm
