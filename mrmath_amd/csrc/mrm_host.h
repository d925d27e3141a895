// mrm_host.h -- host-side helpers shared by the C-ABI translation units.
#pragma once

#include <mutex>

#include "mrm_common.h"

// Staging slots for host-pointer entries: 0 = A, 1 = B, 2 = C/X, 3.. = extra.
constexpr int MRM_NUM_SLOTS = 8;

int mrm_ensure_init();
hipStream_t mrm_stream();
hipStream_t mrm_resolve_stream(void* s);
std::mutex& mrm_host_mutex();
// Grow-only device buffer for `slot`; callers hold mrm_host_mutex().
int mrm_staging(int slot, size_t bytes, void** out);
// Host <-> device copies of `rows` rows of `width` bytes with byte pitches,
// enqueued on `s` (pageable host memory: one 1-D transfer of the host span,
// packed or unpacked on the device, instead of a row-by-row 2-D copy).
int mrm_copy_h2d(void* dst, size_t dpitch, const void* src, size_t spitch, size_t width, size_t rows,
                 hipStream_t s);
int mrm_copy_d2h(void* dst, size_t dpitch, const void* src, size_t spitch, size_t width, size_t rows,
                 hipStream_t s);

namespace mrm {
// Free the library-owned workspaces of the LU and the refinement; callers
// hold mrm_host_mutex() and no work using them is in flight.
void lu_release_workspace();
void ir_release_workspace();
void gmres_release_workspace();
}  // namespace mrm
