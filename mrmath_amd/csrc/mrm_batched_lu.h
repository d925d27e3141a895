// mrm_batched_lu.h -- the small linear solves of the batched plans: one system
// A x = b (n <= 128, one right-hand side) per wave or per workgroup, the
// matrix held in registers (mrm_batched_lu_wave.hip, mrm_batched_lu_block.h).
#pragma once

#include <cstdint>

#include <type_traits>

#include "mrm_common.h"

namespace mrm {

// f(integral_constant<int, J>) for J = B .. E-1 in order: the loop index is a
// compile-time constant in every copy, so register arrays indexed by it stay
// in registers.
template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

struct BatchLuP {  // element stride
    int64_t n, lda;
    const double* A;
    const double* b;
    double* x;
    int idx;  // position in the caller's list (status output)
};

// Threads per workgroup: four one-system waves, or one system on eight waves.
constexpr int BL_THREADS = 256;
constexpr int BL_BLOCK_THREADS = 512;

// Persistent kernel (args: const BatchLuP* sys, int count, unsigned* counter,
// int* status, int refine) for systems of at most nmax unknowns: one system
// per wave for nmax 32 / 64 (BL_THREADS per workgroup), one per workgroup of
// BL_BLOCK_THREADS for nmax 96 / 128.
// mode 0: FP64 factors; mode 1: FP32 factors and `refine` FP64 refinement steps.
const void* batched_lu_wave_fn(int mode, int nmax);
template <int MODE, int NMAX>
const void* batched_lu_block_kernel();  // one source file per (MODE, NMAX in 96, 128)

}  // namespace mrm
