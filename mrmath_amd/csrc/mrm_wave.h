// mrm_wave.h -- wave-level argmax shared by the LU panel and the batched solves.
#pragma once

#include <cstdint>

#include <hip/hip_runtime.h>

namespace mrm {
namespace wave {

// Larger |value| wins; on equal values the smaller row index (earlier row) wins.
template <typename T>
__device__ __forceinline__ bool better(T va, int ra, T vb, int rb) {
    return va > vb || (va == vb && ra < rb);
}

// 32-bit DPP lane permutation (all 64 lanes active).
template <int CTRL>
__device__ __forceinline__ int dpp_i(int x) {
    return __builtin_amdgcn_update_dpp(x, x, CTRL, 0xF, 0xF, false);
}
template <int CTRL, typename T>
__device__ __forceinline__ T dpp_v(T v) {
    if constexpr (sizeof(T) == 8) {
        const long long b = __double_as_longlong(v);
        const unsigned lo = unsigned(dpp_i<CTRL>(int(b))), hi = unsigned(dpp_i<CTRL>(int(b >> 32)));
        return __longlong_as_double((long long)((uint64_t(hi) << 32) | lo));
    } else {
        return __int_as_float(dpp_i<CTRL>(__float_as_int(v)));
    }
}
template <typename T>
__device__ __forceinline__ T readlane_v(T v, int l) {
    if constexpr (sizeof(T) == 8) {
        const long long b = __double_as_longlong(v);
        const unsigned lo = unsigned(__builtin_amdgcn_readlane(int(b), l));
        const unsigned hi = unsigned(__builtin_amdgcn_readlane(int(b >> 32), l));
        return __longlong_as_double((long long)((uint64_t(hi) << 32) | lo));
    } else {
        return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
    }
}
template <int CTRL, typename T>
__device__ __forceinline__ void argmax_step(T& v, int& r) {
    const T ov = dpp_v<CTRL>(v);
    const int orow = dpp_i<CTRL>(r);
    if (better(ov, orow, v, r)) {
        v = ov;
        r = orow;
    }
}

// Best (value, row) of the wave, in every lane.  Within each row of 16 lanes
// by DPP (quad swaps, then half-row and row mirrors: register-to-register,
// no LDS round trip), then across the four rows by readlane.  Needs all 64
// lanes active.
template <typename T>
__device__ __forceinline__ void argmax(T& v, int& r) {
    argmax_step<0xB1>(v, r);   // quad_perm [1,0,3,2]
    argmax_step<0x4E>(v, r);   // quad_perm [2,3,0,1]
    argmax_step<0x141>(v, r);  // row_half_mirror
    argmax_step<0x140>(v, r);  // row_mirror
    T bv = readlane_v(v, 0);
    int br = __builtin_amdgcn_readlane(r, 0);
#pragma unroll
    for (int q = 1; q < 4; ++q) {
        const T qv = readlane_v(v, 16 * q);
        const int qr = __builtin_amdgcn_readlane(r, 16 * q);
        if (better(qv, qr, bv, br)) {
            bv = qv;
            br = qr;
        }
    }
    v = bv;
    r = br;
}

}  // namespace wave
}  // namespace mrm
