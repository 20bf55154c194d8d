// lane_xor.hpp - cross-lane exchanges of a wave of 64 as VALU operations
// (DPP and the gfx950 lane swaps) instead of LDS permutes, and the
// butterfly sums built on them.
#pragma once
#include <hip/hip_runtime.h>

namespace imago {

// x of lane (this lane xor O): xor 1 and 2 by quad permutes, xor 4 as
// xor 3 (quad permute) then xor 7 (half-row mirror), xor 8 by a row
// rotation of 8, xor 16 and 32 by the row and half-wave swaps.  Every lane
// taking part in the exchange must be active.
template <int O>
__device__ __forceinline__ unsigned xor_lane_u32(unsigned x)
{
    if constexpr (O == 1) {
        return (unsigned)__builtin_amdgcn_mov_dpp((int)x, 0xB1, 0xF, 0xF, false);
    } else if constexpr (O == 2) {
        return (unsigned)__builtin_amdgcn_mov_dpp((int)x, 0x4E, 0xF, 0xF, false);
    } else if constexpr (O == 4) {
        const int y = __builtin_amdgcn_mov_dpp((int)x, 0x1B, 0xF, 0xF, false);
        return (unsigned)__builtin_amdgcn_mov_dpp(y, 0x141, 0xF, 0xF, false);
    } else if constexpr (O == 8) {
        return (unsigned)__builtin_amdgcn_mov_dpp((int)x, 0x128, 0xF, 0xF, false);
    } else if constexpr (O == 16) {
        const auto r = __builtin_amdgcn_permlane16_swap(x, x, false, false);
        return (threadIdx.x & 16) ? r[0] : r[1];
    } else {
        static_assert(O == 32, "lane exchange");
        const auto r = __builtin_amdgcn_permlane32_swap(x, x, false, false);
        return (threadIdx.x & 32) ? r[0] : r[1];
    }
}

template <int O>
__device__ __forceinline__ double xor_lane(double v)
{
    const unsigned lo = xor_lane_u32<O>((unsigned)__double2loint(v));
    const unsigned hi = xor_lane_u32<O>((unsigned)__double2hiint(v));
    return __hiloint2double((int)hi, (int)lo);
}

// Butterfly sums and products, xor 32 first: every lane of the group ends
// with the same bits (each step combines the same two values in every lane
// of a pair).
__device__ __forceinline__ double sum16_xor(double v)
{
    v += xor_lane<8>(v);
    v += xor_lane<4>(v);
    v += xor_lane<2>(v);
    return v + xor_lane<1>(v);
}
__device__ __forceinline__ double wave_sum_xor(double v)
{
    v += xor_lane<32>(v);
    v += xor_lane<16>(v);
    return sum16_xor(v);
}
__device__ __forceinline__ double wave_prod_xor(double v)
{
    v *= xor_lane<32>(v);
    v *= xor_lane<16>(v);
    v *= xor_lane<8>(v);
    v *= xor_lane<4>(v);
    v *= xor_lane<2>(v);
    return v * xor_lane<1>(v);
}

}  // namespace imago
