// Cross-lane reductions on CDNA wavefronts (64 lanes) without the LDS crossbar:
// DPP row permutations inside each row of 16 lanes, readlane across the rows.
#pragma once

#include "common.hpp"

namespace mslr {

// DPP controls (row = 16 lanes): quad_perm patterns, row_half_mirror (lane i of
// each 8 reads lane 7-i), row_mirror (lane i of each 16 reads lane 15-i).
constexpr int DPP_QP_1032 = 0xB1;  // quad_perm [1,0,3,2]: lane i reads i^1
constexpr int DPP_QP_2301 = 0x4E;  // quad_perm [2,3,0,1]: lane i reads i^2
constexpr int DPP_ROW_HALF_MIRROR = 0x141;
constexpr int DPP_ROW_MIRROR = 0x140;

// (the row permutations below read a valid lane for every lane, so no "old"
// value is needed: mov_dpp with bound_ctrl, no zeroing move first)
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Sum over each row of 16 lanes; every lane of the row receives the row sum.
// (pairs i^1, then i^2, then the mirrored half, then the mirrored row)
__device__ __forceinline__ double sum16(double v) {
    v += dpp_d<DPP_QP_1032>(v);
    v += dpp_d<DPP_QP_2301>(v);
    v += dpp_d<DPP_ROW_HALF_MIRROR>(v);
    v += dpp_d<DPP_ROW_MIRROR>(v);
    return v;
}
__device__ __forceinline__ C2<double> sum16(C2<double> v) { return {sum16(v.re), sum16(v.im)}; }

// v + v[lane ^ 16] and v + v[lane ^ 32] through v_permlane16_swap / v_permlane32_swap
// (no LDS crossbar): swapping a register with itself gives every lane its own value
// and its partner's, in an order that makes the sum the same in both lanes.
__device__ __forceinline__ double mk_double(int lo, int hi) {
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double xor16_sum(double v) {
    const long long b = __double_as_longlong(v);
    const auto lo = __builtin_amdgcn_permlane16_swap((int)b, (int)b, false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap((int)(b >> 32), (int)(b >> 32), false, false);
    return mk_double(lo[0], hi[0]) + mk_double(lo[1], hi[1]);
}
__device__ __forceinline__ double xor32_sum(double v) {
    const long long b = __double_as_longlong(v);
    const auto lo = __builtin_amdgcn_permlane32_swap((int)b, (int)b, false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap((int)(b >> 32), (int)(b >> 32), false, false);
    return mk_double(lo[0], hi[0]) + mk_double(lo[1], hi[1]);
}
// Sum over the four rows of 16 lanes (lanes l, l ^ 16, l ^ 32, l ^ 48), in every lane.
__device__ __forceinline__ C2<double> sum_rows4(C2<double> v) {
    return {xor32_sum(xor16_sum(v.re)), xor32_sum(xor16_sum(v.im))};
}

__device__ __forceinline__ double readlane_d(double v, int lane) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Row r's lane 15 (row_bcast:15) or lane 31 (row_bcast:31) into the rows of ROWMASK;
// the other rows read 0.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double bcast_d(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, ROWMASK, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROWMASK, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
constexpr int DPP_ROW_BCAST15 = 0x142;
constexpr int DPP_ROW_BCAST31 = 0x143;

// Sum over the wave (uniform result): row sums by DPP, then rows 0 + 1 and 2 + 3 by
// row_bcast:15 and the two halves by row_bcast:31, so that lane 63 holds the sum
// (one readlane instead of four).
__device__ __forceinline__ double wave_sum64(double v) {
    v = sum16(v);
    v += bcast_d<DPP_ROW_BCAST15, 0xA>(v);
    v += bcast_d<DPP_ROW_BCAST31, 0xC>(v);
    return readlane_d(v, 63);
}
__device__ __forceinline__ C2<double> wave_sum64(C2<double> v) {
    return {wave_sum64(v.re), wave_sum64(v.im)};
}

}  // namespace mslr
