// Ozaki slices of FP32 complex data for the split-BF16 products: per block of
// SPL_ROWS consecutive values of a column one exponent E, 2^E > max |x| over the
// block (real and imaginary parts), and x = s1 + s2 + r with s1 a multiple of
// 2^(E-8) (|s1| <= 2^8 of those), s2 a multiple of 2^(E-16) (<= 2^7 of those),
// |r| <= 2^(E-17); both slices exact in BF16.  Written as two words per element
// (BF16 re in the low half, im in the high half): wh = s1, wl = s2.  Shared by the
// SVST's split passes and the POGM update, which writes the words of the
// whole-image patch itself.
#pragma once

#include "common.hpp"

namespace mslr {

constexpr int SPL_ROWS = 128;

typedef __bf16 bf2v __attribute__((ext_vector_type(2)));

// (s1, s2) of x for the quanta set by sg1 = 3*2^(E+14), sg2 = 3*2^(E+6)
__device__ __forceinline__ float2 ozaki_split(float x, float sg1, float sg2) {
    const float s1 = (x + sg1) - sg1;
    const float r1 = x - s1;
    const float s2 = (r1 + sg2) - sg2;
    return make_float2(s1, s2);
}

__device__ __forceinline__ void ozaki_words(c64 v, float sg1, float sg2, unsigned &wh, unsigned &wl) {
    const float2 sr = ozaki_split(v.re, sg1, sg2);
    const float2 si = ozaki_split(v.im, sg1, sg2);
    const bf2v hv = {(__bf16)sr.x, (__bf16)si.x};
    const bf2v lv = {(__bf16)sr.y, (__bf16)si.y};
    wh = __builtin_bit_cast(unsigned, hv);
    wl = __builtin_bit_cast(unsigned, lv);
}

// The quanta (sg1, sg2) of a block from the largest |part| of the lane's values,
// reduced over the wave (every lane of the wave calls it, each holding its share
// of the block; lanes without values pass 0).
__device__ __forceinline__ void ozaki_block_quanta(float mx, float &sg1, float &sg2) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
    int E;
    (void)frexpf(mx, &E);  // mx < 2^E (E = 0 for a zero block: all slices 0)
    sg1 = ldexpf(3.f, E + 14);
    sg2 = ldexpf(3.f, E + 6);
}

}  // namespace mslr
