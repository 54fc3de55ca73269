// Deterministic block reductions in FP64.
#pragma once

#include "common.hpp"

namespace mslr {

// Sum `v` over the workgroup; the result is valid in thread 0.  Fixed order,
// so results are bitwise reproducible for a given launch shape.
__device__ __forceinline__ double block_sum(double v, double *smem /* >= blockDim.x/64 */) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) smem[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
        for (int i = 0; i < nw; ++i) s += smem[i];
    }
    return s;
}

// Enqueue *out_dev = sum(partials[0..n)) (fixed order, one workgroup).
void sum_partials(mslr_ctx_s *ctx, const double *partials, int64_t n, double *out_dev);

// ||x||^2 of a device vector of n complex values (synchronous, FP64 accumulation).
template <class T>
double device_norm2(mslr_ctx_s *ctx, const C2<T> *x, int64_t n);

// x *= s (in place, n complex values).
template <class T>
void device_scale(mslr_ctx_s *ctx, C2<T> *x, int64_t n, T s);

}  // namespace mslr
