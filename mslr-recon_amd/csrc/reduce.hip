#include "reduce.hpp"

namespace mslr {

__global__ void k_sum_partials(const double *__restrict__ partials, int64_t n, double *out) {
    __shared__ double smem[16];
    double acc = 0.0;
    // Each thread sums a contiguous run so the order is fixed.
    const int64_t per = (n + blockDim.x - 1) / blockDim.x;
    const int64_t b = (int64_t)threadIdx.x * per;
    const int64_t e = b + per < n ? b + per : n;
    for (int64_t i = b; i < e; ++i) acc += partials[i];
    double s = block_sum(acc, smem);
    if (threadIdx.x == 0) *out = s;
}

template <class T>
__global__ void k_norm2_partial(const C2<T> *__restrict__ x, int64_t n, double *partials) {
    __shared__ double smem[16];
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        C2<T> v = x[i];
        acc += (double)v.re * (double)v.re + (double)v.im * (double)v.im;
    }
    double s = block_sum(acc, smem);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

template <class T>
__global__ void k_scale(C2<T> *__restrict__ x, int64_t n, T s) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        C2<T> v = x[i];
        x[i] = C2<T>{v.re * s, v.im * s};
    }
}

void sum_partials(mslr_ctx_s *ctx, const double *partials, int64_t n, double *out_dev) {
    k_sum_partials<<<1, 256, 0, ctx->stream>>>(partials, n, out_dev);
    HIP_LAUNCH_CHECK();
}

template <class T>
double device_norm2(mslr_ctx_s *ctx, const C2<T> *x, int64_t n) {
    const int nb = 1024;
    double *part = (double *)ctx->scratch_get(SLOT_REDUCE, (nb + 1) * sizeof(double));
    k_norm2_partial<T><<<nb, 256, 0, ctx->stream>>>(x, n, part);
    HIP_LAUNCH_CHECK();
    k_sum_partials<<<1, 256, 0, ctx->stream>>>(part, nb, part + nb);
    HIP_LAUNCH_CHECK();
    double h = 0.0;
    HIP_CHECK(hipMemcpyAsync(&h, part + nb, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return h;
}

template <class T>
void device_scale(mslr_ctx_s *ctx, C2<T> *x, int64_t n, T s) {
    int64_t nb = cdiv(n, 256);
    if (nb > 65536) nb = 65536;
    if (nb < 1) nb = 1;
    k_scale<T><<<(unsigned)nb, 256, 0, ctx->stream>>>(x, n, s);
    HIP_LAUNCH_CHECK();
}

template double device_norm2<float>(mslr_ctx_s *, const c64 *, int64_t);
template double device_norm2<double>(mslr_ctx_s *, const c128 *, int64_t);
template void device_scale<float>(mslr_ctx_s *, c64 *, int64_t, float);
template void device_scale<double>(mslr_ctx_s *, c128 *, int64_t, double);

}  // namespace mslr
