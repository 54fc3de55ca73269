// Post-reconstruction analysis of a 4D image series (src/analysis.jl): the tSNR
// map and the voxel-wise linear detrend.  Both read a voxel's time course down
// the frame stride V (consecutive threads, consecutive voxels: coalesced) and are
// HBM-bound: tSNR reads V*Nt complex values once and writes V floats; detrend
// reads them twice (slope, then update) and writes them once.
#include <cfloat>
#include <cmath>

#include "common.hpp"

namespace mslr {

// tSNR = mean(|x|) / (std(|x|) + eps(Float32)) over time, std with n - 1
// (src/analysis.jl:39-44).  FP64 sums, two passes (mean, then squared deviations).
__global__ __launch_bounds__(256) void k_tsnr(const c64 *__restrict__ img, int64_t V, int Nt,
                                              float *__restrict__ out) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
         v += (int64_t)gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int t = 0; t < Nt; ++t) {
            const c64 a = img[v + V * t];
            s += (double)hypotf(a.re, a.im);
        }
        const double mean = s / Nt;
        double q = 0.0;
        for (int t = 0; t < Nt; ++t) {
            const c64 a = img[v + V * t];
            const double d = (double)hypotf(a.re, a.im) - mean;
            q += d * d;
        }
        const double sd = Nt > 1 ? sqrt(q / (Nt - 1)) : NAN;
        out[v] = (float)(mean / (sd + (double)FLT_EPSILON));
    }
}

// x_t <- x_t - t * b per voxel, b the least-squares slope of x over t = 1..Nt
// (pinv([1 t]) x, src/analysis.jl:122-132): b = sum (t - tm) x_t / sum (t - tm)^2,
// FP64 arithmetic, stored as ComplexF32.
__global__ __launch_bounds__(256) void k_detrend(c64 *__restrict__ img, int64_t V, int Nt) {
    const double tm = 0.5 * (Nt + 1);
    double den = 0.0;
    for (int t = 1; t <= Nt; ++t) den += (t - tm) * (t - tm);
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
         v += (int64_t)gridDim.x * blockDim.x) {
        double br = 0.0, bi = 0.0;
        for (int t = 1; t <= Nt; ++t) {
            const c64 a = img[v + V * (t - 1)];
            br += (t - tm) * (double)a.re;
            bi += (t - tm) * (double)a.im;
        }
        br /= den;
        bi /= den;
        for (int t = 1; t <= Nt; ++t) {
            c64 &a = img[v + V * (t - 1)];
            a = c64{(float)((double)a.re - t * br), (float)((double)a.im - t * bi)};
        }
    }
}

static unsigned voxel_blocks(int64_t V) {
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(cdiv(V, 256), 1 << 20));
}

}  // namespace mslr

using namespace mslr;

extern "C" {

mslr_status_t mslr_tsnr(mslr_ctx ctx, const void *img_dev, int64_t V, int64_t Nt, void *out_dev) {
    return guarded([&] {
        mslr_ctx_s *c = check_ctx(ctx);
        MSLR_REQUIRE(img_dev && out_dev, MSLR_ERR_ARG, "null argument");
        MSLR_REQUIRE(V >= 0 && Nt >= 1 && Nt < (1 << 30), MSLR_ERR_ARG, "bad sizes");
        if (V == 0) return;
        MSLR_PROF(c, "tsnr");
        k_tsnr<<<voxel_blocks(V), 256, 0, c->stream>>>((const c64 *)img_dev, V, (int)Nt,
                                                       (float *)out_dev);
        HIP_LAUNCH_CHECK();
    });
}

mslr_status_t mslr_detrend(mslr_ctx ctx, void *img_dev, int64_t V, int64_t Nt) {
    return guarded([&] {
        mslr_ctx_s *c = check_ctx(ctx);
        MSLR_REQUIRE(img_dev, MSLR_ERR_ARG, "null argument");
        MSLR_REQUIRE(V >= 0 && Nt >= 1 && Nt < (1 << 30), MSLR_ERR_ARG, "bad sizes");
        if (V == 0) return;
        MSLR_PROF(c, "detrend");
        k_detrend<<<voxel_blocks(V), 256, 0, c->stream>>>((c64 *)img_dev, V, (int)Nt);
        HIP_LAUNCH_CHECK();
    });
}

}  // extern "C"
