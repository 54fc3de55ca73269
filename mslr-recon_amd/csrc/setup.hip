// Setup of run_recon on the device (scripts/reconstruct.jl:88-118,145-147):
// coil-map normalisation, sampling-mask inference and validation, the k-space
// flatten, the 0.99-quantile scale factor of the SENSE combination
// (src/recon.jl:312-323) and the k-space division by it.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"
#include "sense.hpp"

namespace mslr {

// ------------------------------------------------------------------ coil maps
// smaps ./ (sqrt.(sum(abs2.(smaps); dims=4)) .+ eps(Float32)): per voxel the squared
// magnitudes (re*re + im*im, each product rounded) summed over the coils in order, the
// square root, + eps, then each part divided (Complex(re/x, im/x)); all FP32 with
// correctly rounded division and square root, and no contraction into FMAs (which
// round differently): the pragma applies to the operations written here.
__global__ __launch_bounds__(256) void k_normalise_smaps(c64 *__restrict__ S, int64_t V, int Nc) {
#pragma clang fp contract(off)
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
         v += (int64_t)gridDim.x * blockDim.x) {
        float acc = 0.0f;
        for (int c = 0; c < Nc; ++c) {
            const c64 s = S[v + (int64_t)c * V];
            const float p = s.re * s.re, q = s.im * s.im;
            acc = acc + (p + q);
        }
        const float den = sqrtf(acc) + 1.1920928955078125e-7f;
        for (int c = 0; c < Nc; ++c) {
            const c64 s = S[v + (int64_t)c * V];
            S[v + (int64_t)c * V] = c64{s.re / den, s.im / den};
        }
    }
}

// ------------------------------------------------------------------ mask + flatten
constexpr int FL_CHUNK = 4096;  // voxels per workgroup of the mask and flatten passes
constexpr int FL_NT = 256;

__device__ __forceinline__ bool nonzero(c64 a) { return a.re != 0.0f || a.im != 0.0f; }

// Workgroup (chunk, frame): mask[v, t] = ksp0[v, 1, t] != 0, the first coil (0-based)
// whose nonzeros differ anywhere (atomicMin into *bad_coil), and the chunk's count.
__global__ __launch_bounds__(FL_NT) void k_mask_check(const c64 *__restrict__ ksp0, int64_t V,
                                                      int Nvc, uint8_t *__restrict__ mask,
                                                      int *__restrict__ chunk_cnt,
                                                      int *__restrict__ bad_coil) {
    __shared__ int red[FL_NT / 64];
    const int nch = (int)((V + FL_CHUNK - 1) / FL_CHUNK);
    const int ch = blockIdx.x;
    const int t = blockIdx.y;
    const int64_t v0 = (int64_t)ch * FL_CHUNK;
    const int64_t v1 = std::min<int64_t>(V, v0 + FL_CHUNK);
    const c64 *kt = ksp0 + (int64_t)t * Nvc * V;
    int cnt = 0, bad = INT32_MAX;
    for (int64_t v = v0 + threadIdx.x; v < v1; v += FL_NT) {
        const bool m = nonzero(kt[v]);
        mask[v + (int64_t)t * V] = m ? 1 : 0;
        cnt += m ? 1 : 0;
        for (int c = 1; c < Nvc && c < bad; ++c)
            if (nonzero(kt[v + (int64_t)c * V]) != m) bad = c;
    }
    if (bad != INT32_MAX) atomicMin(bad_coil, bad);
    for (int o = 32; o > 0; o >>= 1) cnt += __shfl_down(cnt, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        int s = 0;
        for (int w = 0; w < FL_NT / 64; ++w) s += red[w];
        chunk_cnt[(int64_t)t * nch + ch] = s;
    }
}

// One workgroup per frame: exclusive scan of the chunk counts and the frame total.
__global__ __launch_bounds__(64) void k_chunk_scan(int *__restrict__ chunk_cnt, int nch,
                                                   int *__restrict__ frame_cnt) {
    const int t = blockIdx.x;
    if (threadIdx.x != 0) return;
    int *c = chunk_cnt + (int64_t)t * nch;
    int s = 0;
    for (int i = 0; i < nch; ++i) {
        const int n = c[i];
        c[i] = s;
        s += n;
    }
    frame_cnt[t] = s;
}

// Workgroup (chunk, frame): the chunk's sampled voxels in column-major order get
// positions off + rank, and y[pos, c, t] = ksp0[v, c, t] for every coil.
__global__ __launch_bounds__(FL_NT) void k_flatten(const c64 *__restrict__ ksp0, int64_t V,
                                                   int Nvc, const uint8_t *__restrict__ mask,
                                                   const int *__restrict__ chunk_off, int64_t K,
                                                   c64 *__restrict__ y) {
    __shared__ int wsum[FL_NT / 64];
    const int nch = (int)((V + FL_CHUNK - 1) / FL_CHUNK);
    const int ch = blockIdx.x;
    const int t = blockIdx.y;
    const int64_t v0 = (int64_t)ch * FL_CHUNK;
    const int64_t v1 = std::min<int64_t>(V, v0 + FL_CHUNK);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int base = chunk_off[(int64_t)t * nch + ch];
    const c64 *kt = ksp0 + (int64_t)t * Nvc * V;
    c64 *yt = y + (int64_t)t * Nvc * K;
    for (int64_t r0 = v0; r0 < v1; r0 += FL_NT) {
        const int64_t v = r0 + threadIdx.x;
        const bool m = v < v1 && mask[v + (int64_t)t * V] != 0;
        const uint64_t b = __ballot(m);
        const int below = __popcll(b & ((1ull << lane) - 1ull));
        if (lane == 0) wsum[w] = __popcll(b);
        __syncthreads();
        int wbase = 0, tot = 0;
        for (int i = 0; i < FL_NT / 64; ++i) {
            if (i < w) wbase += wsum[i];
            tot += wsum[i];
        }
        if (m) {
            const int64_t pos = base + wbase + below;
            for (int c = 0; c < Nvc; ++c) yt[pos + (int64_t)c * K] = kt[v + (int64_t)c * V];
        }
        base += tot;
        __syncthreads();
    }
}

// ------------------------------------------------------------------ scale factor
// mag[v, t] = |Complex{Float32}(img[v, t] / (sqrt(V) * den[v]))| with
// den[v] = sum_c |S_c[v]|^2 (FP32, in coil order) + eps(Float64): the magnitude of
// sense_comb (src/recon.jl:312-323), whose numerator is A'y / sqrt(V) for the plan's
// exact adjoint.  The maps are the plan's (plane-major, pre-phase folded in: the
// phase has modulus one).
__global__ __launch_bounds__(256) void k_comb_mag(const c64 *__restrict__ img,
                                                  const c64 *__restrict__ smaps_pm, int Nx,
                                                  int NYZ, int Nc, int Nt, double sqrtV,
                                                  uint32_t *__restrict__ mag) {
#pragma clang fp contract(off)
    const int64_t V = (int64_t)Nx * NYZ;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
         v += (int64_t)gridDim.x * blockDim.x) {
        const int x = (int)(v % Nx);
        const int64_t yz = v / Nx;
        float acc = 0.0f;
        for (int c = 0; c < Nc; ++c) {
            const c64 s = smaps_pm[((int64_t)c * Nx + x) * NYZ + yz];
            const float p = s.re * s.re, q = s.im * s.im;
            acc = acc + (p + q);
        }
        const double d = sqrtV * ((double)acc + 2.220446049250313e-16);
        for (int t = 0; t < Nt; ++t) {
            const c64 a = img[v + (int64_t)t * V];
            const float re = (float)((double)a.re / d), im = (float)((double)a.im / d);
            const float m = hypotf(re, im);
            mag[v + (int64_t)t * V] = __float_as_uint(m);
        }
    }
}

// Radix select on the bit patterns of non-negative floats (their order is the
// floats' order).  state: [0] prefix bits, [1] mask of the bits fixed so far,
// [2] rank within the values matching the prefix, [3] their count after the last
// pass, [4] min pattern above the selected one, [5] max pattern; hist[256].
struct SelState {
    uint32_t prefix, fixed;
    unsigned long long rank, count;
    uint32_t above, maxbits;
    unsigned long long hist[256];
};

__global__ __launch_bounds__(256) void k_radix_hist(const uint32_t *__restrict__ vals, int64_t n,
                                                    int shift, SelState *st) {
    __shared__ unsigned int h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t prefix = st->prefix, fixed = st->fixed;
    uint32_t mx = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t b = vals[i];
        mx = b > mx ? b : mx;
        if ((b & fixed) == prefix) atomicAdd(&h[(b >> shift) & 255u], 1u);
    }
    if (shift == 24) atomicMax(&st->maxbits, mx);
    __syncthreads();
    if (h[threadIdx.x]) atomicAdd(&st->hist[threadIdx.x], (unsigned long long)h[threadIdx.x]);
}

__global__ void k_radix_pick(int shift, SelState *st) {
    if (threadIdx.x != 0) return;
    unsigned long long r = st->rank, below = 0;
    int b = 0;
    for (; b < 256; ++b) {
        if (below + st->hist[b] > r) break;
        below += st->hist[b];
    }
    b = b < 256 ? b : 255;
    st->count = st->hist[b];
    st->rank = r - below;
    st->prefix |= (uint32_t)b << shift;
    st->fixed |= 255u << shift;
    for (int i = 0; i < 256; ++i) st->hist[i] = 0;
}

__global__ __launch_bounds__(256) void k_min_above(const uint32_t *__restrict__ vals, int64_t n,
                                                   SelState *st) {
    const uint32_t p = st->prefix;
    uint32_t mn = 0xFFFFFFFFu;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t b = vals[i];
        if (b > p && b < mn) mn = b;
    }
    for (int o = 32; o > 0; o >>= 1) {
        const uint32_t q = __shfl_down(mn, o, 64);
        mn = q < mn ? q : mn;
    }
    if ((threadIdx.x & 63) == 0 && mn != 0xFFFFFFFFu) atomicMin(&st->above, mn);
}

// y <- Complex{Float32}(y / d) with the division in FP64 (ComplexF32 ./= Float64).
__global__ __launch_bounds__(256) void k_div_kspace(c64 *__restrict__ y, int64_t n, double d) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const c64 a = y[i];
        y[i] = c64{(float)((double)a.re / d), (float)((double)a.im / d)};
    }
}

static int grid_for(int64_t n, int nt = 256, int maxg = 8192) {
    return (int)std::max<int64_t>(1, std::min<int64_t>(maxg, (n + nt - 1) / nt));
}

}  // namespace mslr

using namespace mslr;

extern "C" {

mslr_status_t mslr_normalise_smaps(mslr_ctx ctx, void *smaps_dev, const int64_t N[3], int64_t Nc) {
    return guarded([&] {
        mslr_ctx_s *c = check_ctx(ctx);
        MSLR_REQUIRE(smaps_dev && N, MSLR_ERR_ARG, "null argument");
        MSLR_REQUIRE(N[0] > 0 && N[1] > 0 && N[2] > 0 && Nc > 0, MSLR_ERR_ARG, "sizes must be positive");
        const int64_t V = N[0] * N[1] * N[2];
        MSLR_PROF(c, "normalise_smaps");
        k_normalise_smaps<<<grid_for(V), 256, 0, c->stream>>>((c64 *)smaps_dev, V, (int)Nc);
        HIP_LAUNCH_CHECK();
    });
}

mslr_status_t mslr_kspace_flatten(mslr_ctx ctx, const void *ksp0_dev, const int64_t N[3],
                                  int64_t Nvc, int64_t Nt, uint8_t *mask_dev, void *y_dev,
                                  int64_t y_capacity, int64_t *K) {
    return guarded([&] {
        mslr_ctx_s *c = check_ctx(ctx);
        MSLR_REQUIRE(ksp0_dev && N && mask_dev && K, MSLR_ERR_ARG, "null argument");
        MSLR_REQUIRE(N[0] > 0 && N[1] > 0 && N[2] > 0 && Nvc > 0 && Nt > 0, MSLR_ERR_ARG,
                     "sizes must be positive");
        const int64_t V = N[0] * N[1] * N[2];
        MSLR_REQUIRE(V < ((int64_t)1 << 31) && Nt <= 65535, MSLR_ERR_UNSUPPORTED,
                     "volume or frame count too large");
        const int nch = (int)((V + FL_CHUNK - 1) / FL_CHUNK);
        int *w = (int *)c->scratch_get(SLOT_SETUP, sizeof(int) * ((size_t)nch * Nt + Nt + 1));
        int *chunk = w, *frame = w + (size_t)nch * Nt, *bad = frame + Nt;
        const int big = INT32_MAX;
        HIP_CHECK(hipMemcpyAsync(bad, &big, sizeof(int), hipMemcpyHostToDevice, c->stream));
        {
            MSLR_PROF(c, "mask_check");
            k_mask_check<<<dim3(nch, (unsigned)Nt), FL_NT, 0, c->stream>>>(
                (const c64 *)ksp0_dev, V, (int)Nvc, mask_dev, chunk, bad);
            HIP_LAUNCH_CHECK();
        }
        k_chunk_scan<<<(unsigned)Nt, 64, 0, c->stream>>>(chunk, nch, frame);
        HIP_LAUNCH_CHECK();
        std::vector<int> counts((size_t)Nt + 1);
        HIP_CHECK(hipMemcpyAsync(counts.data(), frame, sizeof(int) * (Nt + 1), hipMemcpyDeviceToHost,
                                 c->stream));
        HIP_CHECK(hipStreamSynchronize(c->stream));
        // the reference's assertions, in its order (scripts/reconstruct.jl:113-118)
        // (the reference's messages exactly, without the entry point's name)
        if (counts[(size_t)Nt] != big)
            throw Error(MSLR_ERR_SHAPE, "Coil " + std::to_string(counts[(size_t)Nt] + 1) +
                                            " has a different sampling pattern");
        for (int64_t t = 1; t < Nt; ++t)
            if (counts[(size_t)t] != counts[(size_t)t - 1])
                throw Error(MSLR_ERR_SHAPE,
                            "Frame " + std::to_string(t + 1) + " has a different sample count");
        *K = counts[0];
        if (!y_dev) return;
        MSLR_REQUIRE(y_capacity >= *K * Nvc * Nt, MSLR_ERR_ARG,
                     "y_dev holds fewer than K * Nvc * Nt values");
        MSLR_PROF(c, "flatten");
        k_flatten<<<dim3(nch, (unsigned)Nt), FL_NT, 0, c->stream>>>(
            (const c64 *)ksp0_dev, V, (int)Nvc, mask_dev, chunk, *K, (c64 *)y_dev);
        HIP_LAUNCH_CHECK();
    });
}

mslr_status_t mslr_sense_scale_factor(mslr_sense plan, const void *y_dev, double q, double *sf) {
    return guarded([&] {
        MSLR_REQUIRE(plan && y_dev && sf, MSLR_ERR_ARG, "null argument");
        MSLR_REQUIRE(q >= 0.0 && q <= 1.0, MSLR_ERR_ARG, "quantile level must be in [0, 1]");
        mslr_sense_s *P = plan;
        mslr_ctx_s *c = P->ctx;
        c->activate();
        const int64_t n = P->V * (int64_t)P->Nt;
        c64 *img = (c64 *)c->scratch_get(SLOT_PATCH_OUT, (size_t)n * sizeof(c64));
        MSLR_REQUIRE(mslr_sense_adjoint(plan, y_dev, img, 0, P->Nt) == MSLR_OK, MSLR_ERR_DEVICE,
                     "adjoint failed");
        char *w = (char *)c->scratch_get(SLOT_SETUP, (size_t)n * 4 + 256 + sizeof(SelState));
        uint32_t *mag = (uint32_t *)w;
        SelState *st = (SelState *)(w + (((size_t)n * 4 + 255) & ~(size_t)255));
        {
            MSLR_PROF(c, "comb_mag");
            k_comb_mag<<<grid_for(P->V, 256, 4096), 256, 0, c->stream>>>(
                img, P->smaps32, P->N[0], P->NYZ, P->Nc, P->Nt, std::sqrt((double)P->V), mag);
            HIP_LAUNCH_CHECK();
        }
        // type 7 (Statistics.quantile): h = (n - 1) q + 1, j = trunc(h) (1-based), value
        // a + (h - j)(b - a) with a, b the j-th and (j+1)-th smallest
        const double aleph = (double)n * q + (1.0 - q);
        const int64_t j = std::min<int64_t>(std::max<int64_t>((int64_t)aleph, 1), std::max<int64_t>(n - 1, 1));
        const double gamma = std::min(std::max(aleph - (double)j, 0.0), 1.0);
        SelState h{};
        h.rank = (unsigned long long)(j - 1);
        h.above = 0xFFFFFFFFu;
        HIP_CHECK(hipMemcpyAsync(st, &h, sizeof h, hipMemcpyHostToDevice, c->stream));
        const int g = grid_for(n, 256, 4096);
        for (int shift = 24; shift >= 0; shift -= 8) {
            k_radix_hist<<<g, 256, 0, c->stream>>>(mag, n, shift, st);
            HIP_LAUNCH_CHECK();
            k_radix_pick<<<1, 64, 0, c->stream>>>(shift, st);
            HIP_LAUNCH_CHECK();
        }
        k_min_above<<<g, 256, 0, c->stream>>>(mag, n, st);
        HIP_LAUNCH_CHECK();
        HIP_CHECK(hipMemcpyAsync(&h, st, sizeof h, hipMemcpyDeviceToHost, c->stream));
        HIP_CHECK(hipStreamSynchronize(c->stream));
        MSLR_REQUIRE(h.maxbits <= 0x7F800000u, MSLR_ERR_NUMERIC,
                     "the SENSE combination has NaN values (quantile undefined)");
        uint32_t bb = h.rank + 1 < h.count || n == 1 ? h.prefix : h.above;
        float a, b;
        std::memcpy(&a, &h.prefix, 4);
        std::memcpy(&b, &bb, 4);
        *sf = (double)a + gamma * ((double)b - (double)a);
    });
}

mslr_status_t mslr_scale_kspace(mslr_ctx ctx, void *y_dev, int64_t n, double sf) {
    return guarded([&] {
        mslr_ctx_s *c = check_ctx(ctx);
        MSLR_REQUIRE(y_dev || n == 0, MSLR_ERR_ARG, "null argument");
        MSLR_REQUIRE(n >= 0, MSLR_ERR_ARG, "n < 0");
        const double d = std::max(sf, (double)1.1920928955078125e-7);
        if (n == 0) return;
        MSLR_PROF(c, "scale_kspace");
        k_div_kspace<<<grid_for(n), 256, 0, c->stream>>>((c64 *)y_dev, n, d);
        HIP_LAUNCH_CHECK();
    });
}

}  // extern "C"
