// SENSE forward / adjoint over all frames (block-diagonal Asense stack).
//
// Forward for frame t, coil c (src/sense_gpu.jl:71-81):
//   y = scale * (fftshift o FFT3 o ifftshift)(S_c .* x_t) restricted to mask_t.
// The centred DFT is written as post_phase .* DFT(pre_phase .* v): the
// pre-phase is folded into the stored coil maps and the post-phase into the
// k-space gather, so no shift copies are made.  Two passes per frame chunk:
//   pass X : scale sum + coil multiply + x-axis FFT of a block of yz-lines,
//            written as per-coil planes T[f][c][kx][y + Ny*z];
//   pass YZ: one (kx, coil, frame) plane in LDS, y- and z-axis FFTs, gather
//            of the sampled (ky,kz) points of that plane (+ residual).
// The adjoint runs the mirror image (scatter -> inverse yz FFT -> inverse x FFT
// -> conj-coil combine) and is the exact conjugate transpose of the forward.
#include <algorithm>
#include <cmath>
#include <type_traits>

#include "fft.hpp"
#include "fft_fixed.hpp"
#include "reduce.hpp"
#include "sense.hpp"

namespace mslr {

constexpr int NTX = 256;     // threads of the x-axis passes
constexpr int MAXE_X = 16;   // elements per thread in the x-axis passes
constexpr int MAXE_YZ = 16;  // elements per thread in the yz-plane passes
constexpr int NTYZ_MAX = 512;  // threads of the yz-plane passes (at most)
constexpr int MAX_NYZ = MAXE_YZ * NTYZ_MAX;

// ------------------------------------------------------------------ FFT policies
// A policy transforms `nlines` lines in LDS: runtime lengths (any 2/3/5 length)
// or one compile-time length.
template <class T, int MAXE>
struct FftRt {
    static constexpr bool kPipe = false;  // x kernels: no next-coil register prefetch
    FftDesc d;
    const C2<T> *tw;
    int sgn;
    __device__ __forceinline__ void operator()(C2<T> *buf, int nlines, int lstride,
                                               int estride) const {
        fft_lines<T, MAXE>(buf, d, nlines, lstride, estride, tw, sgn);
    }
};
template <class T, int N, int SGN, int MAXLINES, int NT>
struct FftFix {
    static constexpr bool kPipe = true;   // x kernels: next-coil register prefetch
    const C2<T> *tw;
    __device__ __forceinline__ void operator()(C2<T> *buf, int nlines, int lstride,
                                               int estride) const {
        fix_fft<T, N, SGN, MAXLINES, NT>(buf, nlines, lstride, estride, tw);
    }
};

// ------------------------------------------------------------------ kernels

template <class T, class FX>
__global__ __launch_bounds__(NTX) void k_fwd_x(const C2<T> *__restrict__ X, int nscales,
                                               int64_t sstride, const C2<T> *__restrict__ smaps,
                                               int Nc, int64_t V, int Nx, int NYZ, int LB, FX fx,
                                               C2<T> *__restrict__ Tb) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C2<T> *buf = reinterpret_cast<C2<T> *>(smem_raw);
    const int f = blockIdx.y;
    const int l0 = blockIdx.x * LB;
    const int nl = min(LB, NYZ - l0);
    const int nel = nl * Nx;
    const int tid = threadIdx.x;
    const int64_t base = (int64_t)f * V + (int64_t)l0 * Nx;
    C2<T> s[MAXE_X];
#pragma unroll
    for (int q = 0; q < MAXE_X; ++q) {
        const int e = tid + q * NTX;
        C2<T> acc = {T(0), T(0)};
        if (e < nel) {
            acc = X[base + e];
            for (int k = 1; k < nscales; ++k) acc = acc + X[(int64_t)k * sstride + base + e];
        }
        s[q] = acc;
    }
    // coil maps of the next coil are loaded before this coil's FFT (latency overlap)
    constexpr int PX = FX::kPipe ? MAXE_X : 1;
    C2<T> smn[PX];
    if constexpr (FX::kPipe) {
        const C2<T> *sm = smaps + (int64_t)l0 * Nx;
#pragma unroll
        for (int q = 0; q < PX; ++q) {
            const int e = tid + q * NTX;
            smn[q] = e < nel ? sm[e] : C2<T>{T(0), T(0)};
        }
    }
    for (int c = 0; c < Nc; ++c) {
        if constexpr (FX::kPipe) {
#pragma unroll
            for (int q = 0; q < PX; ++q) {
                const int e = tid + q * NTX;
                if (e < nel) buf[e] = s[q] * smn[q];
            }
            if (c + 1 < Nc) {
                const C2<T> *sm = smaps + (int64_t)(c + 1) * V + (int64_t)l0 * Nx;
#pragma unroll
                for (int q = 0; q < PX; ++q) {
                    const int e = tid + q * NTX;
                    if (e < nel) smn[q] = sm[e];
                }
            }
        } else {
            const C2<T> *sm = smaps + (int64_t)c * V + (int64_t)l0 * Nx;
#pragma unroll
            for (int q = 0; q < MAXE_X; ++q) {
                const int e = tid + q * NTX;
                if (e < nel) buf[e] = s[q] * sm[e];
            }
        }
        __syncthreads();
        fx(buf, nl, Nx, 1);
        C2<T> *out = Tb + ((int64_t)f * Nc + c) * V + l0;
        for (int idx = tid; idx < nel; idx += NTX) {
            const int kx = idx / nl;
            const int l = idx - kx * nl;
            out[(int64_t)kx * NYZ + l] = buf[l * Nx + kx];
        }
        __syncthreads();
    }
}

// One workgroup per (kx, coil, frame): 2D FFT of one plane, then gather of the
// sampled (ky, kz) with the post-phases, optional residual and its squared norm.
template <class T, class FY, class FZ>
__global__ __launch_bounds__(NTYZ_MAX) void k_fwd_yz(
    const C2<T> *__restrict__ Tb, int Nc, int Nx, int Ny, int Nz, FY fy, FZ fz,
    const int *__restrict__ ptr, const int *__restrict__ off, const int *__restrict__ perm,
    int64_t K, const C2<T> *__restrict__ phx, const C2<T> *__restrict__ phy,
    const C2<T> *__restrict__ phz, T scale, const C2<T> *__restrict__ ysub,
    C2<T> *__restrict__ out, double *__restrict__ partial) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C2<T> *buf = reinterpret_cast<C2<T> *>(smem_raw);
    __shared__ double red[16];
    const int kx = blockIdx.x, c = blockIdx.y, f = blockIdx.z;
    const int NYZ = Ny * Nz;
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int *pf = ptr + (int64_t)f * (Nx + 1);
    const int j0 = pf[kx], j1 = pf[kx + 1];
    const int *of = off + (int64_t)f * K;
    const int *pm = perm ? perm + (int64_t)f * K : nullptr;
    const C2<T> px = phx[kx];
    const C2<T> *src = Tb + (((int64_t)f * Nc + c) * Nx + kx) * NYZ;
    for (int i = tid; i < NYZ; i += nthr) buf[i] = src[i];
    __syncthreads();
    fy(buf, Nz, Ny, 1);
    fz(buf, Ny, 1, Ny);
    const int64_t obase = ((int64_t)f * Nc + c) * K;
    double acc = 0.0;
    for (int j = j0 + tid; j < j1; j += nthr) {
        const int o = of[j];
        const int ky = o % Ny;
        const int kz = o / Ny;
        const C2<T> ph = px * phy[ky] * phz[kz];
        C2<T> v = scale * (buf[o] * ph);
        const int64_t oi = obase + (pm ? pm[j] : j);
        if (ysub) v = v - ysub[oi];
        out[oi] = v;
        acc += (double)v.re * (double)v.re + (double)v.im * (double)v.im;
    }
    if (partial) {
        const double sres = block_sum(acc, red);
        if (tid == 0) partial[((int64_t)f * Nc + c) * Nx + kx] = sres;
    }
}

// One workgroup per (kx, coil, frame): scatter of the sampled (ky, kz) times the
// conjugate post-phases into a zeroed plane, then its inverse-direction 2D FFT.
template <class T, class FY, class FZ>
__global__ __launch_bounds__(NTYZ_MAX) void k_adj_yz(
    const C2<T> *__restrict__ yin, int Nc, int Nx, int Ny, int Nz, FY fy, FZ fz,
    const int *__restrict__ ptr, const int *__restrict__ off, const int *__restrict__ perm,
    int64_t K, const C2<T> *__restrict__ phx, const C2<T> *__restrict__ phy,
    const C2<T> *__restrict__ phz, C2<T> *__restrict__ Tb) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C2<T> *buf = reinterpret_cast<C2<T> *>(smem_raw);
    const int kx = blockIdx.x, c = blockIdx.y, f = blockIdx.z;
    const int NYZ = Ny * Nz;
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int *pf = ptr + (int64_t)f * (Nx + 1);
    const int j0 = pf[kx], j1 = pf[kx + 1];
    const int *of = off + (int64_t)f * K;
    const int *pm = perm ? perm + (int64_t)f * K : nullptr;
    const C2<T> px = phx[kx];
    for (int i = tid; i < NYZ; i += nthr) buf[i] = C2<T>{T(0), T(0)};
    __syncthreads();
    const int64_t ibase = ((int64_t)f * Nc + c) * K;
    for (int j = j0 + tid; j < j1; j += nthr) {
        const int o = of[j];
        const int ky = o % Ny;
        const int kz = o / Ny;
        const C2<T> ph = px * phy[ky] * phz[kz];
        buf[o] = cmulc(ph, yin[ibase + (pm ? pm[j] : j)]);
    }
    __syncthreads();
    fz(buf, Ny, 1, Ny);
    fy(buf, Nz, Ny, 1);
    C2<T> *dst = Tb + (((int64_t)f * Nc + c) * Nx + kx) * NYZ;
    for (int i = tid; i < NYZ; i += nthr) dst[i] = buf[i];
}

template <class T, class FX>
__global__ __launch_bounds__(NTX) void k_adj_x(const C2<T> *__restrict__ Tb,
                                               const C2<T> *__restrict__ smaps, int Nc, int64_t V,
                                               int Nx, int NYZ, int LB, FX fx, T scale,
                                               C2<T> *__restrict__ xout) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C2<T> *buf = reinterpret_cast<C2<T> *>(smem_raw);
    const int f = blockIdx.y;
    const int l0 = blockIdx.x * LB;
    const int nl = min(LB, NYZ - l0);
    const int nel = nl * Nx;
    const int tid = threadIdx.x;
    C2<T> acc[MAXE_X];
#pragma unroll
    for (int q = 0; q < MAXE_X; ++q) acc[q] = C2<T>{T(0), T(0)};
    // plane block of the next coil (transposed read) is loaded before this coil's FFT
    constexpr int PX = FX::kPipe ? MAXE_X : 1;
    C2<T> tn[PX];
    if constexpr (FX::kPipe) {
        const C2<T> *src = Tb + (int64_t)f * Nc * V + l0;
#pragma unroll
        for (int q = 0; q < PX; ++q) {
            const int idx = tid + q * NTX;
            if (idx < nel) {
                const int kx = idx / nl;
                const int l = idx - kx * nl;
                tn[q] = src[(int64_t)kx * NYZ + l];
            }
        }
    }
    for (int c = 0; c < Nc; ++c) {
        if constexpr (FX::kPipe) {
#pragma unroll
            for (int q = 0; q < PX; ++q) {
                const int idx = tid + q * NTX;
                if (idx < nel) {
                    const int kx = idx / nl;
                    const int l = idx - kx * nl;
                    buf[l * Nx + kx] = tn[q];
                }
            }
        } else {
            const C2<T> *src = Tb + ((int64_t)f * Nc + c) * V + l0;
            for (int idx = tid; idx < nel; idx += NTX) {
                const int kx = idx / nl;
                const int l = idx - kx * nl;
                buf[l * Nx + kx] = src[(int64_t)kx * NYZ + l];
            }
        }
        C2<T> smc[PX];
        if constexpr (FX::kPipe) {
            const C2<T> *sm = smaps + (int64_t)c * V + (int64_t)l0 * Nx;
#pragma unroll
            for (int q = 0; q < PX; ++q) {
                const int e = tid + q * NTX;
                if (e < nel) smc[q] = sm[e];
            }
            if (c + 1 < Nc) {
                const C2<T> *src = Tb + ((int64_t)f * Nc + c + 1) * V + l0;
#pragma unroll
                for (int q = 0; q < PX; ++q) {
                    const int idx = tid + q * NTX;
                    if (idx < nel) {
                        const int kx = idx / nl;
                        const int l = idx - kx * nl;
                        tn[q] = src[(int64_t)kx * NYZ + l];
                    }
                }
            }
        }
        __syncthreads();
        fx(buf, nl, Nx, 1);
        if constexpr (FX::kPipe) {
#pragma unroll
            for (int q = 0; q < PX; ++q) {
                const int e = tid + q * NTX;
                if (e < nel) acc[q] = acc[q] + cmulc(smc[q], buf[e]);
            }
        } else {
            const C2<T> *sm = smaps + (int64_t)c * V + (int64_t)l0 * Nx;
#pragma unroll
            for (int q = 0; q < MAXE_X; ++q) {
                const int e = tid + q * NTX;
                if (e < nel) acc[q] = acc[q] + cmulc(sm[e], buf[e]);
            }
        }
        __syncthreads();
    }
    C2<T> *xo = xout + (int64_t)f * V + (int64_t)l0 * Nx;
#pragma unroll
    for (int q = 0; q < MAXE_X; ++q) {
        const int e = tid + q * NTX;
        if (e < nel) xo[e] = scale * acc[q];
    }
}

// y_int[f][c][j] = y_ref[f][c][perm[f][j]]
__global__ void k_perm_gather(const c64 *__restrict__ yref, const int *__restrict__ perm,
                              int64_t K, int Nc, int Nt, c64 *__restrict__ yint) {
    const int64_t n = K * Nc * (int64_t)Nt;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = i % K;
        const int64_t fc = i / K;
        const int64_t f = fc / Nc;
        yint[i] = yref[fc * K + perm[f * K + j]];
    }
}

// ------------------------------------------------------------------ host side

template <class T>
struct Dev;
template <>
struct Dev<float> {
    static c64 *smaps(mslr_sense_s *P) { return P->smaps32; }
    static c64 *tw(mslr_sense_s *P, int a) { return P->tw32[a]; }
    static c64 *ph(mslr_sense_s *P, int a) { return P->ph32[a]; }
};
template <>
struct Dev<double> {
    static c128 *smaps(mslr_sense_s *P) { return P->smaps64; }
    static c128 *tw(mslr_sense_s *P, int a) { return P->tw64[a]; }
    static c128 *ph(mslr_sense_s *P, int a) { return P->ph64[a]; }
};

// Kernels with more than 64 KB of dynamic LDS need the limit raised once.
static void allow_lds(const void *fn) {
    static std::vector<const void *> done;
    for (const void *d : done)
        if (d == fn) return;
    HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 256));
    done.push_back(fn);
}

template <class T>
static int64_t frames_per_chunk(mslr_sense_s *P, int64_t nframes) {
    const int64_t per = (int64_t)P->Nc * P->V * (int64_t)sizeof(C2<T>);
    int64_t F = (int64_t)(P->chunk_bytes / (size_t)per);
    if (F < 1) F = 1;
    if (F > nframes) F = nframes;
    if (F > 65535) F = 65535;
    return F;
}

template <class T>
static void check_lds(mslr_sense_s *P) {
    const size_t need = (size_t)P->NYZ * sizeof(C2<T>);
    MSLR_REQUIRE(need <= 160 * 1024 - 256, MSLR_ERR_UNSUPPORTED,
                 "Ny*Nz plane does not fit in LDS at this precision");
}

// ---- launch wrappers (one chunk of nf frames; ta = absolute first frame)
template <class T, class FX>
static void launch_fwd_x(mslr_sense_s *P, const C2<T> *X, int nscales, int64_t sstride, int nf,
                         C2<T> *Tb, FX fx) {
    mslr_ctx_s *ctx = P->ctx;
    const int LB = P->LBx, Nx = P->N[0];
    allow_lds((const void *)k_fwd_x<T, FX>);
    dim3 gx((unsigned)cdiv(P->NYZ, LB), (unsigned)nf);
    MSLR_PROF(ctx, "fwd_x");
    k_fwd_x<T, FX><<<gx, NTX, (size_t)LB * Nx * sizeof(C2<T>), ctx->stream>>>(
        X, nscales, sstride, Dev<T>::smaps(P), P->Nc, P->V, Nx, P->NYZ, LB, fx, Tb);
    HIP_LAUNCH_CHECK();
}

template <class T, class FY, class FZ>
static void launch_fwd_yz(mslr_sense_s *P, int64_t ta, int nf, const C2<T> *ysub, C2<T> *out,
                          bool ref_order, double *partial, const C2<T> *Tb, FY fy, FZ fz, int nt) {
    mslr_ctx_s *ctx = P->ctx;
    const int Nx = P->N[0];
    allow_lds((const void *)k_fwd_yz<T, FY, FZ>);
    dim3 gyz((unsigned)Nx, (unsigned)P->Nc, (unsigned)nf);
    MSLR_PROF(ctx, "fwd_yz");
    k_fwd_yz<T, FY, FZ><<<gyz, nt, (size_t)P->NYZ * sizeof(C2<T>), ctx->stream>>>(
        Tb, P->Nc, Nx, P->N[1], P->N[2], fy, fz, P->ptr + ta * (Nx + 1), P->off + ta * P->K,
        ref_order ? P->perm + ta * P->K : nullptr, P->K, Dev<T>::ph(P, 0), Dev<T>::ph(P, 1),
        Dev<T>::ph(P, 2), (T)P->scale, ysub, out, partial);
    HIP_LAUNCH_CHECK();
}

template <class T, class FY, class FZ>
static void launch_adj_yz(mslr_sense_s *P, int64_t ta, int nf, const C2<T> *y, bool ref_order,
                          C2<T> *Tb, FY fy, FZ fz, int nt) {
    mslr_ctx_s *ctx = P->ctx;
    const int Nx = P->N[0];
    allow_lds((const void *)k_adj_yz<T, FY, FZ>);
    dim3 gyz((unsigned)Nx, (unsigned)P->Nc, (unsigned)nf);
    MSLR_PROF(ctx, "adj_yz");
    k_adj_yz<T, FY, FZ><<<gyz, nt, (size_t)P->NYZ * sizeof(C2<T>), ctx->stream>>>(
        y, P->Nc, Nx, P->N[1], P->N[2], fy, fz, P->ptr + ta * (Nx + 1), P->off + ta * P->K,
        ref_order ? P->perm + ta * P->K : nullptr, P->K, Dev<T>::ph(P, 0), Dev<T>::ph(P, 1),
        Dev<T>::ph(P, 2), Tb);
    HIP_LAUNCH_CHECK();
}

template <class T, class FX>
static void launch_adj_x(mslr_sense_s *P, int nf, const C2<T> *Tb, C2<T> *xout, FX fx) {
    mslr_ctx_s *ctx = P->ctx;
    const int LB = P->LBx, Nx = P->N[0];
    allow_lds((const void *)k_adj_x<T, FX>);
    dim3 gx((unsigned)cdiv(P->NYZ, LB), (unsigned)nf);
    MSLR_PROF(ctx, "adj_x");
    k_adj_x<T, FX><<<gx, NTX, (size_t)LB * Nx * sizeof(C2<T>), ctx->stream>>>(
        Tb, Dev<T>::smaps(P), P->Nc, P->V, Nx, P->NYZ, LB, fx, (T)P->scale, xout);
    HIP_LAUNCH_CHECK();
}

// ---- dispatch: compile-time FFT lengths for the production / test sizes in
// FP32 with the default exponent sign, runtime lengths otherwise.
constexpr int fix_nt_yz(int ny, int nz) { return ny * nz > 2048 ? 512 : 256; }

template <int NX, int SGN>
using FixX = FftFix<float, NX, SGN, (MAXE_X * NTX) / NX, NTX>;
template <int NY, int NZ, int SGN>
using FixY = FftFix<float, NY, SGN, NZ, fix_nt_yz(NY, NZ)>;
template <int NY, int NZ, int SGN>
using FixZ = FftFix<float, NZ, SGN, NY, fix_nt_yz(NY, NZ)>;

#define MSLR_X_SIZES(M) M(90) M(64) M(32) M(12)
#define MSLR_YZ_SIZES(M) M(90, 60) M(64, 40) M(32, 24) M(12, 10)

template <class T>
static void fwd_x_any(mslr_sense_s *P, const C2<T> *X, int nscales, int64_t sstride, int nf,
                      C2<T> *Tb) {
    if constexpr (std::is_same<T, float>::value) {
        if (P->sgn == -1 && P->use_fixed) {
            switch (P->N[0]) {
#define CASE_X(n)                                                              \
    case n:                                                                    \
        launch_fwd_x<float>(P, X, nscales, sstride, nf, Tb, FixX<n, -1>{P->tw32[0]}); \
        return;
                MSLR_X_SIZES(CASE_X)
#undef CASE_X
                default: break;
            }
        }
    }
    launch_fwd_x<T>(P, X, nscales, sstride, nf, Tb,
                    FftRt<T, MAXE_X>{P->d[0], Dev<T>::tw(P, 0), P->sgn});
}

template <class T>
static void fwd_yz_any(mslr_sense_s *P, int64_t ta, int nf, const C2<T> *ysub, C2<T> *out,
                       bool ref_order, double *partial, const C2<T> *Tb) {
    if constexpr (std::is_same<T, float>::value) {
        if (P->sgn == -1 && P->use_fixed) {
            const int key = P->N[1] * 10000 + P->N[2];
            switch (key) {
#define CASE_YZ(ny, nz)                                                                   \
    case ny * 10000 + nz:                                                                 \
        launch_fwd_yz<float>(P, ta, nf, ysub, out, ref_order, partial, Tb,                \
                             FixY<ny, nz, -1>{P->tw32[1]}, FixZ<ny, nz, -1>{P->tw32[2]},   \
                             fix_nt_yz(ny, nz));                                          \
        return;
                MSLR_YZ_SIZES(CASE_YZ)
#undef CASE_YZ
                default: break;
            }
        }
    }
    launch_fwd_yz<T>(P, ta, nf, ysub, out, ref_order, partial, Tb,
                     FftRt<T, MAXE_YZ>{P->d[1], Dev<T>::tw(P, 1), P->sgn},
                     FftRt<T, MAXE_YZ>{P->d[2], Dev<T>::tw(P, 2), P->sgn}, P->nt_yz);
}

template <class T>
static void adj_yz_any(mslr_sense_s *P, int64_t ta, int nf, const C2<T> *y, bool ref_order,
                       C2<T> *Tb) {
    if constexpr (std::is_same<T, float>::value) {
        if (P->sgn == -1 && P->use_fixed) {
            const int key = P->N[1] * 10000 + P->N[2];
            switch (key) {
#define CASE_YZ(ny, nz)                                                                   \
    case ny * 10000 + nz:                                                                 \
        launch_adj_yz<float>(P, ta, nf, y, ref_order, Tb, FixY<ny, nz, 1>{P->tw32[1]},    \
                             FixZ<ny, nz, 1>{P->tw32[2]}, fix_nt_yz(ny, nz));              \
        return;
                MSLR_YZ_SIZES(CASE_YZ)
#undef CASE_YZ
                default: break;
            }
        }
    }
    launch_adj_yz<T>(P, ta, nf, y, ref_order, Tb,
                     FftRt<T, MAXE_YZ>{P->d[1], Dev<T>::tw(P, 1), -P->sgn},
                     FftRt<T, MAXE_YZ>{P->d[2], Dev<T>::tw(P, 2), -P->sgn}, P->nt_yz);
}

template <class T>
static void adj_x_any(mslr_sense_s *P, int nf, const C2<T> *Tb, C2<T> *xout) {
    if constexpr (std::is_same<T, float>::value) {
        if (P->sgn == -1 && P->use_fixed) {
            switch (P->N[0]) {
#define CASE_X(n)                                                 \
    case n:                                                       \
        launch_adj_x<float>(P, nf, Tb, xout, FixX<n, 1>{P->tw32[0]}); \
        return;
                MSLR_X_SIZES(CASE_X)
#undef CASE_X
                default: break;
            }
        }
    }
    launch_adj_x<T>(P, nf, Tb, xout, FftRt<T, MAXE_X>{P->d[0], Dev<T>::tw(P, 0), -P->sgn});
}

// Forward over frames [t0, t0+nframes) of the plan; X holds local frames.
template <class T>
static void run_forward(mslr_sense_s *P, const C2<T> *X, int nscales, int64_t sstride,
                        int64_t t0, int64_t nframes, const C2<T> *ysub, C2<T> *out,
                        bool ref_order, double *partial) {
    mslr_ctx_s *ctx = P->ctx;
    check_lds<T>(P);
    const int64_t F = frames_per_chunk<T>(P, nframes);
    C2<T> *Tb = (C2<T> *)ctx->scratch_get(SLOT_FFT, (size_t)F * P->Nc * P->V * sizeof(C2<T>));
    for (int64_t f0 = 0; f0 < nframes; f0 += F) {
        const int nf = (int)std::min<int64_t>(F, nframes - f0);
        fwd_x_any<T>(P, X + f0 * P->V, nscales, sstride, nf, Tb);
        fwd_yz_any<T>(P, t0 + f0, nf, ysub ? ysub + f0 * P->Nc * P->K : nullptr,
                      out + f0 * P->Nc * P->K, ref_order,
                      partial ? partial + f0 * P->Nc * P->N[0] : nullptr, Tb);
    }
}

template <class T>
static void run_adjoint(mslr_sense_s *P, const C2<T> *y, int64_t t0, int64_t nframes,
                        C2<T> *xout, bool ref_order) {
    mslr_ctx_s *ctx = P->ctx;
    check_lds<T>(P);
    const int64_t F = frames_per_chunk<T>(P, nframes);
    C2<T> *Tb = (C2<T> *)ctx->scratch_get(SLOT_FFT, (size_t)F * P->Nc * P->V * sizeof(C2<T>));
    for (int64_t f0 = 0; f0 < nframes; f0 += F) {
        const int nf = (int)std::min<int64_t>(F, nframes - f0);
        adj_yz_any<T>(P, t0 + f0, nf, y + f0 * P->Nc * P->K, ref_order, Tb);
        adj_x_any<T>(P, nf, Tb, xout + f0 * P->V);
    }
}

void sense_ref_to_internal(mslr_sense_s *P, const c64 *y_ref, c64 *y_int) {
    const int64_t n = P->K * P->Nc * (int64_t)P->Nt;
    if (n == 0) return;
    int64_t nb = std::min<int64_t>(cdiv(n, 256), 65536);
    k_perm_gather<<<(unsigned)nb, 256, 0, P->ctx->stream>>>(y_ref, P->perm, P->K, P->Nc, P->Nt,
                                                            y_int);
    HIP_LAUNCH_CHECK();
}

void sense_grad_internal(mslr_sense_s *P, const c64 *X, int nscales, int64_t sstride,
                         const c64 *y_int, c64 *g, double *dc_partial) {
    mslr_ctx_s *ctx = P->ctx;
    // Residuals of one frame chunk live in scratch; A and A' alternate per chunk.
    const int64_t F = frames_per_chunk<float>(P, P->Nt);
    c64 *r = (c64 *)ctx->scratch_get(SLOT_RESID, (size_t)std::max<int64_t>(1, F * P->Nc * P->K) *
                                                     sizeof(c64));
    for (int64_t f0 = 0; f0 < P->Nt; f0 += F) {
        const int64_t nf = std::min<int64_t>(F, P->Nt - f0);
        run_forward<float>(P, X + f0 * P->V, nscales, sstride, f0, nf,
                           y_int + f0 * P->Nc * P->K, r, false,
                           dc_partial ? dc_partial + f0 * P->Nc * P->N[0] : nullptr);
        run_adjoint<float>(P, r, f0, nf, g + f0 * P->V, false);
    }
}

void sense_adjoint_internal(mslr_sense_s *P, const c64 *y_int, c64 *x) {
    run_adjoint<float>(P, y_int, 0, P->Nt, x, false);
}

// exp(2 pi i num/den) with exact values on quarter turns.
static c128 unit_phase(int64_t num, int64_t den) {
    int64_t m = ((num % den) + den) % den;
    if ((4 * m) % den == 0) {
        switch ((4 * m) / den) {
            case 0: return {1.0, 0.0};
            case 1: return {0.0, 1.0};
            case 2: return {-1.0, 0.0};
            default: return {0.0, -1.0};
        }
    }
    const double a = 2.0 * M_PI * (double)m / (double)den;
    return {std::cos(a), std::sin(a)};
}

}  // namespace mslr

using namespace mslr;

static void sense_free(mslr_sense_s *P) {
    if (!P) return;
    for (void *p : P->allocs) hipFree(p);
    delete P;
}

template <class U>
static U *upload_vec(mslr_sense_s *P, const std::vector<U> &h) {
    U *d = nullptr;
    const size_t bytes = std::max<size_t>(1, h.size()) * sizeof(U);
    HIP_CHECK(hipMalloc((void **)&d, bytes));
    P->allocs.push_back(d);
    if (!h.empty())
        HIP_CHECK(hipMemcpy(d, h.data(), h.size() * sizeof(U), hipMemcpyHostToDevice));
    return d;
}

extern "C" {

mslr_status_t mslr_sense_create(mslr_ctx ctx, const int64_t N[3], int64_t Nc, int64_t Nt,
                                const float *smaps_host, const uint8_t *mask_host,
                                int fft_forward, int unitary, mslr_sense *out) {
    return guarded([&] {
        MSLR_REQUIRE(out != nullptr, MSLR_ERR_ARG, "out is null");
        *out = nullptr;
        mslr_ctx_s *c = check_ctx(ctx);
        MSLR_REQUIRE(N && smaps_host && mask_host, MSLR_ERR_ARG, "null argument");
        MSLR_REQUIRE(N[0] >= 1 && N[1] >= 1 && N[2] >= 1 && Nc >= 1 && Nt >= 1, MSLR_ERR_ARG,
                     "sizes must be positive");
        auto *P = new mslr_sense_s();
        try {
            P->ctx = c;
            for (int a = 0; a < 3; ++a) {
                P->N[a] = (int)N[a];
                MSLR_REQUIRE(fft_make_desc(P->N[a], P->d[a]), MSLR_ERR_UNSUPPORTED,
                             "FFT length " + std::to_string(N[a]) +
                                 " must factor into 2, 3 and 5");
            }
            const int Nx = P->N[0], Ny = P->N[1], Nz = P->N[2];
            P->V = (int64_t)Nx * Ny * Nz;
            P->NYZ = Ny * Nz;
            MSLR_REQUIRE(P->NYZ <= MAX_NYZ, MSLR_ERR_UNSUPPORTED, "Ny*Nz must be <= 8192");
            MSLR_REQUIRE(Nx <= MAXE_X * NTX, MSLR_ERR_UNSUPPORTED, "Nx must be <= 4096");
            P->Nc = (int)Nc;
            P->Nt = (int)Nt;
            P->sgn = fft_forward ? -1 : 1;
            P->scale = unitary ? 1.0 / std::sqrt((double)P->V) : 1.0;
            // fft_forward = false: the forward applies the normalised inverse DFT
            // (src/sense_gpu.jl:77), so the 1/prod(N) joins the scale.
            if (!fft_forward) P->scale /= (double)P->V;
            P->LBx = std::max(1, std::min(P->NYZ, (MAXE_X * NTX) / Nx));
            int nt = (int)cdiv(P->NYZ, MAXE_YZ);
            nt = (int)(cdiv(nt, 64) * 64);
            P->nt_yz = std::max(64, std::min(NTYZ_MAX, nt));
            P->chunk_bytes = (size_t)512 << 20;
            P->use_fixed = getenv("MSLR_FFT_GENERIC") == nullptr;
            if (const char *e = getenv("MSLR_FFT_CHUNK_MB")) {
                long mb = atol(e);
                if (mb > 0) P->chunk_bytes = (size_t)mb << 20;
            }

            // ---- mask -> sample tables
            const int64_t V = P->V;
            int64_t K = -1;
            std::vector<int> ptr((size_t)(Nx + 1) * Nt);
            std::vector<int64_t> counts(Nt, 0);
            for (int64_t t = 0; t < Nt; ++t) {
                const uint8_t *m = mask_host + t * V;
                int64_t k = 0;
                for (int64_t i = 0; i < V; ++i) k += m[i] != 0;
                counts[t] = k;
                if (t == 0) K = k;
                MSLR_REQUIRE(k == K, MSLR_ERR_SHAPE,
                             "Frame " + std::to_string(t + 1) + " has a different sample count");
            }
            MSLR_REQUIRE(K <= INT32_MAX, MSLR_ERR_UNSUPPORTED, "too many samples per frame");
            P->K = K;
            std::vector<int> off((size_t)std::max<int64_t>(1, K * Nt));
            std::vector<int> perm((size_t)std::max<int64_t>(1, K * Nt));
            std::vector<int> cnt(Nx);
            for (int64_t t = 0; t < Nt; ++t) {
                const uint8_t *m = mask_host + t * V;
                std::fill(cnt.begin(), cnt.end(), 0);
                for (int64_t i = 0; i < V; ++i)
                    if (m[i]) cnt[i % Nx]++;
                int *pt = ptr.data() + t * (Nx + 1);
                pt[0] = 0;
                for (int x = 0; x < Nx; ++x) pt[x + 1] = pt[x] + cnt[x];
                std::vector<int> fill(pt, pt + Nx);
                int kref = 0;
                for (int64_t i = 0; i < V; ++i) {
                    if (!m[i]) continue;
                    const int x = (int)(i % Nx);
                    const int o = (int)(i / Nx);
                    const int j = fill[x]++;
                    off[t * K + j] = o;
                    perm[t * K + j] = kref++;
                }
            }
            P->ptr = upload_vec(P, ptr);
            P->off = upload_vec(P, off);
            P->perm = upload_vec(P, perm);

            // ---- twiddles and phases
            const int s = P->sgn;
            std::vector<c128> pre[3];
            for (int a = 0; a < 3; ++a) {
                const int n = P->N[a];
                const int64_t cc = n / 2;
                std::vector<c128> tw(n), ph(n);
                pre[a].resize(n);
                for (int m = 0; m < n; ++m) tw[m] = unit_phase(m, n);
                for (int k = 0; k < n; ++k) {
                    // post[k] = exp(2 pi i s (c^2 - c k)/n), pre[k] = exp(-2 pi i s c k / n)
                    ph[k] = unit_phase((int64_t)s * (cc * cc - cc * k), n);
                    pre[a][k] = unit_phase(-(int64_t)s * cc * k, n);
                }
                std::vector<c64> tw32(n), ph32(n);
                for (int m = 0; m < n; ++m) {
                    tw32[m] = {(float)tw[m].re, (float)tw[m].im};
                    ph32[m] = {(float)ph[m].re, (float)ph[m].im};
                }
                P->tw64[a] = upload_vec(P, tw);
                P->ph64[a] = upload_vec(P, ph);
                P->tw32[a] = upload_vec(P, tw32);
                P->ph32[a] = upload_vec(P, ph32);
            }
            std::vector<c128> sm64((size_t)V * Nc);
            std::vector<c64> sm32((size_t)V * Nc);
            for (int64_t ci = 0; ci < Nc; ++ci)
                for (int z = 0; z < Nz; ++z)
                    for (int y = 0; y < Ny; ++y) {
                        const c128 pyz = pre[1][y] * pre[2][z];
                        for (int x = 0; x < Nx; ++x) {
                            const int64_t i = x + (int64_t)Nx * (y + (int64_t)Ny * z);
                            const c128 ph = pre[0][x] * pyz;
                            const float *sp = smaps_host + 2 * (ci * V + i);
                            const c128 v = c128{(double)sp[0], (double)sp[1]} * ph;
                            sm64[ci * V + i] = v;
                            sm32[ci * V + i] = {(float)v.re, (float)v.im};
                        }
                    }
            P->smaps64 = upload_vec(P, sm64);
            P->smaps32 = upload_vec(P, sm32);
            *out = P;
        } catch (...) {
            sense_free(P);
            throw;
        }
    });
}

mslr_status_t mslr_sense_destroy(mslr_sense plan) {
    return guarded([&] {
        if (!plan) return;
        plan->ctx->activate();
        HIP_CHECK(hipStreamSynchronize(plan->ctx->stream));
        sense_free(plan);
    });
}

int64_t mslr_sense_nsamples(mslr_sense plan) { return plan ? plan->K : -1; }

static void check_frames(mslr_sense_s *P, int64_t t0, int64_t nframes) {
    MSLR_REQUIRE(P, MSLR_ERR_ARG, "null plan");
    MSLR_REQUIRE(t0 >= 0 && nframes >= 0 && t0 + nframes <= P->Nt, MSLR_ERR_ARG,
                 "frame range outside the plan");
    P->ctx->activate();
}

mslr_status_t mslr_sense_forward(mslr_sense plan, const void *x_dev, void *y_dev, int64_t t0,
                                 int64_t nframes) {
    return guarded([&] {
        check_frames(plan, t0, nframes);
        if (nframes == 0) return;
        run_forward<float>(plan, (const c64 *)x_dev, 1, 0, t0, nframes, nullptr, (c64 *)y_dev,
                           true, nullptr);
    });
}

mslr_status_t mslr_sense_adjoint(mslr_sense plan, const void *y_dev, void *x_dev, int64_t t0,
                                 int64_t nframes) {
    return guarded([&] {
        check_frames(plan, t0, nframes);
        if (nframes == 0) return;
        run_adjoint<float>(plan, (const c64 *)y_dev, t0, nframes, (c64 *)x_dev, true);
    });
}

mslr_status_t mslr_sense_forward_f64(mslr_sense plan, const void *x_dev, void *y_dev, int64_t t0,
                                     int64_t nframes) {
    return guarded([&] {
        check_frames(plan, t0, nframes);
        if (nframes == 0) return;
        run_forward<double>(plan, (const c128 *)x_dev, 1, 0, t0, nframes, nullptr,
                            (c128 *)y_dev, true, nullptr);
    });
}

mslr_status_t mslr_sense_adjoint_f64(mslr_sense plan, const void *y_dev, void *x_dev, int64_t t0,
                                     int64_t nframes) {
    return guarded([&] {
        check_frames(plan, t0, nframes);
        if (nframes == 0) return;
        run_adjoint<double>(plan, (const c128 *)y_dev, t0, nframes, (c128 *)x_dev, true);
    });
}

mslr_status_t mslr_sense_grad(mslr_sense plan, const void *X_dev, int64_t nscales,
                              const void *y_dev, void *g_dev, double *dc_cost) {
    return guarded([&] {
        check_frames(plan, 0, plan ? plan->Nt : 0);
        MSLR_REQUIRE(nscales >= 1, MSLR_ERR_ARG, "nscales must be >= 1");
        mslr_sense_s *P = plan;
        mslr_ctx_s *ctx = P->ctx;
        const int64_t ny = P->K * P->Nc * (int64_t)P->Nt;
        // Residual in reference order for the whole run (public entry point).
        c64 *r = (c64 *)ctx->scratch_get(SLOT_RESID, (size_t)std::max<int64_t>(1, ny) * sizeof(c64));
        const int64_t npart = (int64_t)P->Nt * P->Nc * P->N[0];
        double *part = (double *)ctx->scratch_get(SLOT_REDUCE, (size_t)(npart + 1) * sizeof(double));
        run_forward<float>(P, (const c64 *)X_dev, (int)nscales, P->V * (int64_t)P->Nt, 0, P->Nt,
                           (const c64 *)y_dev, r, true, part);
        run_adjoint<float>(P, r, 0, P->Nt, (c64 *)g_dev, true);
        if (dc_cost) {
            sum_partials(ctx, part, npart, part + npart);
            double h = 0.0;
            HIP_CHECK(hipMemcpyAsync(&h, part + npart, sizeof(double), hipMemcpyDeviceToHost,
                                     ctx->stream));
            HIP_CHECK(hipStreamSynchronize(ctx->stream));
            *dc_cost = 0.5 * h;
        }
    });
}

mslr_status_t mslr_sense_poweriter(mslr_sense plan, int niter, double tol, double *sigma1,
                                   int *iters) {
    return guarded([&] {
        check_frames(plan, 0, plan ? plan->Nt : 0);
        MSLR_REQUIRE(sigma1 != nullptr, MSLR_ERR_ARG, "sigma1 is null");
        MSLR_REQUIRE(niter >= 1, MSLR_ERR_ARG, "niter must be >= 1");
        mslr_sense_s *P = plan;
        mslr_ctx_s *ctx = P->ctx;
        const int64_t nx = P->V * (int64_t)P->Nt;
        const int64_t ny = P->K * P->Nc * (int64_t)P->Nt;
        c128 *x = (c128 *)ctx->scratch_get(SLOT_PWR, (size_t)(nx + ny) * sizeof(c128));
        c128 *Ax = x + nx;
        {
            std::vector<c128> ones((size_t)nx, c128{1.0, 0.0});
            HIP_CHECK(hipMemcpy(x, ones.data(), (size_t)nx * sizeof(c128), hipMemcpyHostToDevice));
        }
        double ratio_old = INFINITY;
        int it = 0;
        double xn2 = (double)nx;  // ||ones||^2
        for (it = 1; it <= niter; ++it) {
            run_forward<double>(P, x, 1, 0, 0, P->Nt, nullptr, Ax, true, nullptr);
            const double an = std::sqrt(device_norm2<double>(ctx, Ax, ny));
            const double ratio = an / std::sqrt(xn2);
            if (std::fabs(ratio - ratio_old) / ratio < tol) break;
            ratio_old = ratio;
            run_adjoint<double>(P, Ax, 0, P->Nt, x, true);
            const double n2 = device_norm2<double>(ctx, x, nx);
            MSLR_REQUIRE(n2 > 0 && std::isfinite(n2), MSLR_ERR_NUMERIC,
                         "power iteration produced a zero or non-finite vector");
            device_scale<double>(ctx, x, nx, 1.0 / std::sqrt(n2));
            xn2 = device_norm2<double>(ctx, x, nx);
        }
        if (it > niter) it = niter;
        // sigma1 = ||A x|| / ||x|| at the returned x.
        run_forward<double>(P, x, 1, 0, 0, P->Nt, nullptr, Ax, true, nullptr);
        const double an = std::sqrt(device_norm2<double>(ctx, Ax, ny));
        *sigma1 = an / std::sqrt(device_norm2<double>(ctx, x, nx));
        if (iters) *iters = it;
    });
}

}  // extern "C"
