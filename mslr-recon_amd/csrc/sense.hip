// SENSE forward / adjoint over all frames (block-diagonal Asense stack).
//
// Forward for frame t, coil c (src/sense_gpu.jl:71-81):
//   y = scale * (fftshift o FFT3 o ifftshift)(S_c .* x_t) restricted to mask_t.
// The centred DFT is written as post_phase .* DFT(pre_phase .* v): the
// pre-phase is folded into the stored coil maps and the post-phase into the
// k-space gather, so no shift copies are made.  The mask samples whole or
// partial lines along kx at a subset of (ky, kz); only those lines are carried
// between the passes.  Per frame chunk:
//   sum_planes: scale sum, image -> plane-major [f][x][y + Ny*z];
//   fwd_yz    : per (x, frame), for each coil: map multiply, y- and z-axis FFT
//               of the plane in LDS, values on the sampled lines kept;
//   fwd_x     : per block of lines: x-axis FFT, gather of the sampled kx with
//               the post-phases (+ residual and its squared norm).
// The adjoint runs the mirror image (scatter -> inverse x FFT -> inverse yz
// FFT -> conj-coil combine -> usual layout) and is the exact conjugate
// transpose of the forward.
#include <algorithm>
#include <cmath>
#include <functional>
#include <mutex>
#include <string>
#include <type_traits>

#include "fft.hpp"
#include "fft2d.hpp"
#include "reduce.hpp"
#include "sense.hpp"

namespace mslr {

#ifndef MSLR_NTX
#define MSLR_NTX 256
#endif
constexpr int NTX = MSLR_NTX;  // threads of the x-axis passes
constexpr int MAXE_X = 16;   // elements per thread in the x-axis passes
constexpr int MAXE_YZ = 16;  // elements per thread in the yz-plane passes
constexpr int NTYZ_MAX = 512;  // threads of the runtime-length yz-plane passes (at most)
constexpr int MAX_NYZ = MAXE_YZ * NTYZ_MAX;

// ------------------------------------------------------------------ FFT policies
// A policy transforms `nlines` lines in LDS for runtime lengths (any 2/3/5 length);
// the compiled lengths use the two-stage transforms of fft2d.hpp instead.
template <class T, int MAXE>
struct FftRt {
    FftDesc d;
    const C2<T> *tw;
    int sgn;
    __device__ __forceinline__ void operator()(C2<T> *buf, int nlines, int lstride, int estride,
                                               const C2<T> *twp) const {
        fft_lines<T, MAXE>(buf, d, nlines, lstride, estride, twp, sgn);
    }
    __device__ __forceinline__ void operator()(C2<T> *buf, int nlines, int lstride,
                                               int estride) const {
        (*this)(buf, nlines, lstride, estride, tw);
    }
};

// A pointer into global memory hidden from loop-invariant code motion: in a loop
// over coils the per-thread loads through it then stay inside the loop instead of
// occupying registers across it.  It keeps the global address space, so the loads
// stay global (not flat) loads.  (The host pass has no address spaces; there the
// type is a plain pointer.)
#if defined(__HIP_DEVICE_COMPILE__)
using gc64p = const __attribute__((address_space(1))) c64 *;
#else
using gc64p = const c64 *;
#endif
__device__ __forceinline__ gc64p opaque_gptr(const c64 *p) {
    uint64_t v = (uint64_t)p;
    asm volatile("" : "+s"(v));
    return (gc64p)v;
}

// ------------------------------------------------------------------ kernels

// Image frames in the usual layout (x fastest) -> plane-major layout
// P[f][x][y + Ny*z], summing nscales images spaced sstride apart.  One
// workgroup per (block of TB yz-lines, frame), tile transposed through LDS.
template <class T>
__global__ __launch_bounds__(256) void k_sum_planes(const C2<T> *__restrict__ X, int nscales,
                                                    int64_t sstride, int Nx, int NYZ, int TB,
                                                    C2<T> *__restrict__ Pm) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C2<T> *tile = reinterpret_cast<C2<T> *>(smem_raw);
    const int f = blockIdx.y;
    const int l0 = blockIdx.x * TB;
    const int nl = min(TB, NYZ - l0);
    const int64_t V = (int64_t)Nx * NYZ;
    const C2<T> *src = X + (int64_t)f * V + (int64_t)l0 * Nx;
    for (int e = threadIdx.x; e < nl * Nx; e += blockDim.x) {
        C2<T> s = src[e];
        for (int k = 1; k < nscales; ++k) s = s + src[(int64_t)k * sstride + e];
        const int l = e / Nx;
        const int x = e - l * Nx;
        tile[x * (TB + 1) + l] = s;
    }
    __syncthreads();
    C2<T> *dst = Pm + (int64_t)f * V + l0;
    for (int e = threadIdx.x; e < nl * Nx; e += blockDim.x) {
        const int x = e / nl;
        const int l = e - x * nl;
        dst[(int64_t)x * NYZ + l] = tile[x * (TB + 1) + l];
    }
}

// Plane-major frames -> usual layout (the inverse of k_sum_planes, one scale).
template <class T>
__global__ __launch_bounds__(256) void k_planes_to_image(const C2<T> *__restrict__ Pm, int Nx,
                                                         int NYZ, int TB, C2<T> *__restrict__ X) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C2<T> *tile = reinterpret_cast<C2<T> *>(smem_raw);
    const int f = blockIdx.y;
    const int l0 = blockIdx.x * TB;
    const int nl = min(TB, NYZ - l0);
    const int64_t V = (int64_t)Nx * NYZ;
    const C2<T> *src = Pm + (int64_t)f * V + l0;
    for (int e = threadIdx.x; e < nl * Nx; e += blockDim.x) {
        const int x = e / nl;
        const int l = e - x * nl;
        tile[x * (TB + 1) + l] = src[(int64_t)x * NYZ + l];
    }
    __syncthreads();
    C2<T> *dst = X + (int64_t)f * V + (int64_t)l0 * Nx;
    for (int e = threadIdx.x; e < nl * Nx; e += blockDim.x) {
        const int l = e / Nx;
        const int x = e - l * Nx;
        dst[e] = tile[x * (TB + 1) + l];
    }
}

// Sampling tables of a frame chunk (absolute frame ta + f).
struct LineTabs {
    const int *lptr;     // (Nt+1) first global line of each frame
    const int *line_yz;  // (lines) ky + Ny*kz of each line, ascending within a frame
    const int *kbeg;     // (lines) first internal sample of the line within its frame
    const int *lk;       // (K, Nt) (l << 16) | kx of sample j, l = line within frame
};

// Forward pass 1 for runtime FFT lengths (any 2/3/5 size, FP32 or FP64): one
// workgroup per (x, coil, frame).  The summed image plane x is multiplied by the
// coil's map, transformed along y and z in LDS, and the values on the frame's
// sampled (ky, kz) lines are kept:
//   T[f][c][x][l] (l = line of the frame), the only part of k-space pass 2 needs.
template <class T, class FY, class FZ>
__global__ __launch_bounds__(NTYZ_MAX) void k_fwd_yz(
    const C2<T> *__restrict__ Pm, const C2<T> *__restrict__ smaps, int Nc, int Nx, int Ny,
    int Nz, FY fy, FZ fz, const int *__restrict__ lptr, const int *__restrict__ line_yz,
    int64_t ta, int nf, C2<T> *__restrict__ Tb) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C2<T> *buf = reinterpret_cast<C2<T> *>(smem_raw);
    // (frame, x, coil) with the frame fastest, numbered onto XCDs: the work-groups
    // resident on one XCD share one or two x-planes of coil maps
    const int64_t wid = xcd_order(blockIdx.x, gridDim.x);
    const int f = (int)(wid % nf);
    const int x = (int)((wid / nf) % Nx);
    const int c = (int)(wid / ((int64_t)nf * Nx));
    const int NYZ = Ny * Nz;
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int g0 = lptr[ta + f];
    const int Lf = lptr[ta + f + 1] - g0;
    const int *lines = line_yz + g0;
    C2<T> *Tf = Tb + (int64_t)Nc * Nx * (g0 - lptr[ta]);
    const C2<T> *pl = Pm + ((int64_t)f * Nx + x) * NYZ;
    const C2<T> *sm = smaps + ((int64_t)c * Nx + x) * NYZ;
    for (int i = tid; i < NYZ; i += nthr) buf[i] = pl[i] * sm[i];
    __syncthreads();
    fy(buf, Nz, Ny, 1);
    fz(buf, Ny, 1, Ny);
    C2<T> *dst = Tf + ((int64_t)c * Nx + x) * Lf;
    for (int l = tid; l < Lf; l += nthr) dst[l] = buf[lines[l]];
}

// Forward pass 2: one workgroup per (block of LB lines, coil, frame).  The
// lines are transformed along x in LDS and the sampled kx of each line are
// gathered with the post-phases and the scale (+ residual and its squared norm).
template <class T, class FX>
__global__ __launch_bounds__(NTX) void k_fwd_x(
    const C2<T> *__restrict__ Tb, int Nc, int Nx, int Ny, int LB, int nlb, FX fx, LineTabs tb,
    int64_t K, int64_t ta, const C2<T> *__restrict__ phx, const C2<T> *__restrict__ phy,
    const C2<T> *__restrict__ phz, T scale, const C2<T> *__restrict__ ysub,
    C2<T> *__restrict__ out, double *__restrict__ partial) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C2<T> *buf = reinterpret_cast<C2<T> *>(smem_raw);
    __shared__ double red[16];
    const int lb = blockIdx.x, c = blockIdx.y, f = blockIdx.z;
    const int tid = threadIdx.x;
    const int g0 = tb.lptr[ta + f];
    const int Lf = tb.lptr[ta + f + 1] - g0;
    const int l0 = lb * LB;
    if (l0 >= Lf) {  // uniform per workgroup
        if (partial && tid == 0) partial[((int64_t)f * Nc + c) * nlb + lb] = 0.0;
        return;
    }
    const int nl = min(LB, Lf - l0);
    C2<T> *phl = buf + LB * Nx;  // per line of the block: y and z post-phases times scale
    const int *lines = tb.line_yz + g0;
    for (int li = tid; li < nl; li += NTX) {
        const int yz = lines[l0 + li];
        phl[li] = scale * (phy[yz % Ny] * phz[yz / Ny]);
    }
    const C2<T> *src = Tb + (int64_t)Nc * Nx * (g0 - tb.lptr[ta]) + (int64_t)c * Nx * Lf + l0;
    {   // thread (line lq, x = x0, x0 + xs, ...): independent loads, one division
        const int lq = tid % nl, x0 = tid / nl, xs = NTX / nl;
        if (x0 < xs) {
#pragma unroll 4
            for (int x = x0; x < Nx; x += xs) buf[lq * Nx + x] = src[(int64_t)x * Lf + lq];
        }
    }
    __syncthreads();
    fx(buf, nl, Nx, 1);
    __syncthreads();
    const int *kb = tb.kbeg + g0;
    const int jb = kb[l0];
    const int je = l0 + nl < Lf ? kb[l0 + nl] : (int)K;
    const int64_t fa = (ta + f) * K;
    const int64_t obase = ((int64_t)f * Nc + c) * K;
    double acc = 0.0;
    for (int j = jb + tid; j < je; j += NTX) {
        const int v = tb.lk[fa + j];  // (line << 16) | kx
        const int l = v >> 16;
        const int kx = v & 0xFFFF;
        const C2<T> ph = phx[kx] * phl[l - l0];
        C2<T> val = buf[(l - l0) * Nx + kx] * ph;
        const int64_t oi = obase + j;
        if (ysub) val = val - ysub[oi];
        out[oi] = val;
        acc += (double)val.re * (double)val.re + (double)val.im * (double)val.im;
    }
    if (partial) {
        const double s = block_sum(acc, red);
        if (tid == 0) partial[((int64_t)f * Nc + c) * nlb + lb] = s;
    }
}

// Adjoint pass A: one workgroup per (block of LB lines, coil, frame).  The
// samples of the lines (times the conjugate post-phases) are scattered into
// zeroed lines, transformed along x, and written as T[f][c][x][l].
template <class T, class FX>
__global__ __launch_bounds__(NTX) void k_adj_x(
    const C2<T> *__restrict__ yin, int Nc, int Nx, int Ny, int LB, FX fx, LineTabs tb, int64_t K,
    int64_t ta, const C2<T> *__restrict__ phx, const C2<T> *__restrict__ phy,
    const C2<T> *__restrict__ phz, C2<T> *__restrict__ Tb) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C2<T> *buf = reinterpret_cast<C2<T> *>(smem_raw);
    const int lb = blockIdx.x, c = blockIdx.y, f = blockIdx.z;
    const int tid = threadIdx.x;
    const int g0 = tb.lptr[ta + f];
    const int Lf = tb.lptr[ta + f + 1] - g0;
    const int l0 = lb * LB;
    if (l0 >= Lf) return;  // uniform per workgroup
    const int nl = min(LB, Lf - l0);
    C2<T> *phl = buf + LB * Nx;  // per line of the block: y and z post-phases
    const int *lines = tb.line_yz + g0;
    for (int li = tid; li < nl; li += NTX) {
        const int yz = lines[l0 + li];
        phl[li] = phy[yz % Ny] * phz[yz / Ny];
    }
    for (int e = tid; e < nl * Nx; e += NTX) buf[e] = C2<T>{T(0), T(0)};
    __syncthreads();
    const int *kb = tb.kbeg + g0;
    const int jb = kb[l0];
    const int je = l0 + nl < Lf ? kb[l0 + nl] : (int)K;
    const int64_t fa = (ta + f) * K;
    const int64_t ibase = ((int64_t)f * Nc + c) * K;
    for (int j = jb + tid; j < je; j += NTX) {
        const int v = tb.lk[fa + j];  // (line << 16) | kx
        const int l = v >> 16;
        const int kx = v & 0xFFFF;
        const C2<T> ph = phx[kx] * phl[l - l0];
        buf[(l - l0) * Nx + kx] = cmulc(ph, yin[ibase + j]);
    }
    __syncthreads();
    fx(buf, nl, Nx, 1);
    __syncthreads();
    C2<T> *dst = Tb + (int64_t)Nc * Nx * (g0 - tb.lptr[ta]) + (int64_t)c * Nx * Lf + l0;
    {   // thread (line lq, x = x0, x0 + xs, ...), one division
        const int lq = tid % nl, x0 = tid / nl, xs = NTX / nl;
        if (x0 < xs) {
#pragma unroll 4
            for (int x = x0; x < Nx; x += xs) dst[(int64_t)x * Lf + lq] = buf[lq * Nx + x];
        }
    }
}

// The x passes of the gradient fused: one workgroup per (block of LB lines, coil,
// frame).  The lines are transformed along x, the sampled kx gathered with the
// post-phases and the scale, y subtracted (the residual and its squared norm), the
// residual scattered with the conjugate post-phases into the zeroed lines,
// transformed back along x and written over the lines it read: k_fwd_x (with ysub)
// followed by k_adj_x on its residual, without the residual's round trip through
// memory.  A block holds at most MAXE_X * NTX samples (LB * Nx <= MAXE_X * NTX), so
// each thread keeps its residuals in registers while the lines are cleared: one
// line buffer in LDS, four workgroups per CU.
template <class T, class FXF, class FXA>
__global__ __launch_bounds__(NTX) void k_x_resid(
    C2<T> *__restrict__ Tb, int Nc, int Nx, int Ny, int LB, int nlb, FXF fxf, FXA fxa, LineTabs tb,
    int64_t K, int64_t ta, const C2<T> *__restrict__ phx, const C2<T> *__restrict__ phy,
    const C2<T> *__restrict__ phz, T scale, const C2<T> *__restrict__ ysub,
    double *__restrict__ partial) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C2<T> *buf = reinterpret_cast<C2<T> *>(smem_raw);  // the block's lines
    C2<T> *pha = buf + LB * Nx;                         // per line: y, z post-phases
    __shared__ double red[16];
    const int lb = blockIdx.x, c = blockIdx.y, f = blockIdx.z;
    const int tid = threadIdx.x;
    const int g0 = tb.lptr[ta + f];
    const int Lf = tb.lptr[ta + f + 1] - g0;
    const int l0 = lb * LB;
    if (l0 >= Lf) {  // uniform per workgroup
        if (partial && tid == 0) partial[((int64_t)f * Nc + c) * nlb + lb] = 0.0;
        return;
    }
    const int nl = min(LB, Lf - l0);
    const int *lines = tb.line_yz + g0;
    for (int li = tid; li < nl; li += NTX) {
        const int yz = lines[l0 + li];
        pha[li] = phy[yz % Ny] * phz[yz / Ny];
    }
    C2<T> *src = Tb + (int64_t)Nc * Nx * (g0 - tb.lptr[ta]) + (int64_t)c * Nx * Lf + l0;
    const int lq = tid % nl, x0 = tid / nl, xs = NTX / nl;
    if (x0 < xs) {
#pragma unroll 4
        for (int x = x0; x < Nx; x += xs) buf[lq * Nx + x] = src[(int64_t)x * Lf + lq];
    }
    __syncthreads();
    fxf(buf, nl, Nx, 1);
    __syncthreads();
    const int *kb = tb.kbeg + g0;
    const int jb = kb[l0];
    const int je = l0 + nl < Lf ? kb[l0 + nl] : (int)K;
    const int64_t fa = (ta + f) * K;
    const int64_t obase = ((int64_t)f * Nc + c) * K;
    double acc = 0.0;
    C2<T> rv[MAXE_X];
    int rp[MAXE_X];
#pragma unroll
    for (int q = 0; q < MAXE_X; ++q) {
        const int j = jb + tid + q * NTX;
        rp[q] = -1;
        if (j < je) {
            const int v = tb.lk[fa + j];  // (line << 16) | kx
            const int l = v >> 16;
            const int kx = v & 0xFFFF;
            const C2<T> p = phx[kx] * pha[l - l0];
            const C2<T> val = buf[(l - l0) * Nx + kx] * (scale * p) - ysub[obase + j];
            acc += (double)val.re * (double)val.re + (double)val.im * (double)val.im;
            rv[q] = cmulc(p, val);
            rp[q] = (l - l0) * Nx + kx;
        }
    }
    __syncthreads();  // every gather done before the lines are cleared
    for (int e = tid; e < nl * Nx; e += NTX) buf[e] = C2<T>{T(0), T(0)};
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MAXE_X; ++q)
        if (rp[q] >= 0) buf[rp[q]] = rv[q];
    if (partial) {
        const double s2 = block_sum(acc, red);
        if (tid == 0) partial[((int64_t)f * Nc + c) * nlb + lb] = s2;
    }
    __syncthreads();
    fxa(buf, nl, Nx, 1);
    __syncthreads();
    if (x0 < xs) {
#pragma unroll 4
        for (int x = x0; x < Nx; x += xs) src[(int64_t)x * Lf + lq] = buf[lq * Nx + x];
    }
}

// ---- two-stage x passes (fft2d.hpp; CT or PFA split of Nx).  A workgroup of NT
// threads holds a block of nl lines in LDS with line stride Nx + 1 (odd: the lanes of
// a stage step across lines without bank conflicts).  The fused gradient pass and the
// forward-only and adjoint-only passes share these steps, so their results agree to
// the bit.

// Forward: stage A reads its points straight from the lines in memory (lanes over
// lines: contiguous), stage B transforms in LDS.  Ends synchronised.
template <class AX, int NT>
__device__ __forceinline__ void x2_forward(const c64 *__restrict__ src, int64_t Lf, int nl,
                                           c64 *buf, const c64 *__restrict__ twf, int tid) {
    constexpr int N1 = AX::N1, N2 = AX::N2, LS = AX::N + 1;
#pragma unroll 1
    for (int u = tid; u < nl * N2; u += NT) {
        const int l = u % nl, b = u / nl;
        c64 w[N1];
#pragma unroll
        for (int a = 0; a < N1; ++a) w[a] = src[(int64_t)AX::spatial(a, b) * Lf + l];
        DftP<float, N1, -1>::run(w);
        c64 *q = buf + l * LS + b;
#pragma unroll
        for (int a = 0; a < N1; ++a) q[N2 * a] = w[a];
    }
    __syncthreads();
#pragma unroll 1
    for (int u = tid; u < nl * N1; u += NT) {
        const int l = u % nl, a = u / nl;
        c64 *q = buf + l * LS + N2 * a;
        c64 w[N2];
#pragma unroll
        for (int b = 0; b < N2; ++b) w[b] = q[b];
        if constexpr (!AX::PFA) {
#pragma unroll
            for (int b = 1; b < N2; ++b) w[b] = w[b] * twf[b * N1 + a];
        }
        DftP<float, N2, -1>::run(w);
#pragma unroll
        for (int b = 0; b < N2; ++b) q[b] = w[b];
    }
    __syncthreads();
}

// Inverse: stage B in LDS (DFT, then the conjugate twiddles), stage A written back
// over the lines in memory.  Starts after the caller's barrier.
template <class AX, int NT>
__device__ __forceinline__ void x2_inverse(c64 *buf, int nl, const c64 *__restrict__ twi,
                                           c64 *__restrict__ dst, int64_t Lf, int tid) {
    constexpr int N1 = AX::N1, N2 = AX::N2, LS = AX::N + 1;
#pragma unroll 1
    for (int u = tid; u < nl * N1; u += NT) {
        const int l = u % nl, a = u / nl;
        c64 *q = buf + l * LS + N2 * a;
        c64 w[N2];
#pragma unroll
        for (int b = 0; b < N2; ++b) w[b] = q[b];
        DftP<float, N2, 1>::run(w);
        if constexpr (!AX::PFA) {
#pragma unroll
            for (int b = 1; b < N2; ++b) w[b] = w[b] * twi[b * N1 + a];
        }
#pragma unroll
        for (int b = 0; b < N2; ++b) q[b] = w[b];
    }
    __syncthreads();
#pragma unroll 1
    for (int u = tid; u < nl * N2; u += NT) {
        const int l = u % nl, b = u / nl;
        const c64 *q = buf + l * LS + b;
        c64 w[N1];
#pragma unroll
        for (int a = 0; a < N1; ++a) w[a] = q[N2 * a];
        DftP<float, N1, 1>::run(w);
#pragma unroll
        for (int a = 0; a < N1; ++a) dst[(int64_t)AX::spatial(a, b) * Lf + l] = w[a];
    }
}

// The block's frame-local geometry and its lines' y, z post-phases (into pha).
struct X2Block {
    int g0, Lf, l0, nl;
};
__device__ __forceinline__ bool x2_block(const LineTabs &tb, int64_t ta, int f, int lb, int LB,
                                         X2Block &B) {
    B.g0 = tb.lptr[ta + f];
    B.Lf = tb.lptr[ta + f + 1] - B.g0;
    B.l0 = lb * LB;
    if (B.l0 >= B.Lf) return false;
    B.nl = min(LB, B.Lf - B.l0);
    return true;
}
template <int NT>
__device__ __forceinline__ void x2_phases(const LineTabs &tb, const X2Block &B, int Ny,
                                          const c64 *__restrict__ phy,
                                          const c64 *__restrict__ phz, c64 *pha, int tid) {
    const int *lines = tb.line_yz + B.g0;
    for (int li = tid; li < B.nl; li += NT) {
        const int yz = lines[B.l0 + li];
        pha[li] = phy[yz % Ny] * phz[yz / Ny];
    }
}

// The fused x passes of the gradient: one workgroup per (block of LB lines, coil,
// frame).  Forward transform, the residual r = scale * post-phase * X - y at every
// sample (and its squared norm), conj(post-phase) * r back at the sample's frequency
// position, inverse transform written over the lines.  FULL: every line holds all
// Nx samples, so the residual covers every position and the lines need no clearing.
template <class AX, int NT, bool FULL>
__global__ __launch_bounds__(NT) void k_x_resid2(
    c64 *__restrict__ Tb, int Nc, int Ny, int LB, int nlb, const c64 *__restrict__ twf,
    const c64 *__restrict__ twi, LineTabs tb, int64_t K, int64_t ta,
    const c64 *__restrict__ phx, const c64 *__restrict__ phy, const c64 *__restrict__ phz,
    float scale, const c64 *__restrict__ ysub, double *__restrict__ partial) {
    constexpr int Nx = AX::N, LS = Nx + 1;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    c64 *buf = reinterpret_cast<c64 *>(smem_raw);  // the block's lines, stride LS
    c64 *pha = buf + LB * LS;                      // per line: y, z post-phases
    __shared__ double red[16];
    const int lb = blockIdx.x, c = blockIdx.y, f = blockIdx.z;
    const int tid = threadIdx.x;
    X2Block B;
    if (!x2_block(tb, ta, f, lb, LB, B)) {  // uniform per workgroup
        if (partial && tid == 0) partial[((int64_t)f * Nc + c) * nlb + lb] = 0.0;
        return;
    }
    x2_phases<NT>(tb, B, Ny, phy, phz, pha, tid);
    c64 *src = Tb + (int64_t)Nc * Nx * (B.g0 - tb.lptr[ta]) + (int64_t)c * Nx * B.Lf + B.l0;
    x2_forward<AX, NT>(src, B.Lf, B.nl, buf, twf, tid);
    const int *kb = tb.kbeg + B.g0;
    const int jb = kb[B.l0];
    const int je = B.l0 + B.nl < B.Lf ? kb[B.l0 + B.nl] : (int)K;
    const int64_t fa = (ta + f) * K;
    const int64_t obase = ((int64_t)f * Nc + c) * K;
    double acc = 0.0;
    c64 rv[MAXE_X];
    int rp[MAXE_X];
#pragma unroll
    for (int qi = 0; qi < MAXE_X; ++qi) {
        const int j = jb + tid + qi * NT;
        rp[qi] = -1;
        if (j < je) {
            // FULL: sample j of the block is (line (j - jb) / Nx, kx (j - jb) % Nx),
            // no table read; else the table's (line << 16) | kx
            int l, kx;
            if constexpr (FULL) {
                l = (j - jb) / Nx;
                kx = (j - jb) - l * Nx;
            } else {
                const int v = tb.lk[fa + j];
                l = (v >> 16) - B.l0;
                kx = v & 0xFFFF;
            }
            const c64 ph = phx[kx] * pha[l];
            const int pos = l * LS + AX::freq_position(kx);
            const c64 val = buf[pos] * (scale * ph) - ysub[obase + j];
            acc += (double)val.re * (double)val.re + (double)val.im * (double)val.im;
            rv[qi] = cmulc(ph, val);
            rp[qi] = pos;
        }
    }
    if constexpr (!FULL) {
        __syncthreads();  // every gather done before the lines are cleared
        for (int e = tid; e < B.nl * LS; e += NT) buf[e] = c64{0.f, 0.f};
        __syncthreads();
    }
    // (FULL: each position is read and written by the one thread owning its sample)
#pragma unroll
    for (int qi = 0; qi < MAXE_X; ++qi)
        if (rp[qi] >= 0) buf[rp[qi]] = rv[qi];
    if (partial) {
        const double s2 = block_sum(acc, red);
        if (tid == 0) partial[((int64_t)f * Nc + c) * nlb + lb] = s2;
    }
    __syncthreads();
    x2_inverse<AX, NT>(buf, B.nl, twi, src, B.Lf, tid);
}

// Forward x pass alone: the samples scale * post-phase * X (- y with ysub) into out
// (+ the block's squared norm).  Same steps and sample order as k_x_resid2.
template <class AX, int NT>
__global__ __launch_bounds__(NT) void k_fwd_x2(
    const c64 *__restrict__ Tb, int Nc, int Ny, int LB, int nlb, const c64 *__restrict__ twf,
    LineTabs tb, int64_t K, int64_t ta, const c64 *__restrict__ phx,
    const c64 *__restrict__ phy, const c64 *__restrict__ phz, float scale,
    const c64 *__restrict__ ysub, c64 *__restrict__ out, double *__restrict__ partial) {
    constexpr int Nx = AX::N, LS = Nx + 1;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    c64 *buf = reinterpret_cast<c64 *>(smem_raw);
    c64 *pha = buf + LB * LS;
    __shared__ double red[16];
    const int lb = blockIdx.x, c = blockIdx.y, f = blockIdx.z;
    const int tid = threadIdx.x;
    X2Block B;
    if (!x2_block(tb, ta, f, lb, LB, B)) {
        if (partial && tid == 0) partial[((int64_t)f * Nc + c) * nlb + lb] = 0.0;
        return;
    }
    x2_phases<NT>(tb, B, Ny, phy, phz, pha, tid);
    const c64 *src = Tb + (int64_t)Nc * Nx * (B.g0 - tb.lptr[ta]) + (int64_t)c * Nx * B.Lf + B.l0;
    x2_forward<AX, NT>(src, B.Lf, B.nl, buf, twf, tid);
    const int *kb = tb.kbeg + B.g0;
    const int jb = kb[B.l0];
    const int je = B.l0 + B.nl < B.Lf ? kb[B.l0 + B.nl] : (int)K;
    const int64_t fa = (ta + f) * K;
    const int64_t obase = ((int64_t)f * Nc + c) * K;
    double acc = 0.0;
    for (int j = jb + tid; j < je; j += NT) {
        const int v = tb.lk[fa + j];
        const int l = (v >> 16) - B.l0;
        const int kx = v & 0xFFFF;
        const c64 ph = phx[kx] * pha[l];
        c64 val = buf[l * LS + AX::freq_position(kx)] * (scale * ph);
        if (ysub) val = val - ysub[obase + j];
        out[obase + j] = val;
        acc += (double)val.re * (double)val.re + (double)val.im * (double)val.im;
    }
    if (partial) {
        const double s2 = block_sum(acc, red);
        if (tid == 0) partial[((int64_t)f * Nc + c) * nlb + lb] = s2;
    }
}

// Adjoint x pass alone: conj(post-phase) * y scattered into cleared lines, inverse
// transform written as T[f][c][x][l].  Same steps as the second half of k_x_resid2.
template <class AX, int NT>
__global__ __launch_bounds__(NT) void k_adj_x2(
    const c64 *__restrict__ yin, int Nc, int Ny, int LB, const c64 *__restrict__ twi,
    LineTabs tb, int64_t K, int64_t ta, const c64 *__restrict__ phx,
    const c64 *__restrict__ phy, const c64 *__restrict__ phz, c64 *__restrict__ Tb) {
    constexpr int Nx = AX::N, LS = Nx + 1;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    c64 *buf = reinterpret_cast<c64 *>(smem_raw);
    c64 *pha = buf + LB * LS;
    const int lb = blockIdx.x, c = blockIdx.y, f = blockIdx.z;
    const int tid = threadIdx.x;
    X2Block B;
    if (!x2_block(tb, ta, f, lb, LB, B)) return;
    x2_phases<NT>(tb, B, Ny, phy, phz, pha, tid);
    for (int e = tid; e < B.nl * LS; e += NT) buf[e] = c64{0.f, 0.f};
    __syncthreads();
    const int *kb = tb.kbeg + B.g0;
    const int jb = kb[B.l0];
    const int je = B.l0 + B.nl < B.Lf ? kb[B.l0 + B.nl] : (int)K;
    const int64_t fa = (ta + f) * K;
    const int64_t ibase = ((int64_t)f * Nc + c) * K;
    for (int j = jb + tid; j < je; j += NT) {
        const int v = tb.lk[fa + j];
        const int l = (v >> 16) - B.l0;
        const int kx = v & 0xFFFF;
        const c64 ph = phx[kx] * pha[l];
        buf[l * LS + AX::freq_position(kx)] = cmulc(ph, yin[ibase + j]);
    }
    __syncthreads();
    c64 *dst = Tb + (int64_t)Nc * Nx * (B.g0 - tb.lptr[ta]) + (int64_t)c * Nx * B.Lf + B.l0;
    x2_inverse<AX, NT>(buf, B.nl, twi, dst, B.Lf, tid);
}

// Adjoint pass B for runtime FFT lengths: one workgroup per (x, frame).  For each
// coil the frame's lines at x are scattered into a zeroed (y, z) plane, transformed
// along z and y, and accumulated times the conjugate coil map into the output plane
// (each thread owns the same plane elements for every coil); the scaled sum is the
// plane-major image plane.
template <class T, class FY, class FZ>
__global__ __launch_bounds__(NTYZ_MAX) void k_adj_yz(
    const C2<T> *__restrict__ Tb, const C2<T> *__restrict__ smaps, int Nc, int Nx, int Ny, int Nz,
    FY fy, FZ fz, const int *__restrict__ lptr, const int *__restrict__ line_yz, int64_t ta,
    int nf, T scale, C2<T> *__restrict__ Pm) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C2<T> *buf = reinterpret_cast<C2<T> *>(smem_raw);
    // (frame, x) with the frame fastest, numbered onto XCDs (coil-map reuse in L2)
    const int64_t wid = xcd_order(blockIdx.x, gridDim.x);
    const int f = (int)(wid % nf);
    const int x = (int)(wid / nf);
    const int NYZ = Ny * Nz;
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int g0 = lptr[ta + f];
    const int Lf = lptr[ta + f + 1] - g0;
    const int *lines = line_yz + g0;
    const C2<T> *Tf = Tb + (int64_t)Nc * Nx * (g0 - lptr[ta]);
    C2<T> *pl = Pm + ((int64_t)f * Nx + x) * NYZ;
#pragma unroll 1
    for (int c = 0; c < Nc; ++c) {
        for (int i = tid; i < NYZ; i += nthr) buf[i] = C2<T>{T(0), T(0)};
        __syncthreads();
        const C2<T> *src = Tf + ((int64_t)c * Nx + x) * Lf;
        for (int l = tid; l < Lf; l += nthr) buf[lines[l]] = src[l];
        __syncthreads();
        fz(buf, Ny, 1, Ny);
        fy(buf, Nz, Ny, 1);
        const C2<T> *sm = smaps + ((int64_t)c * Nx + x) * NYZ;
        for (int i = tid; i < NYZ; i += nthr) {
            const C2<T> v = cmulc(sm[i], buf[i]);
            pl[i] = c == 0 ? v : pl[i] + v;
        }
        __syncthreads();
    }
    for (int i = tid; i < NYZ; i += nthr) pl[i] = scale * pl[i];
}

// ------------------------------------------------------------------ two-stage y-z passes
// The compiled plane sizes as two-stage in-place transforms (fft2d.hpp): four stages
// per coil (y stage A, y stage B, z stage A, z stage B), each one barrier, every LDS
// offset a compile-time constant from the butterfly's base.  Coil maps come in the
// y-stage-A lane order (smp: for coil c and x, element a * UYA + u holds the map at
// the lane's point a), so map loads are contiguous across lanes; line_pos gives the
// LDS offset of each sampled line's (ky, kz).

// Stage butterflies: lane u of a stage handles butterfly u, u + NT, ... (a runtime
// loop, so a stage with several rounds holds one butterfly's registers at a time).
// y stage B (forward: twiddle, then DFT; inverse: DFT, then twiddle).  twl holds
// W^(a b) in the order b * N1y + a.
template <class PP, int SGN, bool FWD>
__device__ __forceinline__ void yz2_stage_yb(c64 *buf, gc64p twl, int tid) {
    constexpr int N1Y = PP::Y::N1, N2Y = PP::Y::N2;
#pragma unroll 1
    for (int u = tid; u < PP::UYB; u += PP::NT) {
        const int a = u % N1Y;
        c64 *p = buf + PP::lds(N2Y * a, u / N1Y);
        const gc64p t = twl + a;
        c64 w[N2Y];
#pragma unroll
        for (int b = 0; b < N2Y; ++b) w[b] = p[b];
        if constexpr (FWD && !PP::Y::PFA) {
#pragma unroll
            for (int b = 1; b < N2Y; ++b) w[b] = w[b] * t[b * N1Y];
        }
        DftP<float, N2Y, SGN>::run(w);
        if constexpr (!FWD && !PP::Y::PFA) {
#pragma unroll
            for (int b = 1; b < N2Y; ++b) w[b] = w[b] * t[b * N1Y];
        }
#pragma unroll
        for (int b = 0; b < N2Y; ++b) p[b] = w[b];
    }
}
template <class PP, int SGN>
__device__ __forceinline__ void yz2_stage_za(c64 *buf, int tid) {
    constexpr int NY = PP::NY, N1Z = PP::Z::N1, N2Z = PP::Z::N2;
#pragma unroll 1
    for (int u = tid; u < PP::UZA; u += PP::NT) {
        c64 *p = buf + PP::lds(u % NY, u / NY);
        c64 w[N1Z];
#pragma unroll
        for (int a = 0; a < N1Z; ++a) w[a] = p[PP::LY * N2Z * a];
        DftP<float, N1Z, SGN>::run(w);
#pragma unroll
        for (int a = 0; a < N1Z; ++a) p[PP::LY * N2Z * a] = w[a];
    }
}

// z stage B over the butterflies of a list only (the frame's butterflies that hold a
// sampled line): the forward gathers nothing else from the last stage, and in the
// adjoint every other butterfly's inputs are zero, so its outputs (zero) are already
// in place.
template <class PP, int SGN>
__device__ __forceinline__ void yz2_stage_zb_list(c64 *buf, int tid, const int *list, int cnt) {
    constexpr int NY = PP::NY, N2Z = PP::Z::N2;
#pragma unroll 1
    for (int i = tid; i < cnt; i += PP::NT) {
        const int u = list[i];
        c64 *p = buf + PP::lds(u % NY, N2Z * (u / NY));
        c64 w[N2Z];
#pragma unroll
        for (int b = 0; b < N2Z; ++b) w[b] = p[PP::LY * b];
        DftP<float, N2Z, SGN>::run(w);
#pragma unroll
        for (int b = 0; b < N2Z; ++b) p[PP::LY * b] = w[b];
    }
}

// Forward: one workgroup per (x, frame).  Per coil the y-stage-A lanes multiply their
// points of the image plane (held in registers across coils when each lane has one
// butterfly, KEEP; else re-read per coil, L2 hits) by the coil map and transform; the
// transformed plane's sampled lines are gathered into T[f][c][x][l].
template <class PP, int SGN, bool KEEP>
__global__ __launch_bounds__(PP::NT) void k_fwd_yz2(
    const c64 *__restrict__ Pm, const c64 *__restrict__ smp, int Nc, int Nx,
    const c64 *__restrict__ twy, const int *__restrict__ lptr, const int *__restrict__ line_pos,
    const int *__restrict__ zb_list, const int *__restrict__ zb_cnt, int64_t ta, int nf,
    c64 *__restrict__ Tb) {
    constexpr int NT = PP::NT, NY = PP::NY, NYZ = PP::NYZ;
    constexpr int N1Y = PP::Y::N1, N2Y = PP::Y::N2;
    __shared__ c64 buf[PP::LY * PP::NZ];
    // (frame, x) with the frame fastest, numbered onto XCDs (coil-map reuse in L2)
    const int64_t wid = xcd_order(blockIdx.x, gridDim.x);
    const int f = (int)(wid % nf);
    const int x = (int)(wid / nf);
    const int tid = threadIdx.x;
    const int g0 = lptr[ta + f];
    const int Lf = lptr[ta + f + 1] - g0;
    const int *lpos = line_pos + g0;
    const int *zbl = zb_list + g0;
    const int nzb = zb_cnt[ta + f];
    c64 *Tf = Tb + (int64_t)Nc * Nx * (g0 - lptr[ta]);
    const c64 *pl = Pm + ((int64_t)f * Nx + x) * NYZ;
    // KEEP: the y-stage-A lanes' image points in registers across coils (one round)
    c64 xk[KEEP ? N1Y : 1];
    if constexpr (KEEP) {
        static_assert(!KEEP || PP::RYA == 1, "image points kept for one round only");
        if (tid < PP::UYA) {
            const c64 *row = pl + PP::zspat(tid / N2Y) * NY;
#pragma unroll
            for (int a = 0; a < N1Y; ++a) xk[a] = row[PP::Y::spatial(a, tid % N2Y)];
        }
    }
#pragma unroll 1
    for (int c = 0; c < Nc; ++c) {
        const gc64p sm = opaque_gptr(smp + ((int64_t)c * Nx + x) * NYZ);
        const gc64p plc = opaque_gptr(pl);
        __syncthreads();  // the previous coil's gather has read the plane
#pragma unroll 1
        for (int u = tid; u < PP::UYA; u += NT) {
            const int b = u % N2Y, pz = u / N2Y;
            const gc64p row = plc + PP::zspat(pz) * NY;
            c64 v[N1Y];
#pragma unroll
            for (int a = 0; a < N1Y; ++a) {
                if constexpr (KEEP) v[a] = xk[a] * sm[a * PP::UYA + u];
                else v[a] = row[PP::Y::spatial(a, b)] * sm[a * PP::UYA + u];
            }
            DftP<float, N1Y, SGN>::run(v);
            c64 *p = buf + PP::lds(b, pz);
#pragma unroll
            for (int a = 0; a < N1Y; ++a) p[N2Y * a] = v[a];
        }
        __syncthreads();
        yz2_stage_yb<PP, SGN, true>(buf, opaque_gptr(twy), tid);
        __syncthreads();
        yz2_stage_za<PP, SGN>(buf, tid);
        __syncthreads();
        yz2_stage_zb_list<PP, SGN>(buf, tid, zbl, nzb);
        __syncthreads();
        c64 *dst = Tf + ((int64_t)c * Nx + x) * Lf;
        for (int l = tid; l < Lf; l += NT) dst[l] = buf[lpos[l]];
    }
}

// Adjoint: one workgroup per (x, frame).  Per coil the frame's lines at x are
// scattered into the (zeroed) plane, the four stages run in reverse with the
// opposite exponent sign, and the y-stage-A lanes accumulate conj(map) times their
// points in registers, writing zeros back so the plane is clear for the next coil.
template <class PP, int SGN>
__global__ __launch_bounds__(PP::NT) void k_adj_yz2(
    const c64 *__restrict__ Tb, const c64 *__restrict__ smp, int Nc, int Nx,
    const c64 *__restrict__ twy, const int *__restrict__ lptr, const int *__restrict__ line_pos,
    const int *__restrict__ zb_list, const int *__restrict__ zb_cnt, int64_t ta, int nf,
    float scale, c64 *__restrict__ Pm) {
    constexpr int NT = PP::NT, NY = PP::NY, NYZ = PP::NYZ;
    constexpr int N1Y = PP::Y::N1, N2Y = PP::Y::N2;
    __shared__ c64 buf[PP::LY * PP::NZ];
    // (frame, x) with the frame fastest, numbered onto XCDs (coil-map reuse in L2)
    const int64_t wid = xcd_order(blockIdx.x, gridDim.x);
    const int f = (int)(wid % nf);
    const int x = (int)(wid / nf);
    const int tid = threadIdx.x;
    const int g0 = lptr[ta + f];
    const int Lf = lptr[ta + f + 1] - g0;
    const int *lpos = line_pos + g0;
    const int *zbl = zb_list + g0;
    const int nzb = zb_cnt[ta + f];
    const c64 *Tf = Tb + (int64_t)Nc * Nx * (g0 - lptr[ta]);

    for (int i = tid; i < PP::LY * PP::NZ; i += NT) buf[i] = c64{0.f, 0.f};
    c64 acc[PP::RYA][N1Y];
#pragma unroll
    for (int r = 0; r < PP::RYA; ++r)
#pragma unroll
        for (int a = 0; a < N1Y; ++a) acc[r][a] = c64{0.f, 0.f};
#pragma unroll 1
    for (int c = 0; c < Nc; ++c) {
        __syncthreads();  // the plane is clear
        const c64 *src = Tf + ((int64_t)c * Nx + x) * Lf;
        for (int l = tid; l < Lf; l += NT) buf[lpos[l]] = src[l];
        __syncthreads();
        yz2_stage_zb_list<PP, SGN>(buf, tid, zbl, nzb);
        __syncthreads();
        yz2_stage_za<PP, SGN>(buf, tid);
        __syncthreads();
        yz2_stage_yb<PP, SGN, false>(buf, opaque_gptr(twy), tid);
        __syncthreads();
        const gc64p sm = opaque_gptr(smp + ((int64_t)c * Nx + x) * NYZ);
#pragma unroll
        for (int r = 0; r < PP::RYA; ++r) {
            const int u = tid + r * NT;
            if (u < PP::UYA) {
                c64 *p = buf + PP::lds(u % N2Y, u / N2Y);
                c64 w[N1Y];
#pragma unroll
                for (int a = 0; a < N1Y; ++a) w[a] = p[N2Y * a];
#pragma unroll
                for (int a = 0; a < N1Y; ++a) p[N2Y * a] = c64{0.f, 0.f};
                DftP<float, N1Y, SGN>::run(w);
#pragma unroll
                for (int a = 0; a < N1Y; ++a)
                    acc[r][a] = acc[r][a] + cmulc(sm[a * PP::UYA + u], w[a]);
            }
        }
    }
#pragma unroll
    for (int r = 0; r < PP::RYA; ++r) {
        const int u = tid + r * NT;
        if (u < PP::UYA) {
            const int b = u % N2Y, pz = u / N2Y;
            c64 *row = Pm + ((int64_t)f * Nx + x) * NYZ + PP::zspat(pz) * NY;
#pragma unroll
            for (int a = 0; a < N1Y; ++a) row[PP::Y::spatial(a, b)] = scale * acc[r][a];
        }
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

// Frame chunks [t, t+n) of [t0, t0+nframes): the line buffer of a chunk
// (Nc*Nx*lines elements) and its plane-major image (nf*V) each fit chunk_bytes.
template <class T>
static std::vector<std::pair<int64_t, int64_t>> frame_chunks(mslr_sense_s *P, int64_t t0,
                                                             int64_t nframes) {
    std::vector<std::pair<int64_t, int64_t>> ch;
    const size_t el = sizeof(C2<T>);
    const size_t per_line = (size_t)P->Nc * P->N[0] * el;
    const size_t per_plane = (size_t)P->V * el;
    int64_t s = t0, n = 0;
    size_t lb = 0, pb = 0;
    for (int64_t t = t0; t < t0 + nframes; ++t) {
        const size_t l = per_line * (size_t)(P->lptr_h[t + 1] - P->lptr_h[t]);
        if (n > 0 && (lb + l > P->chunk_bytes || pb + per_plane > P->chunk_bytes || n >= 65535)) {
            ch.push_back({s, n});
            s = t;
            n = 0;
            lb = pb = 0;
        }
        lb += l;
        pb += per_plane;
        ++n;
    }
    if (n > 0) ch.push_back({s, n});
    return ch;
}

template <class T>
static size_t chunk_line_bytes(mslr_sense_s *P, int64_t t, int64_t n) {
    return std::max<size_t>(1, (size_t)P->Nc * P->N[0] * sizeof(C2<T>) *
                                   (size_t)(P->lptr_h[t + n] - P->lptr_h[t]));
}

template <class T>
static void check_lds(mslr_sense_s *P) {
    const size_t need = (size_t)P->NYZ * sizeof(C2<T>);
    MSLR_REQUIRE(need <= 160 * 1024 - 256, MSLR_ERR_UNSUPPORTED,
                 "Ny*Nz plane does not fit in LDS at this precision");
}

// yz-lines per workgroup of the transposes (LDS tile Nx x (TB+1))
template <class T>
static int transpose_tb(mslr_sense_s *P) {
    const int cap = (int)((150 * 1024) / ((size_t)P->N[0] * sizeof(C2<T>))) - 1;
    return std::max(1, std::min(64, std::min(cap, P->NYZ)));
}

// yz-lines per workgroup of the scale-sum transpose: 32 (a 90 x 33 tile, 23 KB: six
// workgroups per CU against three with 64 lines; c1 1.21 -> 0.94 ms, 16 lines 1.03)
template <class T>
static int sum_planes_tb(mslr_sense_s *P) {
    return std::min(32, transpose_tb<T>(P));
}

template <class T>
static void launch_sum_planes(mslr_sense_s *P, const C2<T> *X, int nscales, int64_t sstride,
                              int nf, C2<T> *Pm) {
    mslr_ctx_s *ctx = P->ctx;
    const int TB = sum_planes_tb<T>(P), Nx = P->N[0];
    allow_lds((const void *)k_sum_planes<T>);
    dim3 g((unsigned)cdiv(P->NYZ, TB), (unsigned)nf);
    MSLR_PROF(ctx, "sum_planes");
    k_sum_planes<T><<<g, 256, (size_t)Nx * (TB + 1) * sizeof(C2<T>), ctx->stream>>>(
        X, nscales, sstride, Nx, P->NYZ, TB, Pm);
    HIP_LAUNCH_CHECK();
}

template <class T>
static void launch_planes_to_image(mslr_sense_s *P, const C2<T> *Pm, int nf, C2<T> *X) {
    mslr_ctx_s *ctx = P->ctx;
    const int TB = transpose_tb<T>(P), Nx = P->N[0];
    allow_lds((const void *)k_planes_to_image<T>);
    dim3 g((unsigned)cdiv(P->NYZ, TB), (unsigned)nf);
    MSLR_PROF(ctx, "planes_to_image");
    k_planes_to_image<T><<<g, 256, (size_t)Nx * (TB + 1) * sizeof(C2<T>), ctx->stream>>>(
        Pm, Nx, P->NYZ, TB, X);
    HIP_LAUNCH_CHECK();
}

static LineTabs line_tabs(mslr_sense_s *P) {
    return LineTabs{P->lptr, P->line_yz, P->kbeg, P->lk};
}

// ---- launch wrappers (one chunk of nf frames; ta = absolute first frame)
template <class T, class FY, class FZ>
static void launch_fwd_yz(mslr_sense_s *P, int64_t ta, int nf, const C2<T> *Pm, C2<T> *Tb,
                          FY fy, FZ fz, int nt) {
    mslr_ctx_s *ctx = P->ctx;
    const int Nx = P->N[0];
    allow_lds((const void *)k_fwd_yz<T, FY, FZ>);
    const int64_t nwg = (int64_t)Nx * P->Nc * nf;
    MSLR_REQUIRE(nwg * nt < ((int64_t)1 << 32), MSLR_ERR_UNSUPPORTED, "too many frames per chunk");
    const unsigned g = (unsigned)nwg;
    MSLR_PROF(ctx, "fwd_yz");
    k_fwd_yz<T, FY, FZ><<<g, nt, (size_t)P->NYZ * sizeof(C2<T>), ctx->stream>>>(
        Pm, Dev<T>::smaps(P), P->Nc, Nx, P->N[1], P->N[2], fy, fz, P->lptr, P->line_yz, ta, nf,
        Tb);
    HIP_LAUNCH_CHECK();
}

template <class T, class FX>
static void launch_fwd_x(mslr_sense_s *P, int64_t ta, int nf, const C2<T> *Tb, const C2<T> *ysub,
                         C2<T> *out, double *partial, FX fx) {
    mslr_ctx_s *ctx = P->ctx;
    const int Nx = P->N[0];
    allow_lds((const void *)k_fwd_x<T, FX>);
    dim3 g((unsigned)P->nlb, (unsigned)P->Nc, (unsigned)nf);
    MSLR_PROF(ctx, "fwd_x");
    k_fwd_x<T, FX><<<g, NTX, (size_t)P->LBx * (Nx + 1) * sizeof(C2<T>), ctx->stream>>>(
        Tb, P->Nc, Nx, P->N[1], P->LBx, P->nlb, fx, line_tabs(P), P->K, ta,
        Dev<T>::ph(P, 0), Dev<T>::ph(P, 1), Dev<T>::ph(P, 2), (T)P->scale, ysub, out, partial);
    HIP_LAUNCH_CHECK();
}

template <class T, class FX>
static void launch_adj_x(mslr_sense_s *P, int64_t ta, int nf, const C2<T> *y, C2<T> *Tb, FX fx) {
    mslr_ctx_s *ctx = P->ctx;
    const int Nx = P->N[0];
    allow_lds((const void *)k_adj_x<T, FX>);
    dim3 g((unsigned)P->nlb, (unsigned)P->Nc, (unsigned)nf);
    MSLR_PROF(ctx, "adj_x");
    k_adj_x<T, FX><<<g, NTX, (size_t)P->LBx * (Nx + 1) * sizeof(C2<T>), ctx->stream>>>(
        y, P->Nc, Nx, P->N[1], P->LBx, fx, line_tabs(P), P->K, ta, Dev<T>::ph(P, 0),
        Dev<T>::ph(P, 1), Dev<T>::ph(P, 2), Tb);
    HIP_LAUNCH_CHECK();
}

template <class T, class FXF, class FXA>
static void launch_x_resid(mslr_sense_s *P, int64_t ta, int nf, C2<T> *Tb, const C2<T> *ysub,
                           double *partial, FXF fxf, FXA fxa) {
    mslr_ctx_s *ctx = P->ctx;
    const int Nx = P->N[0];
    allow_lds((const void *)k_x_resid<T, FXF, FXA>);
    dim3 g((unsigned)P->nlb, (unsigned)P->Nc, (unsigned)nf);
    MSLR_PROF(ctx, "x_resid");
    MSLR_REQUIRE((int64_t)P->LBx * Nx <= (int64_t)MAXE_X * NTX, MSLR_ERR_UNSUPPORTED,
                 "x-pass line block larger than the residual registers");
    k_x_resid<T, FXF, FXA><<<g, NTX, (size_t)P->LBx * (Nx + 1) * sizeof(C2<T>), ctx->stream>>>(
        Tb, P->Nc, Nx, P->N[1], P->LBx, P->nlb, fxf, fxa, line_tabs(P), P->K, ta, Dev<T>::ph(P, 0),
        Dev<T>::ph(P, 1), Dev<T>::ph(P, 2), (T)P->scale, ysub, partial);
    HIP_LAUNCH_CHECK();
}

template <class T, class FY, class FZ>
static void launch_adj_yz(mslr_sense_s *P, int64_t ta, int nf, const C2<T> *Tb, C2<T> *Pm, FY fy,
                          FZ fz, int nt) {
    mslr_ctx_s *ctx = P->ctx;
    const int Nx = P->N[0];
    allow_lds((const void *)k_adj_yz<T, FY, FZ>);
    const int64_t nwg = (int64_t)Nx * nf;
    MSLR_REQUIRE(nwg * nt < ((int64_t)1 << 32), MSLR_ERR_UNSUPPORTED, "too many frames per chunk");
    const unsigned g = (unsigned)nwg;
    MSLR_PROF(ctx, "adj_yz");
    k_adj_yz<T, FY, FZ><<<g, nt, (size_t)P->NYZ * sizeof(C2<T>), ctx->stream>>>(
        Tb, Dev<T>::smaps(P), P->Nc, Nx, P->N[1], P->N[2], fy, fz, P->lptr, P->line_yz, ta, nf,
        (T)P->scale, Pm);
    HIP_LAUNCH_CHECK();
}

// ---- dispatch: compile-time FFT lengths for the production / test sizes in
// FP32 with the default exponent sign, runtime lengths otherwise.
template <class T>
static bool fixed_ok(mslr_sense_s *P) {
    return std::is_same<T, float>::value && P->sgn == -1 && P->use_fixed;
}

// ---- two-stage y-z passes: the compiled plane plans (y split prime-factor where the
// factors are coprime, Cooley-Tukey for powers of two; z split prime-factor).  The
// 90 x 60 plane has three compiled thread counts per direction; the built-in ones are
// the measured fastest at 90x90x60, 300 frames, 18 coils (forward 768: 4.27 ms, its
// y-stage-A lanes holding their image points across coils, against 5.30 ms with 512,
// whose lanes take two butterflies and re-read the points per coil, and 4.25 ms with
// 1024; adjoint 768: 4.82 ms against 5.30 ms with 512 and 7.0 ms with 1024, whose 83
// VGPRs leave one workgroup per CU).
template <int NT>
using Plane90x60 = PlanePlan<PlaneAxis<10, 9, true>, PlaneAxis<12, 5, true>, NT>;
using Plane64x40 = PlanePlan<PlaneAxis<8, 8, false>, PlaneAxis<8, 5, true>, 512>;
using Plane32x24 = PlanePlan<PlaneAxis<8, 4, false>, PlaneAxis<8, 3, true>, 256>;
using Plane12x10 = PlanePlan<PlaneAxis<4, 3, true>, PlaneAxis<2, 5, true>, 64>;
#define MSLR_YZ2_PLANS(M) M(64, 40, Plane64x40) M(32, 24, Plane32x24) M(12, 10, Plane12x10)
constexpr int YZ2_NT_FWD = 768, YZ2_NT_ADJ = 768;

template <class PP>
static void launch_fwd_yz2(mslr_sense_s *P, int64_t ta, int nf, const c64 *Pm, c64 *Tb) {
    mslr_ctx_s *ctx = P->ctx;
    const int Nx = P->N[0];
    const int64_t nwg = (int64_t)Nx * nf;
    MSLR_REQUIRE(nwg * PP::NT < ((int64_t)1 << 32), MSLR_ERR_UNSUPPORTED, "too many frames per chunk");
    MSLR_PROF(ctx, "fwd_yz");
    k_fwd_yz2<PP, -1, PP::RYA == 1><<<(unsigned)nwg, PP::NT, 0, ctx->stream>>>(
        Pm, P->smaps_yz2, P->Nc, Nx, P->tw_yz2[0], P->lptr, P->line_pos, P->zb_list, P->zb_cnt,
        ta, nf, Tb);
    HIP_LAUNCH_CHECK();
}

template <class PP>
static void launch_adj_yz2(mslr_sense_s *P, int64_t ta, int nf, const c64 *Tb, c64 *Pm) {
    mslr_ctx_s *ctx = P->ctx;
    const int Nx = P->N[0];
    const int64_t nwg = (int64_t)Nx * nf;
    MSLR_REQUIRE(nwg * PP::NT < ((int64_t)1 << 32), MSLR_ERR_UNSUPPORTED, "too many frames per chunk");
    MSLR_PROF(ctx, "adj_yz");
    k_adj_yz2<PP, 1><<<(unsigned)nwg, PP::NT, 0, ctx->stream>>>(
        Tb, P->smaps_yz2, P->Nc, Nx, P->tw_yz2[1], P->lptr, P->line_pos, P->zb_list, P->zb_cnt,
        ta, nf, (float)P->scale, Pm);
    HIP_LAUNCH_CHECK();
}

// The two-stage passes for the plan's plane (tables built at plan creation); the
// 90 x 60 plane's thread count comes from the tuning (0 = built-in).
static bool fwd_yz2(mslr_sense_s *P, int64_t ta, int nf, const c64 *Pm, c64 *Tb) {
    if (!P->yz2) return false;
    if (P->N[1] == 90 && P->N[2] == 60) {
        const int t = P->ctx->tuning.yz_threads_fwd ? P->ctx->tuning.yz_threads_fwd : YZ2_NT_FWD;
        switch (t) {
            case 512: launch_fwd_yz2<Plane90x60<512>>(P, ta, nf, Pm, Tb); break;
            case 1024: launch_fwd_yz2<Plane90x60<1024>>(P, ta, nf, Pm, Tb); break;
            default: launch_fwd_yz2<Plane90x60<768>>(P, ta, nf, Pm, Tb); break;
        }
        return true;
    }
    switch (P->N[1] * 10000 + P->N[2]) {
#define CASE_YZ2(ny, nz, PP)                          \
    case ny * 10000 + nz:                             \
        launch_fwd_yz2<PP>(P, ta, nf, Pm, Tb);        \
        return true;
        MSLR_YZ2_PLANS(CASE_YZ2)
#undef CASE_YZ2
        default: return false;
    }
}

static bool adj_yz2(mslr_sense_s *P, int64_t ta, int nf, const c64 *Tb, c64 *Pm) {
    if (!P->yz2) return false;
    if (P->N[1] == 90 && P->N[2] == 60) {
        const int t = P->ctx->tuning.yz_threads_adj ? P->ctx->tuning.yz_threads_adj : YZ2_NT_ADJ;
        switch (t) {
            case 512: launch_adj_yz2<Plane90x60<512>>(P, ta, nf, Tb, Pm); break;
            case 1024: launch_adj_yz2<Plane90x60<1024>>(P, ta, nf, Tb, Pm); break;
            default: launch_adj_yz2<Plane90x60<768>>(P, ta, nf, Tb, Pm); break;
        }
        return true;
    }
    switch (P->N[1] * 10000 + P->N[2]) {
#define CASE_YZ2(ny, nz, PP)                          \
    case ny * 10000 + nz:                             \
        launch_adj_yz2<PP>(P, ta, nf, Tb, Pm);        \
        return true;
        MSLR_YZ2_PLANS(CASE_YZ2)
#undef CASE_YZ2
        default: return false;
    }
}

// ---- two-stage fused x passes: the compiled line lengths (prime-factor split where
// the factors are coprime, Cooley-Tukey otherwise)
constexpr int X2_NT = 512;
using XAxis90 = PlaneAxis<10, 9, true>;
using XAxis64 = PlaneAxis<8, 8, false>;
using XAxis32 = PlaneAxis<8, 4, false>;
using XAxis12 = PlaneAxis<4, 3, true>;
#define MSLR_X2_AXES(M) M(90, XAxis90) M(64, XAxis64) M(32, XAxis32) M(12, XAxis12)

template <class AX>
static void launch_x_resid2(mslr_sense_s *P, int64_t ta, int nf, c64 *Tb, const c64 *ysub,
                            double *partial) {
    mslr_ctx_s *ctx = P->ctx;
    dim3 g((unsigned)P->nlb, (unsigned)P->Nc, (unsigned)nf);
    MSLR_REQUIRE((int64_t)P->LBx * AX::N <= (int64_t)MAXE_X * X2_NT, MSLR_ERR_UNSUPPORTED,
                 "x-pass line block larger than the residual registers");
    const size_t lds = (size_t)P->LBx * (AX::N + 1) * sizeof(c64) + (size_t)P->LBx * sizeof(c64);
    auto go = [&](auto kern) {
        allow_lds((const void *)kern);
        kern<<<g, X2_NT, lds, ctx->stream>>>(Tb, P->Nc, P->N[1], P->LBx, P->nlb, P->tw_x2[0],
                                             P->tw_x2[1], line_tabs(P), P->K, ta, P->ph32[0],
                                             P->ph32[1], P->ph32[2], (float)P->scale, ysub,
                                             partial);
    };
    MSLR_PROF(ctx, "x_resid");
    if (P->full_lines)
        go(k_x_resid2<AX, X2_NT, true>);
    else
        go(k_x_resid2<AX, X2_NT, false>);
    HIP_LAUNCH_CHECK();
}

template <class AX>
static void launch_fwd_x2(mslr_sense_s *P, int64_t ta, int nf, const c64 *Tb, const c64 *ysub,
                          c64 *out, double *partial) {
    mslr_ctx_s *ctx = P->ctx;
    dim3 g((unsigned)P->nlb, (unsigned)P->Nc, (unsigned)nf);
    const size_t lds = (size_t)P->LBx * (AX::N + 1) * sizeof(c64) + (size_t)P->LBx * sizeof(c64);
    allow_lds((const void *)k_fwd_x2<AX, X2_NT>);
    MSLR_PROF(ctx, "fwd_x");
    k_fwd_x2<AX, X2_NT><<<g, X2_NT, lds, ctx->stream>>>(
        Tb, P->Nc, P->N[1], P->LBx, P->nlb, P->tw_x2[0], line_tabs(P), P->K, ta, P->ph32[0],
        P->ph32[1], P->ph32[2], (float)P->scale, ysub, out, partial);
    HIP_LAUNCH_CHECK();
}

template <class AX>
static void launch_adj_x2(mslr_sense_s *P, int64_t ta, int nf, const c64 *y, c64 *Tb) {
    mslr_ctx_s *ctx = P->ctx;
    dim3 g((unsigned)P->nlb, (unsigned)P->Nc, (unsigned)nf);
    const size_t lds = (size_t)P->LBx * (AX::N + 1) * sizeof(c64) + (size_t)P->LBx * sizeof(c64);
    allow_lds((const void *)k_adj_x2<AX, X2_NT>);
    MSLR_PROF(ctx, "adj_x");
    k_adj_x2<AX, X2_NT><<<g, X2_NT, lds, ctx->stream>>>(y, P->Nc, P->N[1], P->LBx, P->tw_x2[1],
                                                        line_tabs(P), P->K, ta, P->ph32[0],
                                                        P->ph32[1], P->ph32[2], Tb);
    HIP_LAUNCH_CHECK();
}

static bool fwd_x2(mslr_sense_s *P, int64_t ta, int nf, const c64 *Tb, const c64 *ysub, c64 *out,
                   double *partial) {
    if (!P->x2) return false;
    switch (P->N[0]) {
#define CASE_X2(n, AX)                                            \
    case n:                                                       \
        launch_fwd_x2<AX>(P, ta, nf, Tb, ysub, out, partial);     \
        return true;
        MSLR_X2_AXES(CASE_X2)
#undef CASE_X2
        default: return false;
    }
}

static bool adj_x2(mslr_sense_s *P, int64_t ta, int nf, const c64 *y, c64 *Tb) {
    if (!P->x2) return false;
    switch (P->N[0]) {
#define CASE_X2(n, AX)                           \
    case n:                                      \
        launch_adj_x2<AX>(P, ta, nf, y, Tb);     \
        return true;
        MSLR_X2_AXES(CASE_X2)
#undef CASE_X2
        default: return false;
    }
}

static bool x_resid2(mslr_sense_s *P, int64_t ta, int nf, c64 *Tb, const c64 *ysub,
                     double *partial) {
    if (!P->x2) return false;
    switch (P->N[0]) {
#define CASE_X2(n, AX)                                       \
    case n:                                                  \
        launch_x_resid2<AX>(P, ta, nf, Tb, ysub, partial);   \
        return true;
        MSLR_X2_AXES(CASE_X2)
#undef CASE_X2
        default: return false;
    }
}

template <class T>
static void fwd_yz_any(mslr_sense_s *P, int64_t ta, int nf, const C2<T> *Pm, C2<T> *Tb) {
    if constexpr (std::is_same<T, float>::value) {
        if (fixed_ok<T>(P) && fwd_yz2(P, ta, nf, Pm, Tb)) return;
    }
    launch_fwd_yz<T>(P, ta, nf, Pm, Tb, FftRt<T, MAXE_YZ>{P->d[1], Dev<T>::tw(P, 1), P->sgn},
                     FftRt<T, MAXE_YZ>{P->d[2], Dev<T>::tw(P, 2), P->sgn}, P->nt_yz);
}

template <class T>
static void fwd_x_any(mslr_sense_s *P, int64_t ta, int nf, const C2<T> *Tb, const C2<T> *ysub,
                      C2<T> *out, double *partial) {
    if constexpr (std::is_same<T, float>::value) {
        if (fixed_ok<T>(P) && fwd_x2(P, ta, nf, Tb, ysub, out, partial)) return;
    }
    launch_fwd_x<T>(P, ta, nf, Tb, ysub, out, partial,
                    FftRt<T, MAXE_X>{P->d[0], Dev<T>::tw(P, 0), P->sgn});
}

template <class T>
static void adj_x_any(mslr_sense_s *P, int64_t ta, int nf, const C2<T> *y, C2<T> *Tb) {
    if constexpr (std::is_same<T, float>::value) {
        if (fixed_ok<T>(P) && adj_x2(P, ta, nf, y, Tb)) return;
    }
    launch_adj_x<T>(P, ta, nf, y, Tb, FftRt<T, MAXE_X>{P->d[0], Dev<T>::tw(P, 0), -P->sgn});
}

template <class T>
static void x_resid_any(mslr_sense_s *P, int64_t ta, int nf, C2<T> *Tb, const C2<T> *ysub,
                        double *partial) {
    if constexpr (std::is_same<T, float>::value) {
        if (fixed_ok<T>(P) && x_resid2(P, ta, nf, Tb, ysub, partial)) return;
    }
    launch_x_resid<T>(P, ta, nf, Tb, ysub, partial,
                      FftRt<T, MAXE_X>{P->d[0], Dev<T>::tw(P, 0), P->sgn},
                      FftRt<T, MAXE_X>{P->d[0], Dev<T>::tw(P, 0), -P->sgn});
}

template <class T>
static void adj_yz_any(mslr_sense_s *P, int64_t ta, int nf, const C2<T> *Tb, C2<T> *Pm) {
    if constexpr (std::is_same<T, float>::value) {
        if (fixed_ok<T>(P) && adj_yz2(P, ta, nf, Tb, Pm)) return;
    }
    launch_adj_yz<T>(P, ta, nf, Tb, Pm, FftRt<T, MAXE_YZ>{P->d[1], Dev<T>::tw(P, 1), -P->sgn},
                     FftRt<T, MAXE_YZ>{P->d[2], Dev<T>::tw(P, 2), -P->sgn}, P->nt_yz);
}

// Forward of one chunk [ta, ta+nf): X holds the chunk's frames (nscales images
// spaced sstride apart), ysub/out/partial start at the chunk; Pm (nf*V elements)
// and Tb (the chunk's sampled lines) are the intermediates, both fully written
// before they are read.
template <class T>
static void forward_chunk_into(mslr_sense_s *P, const C2<T> *X, int nscales, int64_t sstride,
                               int64_t ta, int nf, const C2<T> *ysub, C2<T> *out,
                               double *partial, C2<T> *Pm, C2<T> *Tb) {
    launch_sum_planes<T>(P, X, nscales, sstride, nf, Pm);
    fwd_yz_any<T>(P, ta, nf, Pm, Tb);
    fwd_x_any<T>(P, ta, nf, Tb, ysub, out, partial);
}

template <class T>
static void forward_chunk(mslr_sense_s *P, const C2<T> *X, int nscales, int64_t sstride,
                          int64_t ta, int nf, const C2<T> *ysub, C2<T> *out, double *partial) {
    mslr_ctx_s *ctx = P->ctx;
    C2<T> *Pm = (C2<T> *)ctx->scratch_get(SLOT_PLANES, (size_t)nf * P->V * sizeof(C2<T>));
    C2<T> *Tb = (C2<T> *)ctx->scratch_get(SLOT_FFT, chunk_line_bytes<T>(P, ta, nf));
    forward_chunk_into<T>(P, X, nscales, sstride, ta, nf, ysub, out, partial, Pm, Tb);
}

template <class T>
static void adjoint_chunk(mslr_sense_s *P, const C2<T> *y, int64_t ta, int nf, C2<T> *xout) {
    mslr_ctx_s *ctx = P->ctx;
    C2<T> *Pm = (C2<T> *)ctx->scratch_get(SLOT_PLANES, (size_t)nf * P->V * sizeof(C2<T>));
    C2<T> *Tb = (C2<T> *)ctx->scratch_get(SLOT_FFT, chunk_line_bytes<T>(P, ta, nf));
    adj_x_any<T>(P, ta, nf, y, Tb);
    adj_yz_any<T>(P, ta, nf, Tb, Pm);
    launch_planes_to_image<T>(P, Pm, nf, xout);
}

// Gradient of one chunk [ta, ta+nf): g = A'(A sum_k X_k - y) with the x passes fused
// (the residual stays in LDS); partial (nullable) gets the chunk's ||r||^2 partials.
// xout == nullptr leaves g in the plane-major scratch (returned) for a consumer that
// reads it there.
template <class T>
static C2<T> *grad_chunk(mslr_sense_s *P, const C2<T> *X, int nscales, int64_t sstride, int64_t ta,
                         int nf, const C2<T> *ysub, double *partial, C2<T> *xout) {
    mslr_ctx_s *ctx = P->ctx;
    C2<T> *Pm = (C2<T> *)ctx->scratch_get(SLOT_PLANES, (size_t)nf * P->V * sizeof(C2<T>));
    C2<T> *Tb = (C2<T> *)ctx->scratch_get(SLOT_FFT, chunk_line_bytes<T>(P, ta, nf));
    launch_sum_planes<T>(P, X, nscales, sstride, nf, Pm);
    fwd_yz_any<T>(P, ta, nf, Pm, Tb);
    x_resid_any<T>(P, ta, nf, Tb, ysub, partial);
    adj_yz_any<T>(P, ta, nf, Tb, Pm);
    if (xout) launch_planes_to_image<T>(P, Pm, nf, xout);
    return Pm;
}

// Forward over frames [t0, t0+nframes) of the plan; X holds local frames.
template <class T>
static void run_forward(mslr_sense_s *P, const C2<T> *X, int nscales, int64_t sstride,
                        int64_t t0, int64_t nframes, const C2<T> *ysub, C2<T> *out,
                        double *partial) {
    check_lds<T>(P);
    for (auto [ta, nf] : frame_chunks<T>(P, t0, nframes)) {
        const int64_t f0 = ta - t0;
        forward_chunk<T>(P, X + f0 * P->V, nscales, sstride, ta, (int)nf,
                         ysub ? ysub + f0 * P->Nc * P->K : nullptr, out + f0 * P->Nc * P->K,
                         partial ? partial + f0 * P->Nc * P->nlb : nullptr);
    }
}

template <class T>
static void run_adjoint(mslr_sense_s *P, const C2<T> *y, int64_t t0, int64_t nframes,
                        C2<T> *xout) {
    check_lds<T>(P);
    for (auto [ta, nf] : frame_chunks<T>(P, t0, nframes)) {
        const int64_t f0 = ta - t0;
        adjoint_chunk<T>(P, y + f0 * P->Nc * P->K, ta, (int)nf, xout + f0 * P->V);
    }
}

void sense_ref_to_internal(mslr_sense_s *P, const c64 *y_ref, c64 *y_int) {
    const int64_t n = P->K * P->Nc * (int64_t)P->Nt;
    if (n == 0 || y_ref == y_int) return;
    HIP_CHECK(hipMemcpyAsync(y_int, y_ref, (size_t)n * sizeof(c64), hipMemcpyDeviceToDevice,
                             P->ctx->stream));
}

void sense_grad_internal(mslr_sense_s *P, const c64 *X, int nscales, int64_t sstride,
                         const c64 *y_int, c64 *g, double *dc_partial) {
    mslr_ctx_s *ctx = P->ctx;
    check_lds<float>(P);
    (void)ctx;
    for (auto [ta, nf] : frame_chunks<float>(P, 0, P->Nt))
        grad_chunk<float>(P, X + ta * P->V, nscales, sstride, ta, (int)nf, y_int + ta * P->Nc * P->K,
                          dc_partial ? dc_partial + ta * P->Nc * P->nlb : nullptr, g + ta * P->V);
}

std::vector<std::pair<int64_t, int64_t>> sense_grad_chunks(mslr_sense_s *P, int nch) {
    std::vector<std::pair<int64_t, int64_t>> out;
    nch = std::max(1, std::min(nch, P->Nt));
    for (int i = 0; i < nch; ++i) {
        const int64_t a = (int64_t)P->Nt * i / nch, b = (int64_t)P->Nt * (i + 1) / nch;
        if (b > a)
            for (auto c : frame_chunks<float>(P, a, b - a)) out.push_back(c);
    }
    return out;
}

void sense_grad_pipelined(mslr_sense_s *P, const c64 *X, int nscales, int64_t sstride,
                          const c64 *y_int, c64 *g, double *dc_partial,
                          const std::vector<std::pair<int64_t, int64_t>> &chunks,
                          const SensePipe &pipe,
                          const std::function<void(size_t, int64_t, int64_t, const c64 *)> &after) {
    mslr_ctx_s *ctx = P->ctx;
    check_lds<float>(P);
    hipStream_t main = ctx->stream;
    const size_t base0 = ctx->slot_base;
    const bool two = pipe.stream != nullptr && chunks.size() > 1;
    if (two) {
        HIP_CHECK(hipEventRecord(pipe.fork, main));
        HIP_CHECK(hipStreamWaitEvent(pipe.stream, pipe.fork, 0));
    }
    try {
        for (size_t i = 0; i < chunks.size(); ++i) {
            const auto [ta, nf] = chunks[i];
            const bool on_pipe = two && (i % 2 == 1);
            ctx->stream = on_pipe ? pipe.stream : main;
            ctx->slot_base = on_pipe ? pipe.slot_base : base0;
            const c64 *planes = grad_chunk<float>(
                P, X + ta * P->V, nscales, sstride, ta, (int)nf, y_int + ta * P->Nc * P->K,
                dc_partial ? dc_partial + ta * P->Nc * P->nlb : nullptr, g ? g + ta * P->V : nullptr);
            after(i, ta, nf, planes);
        }
    } catch (...) {
        ctx->stream = main;
        ctx->slot_base = base0;
        throw;
    }
    ctx->stream = main;
    ctx->slot_base = base0;
    if (two) {
        HIP_CHECK(hipEventRecord(pipe.join, pipe.stream));
        HIP_CHECK(hipStreamWaitEvent(main, pipe.join, 0));
    }
}

int sense_transpose_lines(mslr_sense_s *P) { return transpose_tb<float>(P); }

void sense_resid_internal(mslr_sense_s *P, const c64 *X, int nscales, int64_t sstride,
                          const c64 *y_int, double *dc_partial) {
    mslr_ctx_s *ctx = P->ctx;
    check_lds<float>(P);
    const auto chunks = frame_chunks<float>(P, 0, P->Nt);
    int64_t fmax = 1;
    for (auto [ta, nf] : chunks) fmax = std::max(fmax, nf);
    c64 *r = (c64 *)ctx->scratch_get(SLOT_RESID, (size_t)std::max<int64_t>(1, fmax * P->Nc * P->K) *
                                                     sizeof(c64));
    for (auto [ta, nf] : chunks)
        forward_chunk<float>(P, X + ta * P->V, nscales, sstride, ta, (int)nf,
                             y_int + ta * P->Nc * P->K, r, dc_partial + ta * P->Nc * P->nlb);
}

void sense_adjoint_internal(mslr_sense_s *P, const c64 *y_int, c64 *x) {
    run_adjoint<float>(P, y_int, 0, P->Nt, x);
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

// ------------------------------------------------ allocation-free forward
// The intermediates of mslr_sense_forward_noalloc live in this file-scope
// array, one copy per device, held until the process ends.  Host code never
// uses the array's name as an address: each device's address comes from
// hipGetSymbolAddress at context creation (sense_noalloc_init).
static __device__ __align__(256) c64 scratch[MSLR_NOALLOC_SCRATCH_BYTES / sizeof(c64)];

static std::mutex g_noalloc_mu;
static std::vector<c64 *> g_noalloc_ptr;  // per device ordinal

void sense_noalloc_init(int device) {
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if ((size_t)device < g_noalloc_ptr.size() && g_noalloc_ptr[(size_t)device]) return;
    c64 *p = nullptr;
    HIP_CHECK(hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch)));
    MSLR_REQUIRE(p != nullptr, MSLR_ERR_DEVICE, "no device address for the scratch array");
    if (g_noalloc_ptr.size() <= (size_t)device) g_noalloc_ptr.resize((size_t)device + 1, nullptr);
    g_noalloc_ptr[(size_t)device] = p;
}

static c64 *noalloc_scratch(int device) {
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    return (size_t)device < g_noalloc_ptr.size() ? g_noalloc_ptr[(size_t)device] : nullptr;
}

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Scratch bytes a frame chunk [t, t+n) needs: plane-major image + sampled lines.
static size_t noalloc_chunk_bytes(mslr_sense_s *P, int64_t t, int64_t n) {
    return align256((size_t)n * P->V * sizeof(c64)) + align256(chunk_line_bytes<float>(P, t, n));
}

// The same frames as mslr_sense_forward in chunks that fit the scratch array.
static void run_forward_noalloc(mslr_sense_s *P, const c64 *X, int64_t t0, int64_t nframes,
                                c64 *out) {
    check_lds<float>(P);
    c64 *base = noalloc_scratch(P->ctx->device);
    MSLR_REQUIRE(base != nullptr, MSLR_ERR_DEVICE,
                 "the scratch array of this device has no address (context not created)");
    for (int64_t t = t0; t < t0 + nframes; ++t) {
        const size_t need = noalloc_chunk_bytes(P, t, 1);
        if (need > (size_t)MSLR_NOALLOC_SCRATCH_BYTES)
            MSLR_FAIL(MSLR_ERR_UNSUPPORTED,
                      "frame " + std::to_string(t) + " needs " + std::to_string(need) +
                          " bytes of scratch; the allocation-free forward has " +
                          std::to_string((size_t)MSLR_NOALLOC_SCRATCH_BYTES));
    }
    // no profiling records here: they would create events
    mslr_ctx_s *ctx = P->ctx;
    const bool prof = ctx->prof;
    ctx->prof = false;
    try {
        int64_t ta = t0;
        while (ta < t0 + nframes) {
            int64_t nf = 1;
            while (ta + nf < t0 + nframes && nf < 65535 &&
                   noalloc_chunk_bytes(P, ta, nf + 1) <= (size_t)MSLR_NOALLOC_SCRATCH_BYTES)
                ++nf;
            c64 *Pm = base;
            c64 *Tb = base + align256((size_t)nf * P->V * sizeof(c64)) / sizeof(c64);
            forward_chunk_into<float>(P, X + (ta - t0) * P->V, 1, 0, ta, (int)nf, nullptr,
                                      out + (ta - t0) * P->Nc * P->K, nullptr, Pm, Tb);
            ta += nf;
        }
    } catch (...) {
        ctx->prof = prof;
        throw;
    }
    ctx->prof = prof;
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

// Tables of the two-stage y-z passes for a compiled plane plan: the coil maps in
// the y-stage-A lane order and the LDS offset of every sampled line.
template <class PP>
static void yz2_tables(mslr_sense_s *P, const std::vector<c64> &sm32,
                       const std::vector<int> &line_yz) {
    using AY = typename PP::Y;
    using AZ = typename PP::Z;
    const int Nx = P->N[0], Nc = P->Nc, NY = PP::NY, NYZ = PP::NYZ;
    std::vector<c64> smp(sm32.size());
    for (int64_t cx = 0; cx < (int64_t)Nc * Nx; ++cx) {
        const c64 *src = sm32.data() + cx * NYZ;
        c64 *dst = smp.data() + cx * NYZ;
        for (int u = 0; u < PP::UYA; ++u) {
            const int b = u % AY::N2, z = PP::zspat(u / AY::N2);
            for (int a = 0; a < AY::N1; ++a) dst[a * PP::UYA + u] = src[AY::spatial(a, b) + NY * z];
        }
    }
    std::vector<int> lp(line_yz.size());
    for (size_t l = 0; l < line_yz.size(); ++l) {
        const int ky = line_yz[l] % NY, kz = line_yz[l] / NY;
        lp[l] = PP::lds(AY::freq_position(ky), AZ::freq_position(kz));
    }
    // per frame, the z-stage-B butterflies u = (pz / N2z) NY + py of its lines, ascending
    const int Nt = P->Nt;
    std::vector<int> zbl(std::max<size_t>(line_yz.size(), 1)), zbc(Nt);
    std::vector<char> seen(PP::UZB);
    for (int t = 0; t < Nt; ++t) {
        std::fill(seen.begin(), seen.end(), 0);
        for (int l = P->lptr_h[t]; l < P->lptr_h[t + 1]; ++l) {
            const int py = AY::freq_position(line_yz[l] % NY);
            const int pz = AZ::freq_position(line_yz[l] / NY);
            seen[(pz / AZ::N2) * NY + py] = 1;
        }
        int n = 0;
        for (int u = 0; u < PP::UZB; ++u)
            if (seen[u]) zbl[P->lptr_h[t] + n++] = u;
        zbc[t] = n;
    }
    // twiddles W^(a b) = exp(-+ 2 pi i a b / Ny) in the order b * N1y + a
    for (int sgn = 0; sgn < 2; ++sgn) {
        std::vector<c64> tl((size_t)AY::N1 * AY::N2);
        for (int b = 0; b < AY::N2; ++b)
            for (int a = 0; a < AY::N1; ++a) {
                const c128 w = unit_phase((sgn == 0 ? -1 : 1) * (int64_t)a * b, NY);
                tl[(size_t)b * AY::N1 + a] = {(float)w.re, (float)w.im};
            }
        P->tw_yz2[sgn] = upload_vec(P, tl);
    }
    P->smaps_yz2 = upload_vec(P, smp);
    P->line_pos = upload_vec(P, lp);
    P->zb_list = upload_vec(P, zbl);
    P->zb_cnt = upload_vec(P, zbc);
    P->yz2 = 1;
}

// Tables of the two-stage x passes: twiddles W^(a b) of the Cooley-Tukey split in the
// order b * N1 + a (forward and inverse sign).
template <class AX>
static void x2_tables(mslr_sense_s *P) {
    for (int sgn = 0; sgn < 2; ++sgn) {
        std::vector<c64> tl((size_t)AX::N1 * AX::N2);
        for (int b = 0; b < AX::N2; ++b)
            for (int a = 0; a < AX::N1; ++a) {
                const c128 w = unit_phase((sgn == 0 ? -1 : 1) * (int64_t)a * b, AX::N);
                tl[(size_t)b * AX::N1 + a] = {(float)w.re, (float)w.im};
            }
        P->tw_x2[sgn] = upload_vec(P, tl);
    }
    P->x2 = 1;
}

static void build_x2_tables(mslr_sense_s *P) {
    switch (P->N[0]) {
#define CASE_X2(n, AX)        \
    case n:                   \
        x2_tables<AX>(P);     \
        break;
        MSLR_X2_AXES(CASE_X2)
#undef CASE_X2
        default: break;
    }
}

static void build_yz2_tables(mslr_sense_s *P, const std::vector<c64> &sm32,
                             const std::vector<int> &line_yz) {
    switch (P->N[1] * 10000 + P->N[2]) {
#define CASE_YZ2(ny, nz, PP)                  \
    case ny * 10000 + nz:                     \
        yz2_tables<PP>(P, sm32, line_yz);     \
        break;
        MSLR_YZ2_PLANS(CASE_YZ2)
        CASE_YZ2(90, 60, Plane90x60<512>)
#undef CASE_YZ2
        default: break;
    }
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
            int nt = (int)cdiv(P->NYZ, MAXE_YZ);
            nt = (int)(cdiv(nt, 64) * 64);
            P->nt_yz = std::max(64, std::min(NTYZ_MAX, nt));
            P->chunk_bytes = (size_t)4 << 30;
            P->use_fixed = getenv("MSLR_FFT_GENERIC") == nullptr;
            if (const char *e = getenv("MSLR_FFT_CHUNK_MB")) {
                long mb = atol(e);
                if (mb > 0) P->chunk_bytes = (size_t)mb << 20;
            }

            // ---- mask -> sampled lines and sample tables.  Samples of a frame are
            // ordered by line (ascending ky + Ny*kz) then kx, which is the column-major
            // order of findall(vec(mask_t)) (scripts/reconstruct.jl:135-142).
            const int64_t V = P->V;
            int64_t K = -1;
            for (int64_t t = 0; t < Nt; ++t) {
                const uint8_t *m = mask_host + t * V;
                int64_t k = 0;
                for (int64_t i = 0; i < V; ++i) k += m[i] != 0;
                if (t == 0) K = k;
                MSLR_REQUIRE(k == K, MSLR_ERR_SHAPE,
                             "Frame " + std::to_string(t + 1) + " has a different sample count");
            }
            MSLR_REQUIRE(K <= INT32_MAX, MSLR_ERR_UNSUPPORTED, "too many samples per frame");
            MSLR_REQUIRE(Nx < 65536 && P->NYZ < 32768, MSLR_ERR_UNSUPPORTED,
                         "sample table packs (line, kx) in 15 + 16 bits");
            MSLR_REQUIRE((int64_t)P->NYZ * Nt <= INT32_MAX, MSLR_ERR_UNSUPPORTED,
                         "too many k-space lines");
            P->K = K;
            std::vector<int> lptr((size_t)Nt + 1, 0), line_yz, kbeg;
            std::vector<int> lk((size_t)std::max<int64_t>(1, K * Nt));
            int lmax = 1;
            for (int64_t t = 0; t < Nt; ++t) {
                const uint8_t *m = mask_host + t * V;
                int j = 0, l = 0;
                for (int yz = 0; yz < P->NYZ; ++yz) {
                    const uint8_t *row = m + (int64_t)yz * Nx;
                    bool any = false;
                    for (int x = 0; x < Nx; ++x) {
                        if (!row[x]) continue;
                        if (!any) {
                            line_yz.push_back(yz);
                            kbeg.push_back(j);
                            any = true;
                        }
                        lk[t * K + j] = (l << 16) | x;
                        ++j;
                    }
                    l += any;
                }
                lptr[t + 1] = lptr[t] + l;
                lmax = std::max(lmax, l);
            }
            P->LBx = std::max(1, std::min(lmax, (MAXE_X * NTX) / Nx));
            P->nlb = (int)cdiv(lmax, P->LBx);
            P->full_lines = true;
            for (int64_t t = 0; t < Nt; ++t)
                P->full_lines = P->full_lines && K == (int64_t)Nx * (lptr[t + 1] - lptr[t]);
            P->lptr_h = lptr;
            P->lptr = upload_vec(P, lptr);
            P->line_yz = upload_vec(P, line_yz);
            P->kbeg = upload_vec(P, kbeg);
            P->lk = upload_vec(P, lk);

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
            // coil maps times the pre-phase, plane-major: [c][x][y + Ny*z]
            std::vector<c128> sm64((size_t)V * Nc);
            std::vector<c64> sm32((size_t)V * Nc);
            for (int64_t ci = 0; ci < Nc; ++ci)
                for (int z = 0; z < Nz; ++z)
                    for (int y = 0; y < Ny; ++y) {
                        const c128 pyz = pre[1][y] * pre[2][z];
                        const int64_t yz = y + (int64_t)Ny * z;
                        for (int x = 0; x < Nx; ++x) {
                            const int64_t i = x + (int64_t)Nx * yz;
                            const c128 ph = pre[0][x] * pyz;
                            const float *sp = smaps_host + 2 * (ci * V + i);
                            const c128 v = c128{(double)sp[0], (double)sp[1]} * ph;
                            const int64_t o = (ci * Nx + x) * (int64_t)P->NYZ + yz;
                            sm64[o] = v;
                            sm32[o] = {(float)v.re, (float)v.im};
                        }
                    }
            P->smaps64 = upload_vec(P, sm64);
            P->smaps32 = upload_vec(P, sm32);
            if (P->use_fixed && P->sgn == -1) {
                build_yz2_tables(P, sm32, line_yz);
                build_x2_tables(P);
            }
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
                           nullptr);
    });
}

mslr_status_t mslr_sense_forward_noalloc(mslr_sense plan, const void *x_dev, void *y_dev,
                                         int64_t t0, int64_t nframes) {
    return guarded([&] {
        check_frames(plan, t0, nframes);
        if (nframes == 0) return;
        MSLR_REQUIRE(x_dev && y_dev, MSLR_ERR_ARG, "null argument");
        run_forward_noalloc(plan, (const c64 *)x_dev, t0, nframes, (c64 *)y_dev);
    });
}

mslr_status_t mslr_sense_adjoint(mslr_sense plan, const void *y_dev, void *x_dev, int64_t t0,
                                 int64_t nframes) {
    return guarded([&] {
        check_frames(plan, t0, nframes);
        if (nframes == 0) return;
        run_adjoint<float>(plan, (const c64 *)y_dev, t0, nframes, (c64 *)x_dev);
    });
}

mslr_status_t mslr_sense_forward_f64(mslr_sense plan, const void *x_dev, void *y_dev, int64_t t0,
                                     int64_t nframes) {
    return guarded([&] {
        check_frames(plan, t0, nframes);
        if (nframes == 0) return;
        run_forward<double>(plan, (const c128 *)x_dev, 1, 0, t0, nframes, nullptr,
                            (c128 *)y_dev, nullptr);
    });
}

mslr_status_t mslr_sense_adjoint_f64(mslr_sense plan, const void *y_dev, void *x_dev, int64_t t0,
                                     int64_t nframes) {
    return guarded([&] {
        check_frames(plan, t0, nframes);
        if (nframes == 0) return;
        run_adjoint<double>(plan, (const c128 *)y_dev, t0, nframes, (c128 *)x_dev);
    });
}

mslr_status_t mslr_sense_grad(mslr_sense plan, const void *X_dev, int64_t nscales,
                              const void *y_dev, void *g_dev, double *dc_cost) {
    return guarded([&] {
        check_frames(plan, 0, plan ? plan->Nt : 0);
        MSLR_REQUIRE(nscales >= 1, MSLR_ERR_ARG, "nscales must be >= 1");
        mslr_sense_s *P = plan;
        mslr_ctx_s *ctx = P->ctx;
        const int64_t npart = (int64_t)P->Nt * P->Nc * P->nlb;
        double *part = (double *)ctx->scratch_get(SLOT_REDUCE, (size_t)(npart + 1) * sizeof(double));
        sense_grad_internal(P, (const c64 *)X_dev, (int)nscales, P->V * (int64_t)P->Nt,
                            (const c64 *)y_dev, (c64 *)g_dev, part);
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

mslr_status_t mslr_sense_residual(mslr_sense plan, const void *X_dev, int64_t nscales,
                                  const void *y_dev, void *r_dev, double *rnorm2) {
    return guarded([&] {
        check_frames(plan, 0, plan ? plan->Nt : 0);
        MSLR_REQUIRE(nscales >= 1, MSLR_ERR_ARG, "nscales must be >= 1");
        MSLR_REQUIRE(X_dev && y_dev && rnorm2, MSLR_ERR_ARG, "null argument");
        mslr_sense_s *P = plan;
        mslr_ctx_s *ctx = P->ctx;
        const int64_t ny = P->K * P->Nc * (int64_t)P->Nt;
        c64 *r = r_dev ? (c64 *)r_dev
                       : (c64 *)ctx->scratch_get(SLOT_RESID, (size_t)std::max<int64_t>(1, ny) *
                                                                 sizeof(c64));
        const int64_t npart = (int64_t)P->Nt * P->Nc * P->nlb;
        double *part = (double *)ctx->scratch_get(SLOT_REDUCE, (size_t)(npart + 1) * sizeof(double));
        run_forward<float>(P, (const c64 *)X_dev, (int)nscales, P->V * (int64_t)P->Nt, 0, P->Nt,
                           (const c64 *)y_dev, r, part);
        sum_partials(ctx, part, npart, part + npart);
        double h = 0.0;
        HIP_CHECK(hipMemcpyAsync(&h, part + npart, sizeof(double), hipMemcpyDeviceToHost,
                                 ctx->stream));
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
        *rnorm2 = h;
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
            run_forward<double>(P, x, 1, 0, 0, P->Nt, nullptr, Ax, nullptr);
            const double an = std::sqrt(device_norm2<double>(ctx, Ax, ny));
            const double ratio = an / std::sqrt(xn2);
            if (std::fabs(ratio - ratio_old) / ratio < tol) break;
            ratio_old = ratio;
            run_adjoint<double>(P, Ax, 0, P->Nt, x);
            const double n2 = device_norm2<double>(ctx, x, nx);
            MSLR_REQUIRE(n2 > 0 && std::isfinite(n2), MSLR_ERR_NUMERIC,
                         "power iteration produced a zero or non-finite vector");
            device_scale<double>(ctx, x, nx, 1.0 / std::sqrt(n2));
            xn2 = device_norm2<double>(ctx, x, nx);
        }
        if (it > niter) it = niter;
        // sigma1 = ||A x|| / ||x|| at the returned x.
        run_forward<double>(P, x, 1, 0, 0, P->Nt, nullptr, Ax, nullptr);
        const double an = std::sqrt(device_norm2<double>(ctx, Ax, ny));
        *sigma1 = an / std::sqrt(device_norm2<double>(ctx, x, nx));
        if (iters) *iters = it;
    });
}

}  // extern "C"
