// Two-stage Hermitian tridiagonalisation for the batched patch SVST.
//
// One-stage Householder tridiagonalisation re-reads the whole trailing matrix
// for every column (its matrix-vector product), which for hundreds of 300x300
// FP64 matrices streams ~n/3 matrix volumes from HBM.  Here:
//   stage 1 (k_sy2sb): dense -> band of width B.  Per panel of B columns a
//     Householder QR (one wave, in LDS) gives Q = I - V T V^H, and the trailing
//     matrix gets A <- Q^H A Q = A - V W^H - W V^H with Y = A V T,
//     W = Y - 1/2 V (T^H V^H Y): per B columns one read of the trailing matrix
//     for Y and a read and write of its lower triangle for the update (the
//     upper triangle is never read again), instead of a read per column.
//   stage 2 (k_sb2st): band -> real tridiagonal by bulge chasing, the band
//     (2B diagonals) held in LDS; each wave runs one sweep at a time, sweeps
//     start STAGGER = 2 time steps apart (one workgroup barrier per time step).
//     Task (s, k) touches entries whose column lies in I_k(s) (plus column s at
//     k = 0) and reads what tasks (s-1, k) and (s-1, k+1) wrote; the task of
//     sweep s-1 running beside it, (s-1, k+2), has columns I_{k+2}(s-1), which
//     begin after I_k(s) ends, so concurrent tasks never share an entry.
//   back-transformation (k_sbr_bt): U = Q1 Q2 Z for the wanted eigenvectors.
// Reflectors follow zlarfg: H = I - tau v v^H, H^H x = beta e1, beta real.
#include <algorithm>
#include <cfloat>
#include <cmath>

#include "sbr.hpp"
#include "wave.hpp"

namespace mslr {

constexpr int B = SBR_B;
constexpr int S1T = SBR_S1_WAVES * 64;  // threads of the dense -> band kernel
constexpr int S2T = 768;       // threads of the band -> tridiagonal kernel
constexpr int S2W = S2T / 64;  // concurrent sweeps
// the band -> tridiagonal kernel with the band in global memory: threads and
// time steps between sweep starts (a wave's sweep, up to 30 tasks at m = 480, must
// end before its next one starts: tasks <= stagger x waves).  12 waves 3 steps
// apart (170 VGPRs, no spills): c2 1.58 iter/s against 1.49 with 16 waves 2 steps
// apart (128 VGPRs, ~40 spilled).
#ifndef MSLR_SB2ST_G_THREADS
#define MSLR_SB2ST_G_THREADS 768
#endif
#ifndef MSLR_SB2ST_G_STAGGER
#define MSLR_SB2ST_G_STAGGER 3
#endif
constexpr int S2TG = MSLR_SB2ST_G_THREADS;
constexpr int STAGGER_G = MSLR_SB2ST_G_STAGGER;
static_assert((SBR_MAXG - 1 + B - 1) / B <= STAGGER_G * (S2TG / 64),
              "a sweep must end before the wave's next one starts");
constexpr int STAGGER = 2;     // time steps between the starts of consecutive sweeps
constexpr int BTT = 256;       // threads of the back-transformation
constexpr int BTC = 8;         // eigenvectors per back-transformation workgroup: one thread per
                               // (reflector, eigenvector) of a sweep, LDS for two workgroups per CU
constexpr int BT_PD = 4;       // sweeps of bulge-chasing reflectors prefetched into registers

__device__ __forceinline__ c128 c_div(c128 a, c128 b) {
    const double den = b.re * b.re + b.im * b.im;
    return c128{(a.re * b.re + a.im * b.im) / den, (a.im * b.re - a.re * b.im) / den};
}

// An LDS table's address hidden from loop-invariant code motion, so that its
// entries are re-read per loop iteration instead of held in registers.
template <class P>
__device__ __forceinline__ P *lds_opaque(P *p) {
    asm volatile("" : "+v"(p));
    return p;
}

// Orders LDS traffic between the lanes of one wave (a wave-level barrier).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct Hh {
    c128 tau, scal;
    double beta;
};
// Reflector for x = (alpha, x1) with ||x1||^2 = xn (zlarfg without rescaling).
__device__ __forceinline__ Hh householder(c128 alpha, double xn) {
    Hh h;
    if (xn == 0.0 && alpha.im == 0.0) {
        h.tau = c128{0.0, 0.0};
        h.scal = c128{0.0, 0.0};
        h.beta = alpha.re;
    } else {
        const double nrm = sqrt(alpha.re * alpha.re + alpha.im * alpha.im + xn);
        const double beta = alpha.re >= 0.0 ? -nrm : nrm;
        // tau = (beta - alpha) / beta, scal = 1 / (alpha - beta): two divisions
        const double ib = 1.0 / beta;
        h.tau = c128{(beta - alpha.re) * ib, -alpha.im * ib};
        const double dr = alpha.re - beta, di = alpha.im;
        const double id = 1.0 / (dr * dr + di * di);
        h.scal = c128{dr * id, -di * id};
        h.beta = beta;
    }
    return h;
}
__device__ __forceinline__ bool nonzero(c128 t) { return t.re != 0.0 || t.im != 0.0; }

typedef double md4 __attribute__((ext_vector_type(4)));

// d + conj(a) b and w - f v, four FMAs each
__device__ __forceinline__ c128 fma_conj(c128 a, c128 b, c128 d) {
    return c128{fma(a.re, b.re, fma(a.im, b.im, d.re)), fma(a.re, b.im, fma(-a.im, b.re, d.im))};
}
__device__ __forceinline__ c128 fms_c(c128 f, c128 v, c128 w) {
    return c128{fma(-f.re, v.re, fma(f.im, v.im, w.re)), fma(-f.re, v.im, fma(-f.im, v.re, w.im))};
}
// w - a conj(b), four FMAs
__device__ __forceinline__ c128 fms_cc(c128 a, c128 b, c128 w) {
    return c128{fma(-a.re, b.re, fma(-a.im, b.im, w.re)), fma(-a.im, b.re, fma(a.re, b.im, w.im))};
}

// v when c, else 0: with a load from an address clamped into the array, the mask is a
// select after it, so loads in a loop carry no branches, issue together and wait
// once (a load under a branch waits before the next one can issue).
__device__ __forceinline__ c128 sel(bool c, c128 v) {
    return c128{c ? v.re : 0.0, c ? v.im : 0.0};
}

// Value of *p when c, else 0, through scalar loads (a conditionally assigned
// struct otherwise ends up in private memory).
__device__ __forceinline__ c128 ld_if(bool c, const c128 *p) {
    double r = 0.0, i = 0.0;
    if (c) {
        r = p->re;
        i = p->im;
    }
    return c128{r, i};
}


// ------------------------------------------------------------------ stage 1
// One workgroup per matrix, two per CU (a batch of a few hundred matrices runs in
// one round): V ((m - B) x B, column-major), T and one B x B block (S = V^H V, then
// Z = V^H Y turned into M = T^H Z T in place) in LDS, at most 80 KB; Y, W and the
// waves' partials of Z in global memory (L2); at most 128 VGPRs (four waves per SIMD).
__global__ __launch_bounds__(S1T, 4) void k_sy2sb(c128 *__restrict__ Gall, int m, int npanel,
                                                  c128 *__restrict__ Tall,
                                                  c128 *__restrict__ W1all, int hi) {
    if (hi) __builtin_amdgcn_s_setprio(3);
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ c128 Ts[B][B + 1];
    __shared__ c128 Ms[B][B + 1];
    __shared__ c128 taus[B];
    __shared__ double betas[B];
    const int Lmax = m - B;
    c128 *V = reinterpret_cast<c128 *>(smem_raw);
    c128 *Y = W1all + (int64_t)blockIdx.x * sbr_w1_elems(m);
    c128 *W = Y + (size_t)Lmax * B;
    c128 *Zp = W + (size_t)Lmax * B;  // the waves' partials of V^H Y
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    c128 *A = Gall + (int64_t)blockIdx.x * m * m;
    c128 *Tg = Tall + (int64_t)blockIdx.x * npanel * B * B;
    int p = 0;
    for (int c0 = 0; m - c0 - B >= 2; c0 += B, ++p) {
        const int L = m - c0 - B, r0 = c0 + B;
        const int kq = min(B, L - 1);
        // 1. panel rows r0.. of columns c0 .. c0+B-1 -> V
#pragma unroll 4
        for (int idx = tid; idx < L * B; idx += S1T) {
            const int i = idx % L, j = idx / L;
            V[j * L + i] = A[(r0 + i) + (int64_t)m * (c0 + j)];
        }
        __syncthreads();
        // 2. Householder QR of the panel.  Per column j all waves apply reflector j
        //    to the later columns, one column per wave; wave 0 takes column j + 1 and
        //    forms reflector j + 1 right after updating it (look-ahead), so a column
        //    costs one workgroup barrier.  Lane l owns rows l + 64t of a column.
        auto make_reflector = [&](int j) {  // wave 0
            c128 *vj = V + j * L;
            double xn = 0.0;
            c128 own = {0.0, 0.0};
            for (int i = lane; i < L; i += 64) {
                if (i > j) xn += abs2(vj[i]);
                if (i == j) own = vj[i];
            }
            xn = wave_sum64(xn);
            const c128 alpha = {readlane_d(own.re, j & 63), readlane_d(own.im, j & 63)};
            const Hh h = householder(alpha, xn);
            if (nonzero(h.tau))
                for (int i = lane; i < L; i += 64)
                    if (i > j) vj[i] = vj[i] * h.scal;
            if (lane == 0) {
                taus[j] = h.tau;
                betas[j] = h.beta;
            }
        };
        if (wave == 0 && kq > 0) make_reflector(0);
        __syncthreads();
        for (int j = 0; j < kq; ++j) {
            const c128 *vj = V + j * L;
            const c128 tj = taus[j];
            // column c > j: x <- H^H x = x - conj(tau) v (v^H x)
            const c128 ct = conj(tj);
            for (int c = j + 1 + wave; c < B; c += S1T / 64) {
                if (nonzero(tj)) {
                    c128 *pc = V + c * L;
                    c128 sacc = {0.0, 0.0};
                    for (int i = j + lane; i < L; i += 64) {
                        c128 vi = ld_if(i != j, vj + i);
                        if (i == j) vi.re = 1.0;
                        sacc = sacc + cmulc(vi, pc[i]);
                    }
                    sacc = wave_sum64(sacc);
                    const c128 f = ct * sacc;
                    for (int i = j + lane; i < L; i += 64) {
                        c128 vi = ld_if(i != j, vj + i);
                        if (i == j) vi.re = 1.0;
                        pc[i] = pc[i] - vi * f;
                    }
                }
                if (c == j + 1 && j + 1 < kq) {  // wave 0: the next reflector
                    wave_lds_sync();
                    make_reflector(j + 1);
                }
            }
            __syncthreads();
        }
        for (int j = kq + tid; j < B; j += S1T) taus[j] = c128{0.0, 0.0};
        __syncthreads();
        // 3. R -> the band, reflector tails -> below the band; V -> unit lower form
        for (int idx = tid; idx < L * B; idx += S1T) {
            const int i = idx % L, j = idx / L;
            const c128 val = V[j * L + i];
            c128 *g = A + (r0 + i) + (int64_t)m * (c0 + j);
            if (j < kq) {
                if (i < j) {
                    *g = val;
                    V[j * L + i] = c128{0.0, 0.0};
                } else if (i == j) {
                    *g = c128{betas[j], 0.0};
                    V[j * L + i] = c128{1.0, 0.0};
                } else {
                    *g = val;
                }
            } else {
                *g = val;
                V[j * L + i] = c128{0.0, 0.0};
            }
        }
        __syncthreads();
        // 4. T (upper triangular, Q = I - V T V^H): S = V^H V (entries a < c, four
        //    threads each over interleaved rows, two partial sums per thread; V_c is
        //    zero above row c), then the zlarft recurrence
        {
            const int pr = tid >> 2, part = tid & 3;
            int c = 1;
            while (c * (c + 1) / 2 <= pr) ++c;
            const int a = pr - c * (c - 1) / 2;
            c128 s0 = {0.0, 0.0}, s1 = {0.0, 0.0};
            if (c < B) {
                const c128 *va = V + a * L, *vc = V + c * L;
                int i = c + part;
                for (; i + 4 < L; i += 8) {
                    s0 = s0 + cmulc(va[i], vc[i]);
                    s1 = s1 + cmulc(va[i + 4], vc[i + 4]);
                }
                if (i < L) s0 = s0 + cmulc(va[i], vc[i]);
            }
            s0 = s0 + s1;
            s0.re += __shfl_xor(s0.re, 1, 64);
            s0.im += __shfl_xor(s0.im, 1, 64);
            s0.re += __shfl_xor(s0.re, 2, 64);
            s0.im += __shfl_xor(s0.im, 2, 64);
            if (c < B && part == 0) Ms[a][c] = s0;
        }
        __syncthreads();
        if (wave == 0) {
            // T[0:j, j] = -tau_j T[0:j, 0:j] S[0:j, j], column by column; row `lane` of
            // T is read back from LDS (zero left of the diagonal), the S entries are
            // broadcasts, two partial sums
            for (int j = 0; j < B; ++j) {
                const c128 tj = taus[j];
                c128 a0 = {0.0, 0.0}, a1 = {0.0, 0.0};
                const int row = lane < B ? lane : 0;
#pragma unroll 4
                for (int q = 0; q < j; ++q) {
                    const c128 tq = (-1.0) * (tj * Ms[q][j]);
                    const c128 t = Ts[row][q];
                    if (q & 1) a1 = a1 + t * tq;
                    else a0 = a0 + t * tq;
                }
                const c128 acc = a0 + a1;
                if (lane < j) Ts[lane][j] = acc;
                else if (lane == j) Ts[j][j] = tj;
                else if (lane < B) Ts[lane][j] = c128{0.0, 0.0};
                wave_lds_sync();
            }
        }
        __syncthreads();
        for (int pr = tid; pr < B * B; pr += S1T) Tg[(int64_t)p * B * B + pr] = Ts[pr % B][pr / B];
        // 5. Y = A22 V on the FP64 matrix cores, one 16-row block per wave at a time;
        //    the A22 operand is read from global memory once.  Only the lower triangle
        //    of A22 is kept up to date: an entry above the diagonal is read as the
        //    conjugate of its mirror.  v_mfma_f64_16x16x4f64 (D = P Q): lane l holds
        //    P[l%16][l/16], Q[l/16][l%16] and D[l/16 + 4v][l%16] -- the accumulator
        //    layout is the Q operand's over four k-steps, so the same pass also
        //    accumulates the wave's share of Z = V^H Y from the Y block in registers.
        //    Y is stored row-major (Y[i B + c]) for step 7.
        {
            const int nrb = (L + 15) / 16;
            const int rl = lane & 15, kl = lane >> 4;
            md4 zr = {0.0, 0.0, 0.0, 0.0}, zi = {0.0, 0.0, 0.0, 0.0};
            for (int rb = wave; rb < nrb; rb += S1T / 64) {
                // four independent accumulator chains (summed below) keep the
                // matrix pipe busy between dependent MFMAs
                md4 yr = {0.0, 0.0, 0.0, 0.0}, yi = {0.0, 0.0, 0.0, 0.0};
                md4 yr2 = {0.0, 0.0, 0.0, 0.0}, yi2 = {0.0, 0.0, 0.0, 0.0};
                const int i = rb * 16 + rl;
                const int ic = min(i, L - 1);
                // A(i, k) from the lower triangle (or the conjugate of its mirror), from
                // rows and columns clamped into the matrix; YP k-steps of loads are in
                // flight before their products
                auto aload = [&](int k) {
                    const int kc = min(k, L - 1);
                    return A[i >= k ? (r0 + ic) + (int64_t)m * (r0 + kc)
                                    : (r0 + kc) + (int64_t)m * (r0 + ic)];
                };
                constexpr int YP = 4;
                for (int ks = 0; ks < L; ks += 4 * YP) {
                    c128 araw[YP];
#pragma unroll
                    for (int u = 0; u < YP; ++u) araw[u] = aload(ks + 4 * u + kl);
#pragma unroll
                    for (int u = 0; u < YP; ++u) {
                        const int k = ks + 4 * u + kl;
                        c128 a = sel(i < L && k < L, araw[u]);
                        if (i < k) a.im = -a.im;
                        const c128 v = sel(k < L, V[rl * L + min(k, L - 1)]);
                        // P[i][k] = A, Q[k][c] = V: the block of Y = A V
                        yr = __builtin_amdgcn_mfma_f64_16x16x4f64(a.re, v.re, yr, 0, 0, 0);
                        yi = __builtin_amdgcn_mfma_f64_16x16x4f64(a.re, v.im, yi, 0, 0, 0);
                        yr2 = __builtin_amdgcn_mfma_f64_16x16x4f64(-a.im, v.im, yr2, 0, 0, 0);
                        yi2 = __builtin_amdgcn_mfma_f64_16x16x4f64(a.im, v.re, yi2, 0, 0, 0);
                    }
                }
                yr += yr2;
                yi += yi2;
                // lane (rl, kl) holds Y[rb 16 + kl + 4q][rl]
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int ii = rb * 16 + kl + 4 * q;
                    if (ii < L) Y[ii * B + rl] = c128{yr[q], yi[q]};
                }
                // Z += V_blk^H Y_blk over the block's rows: k-step s covers rows
                // rb 16 + 4s .. +3, P[a][i] = conj(V[i][a]), Q[i][c] = Y[i][c] (yr[s], yi[s])
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int ii = rb * 16 + 4 * q + kl;
                    const c128 vv = sel(ii < L, V[rl * L + min(ii, L - 1)]);
                    zr = __builtin_amdgcn_mfma_f64_16x16x4f64(vv.re, yr[q], zr, 0, 0, 0);
                    zr = __builtin_amdgcn_mfma_f64_16x16x4f64(vv.im, yi[q], zr, 0, 0, 0);
                    zi = __builtin_amdgcn_mfma_f64_16x16x4f64(vv.re, yi[q], zi, 0, 0, 0);
                    zi = __builtin_amdgcn_mfma_f64_16x16x4f64(-vv.im, yr[q], zi, 0, 0, 0);
                }
            }
            // the wave's partial Z (lane holds Z[kl + 4q][rl]; zero for a wave without
            // blocks) -> its slot of the workspace
#pragma unroll
            for (int q = 0; q < 4; ++q) Zp[wave * B * B + (kl + 4 * q) * B + rl] = c128{zr[q], zi[q]};
        }
        __syncthreads();
        // 6. Z = the sum of the wave partials (fixed order), then M = T^H X with
        //    X = V^H W = V^H Y T, i.e. M = T^H Z T (Hermitian), one entry per thread:
        //    Z -> Ms, Z T -> Ms, T^H (Z T) -> Ms (T upper triangular)
        {
            const int a = tid % B, c = tid / B;
            if (tid < B * B) {
                // the partials loaded together, then summed in wave order
                c128 zp[S1T / 64];
#pragma unroll
                for (int w = 0; w < S1T / 64; ++w) zp[w] = Zp[w * B * B + a * B + c];
                c128 z = {0.0, 0.0};
#pragma unroll
                for (int w = 0; w < S1T / 64; ++w) z = z + zp[w];
                Ms[a][c] = z;
            }
            __syncthreads();
            c128 acc = {0.0, 0.0};
            if (tid < B * B)
                for (int q = 0; q <= c; ++q) acc = acc + Ms[a][q] * Ts[q][c];
            __syncthreads();
            if (tid < B * B) Ms[a][c] = acc;
            __syncthreads();
            acc = c128{0.0, 0.0};
            if (tid < B * B)
                for (int q = 0; q <= a; ++q) acc = acc + cmulc(Ts[q][a], Ms[q][c]);
            __syncthreads();
            if (tid < B * B) Ms[a][c] = acc;
        }
        __syncthreads();
        // 7. W = Y T - 1/2 V M on the matrix cores, one 16-row block per wave at a time
        //    (the block computed is W^T = T^T Y^T - 1/2 M^T V^T: lane (rl, kl) holds
        //    W[rb 16 + rl][kl + 4q])
        {
            const int nrb = (L + 15) / 16;
            const int rl = lane & 15, kl = lane >> 4;
            for (int rb = wave; rb < nrb; rb += S1T / 64) {
                const int i = rb * 16 + rl;
                md4 wr = {0.0, 0.0, 0.0, 0.0}, wi = {0.0, 0.0, 0.0, 0.0};
                md4 ur = {0.0, 0.0, 0.0, 0.0}, ui = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 2
                for (int as = 0; as < B; as += 4) {
                    const int a = as + kl;
                    const c128 t = Ts[a][rl];   // P[c][a] = T[a][c], c = rl
                    const c128 mm = Ms[a][rl];  // P[c][a] = M[a][c]
                    const c128 y = sel(i < L, Y[min(i, L - 1) * B + a]);
                    const c128 v = sel(i < L, V[a * L + min(i, L - 1)]);
                    // T^T Y^T: re = tr yr - ti yi, im = tr yi + ti yr
                    wr = __builtin_amdgcn_mfma_f64_16x16x4f64(t.re, y.re, wr, 0, 0, 0);
                    wi = __builtin_amdgcn_mfma_f64_16x16x4f64(t.re, y.im, wi, 0, 0, 0);
                    wr = __builtin_amdgcn_mfma_f64_16x16x4f64(-t.im, y.im, wr, 0, 0, 0);
                    wi = __builtin_amdgcn_mfma_f64_16x16x4f64(t.im, y.re, wi, 0, 0, 0);
                    // -1/2 M^T V^T: re = -(vr mr - vi mi)/2, im = -(vr mi + vi mr)/2
                    ur = __builtin_amdgcn_mfma_f64_16x16x4f64(-0.5 * mm.re, v.re, ur, 0, 0, 0);
                    ui = __builtin_amdgcn_mfma_f64_16x16x4f64(-0.5 * mm.im, v.re, ui, 0, 0, 0);
                    ur = __builtin_amdgcn_mfma_f64_16x16x4f64(0.5 * mm.im, v.im, ur, 0, 0, 0);
                    ui = __builtin_amdgcn_mfma_f64_16x16x4f64(-0.5 * mm.re, v.im, ui, 0, 0, 0);
                }
                wr += ur;
                wi += ui;
                if (i < L) {
#pragma unroll
                    for (int q = 0; q < 4; ++q) W[(kl + 4 * q) * L + i] = c128{wr[q], wi[q]};
                }
            }
        }
        __syncthreads();
        // 8. A22 <- A22 - V W^H - W V^H on the FP64 matrix cores, the 16x16 blocks on
        //    and below the diagonal (the lower triangle is what later steps read),
        //    one block per wave at a time, K = 2B complex.
        {
            const int nb = (L + 15) / 16;
            const int rl = lane & 15, kl = lane >> 4;
            for (int blk = wave; blk < nb * (nb + 1) / 2; blk += S1T / 64) {
                int bi = 0;
                while ((bi + 1) * (bi + 2) / 2 <= blk) ++bi;
                const int bk = blk - bi * (bi + 1) / 2;
                md4 dr, di;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int i = bi * 16 + rl, k = bk * 16 + kl + 4 * q;
                    const c128 a = sel(i < L && k < L, A[(r0 + min(i, L - 1)) +
                                                         (int64_t)m * (r0 + min(k, L - 1))]);
                    dr[q] = a.re;
                    di[q] = a.im;
                }
                const int ii = bi * 16 + rl, kk = bk * 16 + rl;
                // the update as four independent accumulator chains (the product
                // terms split two ways), summed into the block below
                md4 er = {0.0, 0.0, 0.0, 0.0}, ei = {0.0, 0.0, 0.0, 0.0};
                md4 fr = {0.0, 0.0, 0.0, 0.0}, fi = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int js = 0; js < B; js += 4) {
                    const int j = js + kl;
                    const int iic = min(ii, L - 1), kkc = min(kk, L - 1);
                    const c128 vI = sel(ii < L, V[j * L + iic]), wI = sel(ii < L, W[j * L + iic]);
                    const c128 vK = sel(kk < L, V[j * L + kkc]), wK = sel(kk < L, W[j * L + kkc]);
                    // V_I W_K^H + W_I V_K^H: re = vr wr + vi wi, im = vi wr - vr wi.
                    // Operands swapped (K-side first): the block computed is the
                    // transpose, so the output lanes run down the columns of A.
                    er = __builtin_amdgcn_mfma_f64_16x16x4f64(wK.re, vI.re, er, 0, 0, 0);
                    ei = __builtin_amdgcn_mfma_f64_16x16x4f64(wK.re, vI.im, ei, 0, 0, 0);
                    fr = __builtin_amdgcn_mfma_f64_16x16x4f64(vK.re, wI.re, fr, 0, 0, 0);
                    fi = __builtin_amdgcn_mfma_f64_16x16x4f64(vK.re, wI.im, fi, 0, 0, 0);
                    er = __builtin_amdgcn_mfma_f64_16x16x4f64(wK.im, vI.im, er, 0, 0, 0);
                    ei = __builtin_amdgcn_mfma_f64_16x16x4f64(wK.im, -vI.re, ei, 0, 0, 0);
                    fr = __builtin_amdgcn_mfma_f64_16x16x4f64(vK.im, wI.im, fr, 0, 0, 0);
                    fi = __builtin_amdgcn_mfma_f64_16x16x4f64(vK.im, -wI.re, fi, 0, 0, 0);
                }
                dr -= er + fr;
                di -= ei + fi;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int i = bi * 16 + rl, k = bk * 16 + kl + 4 * q;
                    if (i < L && k < L) A[(r0 + i) + (int64_t)m * (r0 + k)] = c128{dr[q], di[q]};
                }
            }
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------ stage 2
// Band element (i, j), i >= j, i - j < 2B, stored column by column at
// Bd[j * 2B + (i - j)]: along a row the stride is 2B - 1 entries, down a column
// 1, both free of LDS bank conflicts for 16 consecutive lanes.
__device__ __forceinline__ c128 &bref(c128 *Bd, int m, int i, int j) { return Bd[j * (2 * B) + (i - j)]; }
// Element (q, r) of the Hermitian block starting at (st, st).
__device__ __forceinline__ c128 bherm(c128 *Bd, int m, int st, int q, int r) {
    if (q >= r) return bref(Bd, m, st + q, st + r);
    return conj(bref(Bd, m, st + r, st + q));
}

// One task of sweep s at step k on the wave's lanes; returns the tau of H_{k+1}.
// H_k (reflector on rows st .. st+len-1, st = s + 1 + k B) is in sv / tau
// (k > 0) or is made here from column s (k = 0).  Then: two-sided H_k on the
// diagonal block D, right H_k on the block below (E, rows st+B ..), a new
// reflector H_{k+1} from E's first column and its left application to the
// rest of E.  D and E stay in registers for the whole task: lane l holds
// column r = l % 16 of rows q = l / 16 + 4t, t = 0..3; row products are DPP
// row sums, column products sums over the four rows of lanes.
__device__ __forceinline__ c128 sb2st_task(c128 *Bd, int m, int s, int k, c128 *sv, c128 *sw,
                                           c128 *sx, c128 tau, c128 *V2, c128 *tau2, int ks) {
    const int lane = threadIdx.x & 63;
    const int r = lane & 15, g = lane >> 4;
    const int st = s + 1 + k * B;
    const int len = min(B, m - st);
    if (k == 0) {
        const c128 x = ld_if(lane < len, &bref(Bd, m, st + lane, s));
        const double xn = wave_sum64(lane >= 1 && lane < len ? abs2(x) : 0.0);
        const c128 alpha = {readlane_d(x.re, 0), readlane_d(x.im, 0)};
        const Hh h = householder(alpha, xn);
        tau = h.tau;
        const bool refl = nonzero(h.tau);
        const c128 xs = x * h.scal;
        const c128 v = {lane == 0 ? 1.0 : (refl ? xs.re : 0.0), lane == 0 ? 0.0 : (refl ? xs.im : 0.0)};
        const c128 vl = {lane < len ? v.re : 0.0, lane < len ? v.im : 0.0};
        wave_lds_sync();
        if (lane < B) sv[lane] = vl;
        if (lane == 0) bref(Bd, m, st, s) = c128{h.beta, 0.0};
        else if (lane < len && refl) bref(Bd, m, st + lane, s) = c128{0.0, 0.0};
        if (lane < B) V2[((int64_t)s * ks) * B + lane] = vl;
        if (lane == 0) tau2[(int64_t)s * ks] = h.tau;
        wave_lds_sync();
    }
    const int st2 = st + B;
    const int len2 = min(B, m - st2);
    // D (Hermitian, upper part from the lower) and E into registers; loads from
    // clamped addresses and selects: entries past the matrix read as zero (sv is
    // zero past len)
    c128 Dr[4], Er[4], vq[4];
    const c128 vr = sv[r];
    const int rc = min(r, len - 1);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int q = g + 4 * t;
        vq[t] = sv[q];
        const int lo = min(min(q, r), len - 1), hi = max(q, r);
        c128 dv = Bd[(st + lo) * (2 * B) + (hi - lo)];
        if (q < r) dv.im = -dv.im;
        Dr[t] = (q < len && r < len) ? dv : c128{0.0, 0.0};
        const c128 ev = Bd[(st + rc) * (2 * B) + (B + q - rc)];
        Er[t] = (q < len2 && r < len) ? ev : c128{0.0, 0.0};
    }
    if (nonzero(tau)) {
        // p = tau D v by columns (D Hermitian: (D v)[c] = sum_q conj(D[q][c]) v[q]):
        // the lane's four rows, then the four row groups
        c128 pc = {0.0, 0.0};
#pragma unroll
        for (int t = 0; t < 4; ++t) pc = fma_conj(Dr[t], vq[t], pc);
        pc = tau * sum_rows4(pc);  // sum_q conj(D[q][c]) v[q] = sum_q D[c][q] v[q] = (D v)[c]
        // e = tau E v by rows, the same way from E transposed in the lanes: lane (r, g)
        // takes row r at columns g + 4t (loads from clamped addresses and selects, as
        // above), then the four row groups; e[r] lands in the lanes of column r
        c128 pt = {0.0, 0.0};
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int c = g + 4 * t;
            const int cc = min(c, len - 1);
            const c128 ev = Bd[(st + cc) * (2 * B) + (B + r - cc)];
            const c128 e = (r < len2 && c < len) ? ev : c128{0.0, 0.0};
            pt = pt + e * vq[t];
        }
        pt = tau * sum_rows4(pt);
        // p^H v over the columns of one row of lanes (every row holds all of p)
        const c128 dot = sum16(cmulc(pc, vr));
        const c128 alpha2 = (-0.5) * (tau * dot);
        // w = p + alpha2 v: w_r for the own column, w_q for the own rows through LDS;
        // e_q for the own rows through LDS as well (sx is free until step (c))
        const c128 wr = pc + alpha2 * vr;
        wave_lds_sync();
        if (g == 0) {
            sw[r] = pc;
            sx[r] = pt;
        }
        wave_lds_sync();
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const c128 wq = sw[g + 4 * t] + alpha2 * vq[t];
            Dr[t] = fms_cc(wq, vr, fms_cc(vq[t], wr, Dr[t]));  // D <- H^H D H
            Er[t] = fms_cc(sx[g + 4 * t], vr, Er[t]);          // E <- E H
        }
        // D back to the band (lower part)
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int q = g + 4 * t;
            if (q < len && r <= q) bref(Bd, m, st + q, st + r) = Dr[t];
        }
    }
    if (len2 > 0) {
        // (c) H_{k+1} from E's first column (exchanged through LDS)
        wave_lds_sync();
        if (r == 0) {
#pragma unroll
            for (int t = 0; t < 4; ++t) sx[g + 4 * t] = Er[t];
        }
        wave_lds_sync();
        const c128 x = ld_if(lane < len2, sx + lane);
        const double xn = wave_sum64(lane >= 1 && lane < len2 ? abs2(x) : 0.0);
        const c128 alpha = {readlane_d(x.re, 0), readlane_d(x.im, 0)};
        const Hh h = householder(alpha, xn);
        const bool refl = nonzero(h.tau);
        const c128 xs = x * h.scal;
        const c128 v = {lane == 0 ? 1.0 : (refl ? xs.re : 0.0), lane == 0 ? 0.0 : (refl ? xs.im : 0.0)};
        const c128 vl = {lane < len2 ? v.re : 0.0, lane < len2 ? v.im : 0.0};
        if (lane < B) sv[lane] = vl;
        const int64_t slot = (int64_t)s * ks + k + 1;
        if (lane < B) V2[slot * B + lane] = vl;
        if (lane == 0) tau2[slot] = h.tau;
        tau = h.tau;
        if (r == 0) {
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int q = g + 4 * t;
                if (q == 0) Er[t] = c128{h.beta, 0.0};
                else if (refl) Er[t] = c128{0.0, 0.0};
            }
        }
        wave_lds_sync();
        // (d) E[:, 1..] <- H_{k+1}^H E = E - conj(tau') v (v^H E) (column sums)
        if (refl) {
            c128 vn[4];
            c128 cs = {0.0, 0.0};
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                vn[t] = sv[g + 4 * t];  // (zero past len2)
                cs = fma_conj(vn[t], Er[t], cs);
            }
            cs = conj(h.tau) * sum_rows4(cs);
            if (r >= 1) {
#pragma unroll
                for (int t = 0; t < 4; ++t) Er[t] = fms_c(cs, vn[t], Er[t]);
            }
        }
        // E back to the band
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int q = g + 4 * t;
            if (q < len2 && r < len) bref(Bd, m, st2 + q, st + r) = Er[t];
        }
    } else if (nonzero(tau)) {
        // no block below: nothing else to store
    }
    wave_lds_sync();
    return tau;
}

// NT threads (NT / 64 concurrent sweeps).  GBAND: the band in global memory
// (Bglob, 2B x m per matrix; L2-resident for the matrices in flight) for sizes
// whose band does not fit in LDS; the workgroup's waves are on one CU, so the
// same wave- and workgroup-level ordering holds as for LDS.
template <int NT, bool GBAND, int STG>
__global__ __launch_bounds__(NT) void k_sb2st(const c128 *__restrict__ Gall, int m, int ks,
                                              double *__restrict__ dall, double *__restrict__ eall,
                                              c128 *__restrict__ V2all,
                                              c128 *__restrict__ tau2all, c128 *__restrict__ Bglob,
                                              int hi) {
    if (hi) __builtin_amdgcn_s_setprio(3);
    constexpr int NW = NT / 64;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int ip = blockIdx.x;
    c128 *Bd = GBAND ? Bglob + (int64_t)ip * 2 * B * m
                     : reinterpret_cast<c128 *>(smem_raw);  // 2B x m
    // per wave: sv[B], sw[B], sx[B]
    c128 *scr = GBAND ? reinterpret_cast<c128 *>(smem_raw) : Bd + (size_t)2 * B * m;
    const int tid = threadIdx.x, wave = tid >> 6;
    const c128 *A = Gall + (int64_t)ip * m * m;
#pragma unroll 4
    for (int idx = tid; idx < 2 * B * m; idx += NT) {
        const int j = idx / (2 * B), dg = idx % (2 * B);
        Bd[idx] = sel(dg <= B && j + dg < m, A[min(j + dg, m - 1) + (int64_t)m * j]);
    }
    __syncthreads();
    c128 *sv = scr + wave * 3 * B;
    c128 *sw = sv + B;
    c128 *sx = sw + B;
    c128 *V2 = V2all + (int64_t)ip * m * ks * B;
    c128 *tau2 = tau2all + (int64_t)ip * m * ks;
    c128 tau = {0.0, 0.0};
    int s = wave;  // this wave's current sweep: s, s + NW, ...
    const int nsweep = m - 1;
    const int tend = STG * (nsweep - 1) + 1;
    for (int t = 0; t < tend; ++t) {
        if (s < nsweep) {
            const int k = t - STG * s;
            const int nst = (m - s - 1 + B - 1) / B;
            if (k >= 0 && k < nst) {
                tau = sb2st_task(Bd, m, s, k, sv, sw, sx, tau, V2, tau2, ks);
                if (k == nst - 1) s += NW;
            }
        }
        __syncthreads();
    }
    double *d = dall + (int64_t)ip * m;
    double *e = eall + (int64_t)ip * m;
    for (int i = tid; i < m; i += NT) {
        d[i] = Bd[i * 2 * B].re;
        e[i] = i + 1 < m ? Bd[i * 2 * B + 1].re : 0.0;
    }
}

// Back-transformation U[:, c0 + c] = Q1 Q2 Z[:, c0 + c] for BTC eigenvectors per
// workgroup, accumulated in FP64 in LDS (BTC x m).
// Q2, the sweep-major product of the bulge-chasing reflectors, is applied in
// reverse sweep order, z <- z - tau v (v^H z); the reflectors of one sweep act
// on disjoint row windows and are applied together, one thread per (reflector,
// eigenvector) with its 16-term product in registers.  A sweep's reflectors are
// staged in LDS; the next BT_PD sweeps are in flight in registers (a ring, the
// loop unrolled by its depth), so a sweep waits for no global load.
// Q1, the panels in reverse order, U <- U - V (T (V^H U)) on the FP64 matrix
// cores with the V operand (unit lower form, the tails stored below the band of
// the reduced matrix) loaded from global memory: X = V^H U row block by row block
// over the four waves (partials summed through LDS in a fixed order), Y = T X,
// then U -= V Y, each row block by one wave.
// KS: reflectors per sweep, at most (the smaller instance keeps two workgroups per
// CU up to m = SBR_MAXM).
constexpr int BT_KS_LDS = (SBR_MAXM - 1 + B - 1) / B;
constexpr int BT_KS_MAX = (SBR_MAXG - 1 + B - 1) / B;
constexpr int BT_KP = BTT / BTC;  // reflectors applied at once
static_assert(BT_KS_MAX * B <= 2 * BTT, "a sweep's reflectors are staged two per thread");
static_assert(BTC <= 16 && BTT == 256, "one 16-column MFMA tile of eigenvectors, four waves");
static_assert(BT_KP >= BT_KS_MAX, "one reflector of a sweep per thread");

template <int KS>
__global__ __launch_bounds__(BTT) void k_sbr_bt(const c128 *__restrict__ Gall, int m, int npanel,
                                                int ks, const c128 *__restrict__ Tall,
                                                const c128 *__restrict__ V2all,
                                                const c128 *__restrict__ tau2all,
                                                const double *__restrict__ Zall,
                                                const int *__restrict__ rall,
                                                c64 *__restrict__ Uall, int hi) {
    if (hi) __builtin_amdgcn_s_setprio(3);
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ c128 Ts[B][B + 1];
    // two sweeps' reflectors (stride BS: room for every thread's two values) / two
    // 16 x 16 partials of X
    constexpr int BS = KS * B > 2 * BTT ? KS * B : 2 * BTT;
    __shared__ c128 Vbuf[2 * BS];
    __shared__ c128 Tb[2][KS];
    // Q1's Y = T X tile, in the second staging buffer (idle during Q1)
    c128 (*XY)[B + 1] = reinterpret_cast<c128 (*)[B + 1]>(Vbuf + BS);
    static_assert(B * (B + 1) <= BS, "the Y tile fits the second staging buffer");
    __shared__ c128 xch[2][BT_KP][BTC];  // window rows handed from thread k to k + 1
    const int ip = blockIdx.x;
    const int r = rall[ip];
    const int c0 = blockIdx.y * BTC;
    if (c0 >= r) return;
    const int ncol = min(BTC, r - c0);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    c128 *Uc = reinterpret_cast<c128 *>(smem_raw);  // BTC x m
    const c128 *A = Gall + (int64_t)ip * m * m;
    const double *Z = Zall + (int64_t)ip * m * m;
#pragma unroll 4
    for (int idx = tid; idx < BTC * m; idx += BTT) {
        const int c = idx / m, i = idx % m;
        const double zr = Z[(int64_t)(c0 + min(c, ncol - 1)) * m + i];
        Uc[idx] = c128{c < ncol ? zr : 0.0, 0.0};
    }
    // Q2.  Thread (k, c) applies reflector k of every sweep to eigenvector c and
    // keeps its 16-row window of the eigenvector in registers: the windows of sweep s,
    // rows s+1+16k .. s+16+16k, tile rows s+1 .. m-1, and from sweep s to s-1 each
    // moves up one row -- the thread takes row s+16k from thread k-1 (row s, never
    // touched by a later sweep, from LDS for k = 0) and hands its last row to thread
    // k+1, through a double-buffered LDS slot.  The sweep loop is unrolled by B so
    // that the window's rotation is a renaming of registers.
    {
        const c128 *V2 = V2all + (int64_t)ip * m * ks * B;
        const c128 *tau2 = tau2all + (int64_t)ip * m * ks;
        auto nsweep_refl = [&](int s) { return (m - s - 1 + B - 1) / B; };
        // ring slot j holds this thread's share of a sweep's reflectors (<= 2
        // values) and tau: sweep s sits in slot (m - 2 - s) % BT_PD
        c128 rv0[BT_PD], rv1[BT_PD], rt[BT_PD];
        // unconditional loads from clamped addresses; the entries past the sweep's
        // reflectors are zeroed when the sweep is staged (put), so no loaded value is
        // consumed before then and the loads of the BT_PD sweeps ahead stay in flight
        // (with the LDS-only barriers of the sweep loop)
        auto fetch = [&](int s, c128 &v0, c128 &v1, c128 &t) {
            const int nr = nsweep_refl(s), n = nr * B;
            const c128 *src = V2 + (int64_t)s * ks * B;
            v0 = src[min(tid, n - 1)];
            v1 = src[min(tid + BTT, n - 1)];
            t = tau2[(int64_t)s * ks + min(tid, nr - 1)];
        };
        auto put = [&](int buf, int s, const c128 &v0, const c128 &v1, const c128 &t) {
            const int nr = nsweep_refl(s), n = nr * B;
            Vbuf[buf * BS + tid] = tid < n ? v0 : c128{0.0, 0.0};
            Vbuf[buf * BS + tid + BTT] = tid + BTT < n ? v1 : c128{0.0, 0.0};
            if (tid < KS) Tb[buf][tid] = tid < nr ? t : c128{0.0, 0.0};
        };
        const int k = tid / BTC, c = tid - k * BTC;  // BT_KP >= KS: one reflector a thread
        const bool col = c < ncol;
        c128 w[B];
#pragma unroll
        for (int q = 0; q < B; ++q) w[q] = c128{0.0, 0.0};
#pragma unroll
        for (int j = 0; j < BT_PD; ++j)
            if (m - 2 - j >= 0) fetch(m - 2 - j, rv0[j], rv1[j], rt[j]);
        if (m >= 2) put(0, m - 2, rv0[0], rv1[0], rt[0]);
        __syncthreads();  // (also orders the initialisation of Uc before the read below)
        if (k == 0 && col && m >= 2) w[0] = Uc[c * m + m - 1];
        int buf = 0;
        static_assert(B % BT_PD == 0, "the ring slot of an unrolled sweep is static");
        for (int s0 = m - 2; s0 >= 0; s0 -= B) {
#pragma clang loop unroll(full)
            for (int j = 0; j < B; ++j) {
                const int s = s0 - j;
                // slot (sweep s) is staged: refill it with sweep s - BT_PD, on every
                // step (clamped to sweep 0; never staged past it), so that the loads
                // in flight are the same on every path and each staging waits only
                // for its own sweep's loads
                fetch(max(s - BT_PD, 0), rv0[j % BT_PD], rv1[j % BT_PD], rt[j % BT_PD]);
                if (s >= 0) {  // (no early exit: a constant trip count unrolls fully)
                    const int nst = nsweep_refl(s);
                    if (k < nst && col) {
                        // the reflector in registers, zero past its length (rows past
                        // m - 1), so that the product and the update need no per-row
                        // condition; a zero tau makes the update a no-op
                        const c128 tk = Tb[buf][k];
                        const int len = min(B, m - (s + 1 + k * B));
                        const c128 *v = &Vbuf[buf * BS + k * B];
                        c128 vr[B];
#pragma unroll
                        for (int q = 0; q < B; ++q) {
                            const c128 t = v[q];
                            vr[q] = q < len ? t : c128{0.0, 0.0};
                        }
                        c128 d0 = {0.0, 0.0}, d1 = {0.0, 0.0};
#pragma unroll
                        for (int q = 0; q < B; q += 2) {
                            d0 = fma_conj(vr[q], w[(q - j + 16 * B) % B], d0);
                            d1 = fma_conj(vr[q + 1], w[(q + 1 - j + 16 * B) % B], d1);
                        }
                        const c128 f = tk * (d0 + d1);
#pragma unroll
                        for (int q = 0; q < B; ++q) {
                            c128 &wq = w[(q - j + 16 * B) % B];
                            wq = fms_c(f, vr[q], wq);
                        }
                    }
                    // the window's last row (logical q = B - 1) to thread k + 1
                    xch[s & 1][k][c] = w[(B - 1 - j + 16 * B) % B];
                    // sweep s - 1 (the next slot) -> the other staging buffer
                    if (s > 0) {
                        const int nx = (j + 1) % BT_PD;
                        put(buf ^ 1, s - 1, rv0[nx], rv1[nx], rt[nx]);
                    }
                    lds_barrier();  // (the ring's loads stay in flight)
                    buf ^= 1;
                    // rotate: logical q of sweep s - 1 is logical q - 1 of sweep s; the
                    // freed register (logical B - 1) becomes logical 0 = row s + 16k
                    if (s > 0) {
                        const c128 from_u = Uc[c * m + s];  // (c < BTC: a valid address)
                        const c128 from_x = xch[s & 1][k > 0 ? k - 1 : 0][c];
                        w[(B - 1 - j + 16 * B) % B] =
                            !col ? c128{0.0, 0.0} : (k == 0 ? from_u : from_x);
                    } else if (col && k < nsweep_refl(0)) {
                        // after sweep 0: the windows (rows 1+16k ..) back to LDS
#pragma unroll
                        for (int q = 0; q < B; ++q) {
                            const int row = 1 + k * B + q;
                            if (row < m) Uc[c * m + row] = w[(q - j + 16 * B) % B];
                        }
                    }
                }
            }
        }
        __syncthreads();
    }
    // Q1 on the matrix cores: v_mfma_f64_16x16x4f64 (D = P Q), lane (rl, kl) holds
    // P[rl][kl], Q[kl][rl] and D[kl + 4q][rl]
    const int rl = lane & 15, kl = lane >> 4;
    for (int p = npanel - 1; p >= 0; --p) {
        const int pc0 = p * B, r0 = pc0 + B, L = m - r0;
        const int kq = min(B, L - 1);
        const int nrb = (L + 15) / 16;
        for (int pr = tid; pr < B * B; pr += BTT)
            Ts[pr % B][pr / B] = Tall[((int64_t)ip * npanel + p) * B * B + pr];
        // V[i][a] of the panel (i < L): 1 on the diagonal, 0 above, the stored tail below
        // (as raw loads from rows clamped into the matrix, then the unit lower form:
        // loads without branches issue together and wait once per row block)
        auto vraw = [&](int i, int a) {
            return A[(r0 + min(i, L - 1)) + (int64_t)m * (pc0 + a)];
        };
        auto vfix = [&](c128 v, int i, int a) {
            const bool below = a < kq && i > a && i < L;
            const bool diag = a < kq && i == a;
            return c128{below ? v.re : (diag ? 1.0 : 0.0), below ? v.im : 0.0};
        };
        // X = V^H U: k-steps over the block's rows, P[a][i] = conj(V[i][a]) (a = rl),
        // Q[i][c] = U[c][r0 + i] (c = rl)
        md4 xr = {0.0, 0.0, 0.0, 0.0}, xi = {0.0, 0.0, 0.0, 0.0};
        for (int rb = wave; rb < nrb; rb += BTT / 64) {
            c128 vv[4], uu[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) vv[q] = vraw(rb * 16 + 4 * q + kl, rl);
#pragma unroll
            for (int q = 0; q < 4; ++q) {  // (clamped into the tile, masked)
                const int i = rb * 16 + 4 * q + kl;
                const c128 u0 = Uc[min(rl, BTC - 1) * m + r0 + min(i, L - 1)];
                uu[q] = (rl < ncol && i < L) ? u0 : c128{0.0, 0.0};
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int i = rb * 16 + 4 * q + kl;
                const c128 v = vfix(vv[q], i, rl);
                const c128 u = uu[q];
                // conj(v) u: re = vr ur + vi ui, im = vr ui - vi ur
                xr = __builtin_amdgcn_mfma_f64_16x16x4f64(v.re, u.re, xr, 0, 0, 0);
                xr = __builtin_amdgcn_mfma_f64_16x16x4f64(v.im, u.im, xr, 0, 0, 0);
                xi = __builtin_amdgcn_mfma_f64_16x16x4f64(v.re, u.im, xi, 0, 0, 0);
                xi = __builtin_amdgcn_mfma_f64_16x16x4f64(-v.im, u.re, xi, 0, 0, 0);
            }
        }
        // the four partials (lane: X[kl + 4q][rl]) summed in a fixed order: waves 0, 1
        // store, waves 2, 3 add theirs, then the two slots are added
        c128 *P0 = Vbuf;
        if (wave < 2) {
#pragma unroll
            for (int q = 0; q < 4; ++q)
                P0[wave * B * B + (kl + 4 * q) * B + rl] = c128{xr[q], xi[q]};
        }
        __syncthreads();
        if (wave >= 2) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                c128 &d = P0[(wave - 2) * B * B + (kl + 4 * q) * B + rl];
                d = d + c128{xr[q], xi[q]};
            }
        }
        __syncthreads();
        {   // Y = T X (T upper), thread (a, c) = (tid / B, tid % B)
            const int a = tid / B, c = tid % B;
            c128 y = {0.0, 0.0};
            for (int q = a; q < B; ++q) y = y + Ts[a][q] * (P0[q * B + c] + P0[B * B + q * B + c]);
            XY[a][c] = y;
        }
        __syncthreads();
        // U -= V Y: P[i][a] = V[rb 16 + rl][a], Q[a][c] = Y[a][c] (a = 4s + kl, c = rl);
        // lane (rl, kl) owns U[c = rl][r0 + rb 16 + kl + 4q]
        for (int rb = wave; rb < nrb; rb += BTT / 64) {
            md4 dr = {0.0, 0.0, 0.0, 0.0}, di = {0.0, 0.0, 0.0, 0.0};
            c128 vv[4];
#pragma unroll
            for (int s4 = 0; s4 < 4; ++s4) vv[s4] = vraw(rb * 16 + rl, 4 * s4 + kl);
#pragma unroll
            for (int s4 = 0; s4 < 4; ++s4) {
                const int a = 4 * s4 + kl;
                const c128 v = vfix(vv[s4], rb * 16 + rl, a);
                const c128 y = XY[a][rl];
                // v y: re = vr yr - vi yi, im = vr yi + vi yr
                dr = __builtin_amdgcn_mfma_f64_16x16x4f64(v.re, y.re, dr, 0, 0, 0);
                dr = __builtin_amdgcn_mfma_f64_16x16x4f64(-v.im, y.im, dr, 0, 0, 0);
                di = __builtin_amdgcn_mfma_f64_16x16x4f64(v.re, y.im, di, 0, 0, 0);
                di = __builtin_amdgcn_mfma_f64_16x16x4f64(v.im, y.re, di, 0, 0, 0);
            }
            if (rl < ncol) {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int i = rb * 16 + kl + 4 * q;
                    if (i < L) {
                        c128 &u = Uc[rl * m + r0 + i];
                        u = u - c128{dr[q], di[q]};
                    }
                }
            }
        }
        __syncthreads();
    }
    c64 *U = Uall + (int64_t)ip * m * m;
    for (int idx = tid; idx < BTC * m; idx += BTT) {
        const int c = idx / m, i = idx % m;
        if (c < ncol) U[(int64_t)(c0 + c) * m + i] = c64{(float)Uc[idx].re, (float)Uc[idx].im};
    }
}

// ------------------------------------------------------------------ host side

// Raises a kernel's dynamic LDS limit to what its static LDS leaves of the
// 160 KB per workgroup, and checks that `need` bytes fit.
static void raise_lds(const void *fn, size_t need) {
    hipFuncAttributes at{};
    HIP_CHECK(hipFuncGetAttributes(&at, fn));
    const size_t cap = (size_t)160 * 1024 - at.sharedSizeBytes;
    MSLR_REQUIRE(need <= cap, MSLR_ERR_UNSUPPORTED, "two-stage reduction LDS exceeds 160 KB");
    static std::vector<const void *> done;
    for (const void *d : done)
        if (d == fn) return;
    HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)cap));
    done.push_back(fn);
}

void sbr_reduce(mslr_ctx_s *ctx, c128 *G, int m, int64_t nb, const SbrWork &w, double *d,
                double *e) {
    MSLR_REQUIRE(sbr_eligible(m), MSLR_ERR_UNSUPPORTED, "two-stage reduction size");
    const bool gband = sbr_global_band(m);
    // a wave must finish its sweep before its next one (one sweep per wave later)
    // starts
    MSLR_REQUIRE(gband || sbr_steps(m) <= STAGGER * S2W, MSLR_ERR_UNSUPPORTED,
                 "sweep longer than the schedule");
    MSLR_REQUIRE(!gband || w.Band, MSLR_ERR_DEVICE, "no band workspace");
    const int npanel = sbr_panels(m), ks = sbr_steps(m);
    const int hi = svst_small_batch(nb);
    {
        const size_t lds = (size_t)(m - B) * B * sizeof(c128);
        raise_lds((const void *)k_sy2sb, lds);
        MSLR_PROF(ctx, "sy2sb");
        k_sy2sb<<<(unsigned)nb, S1T, lds, ctx->stream>>>(G, m, npanel, w.T1, w.W1, hi);
        HIP_LAUNCH_CHECK();
    }
    if (gband) {
        // the schedule: built-in 12 waves 3 steps apart, or one of the tuning's
        // compiled alternatives; each needs a sweep to end before the wave's next one
        auto run = [&](auto kern, int nt, int stagger) {
            MSLR_REQUIRE(ks <= stagger * (nt / 64), MSLR_ERR_UNSUPPORTED, "sweep longer than the schedule");
            const size_t lds = (size_t)(nt / 64) * 3 * B * sizeof(c128);
            raise_lds((const void *)kern, lds);
            MSLR_PROF(ctx, "sb2st");
            kern<<<(unsigned)nb, nt, lds, ctx->stream>>>(G, m, ks, d, e, w.V2, w.tau2, w.Band, hi);
            HIP_LAUNCH_CHECK();
        };
        switch (ctx->tuning.sb2st_schedule) {
            case 1: run(k_sb2st<1024, true, 2>, 1024, 2); break;
            case 2: run(k_sb2st<384, true, 5>, 384, 5); break;
            case 3: run(k_sb2st<256, true, 8>, 256, 8); break;
            default: run(k_sb2st<S2TG, true, STAGGER_G>, S2TG, STAGGER_G); break;
        }
    } else {
        const size_t lds = ((size_t)2 * B * m + (size_t)S2W * 3 * B) * sizeof(c128);
        raise_lds((const void *)k_sb2st<S2T, false, STAGGER>, lds);
        MSLR_PROF(ctx, "sb2st");
        k_sb2st<S2T, false, STAGGER><<<(unsigned)nb, S2T, lds, ctx->stream>>>(G, m, ks, d, e, w.V2, w.tau2,
                                                                      nullptr, hi);
        HIP_LAUNCH_CHECK();
    }
}

void sbr_backtransform(mslr_ctx_s *ctx, const c128 *G, int m, int64_t nb, const SbrWork &w,
                       const double *Z, const int *r, c64 *U) {
    const int npanel = sbr_panels(m), ks = sbr_steps(m);
    const size_t lds = (size_t)BTC * m * sizeof(c128);
    MSLR_REQUIRE(ks <= BT_KS_MAX, MSLR_ERR_UNSUPPORTED, "two-stage back-transformation size");
    const bool small = ks <= BT_KS_LDS;
    raise_lds(small ? (const void *)k_sbr_bt<BT_KS_LDS> : (const void *)k_sbr_bt<BT_KS_MAX>, lds);
    dim3 g((unsigned)nb, (unsigned)cdiv(m, BTC));
    const int hi = svst_small_batch(nb);
    MSLR_PROF(ctx, "backtransform");
    if (small)
        k_sbr_bt<BT_KS_LDS><<<g, BTT, lds, ctx->stream>>>(G, m, npanel, ks, w.T1, w.V2, w.tau2, Z, r, U, hi);
    else
        k_sbr_bt<BT_KS_MAX><<<g, BTT, lds, ctx->stream>>>(G, m, npanel, ks, w.T1, w.V2, w.tau2, Z, r, U, hi);
    HIP_LAUNCH_CHECK();
}

}  // namespace mslr
