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
constexpr int S1T = 512;       // threads of the dense -> band kernel
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
constexpr int BTC = 13;        // eigenvectors per back-transformation workgroup: one thread per
                               // (reflector, eigenvector) of a sweep, LDS for two workgroups per CU

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
        h.tau = c128{(beta - alpha.re) / beta, -alpha.im / beta};
        h.scal = c_div(c128{1.0, 0.0}, c128{alpha.re - beta, alpha.im});
        h.beta = beta;
    }
    return h;
}
__device__ __forceinline__ bool nonzero(c128 t) { return t.re != 0.0 || t.im != 0.0; }

typedef double md4 __attribute__((ext_vector_type(4)));

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
// X = V^H W turned into M = T^H X in place) in LDS, at most 80 KB; Y and W in global
// memory (L2); at most 128 VGPRs (four waves per SIMD).
__global__ __launch_bounds__(S1T, 4) void k_sy2sb(c128 *__restrict__ Gall, int m, int npanel,
                                                  c128 *__restrict__ Tall,
                                                  c128 *__restrict__ W1all) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ c128 Ts[B][B + 1];
    __shared__ c128 Ms[B][B + 1];
    __shared__ c128 taus[B];
    __shared__ double betas[B];
    const int Lmax = m - B;
    c128 *V = reinterpret_cast<c128 *>(smem_raw);
    c128 *Y = W1all + (int64_t)blockIdx.x * sbr_w1_elems(m);
    c128 *W = Y + (size_t)Lmax * B;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    c128 *A = Gall + (int64_t)blockIdx.x * m * m;
    c128 *Tg = Tall + (int64_t)blockIdx.x * npanel * B * B;
    int p = 0;
    for (int c0 = 0; m - c0 - B >= 2; c0 += B, ++p) {
        const int L = m - c0 - B, r0 = c0 + B;
        const int kq = min(B, L - 1);
        // 1. panel rows r0.. of columns c0 .. c0+B-1 -> V
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
        // 5. Y = A22 V -> W on the FP64 matrix cores: one 16-row block per wave at a
        //    time; the A22 operand is read from global memory once.  Only the lower
        //    triangle of A22 is kept up to date: an entry above the diagonal is read
        //    as the conjugate of its mirror.
        //    v_mfma_f64_16x16x4f64 (D = P Q): lane l holds P[l%16][l/16],
        //    Q[l/16][l%16] and D[l/16 + 4v][l%16].
        {
            const int nrb = (L + 15) / 16;
            const int rl = lane & 15, kl = lane >> 4;
            for (int rb = wave; rb < nrb; rb += S1T / 64) {
                // four independent accumulator chains (summed below) keep the
                // matrix pipe busy between dependent MFMAs
                md4 yr = {0.0, 0.0, 0.0, 0.0}, yi = {0.0, 0.0, 0.0, 0.0};
                md4 yr2 = {0.0, 0.0, 0.0, 0.0}, yi2 = {0.0, 0.0, 0.0, 0.0};
                const int i = rb * 16 + rl;
#pragma unroll 4
                for (int ks = 0; ks < L; ks += 4) {
                    const int k = ks + kl;
                    const bool lo = i >= k;
                    c128 a = ld_if(i < L && k < L, lo ? A + (r0 + i) + (int64_t)m * (r0 + k)
                                                      : A + (r0 + k) + (int64_t)m * (r0 + i));
                    if (!lo) a.im = -a.im;
                    const c128 v = ld_if(k < L, V + rl * L + k);
                    // operands swapped: the block computed is Y^T, so the output
                    // lanes run down the columns of W
                    yr = __builtin_amdgcn_mfma_f64_16x16x4f64(v.re, a.re, yr, 0, 0, 0);
                    yi = __builtin_amdgcn_mfma_f64_16x16x4f64(v.im, a.re, yi, 0, 0, 0);
                    yr2 = __builtin_amdgcn_mfma_f64_16x16x4f64(v.im, -a.im, yr2, 0, 0, 0);
                    yi2 = __builtin_amdgcn_mfma_f64_16x16x4f64(v.re, a.im, yi2, 0, 0, 0);
                }
                yr += yr2;
                yi += yi2;
                if (i < L) {
#pragma unroll
                    for (int q = 0; q < 4; ++q) Y[(kl + 4 * q) * L + i] = c128{yr[q], yi[q]};
                }
            }
        }
        __syncthreads();
        // 6. W <- Y T (T upper)
        for (int task = tid; task < L * B; task += S1T) {
            const int i = task % L, c = task / L;
            c128 acc = {0.0, 0.0};
            for (int a = 0; a <= c; ++a) acc = acc + Y[a * L + i] * Ts[a][c];
            W[c * L + i] = acc;
        }
        __syncthreads();
        // 7. X = V^H Y (two threads per entry over interleaved rows, two partial sums
        //    each; V_a is zero above row a), M = T^H X (Hermitian)
        {
            const int pr = tid >> 1, part = tid & 1;
            const int a = pr % B, c = pr / B;
            const c128 *va = V + a * L, *wc = W + c * L;
            c128 s0 = {0.0, 0.0}, s1 = {0.0, 0.0};
            int i = a + part;
            for (; i + 2 < L; i += 4) {
                s0 = s0 + cmulc(va[i], wc[i]);
                s1 = s1 + cmulc(va[i + 2], wc[i + 2]);
            }
            if (i < L) s0 = s0 + cmulc(va[i], wc[i]);
            s0 = s0 + s1;
            s0.re += __shfl_xor(s0.re, 1, 64);
            s0.im += __shfl_xor(s0.im, 1, 64);
            if (part == 0) Ms[a][c] = s0;  // X
        }
        __syncthreads();
        // M = T^H X in place, one column per thread, bottom row first (row a needs
        // rows 0..a of X)
        if (tid < B) {
            const int c = tid;
            for (int a = B - 1; a >= 0; --a) {
                c128 s = {0.0, 0.0};
                for (int q = 0; q <= a; ++q) s = s + cmulc(Ts[q][a], Ms[q][c]);
                Ms[a][c] = s;
            }
        }
        __syncthreads();
        // 8. W = Y - 1/2 V M on the matrix cores, one 16-row block per wave at a time
        //    (the block computed is W^T: lane (g, rl) holds W[rl][g + 4q])
        {
            const int nrb = (L + 15) / 16;
            const int rl = lane & 15, kl = lane >> 4;
            for (int rb = wave; rb < nrb; rb += S1T / 64) {
                const int i = rb * 16 + rl;
                md4 wr, wi;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const c128 w0 = ld_if(i < L, W + (kl + 4 * q) * L + i);
                    wr[q] = w0.re;
                    wi[q] = w0.im;
                }
#pragma unroll
                for (int as = 0; as < B; as += 4) {
                    const int a = as + kl;
                    const c128 mm = Ms[a][rl];  // P[c][a] = M[a][c], c = rl
                    const c128 v = ld_if(i < L, V + a * L + i);
                    // V[i][a] M[a][c]: re = vr mr - vi mi, im = vr mi + vi mr
                    wr = __builtin_amdgcn_mfma_f64_16x16x4f64(-0.5 * mm.re, v.re, wr, 0, 0, 0);
                    wr = __builtin_amdgcn_mfma_f64_16x16x4f64(0.5 * mm.im, v.im, wr, 0, 0, 0);
                    wi = __builtin_amdgcn_mfma_f64_16x16x4f64(-0.5 * mm.im, v.re, wi, 0, 0, 0);
                    wi = __builtin_amdgcn_mfma_f64_16x16x4f64(-0.5 * mm.re, v.im, wi, 0, 0, 0);
                }
                if (i < L) {
#pragma unroll
                    for (int q = 0; q < 4; ++q) W[(kl + 4 * q) * L + i] = c128{wr[q], wi[q]};
                }
            }
        }
        __syncthreads();
        // 9. A22 <- A22 - V W^H - W V^H on the FP64 matrix cores, the 16x16 blocks on
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
                    const c128 a = ld_if(i < L && k < L, A + (r0 + i) + (int64_t)m * (r0 + k));
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
                    const c128 vI = ld_if(ii < L, V + j * L + ii), wI = ld_if(ii < L, W + j * L + ii);
                    const c128 vK = ld_if(kk < L, V + j * L + kk), wK = ld_if(kk < L, W + j * L + kk);
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
    // D (Hermitian, upper part from the lower) and E into registers
    c128 Dr[4], Er[4], vq[4];
    const c128 vr = ld_if(r < len, sv + r);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int q = g + 4 * t;
        vq[t] = ld_if(q < len, sv + q);
        Dr[t] = c128{0.0, 0.0};
        if (q < len && r < len) {
            if (q >= r) Dr[t] = bref(Bd, m, st + q, st + r);
            else Dr[t] = conj(bref(Bd, m, st + r, st + q));
        }
        Er[t] = ld_if(q < len2 && r < len, &bref(Bd, m, st2 + q, st + r));
    }
    if (nonzero(tau)) {
        // p = tau D v and tau E v (row sums)
        c128 pd[4], pe[4];
        c128 part = {0.0, 0.0};
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            pd[t] = tau * sum16(Dr[t] * vr);
            pe[t] = tau * sum16(Er[t] * vr);
            part = part + cmulc(pd[t], vq[t]);
        }
        // p^H v: the rows of lanes hold disjoint q; add the four row results
        const c128 dot = {(readlane_d(part.re, 0) + readlane_d(part.re, 16)) +
                              (readlane_d(part.re, 32) + readlane_d(part.re, 48)),
                          (readlane_d(part.im, 0) + readlane_d(part.im, 16)) +
                              (readlane_d(part.im, 32) + readlane_d(part.im, 48))};
        const c128 alpha2 = (-0.5) * (tau * dot);
        // w_q for the own rows; w_r for the own column through LDS
        wave_lds_sync();
        if (r == 0) {
#pragma unroll
            for (int t = 0; t < 4; ++t) sw[g + 4 * t] = pd[t];
        }
        wave_lds_sync();
        const c128 wr = ld_if(r < len, sw + r) + alpha2 * vr;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const c128 wq = pd[t] + alpha2 * vq[t];
            Dr[t] = Dr[t] - vq[t] * conj(wr) - wq * conj(vr);   // D <- H^H D H
            Er[t] = Er[t] - pe[t] * conj(vr);                   // E <- E H
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
                vn[t] = ld_if(g + 4 * t < len2, sv + g + 4 * t);
                cs = cs + cmulc(vn[t], Er[t]);
            }
            cs.re += __shfl_xor(cs.re, 16, 64);
            cs.im += __shfl_xor(cs.im, 16, 64);
            cs.re += __shfl_xor(cs.re, 32, 64);
            cs.im += __shfl_xor(cs.im, 32, 64);
            cs = conj(h.tau) * cs;
            if (r >= 1) {
#pragma unroll
                for (int t = 0; t < 4; ++t) Er[t] = Er[t] - vn[t] * cs;
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
                                              c128 *__restrict__ tau2all, c128 *__restrict__ Bglob) {
    constexpr int NW = NT / 64;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int ip = blockIdx.x;
    c128 *Bd = GBAND ? Bglob + (int64_t)ip * 2 * B * m
                     : reinterpret_cast<c128 *>(smem_raw);  // 2B x m
    // per wave: sv[B], sw[B], sx[B]
    c128 *scr = GBAND ? reinterpret_cast<c128 *>(smem_raw) : Bd + (size_t)2 * B * m;
    const int tid = threadIdx.x, wave = tid >> 6;
    const c128 *A = Gall + (int64_t)ip * m * m;
    for (int idx = tid; idx < 2 * B * m; idx += NT) {
        const int j = idx / (2 * B), dg = idx % (2 * B);
        Bd[idx] = ld_if(dg <= B && j + dg < m, A + (j + dg) + (int64_t)m * j);
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
// workgroup, accumulated in FP64 in LDS (BTC x m: two workgroups per CU).
// Q2, the sweep-major product of the bulge-chasing reflectors, is applied in
// reverse sweep order, z <- z - tau v (v^H z); the reflectors of one sweep act
// on disjoint row windows and are applied together, one thread per (reflector,
// eigenvector) with its 16-term product in registers.  A sweep's reflectors are
// staged in LDS, the next sweep's loaded while the current one is applied.  Q1,
// the panels in reverse order, as U <- U - V (T (V^H U)) with V (unit lower
// form, tails stored below the band of the reduced matrix) passed through LDS
// in 16 x 16 tiles, again loading the next tile during the current one.
// KS: reflectors per sweep, at most (the staging of a sweep in LDS; the smaller
// instance keeps two workgroups per CU up to m = SBR_MAXM).
constexpr int BT_KS_LDS = (SBR_MAXM - 1 + B - 1) / B;
constexpr int BT_KS_MAX = (SBR_MAXG - 1 + B - 1) / B;
constexpr int BT_KP = BTT / BTC;  // reflectors applied at once
static_assert(BT_KS_MAX * B <= 2 * BTT, "a sweep's reflectors are staged two per thread");
static_assert(BT_KS_LDS * B >= B * B, "the staging buffer also holds a 16 x 16 tile");

template <int KS>
__global__ __launch_bounds__(BTT) void k_sbr_bt(const c128 *__restrict__ Gall, int m, int npanel,
                                                int ks, const c128 *__restrict__ Tall,
                                                const c128 *__restrict__ V2all,
                                                const c128 *__restrict__ tau2all,
                                                const double *__restrict__ Zall,
                                                const int *__restrict__ rall,
                                                c64 *__restrict__ Uall) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ c128 Ts[B][B + 1];
    __shared__ c128 XY[B][BTC + 1];
    __shared__ c128 Vb[2][KS * B];  // a sweep's reflectors / a 16 x 16 tile of V
    __shared__ c128 Tb[2][KS];
    const int ip = blockIdx.x;
    const int r = rall[ip];
    const int c0 = blockIdx.y * BTC;
    if (c0 >= r) return;
    const int ncol = min(BTC, r - c0);
    const int tid = threadIdx.x;
    c128 *Uc = reinterpret_cast<c128 *>(smem_raw);  // BTC x m
    const c128 *A = Gall + (int64_t)ip * m * m;
    const double *Z = Zall + (int64_t)ip * m * m;
    for (int idx = tid; idx < BTC * m; idx += BTT) {
        const int c = idx / m, i = idx % m;
        c128 z = {0.0, 0.0};
        if (c < ncol) z.re = Z[(int64_t)(c0 + c) * m + i];
        Uc[idx] = z;
    }
    // Q2
    {
        const c128 *V2 = V2all + (int64_t)ip * m * ks * B;
        const c128 *tau2 = tau2all + (int64_t)ip * m * ks;
        auto nsweep_refl = [&](int s) { return (m - s - 1 + B - 1) / B; };
        // this thread's share of a sweep's reflectors (<= 2 values) and tau
        c128 pv0, pv1, pt;
        auto fetch = [&](int s) {
            const int n = nsweep_refl(s) * B;
            const c128 *src = V2 + (int64_t)s * ks * B;
            pv0 = ld_if(tid < n, src + tid);
            pv1 = ld_if(tid + BTT < n, src + tid + BTT);
            pt = ld_if(tid < nsweep_refl(s), tau2 + (int64_t)s * ks + tid);
        };
        auto put = [&](int buf) {
            Vb[buf][tid] = pv0;
            if (tid + BTT < KS * B) Vb[buf][tid + BTT] = pv1;
            if (tid < KS) Tb[buf][tid] = pt;
        };
        // (reflector, eigenvector) of this thread; reflectors k, k + BT_KP, ...
        const int k = tid / BTC, c = tid - k * BTC;
        if (m >= 2) {
            fetch(m - 2);
            put(0);
        }
        __syncthreads();
        for (int s = m - 2, buf = 0; s >= 0; --s, buf ^= 1) {
            if (s > 0) fetch(s - 1);
            const int nst = nsweep_refl(s);
            for (int kk = k; k < BT_KP && kk < nst; kk += BT_KP) {
                const c128 tk = Tb[buf][kk];
                if (!(c < ncol && nonzero(tk))) continue;
                const int st = s + 1 + kk * B;
                const int len = min(B, m - st);
                const c128 *v = &Vb[buf][kk * B];
                c128 *u = Uc + c * m + st;
                c128 vv[B], uu[B];
#pragma unroll
                for (int q = 0; q < B; ++q) {
                    vv[q] = ld_if(q < len, v + q);
                    uu[q] = ld_if(q < len, u + q);
                }
                c128 d0 = {0.0, 0.0}, d1 = {0.0, 0.0};
#pragma unroll
                for (int q = 0; q < B; q += 2) {
                    d0 = d0 + cmulc(vv[q], uu[q]);
                    d1 = d1 + cmulc(vv[q + 1], uu[q + 1]);
                }
                const c128 f = tk * (d0 + d1);
#pragma unroll
                for (int q = 0; q < B; ++q)
                    if (q < len) u[q] = uu[q] - f * vv[q];
            }
            if (s > 0) put(buf ^ 1);
            __syncthreads();
        }
    }
    // Q1
    const int ta = tid % B, tr = tid / B;  // tile slot: V[i0 + tr][ta]; also (a, c) / (row, c)
    for (int p = npanel - 1; p >= 0; --p) {
        const int pc0 = p * B, r0 = pc0 + B, L = m - r0;
        const int kq = min(B, L - 1);
        for (int pr = tid; pr < B * B; pr += BTT)
            Ts[pr % B][pr / B] = Tall[((int64_t)ip * npanel + p) * B * B + pr];
        auto vtile = [&](int i0) {
            const int i = i0 + tr;
            c128 v = ld_if(ta < kq && i > ta && i < L, A + (r0 + i) + (int64_t)m * (pc0 + ta));
            if (ta < kq && i == ta) v.re = 1.0;
            return v;
        };
        // X = V^H U: thread (a, c) = (ta, tr)
        c128 x0 = {0.0, 0.0}, x1 = {0.0, 0.0};
        c128 pv = vtile(0);
        for (int i0 = 0, buf = 0; i0 < L; i0 += B, buf ^= 1) {
            Vb[buf][tr * B + ta] = pv;
            __syncthreads();
            if (i0 + B < L) pv = vtile(i0 + B);
            if (tr < BTC) {
                const c128 *uc = Uc + tr * m + r0 + i0;
                const int n = min(B, L - i0);
#pragma unroll
                for (int ii = 0; ii < B; ii += 2) {
                    if (ii < n) x0 = x0 + cmulc(Vb[buf][ii * B + ta], uc[ii]);
                    if (ii + 1 < n) x1 = x1 + cmulc(Vb[buf][(ii + 1) * B + ta], uc[ii + 1]);
                }
            }
        }
        if (tr < BTC) XY[ta][tr] = x0 + x1;
        __syncthreads();
        {   // Y = T X (T upper)
            c128 y = {0.0, 0.0};
            if (tr < BTC)
                for (int q = ta; q < B; ++q) y = y + Ts[ta][q] * XY[q][tr];
            __syncthreads();
            if (tr < BTC) XY[ta][tr] = y;
        }
        // U <- U - V Y: thread (row i0 + ta, c = tr)
        pv = vtile(0);
        for (int i0 = 0, buf = 0; i0 < L; i0 += B, buf ^= 1) {
            Vb[buf][tr * B + ta] = pv;
            __syncthreads();
            if (i0 + B < L) pv = vtile(i0 + B);
            const int i = i0 + ta;
            if (tr < BTC && i < L) {
                c128 y0 = {0.0, 0.0}, y1 = {0.0, 0.0};
#pragma unroll
                for (int a2 = 0; a2 < B; a2 += 2) {
                    y0 = y0 + Vb[buf][ta * B + a2] * XY[a2][tr];
                    y1 = y1 + Vb[buf][ta * B + a2 + 1] * XY[a2 + 1][tr];
                }
                Uc[tr * m + r0 + i] = Uc[tr * m + r0 + i] - (y0 + y1);
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
    {
        const size_t lds = (size_t)(m - B) * B * sizeof(c128);
        raise_lds((const void *)k_sy2sb, lds);
        MSLR_PROF(ctx, "sy2sb");
        k_sy2sb<<<(unsigned)nb, S1T, lds, ctx->stream>>>(G, m, npanel, w.T1, w.W1);
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
            kern<<<(unsigned)nb, nt, lds, ctx->stream>>>(G, m, ks, d, e, w.V2, w.tau2, w.Band);
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
                                                                      nullptr);
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
    MSLR_PROF(ctx, "backtransform");
    if (small)
        k_sbr_bt<BT_KS_LDS><<<g, BTT, lds, ctx->stream>>>(G, m, npanel, ks, w.T1, w.V2, w.tau2, Z, r, U);
    else
        k_sbr_bt<BT_KS_MAX><<<g, BTT, lds, ctx->stream>>>(G, m, npanel, ks, w.T1, w.V2, w.tau2, Z, r, U);
    HIP_LAUNCH_CHECK();
}

}  // namespace mslr
