// Two-stage Hermitian tridiagonalisation for the batched patch SVST.
//
// One-stage Householder tridiagonalisation re-reads the whole trailing matrix
// for every column (its matrix-vector product), which for hundreds of 300x300
// FP64 matrices streams ~n/3 matrix volumes from HBM.  Here:
//   stage 1 (k_sy2sb): dense -> band of width B.  Per panel of B columns a
//     Householder QR (one wave, in LDS) gives Q = I - V T V^H, and the trailing
//     matrix gets A <- Q^H A Q = A - V W^H - W V^H with Y = A V T,
//     W = Y - 1/2 V (T^H V^H Y): three passes over the trailing matrix per B
//     columns instead of one per column.
//   stage 2 (k_sb2st): band -> real tridiagonal by bulge chasing, the band
//     (2B diagonals) held in LDS; each wave runs one sweep at a time, sweeps
//     start STAGGER time steps apart so that concurrent tasks touch disjoint
//     index ranges (one workgroup barrier per time step).
//   back-transformation (k_sbr_bt): U = Q1 Q2 Z for the wanted eigenvectors.
// Reflectors follow zlarfg: H = I - tau v v^H, H^H x = beta e1, beta real.
#include <algorithm>
#include <cfloat>
#include <cmath>

#include "sbr.hpp"

namespace mslr {

constexpr int B = SBR_B;
constexpr int S1T = 512;       // threads of the dense -> band kernel
constexpr int S2T = 512;       // threads of the band -> tridiagonal kernel
constexpr int S2W = S2T / 64;  // concurrent sweeps
constexpr int STAGGER = 3;     // time steps between the starts of consecutive sweeps
constexpr int BTT = 256;       // threads of the back-transformation
constexpr int BTC = BTT / B;   // eigenvectors per back-transformation workgroup

__device__ __forceinline__ c128 c_div(c128 a, c128 b) {
    const double den = b.re * b.re + b.im * b.im;
    return c128{(a.re * b.re + a.im * b.im) / den, (a.im * b.re - a.re * b.im) / den};
}

__device__ __forceinline__ double wsum_d(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ c128 wsum_c(c128 v) {
    for (int o = 32; o > 0; o >>= 1) {
        v.re += __shfl_xor(v.re, o, 64);
        v.im += __shfl_xor(v.im, o, 64);
    }
    return v;
}
__device__ __forceinline__ c128 shfl_c(c128 v, int src) {
    return c128{__shfl(v.re, src, 64), __shfl(v.im, src, 64)};
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

// ------------------------------------------------------------------ stage 1
// One workgroup per matrix.  LDS: V and W, (m - B) x B each, column-major.
__global__ __launch_bounds__(S1T) void k_sy2sb(c128 *__restrict__ Gall, int m, int npanel,
                                               c128 *__restrict__ Tall) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ c128 Ts[B][B + 1];
    __shared__ c128 Ms[B][B + 1];
    __shared__ c128 Xs[B][B + 1];
    __shared__ c128 taus[B];
    __shared__ double betas[B];
    const int Lmax = m - B;
    c128 *V = reinterpret_cast<c128 *>(smem_raw);
    c128 *W = V + (size_t)Lmax * B;
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
        // 2. Householder QR of the panel by wave 0; lane l owns rows l + 64t of
        //    every column, so lanes exchange values only through shuffles.
        if (wave == 0) {
            for (int j = 0; j < B; ++j) {
                c128 *vj = V + j * L;
                if (j >= kq) {
                    if (lane == 0) taus[j] = c128{0.0, 0.0};
                    continue;
                }
                double xn = 0.0;
                c128 own = {0.0, 0.0};
                for (int i = lane; i < L; i += 64) {
                    if (i > j) xn += abs2(vj[i]);
                    if (i == j) own = vj[i];
                }
                xn = wsum_d(xn);
                const c128 alpha = shfl_c(own, j & 63);
                const Hh h = householder(alpha, xn);
                const bool refl = nonzero(h.tau);
                if (refl)
                    for (int i = lane; i < L; i += 64)
                        if (i > j) vj[i] = vj[i] * h.scal;
                if (lane == 0) {
                    taus[j] = h.tau;
                    betas[j] = h.beta;
                }
                if (!refl) continue;
                // columns c > j: x <- H^H x = x - conj(tau) v (v^H x)
                const c128 ct = conj(h.tau);
                for (int c = j + 1; c < B; ++c) {
                    c128 *pc = V + c * L;
                    c128 s = {0.0, 0.0};
                    for (int i = lane; i < L; i += 64) {
                        if (i < j) continue;
                        const c128 vi = i == j ? c128{1.0, 0.0} : vj[i];
                        s = s + cmulc(vi, pc[i]);
                    }
                    s = wsum_c(s);
                    const c128 f = ct * s;
                    for (int i = lane; i < L; i += 64) {
                        if (i < j) continue;
                        const c128 vi = i == j ? c128{1.0, 0.0} : vj[i];
                        pc[i] = pc[i] - vi * f;
                    }
                }
            }
        }
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
        // 4. T (upper triangular, Q = I - V T V^H): S = V^H V, then the zlarft recurrence
        for (int pr = tid; pr < B * B; pr += S1T) {
            const int a = pr % B, c = pr / B;
            if (a < c) {
                c128 s = {0.0, 0.0};
                for (int i = 0; i < L; ++i) s = s + cmulc(V[a * L + i], V[c * L + i]);
                Ms[a][c] = s;
            }
        }
        __syncthreads();
        if (wave == 0) {
            for (int j = 0; j < B; ++j) {
                const c128 tj = taus[j];
                const c128 tmp = lane < j ? (-1.0) * (tj * Ms[lane][j]) : c128{0.0, 0.0};
                c128 acc = {0.0, 0.0};
                for (int q = 0; q < j; ++q) {
                    const c128 tq = shfl_c(tmp, q);
                    if (lane <= q) acc = acc + Ts[lane][q] * tq;
                }
                if (lane < j) Ts[lane][j] = acc;
                else if (lane == j) Ts[j][j] = tj;
                else if (lane < B) Ts[lane][j] = c128{0.0, 0.0};
            }
        }
        __syncthreads();
        for (int pr = tid; pr < B * B; pr += S1T) Tg[(int64_t)p * B * B + pr] = Ts[pr % B][pr / B];
        // 5. Y = A22 V -> W (two threads per row, eight columns each)
        for (int task = tid; task < 2 * L; task += S1T) {
            const int i = task % L, h = task / L;
            c128 acc[8];
#pragma unroll
            for (int t = 0; t < 8; ++t) acc[t] = c128{0.0, 0.0};
            const c128 *arow = A + (r0 + i) + (int64_t)m * r0;
            const c128 *vb = V + (8 * h) * L;
#pragma unroll 4
            for (int k = 0; k < L; ++k) {
                const c128 a = arow[(int64_t)m * k];
#pragma unroll
                for (int t = 0; t < 8; ++t) acc[t] = acc[t] + a * vb[t * L + k];
            }
#pragma unroll
            for (int t = 0; t < 8; ++t) W[(8 * h + t) * L + i] = acc[t];
        }
        __syncthreads();
        // 6. Y <- Y T (row by row, T upper)
        for (int i = tid; i < L; i += S1T) {
            c128 (*T_)[B + 1] = lds_opaque(Ts);
            c128 y[B];
#pragma unroll
            for (int a = 0; a < B; ++a) y[a] = W[a * L + i];
#pragma unroll
            for (int c = 0; c < B; ++c) {
                c128 acc = {0.0, 0.0};
#pragma unroll
                for (int a = 0; a <= c; ++a) acc = acc + y[a] * T_[a][c];
                W[c * L + i] = acc;
            }
        }
        __syncthreads();
        // 7. X = V^H Y, M = T^H X (Hermitian)
        for (int pr = tid; pr < B * B; pr += S1T) {
            const int a = pr % B, c = pr / B;
            c128 s = {0.0, 0.0};
            for (int i = 0; i < L; ++i) s = s + cmulc(V[a * L + i], W[c * L + i]);
            Xs[a][c] = s;
        }
        __syncthreads();
        for (int pr = tid; pr < B * B; pr += S1T) {
            const int a = pr % B, c = pr / B;
            c128 s = {0.0, 0.0};
            for (int q = 0; q <= a; ++q) s = s + cmulc(Ts[q][a], Xs[q][c]);
            Ms[a][c] = s;
        }
        __syncthreads();
        // 8. W = Y - 1/2 V M
        for (int i = tid; i < L; i += S1T) {
            c128 (*M_)[B + 1] = lds_opaque(Ms);
            c128 v[B];
#pragma unroll
            for (int a = 0; a < B; ++a) v[a] = V[a * L + i];
#pragma unroll
            for (int c = 0; c < B; ++c) {
                c128 acc = {0.0, 0.0};
#pragma unroll
                for (int a = 0; a < B; ++a) acc = acc + v[a] * M_[a][c];
                W[c * L + i] = W[c * L + i] - 0.5 * acc;
            }
        }
        __syncthreads();
        // 9. A22 <- A22 - V W^H - W V^H (all L x L entries, 4x4 per thread)
        const int nbl = (L + 3) / 4;
        for (int blk = tid; blk < nbl * nbl; blk += S1T) {
            const int i0 = 4 * (blk % nbl), k0 = 4 * (blk / nbl);
            c128 acc[4][4];
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const int i = i0 + a, k = k0 + c;
                    acc[a][c] = (i < L && k < L) ? A[(r0 + i) + (int64_t)m * (r0 + k)]
                                                 : c128{0.0, 0.0};
                }
            for (int j = 0; j < B; ++j) {
                c128 vi[4], wi[4], vk[4], wk[4];
#pragma unroll
                for (int a = 0; a < 4; ++a) {
                    const int i = min(i0 + a, L - 1), k = min(k0 + a, L - 1);
                    vi[a] = V[j * L + i];
                    wi[a] = W[j * L + i];
                    vk[a] = V[j * L + k];
                    wk[a] = W[j * L + k];
                }
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int c = 0; c < 4; ++c)
                        acc[a][c] = acc[a][c] - vi[a] * conj(wk[c]) - wi[a] * conj(vk[c]);
            }
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const int i = i0 + a, k = k0 + c;
                    if (i < L && k < L) A[(r0 + i) + (int64_t)m * (r0 + k)] = acc[a][c];
                }
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------ stage 2
// Band element (i, j), i >= j, i - j < 2B, stored at Bd[(i - j) * m + j].
__device__ __forceinline__ c128 &bref(c128 *Bd, int m, int i, int j) { return Bd[(i - j) * m + j]; }
// Element (q, r) of the Hermitian block starting at (st, st).
__device__ __forceinline__ c128 bherm(c128 *Bd, int m, int st, int q, int r) {
    return q >= r ? bref(Bd, m, st + q, st + r) : conj(bref(Bd, m, st + r, st + q));
}

// One task of sweep s at step k on the wave's lanes.  H_k (reflector on rows
// st .. st+len-1, st = s + 1 + k B) is in sv / tau (k > 0) or is made here from
// column s (k = 0).  Then: two-sided H_k on the diagonal block, right H_k on the
// block below (E, rows st+B ..), a new reflector H_{k+1} from E's first column,
// and its left application to the rest of E.
__device__ void sb2st_task(c128 *Bd, int m, int s, int k, c128 *sv, c128 *sw, c128 &tau,
                           c128 *V2, c128 *tau2, int ks) {
    const int lane = threadIdx.x & 63;
    const int st = s + 1 + k * B;
    const int len = min(B, m - st);
    if (k == 0) {
        const c128 x = lane < len ? bref(Bd, m, st + lane, s) : c128{0.0, 0.0};
        const double xn = wsum_d(lane >= 1 && lane < len ? abs2(x) : 0.0);
        const c128 alpha = shfl_c(x, 0);
        const Hh h = householder(alpha, xn);
        tau = h.tau;
        const bool refl = nonzero(h.tau);
        const c128 v = lane == 0 ? c128{1.0, 0.0} : (refl ? x * h.scal : c128{0.0, 0.0});
        wave_lds_sync();
        if (lane < B) sv[lane] = lane < len ? v : c128{0.0, 0.0};
        if (lane == 0) bref(Bd, m, st, s) = c128{h.beta, 0.0};
        else if (lane < len && refl) bref(Bd, m, st + lane, s) = c128{0.0, 0.0};
        if (lane < B) V2[((int64_t)s * ks) * B + lane] = lane < len ? v : c128{0.0, 0.0};
        if (lane == 0) tau2[(int64_t)s * ks] = h.tau;
        wave_lds_sync();
    }
    const int st2 = st + B;
    const int len2 = min(B, m - st2);
    if (nonzero(tau)) {
        // (a) D <- H^H D H on the diagonal block: p = tau D v, w = p - 1/2 tau (p^H v) v
        c128 pq = {0.0, 0.0};
        if (lane < len)
            for (int r = 0; r < len; ++r) pq = pq + bherm(Bd, m, st, lane, r) * sv[r];
        pq = tau * pq;
        const c128 vq = lane < len ? sv[lane] : c128{0.0, 0.0};
        const c128 dot = wsum_c(cmulc(pq, vq));
        const c128 alpha2 = (-0.5) * (tau * dot);
        const c128 wq = pq + alpha2 * vq;
        if (lane < B) sw[lane] = wq;
        wave_lds_sync();
        for (int idx = lane; idx < B * B; idx += 64) {
            const int q = idx / B, r = idx % B;
            if (q < len && r <= q) {
                c128 &d = bref(Bd, m, st + q, st + r);
                d = d - sv[q] * conj(sw[r]) - sw[q] * conj(sv[r]);
            }
        }
        wave_lds_sync();
        // (b) E <- E H = E - tau (E v) v^H, E = A[st2 .., st ..]
        if (len2 > 0) {
            c128 qi = {0.0, 0.0};
            if (lane < len2)
                for (int r = 0; r < len; ++r) qi = qi + bref(Bd, m, st2 + lane, st + r) * sv[r];
            qi = tau * qi;
            if (lane < B) sw[lane] = qi;
            wave_lds_sync();
            for (int idx = lane; idx < B * B; idx += 64) {
                const int i = idx / B, r = idx % B;
                if (i < len2 && r < len) {
                    c128 &e = bref(Bd, m, st2 + i, st + r);
                    e = e - sw[i] * conj(sv[r]);
                }
            }
            wave_lds_sync();
        }
    }
    if (len2 <= 0) return;
    // (c) H_{k+1} from E's first column
    const c128 x = lane < len2 ? bref(Bd, m, st2 + lane, st) : c128{0.0, 0.0};
    const double xn = wsum_d(lane >= 1 && lane < len2 ? abs2(x) : 0.0);
    const c128 alpha = shfl_c(x, 0);
    const Hh h = householder(alpha, xn);
    const bool refl = nonzero(h.tau);
    const c128 v = lane == 0 ? c128{1.0, 0.0} : (refl ? x * h.scal : c128{0.0, 0.0});
    wave_lds_sync();
    if (lane < B) sv[lane] = lane < len2 ? v : c128{0.0, 0.0};
    if (lane == 0) bref(Bd, m, st2, st) = c128{h.beta, 0.0};
    else if (lane < len2 && refl) bref(Bd, m, st2 + lane, st) = c128{0.0, 0.0};
    const int64_t slot = (int64_t)s * ks + k + 1;
    if (lane < B) V2[slot * B + lane] = lane < len2 ? v : c128{0.0, 0.0};
    if (lane == 0) tau2[slot] = h.tau;
    tau = h.tau;
    wave_lds_sync();
    // (d) E[:, 1..] <- H_{k+1}^H E[:, 1..] = E - conj(tau') v (v^H E)
    if (refl) {
        c128 sr = {0.0, 0.0};
        if (lane >= 1 && lane < len)
            for (int i = 0; i < len2; ++i) sr = sr + cmulc(sv[i], bref(Bd, m, st2 + i, st + lane));
        sr = conj(h.tau) * sr;
        if (lane < B) sw[lane] = sr;
        wave_lds_sync();
        for (int idx = lane; idx < B * B; idx += 64) {
            const int i = idx / B, r = idx % B;
            if (i < len2 && r >= 1 && r < len) {
                c128 &e = bref(Bd, m, st2 + i, st + r);
                e = e - sv[i] * sw[r];
            }
        }
        wave_lds_sync();
    }
}

__global__ __launch_bounds__(S2T) void k_sb2st(const c128 *__restrict__ Gall, int m, int ks,
                                               double *__restrict__ dall, double *__restrict__ eall,
                                               c128 *__restrict__ V2all,
                                               c128 *__restrict__ tau2all) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    c128 *Bd = reinterpret_cast<c128 *>(smem_raw);  // 2B x m
    c128 *scr = Bd + (size_t)2 * B * m;              // per wave: sv[B], sw[B]
    const int ip = blockIdx.x;
    const int tid = threadIdx.x, wave = tid >> 6;
    const c128 *A = Gall + (int64_t)ip * m * m;
    for (int idx = tid; idx < 2 * B * m; idx += S2T) {
        const int dg = idx / m, j = idx % m;
        Bd[idx] = (dg <= B && j + dg < m) ? A[(j + dg) + (int64_t)m * j] : c128{0.0, 0.0};
    }
    __syncthreads();
    c128 *sv = scr + wave * 2 * B;
    c128 *sw = sv + B;
    c128 *V2 = V2all + (int64_t)ip * m * ks * B;
    c128 *tau2 = tau2all + (int64_t)ip * m * ks;
    c128 tau = {0.0, 0.0};
    int s = wave;  // this wave's current sweep: s, s + S2W, ...
    const int nsweep = m - 1;
    const int tend = STAGGER * (nsweep - 1) + 1;
    for (int t = 0; t < tend; ++t) {
        if (s < nsweep) {
            const int k = t - STAGGER * s;
            const int nst = (m - s - 1 + B - 1) / B;
            if (k >= 0 && k < nst) {
                sb2st_task(Bd, m, s, k, sv, sw, tau, V2, tau2, ks);
                if (k == nst - 1) s += S2W;
            }
        }
        __syncthreads();
    }
    double *d = dall + (int64_t)ip * m;
    double *e = eall + (int64_t)ip * m;
    for (int i = tid; i < m; i += S2T) {
        d[i] = Bd[i].re;
        e[i] = i + 1 < m ? Bd[m + i].re : 0.0;
    }
}

// ------------------------------------------------------------------ back-transformation
// grid (matrices, cdiv(m, BTC)); BTC eigenvectors per workgroup in LDS (column
// c at Uc[c*m ..]).  Q2: B lanes per eigenvector, one reflector at a time in
// reverse order (no workgroup barrier).  Q1: per panel in reverse order,
// U <- U - V (T (V^H U)) with the panel's V staged in LDS.
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
    const int ip = blockIdx.x;
    const int r = rall[ip];
    const int c0 = blockIdx.y * BTC;
    if (c0 >= r) return;
    const int ncol = min(BTC, r - c0);
    const int tid = threadIdx.x;
    c128 *Uc = reinterpret_cast<c128 *>(smem_raw);  // BTC x m
    c128 *Vs = Uc + (size_t)BTC * m;                 // (m - B) x B
    const c128 *A = Gall + (int64_t)ip * m * m;
    const double *Z = Zall + (int64_t)ip * m * m;
    for (int idx = tid; idx < BTC * m; idx += BTT) {
        const int c = idx / m, i = idx % m;
        Uc[idx] = c < ncol ? c128{Z[(int64_t)(c0 + c) * m + i], 0.0} : c128{0.0, 0.0};
    }
    __syncthreads();
    // Q2 (sweep-major product) applied in reverse: z <- z - tau v (v^H z)
    {
        const int q = tid % B, c = tid / B;
        c128 *u = Uc + c * m;
        const c128 *V2 = V2all + (int64_t)ip * m * ks * B;
        const c128 *tau2 = tau2all + (int64_t)ip * m * ks;
        for (int s = m - 2; s >= 0; --s) {
            const int nst = (m - s - 1 + B - 1) / B;
            for (int k = nst - 1; k >= 0; --k) {
                const int64_t slot = (int64_t)s * ks + k;
                const c128 tau = tau2[slot];
                if (!nonzero(tau)) continue;
                const int st = s + 1 + k * B;
                const int len = min(B, m - st);
                const c128 v = q < len ? V2[slot * B + q] : c128{0.0, 0.0};
                const c128 x = q < len ? u[st + q] : c128{0.0, 0.0};
                c128 dot = cmulc(v, x);
#pragma unroll
                for (int o = B / 2; o >= 1; o >>= 1) {
                    dot.re += __shfl_xor(dot.re, o, 64);
                    dot.im += __shfl_xor(dot.im, o, 64);
                }
                if (q < len) u[st + q] = x - (tau * dot) * v;
                wave_lds_sync();
            }
        }
    }
    __syncthreads();
    // Q1: panels in reverse order
    for (int p = npanel - 1; p >= 0; --p) {
        const int pc0 = p * B, r0 = pc0 + B, L = m - r0;
        const int kq = min(B, L - 1);
        for (int idx = tid; idx < L * B; idx += BTT) {
            const int i = idx % L, j = idx / L;
            c128 v = {0.0, 0.0};
            if (j < kq) {
                if (i == j) v = c128{1.0, 0.0};
                else if (i > j) v = A[(r0 + i) + (int64_t)m * (pc0 + j)];
            }
            Vs[j * L + i] = v;
        }
        for (int pr = tid; pr < B * B; pr += BTT)
            Ts[pr % B][pr / B] = Tall[((int64_t)ip * npanel + p) * B * B + pr];
        __syncthreads();
        {   // X = V^H U (B x BTC), one entry per thread
            const int a = tid % B, c = tid / B;
            c128 s = {0.0, 0.0};
            const c128 *uc = Uc + c * m + r0;
            for (int i = 0; i < L; ++i) s = s + cmulc(Vs[a * L + i], uc[i]);
            XY[a][c] = s;
        }
        __syncthreads();
        {   // Y = T X (T upper)
            const int a = tid % B, c = tid / B;
            c128 s = {0.0, 0.0};
            for (int q = a; q < B; ++q) s = s + Ts[a][q] * XY[q][c];
            __syncthreads();
            XY[a][c] = s;
        }
        __syncthreads();
        for (int idx = tid; idx < L * BTC; idx += BTT) {
            const int i = idx % L, c = idx / L;
            c128 s = {0.0, 0.0};
#pragma unroll
            for (int a = 0; a < B; ++a) s = s + Vs[a * L + i] * XY[a][c];
            Uc[c * m + r0 + i] = Uc[c * m + r0 + i] - s;
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
    const int npanel = sbr_panels(m), ks = sbr_steps(m);
    {
        const size_t lds = (size_t)2 * (m - B) * B * sizeof(c128);
        raise_lds((const void *)k_sy2sb, lds);
        MSLR_PROF(ctx, "sy2sb");
        k_sy2sb<<<(unsigned)nb, S1T, lds, ctx->stream>>>(G, m, npanel, w.T1);
        HIP_LAUNCH_CHECK();
    }
    {
        const size_t lds = ((size_t)2 * B * m + (size_t)S2W * 2 * B) * sizeof(c128);
        raise_lds((const void *)k_sb2st, lds);
        MSLR_PROF(ctx, "sb2st");
        k_sb2st<<<(unsigned)nb, S2T, lds, ctx->stream>>>(G, m, ks, d, e, w.V2, w.tau2);
        HIP_LAUNCH_CHECK();
    }
}

void sbr_backtransform(mslr_ctx_s *ctx, const c128 *G, int m, int64_t nb, const SbrWork &w,
                       const double *Z, const int *r, c64 *U) {
    const int npanel = sbr_panels(m), ks = sbr_steps(m);
    const size_t lds = ((size_t)BTC * m + (size_t)(m - B) * B) * sizeof(c128);
    raise_lds((const void *)k_sbr_bt, lds);
    dim3 g((unsigned)nb, (unsigned)cdiv(m, BTC));
    MSLR_PROF(ctx, "backtransform");
    k_sbr_bt<<<g, BTT, lds, ctx->stream>>>(G, m, npanel, ks, w.T1, w.V2, w.tau2, Z, r, U);
    HIP_LAUNCH_CHECK();
}

}  // namespace mslr
