// Batched SVST (src/recon.jl:183-211, 217-241; src/sense_gpu.jl:122-128).
//
// For a patch matrix X (p x Nt) with m = min(p, Nt):
//   1. Gram G = X^H X (p >= Nt) or X X^H (p < Nt): FP64 products of the FP32
//      data on 64x64 tiles, split over the contraction length, partials summed
//      in FP64.
//   2. FP64 Hermitian eigensolve of G, restricted to eigenvalues above
//      tau = beta^2: Householder reduction to a real tridiagonal T (one
//      workgroup per patch), Sturm-count bisection for the eigenvalues above
//      tau (the count is exact, so no singular value above beta is missed),
//      inverse iteration with in-cluster re-orthogonalisation for their
//      eigenvectors, and back-transformation through the reflectors with one
//      wavefront per eigenvector.
//   3. Y = X V_r diag(f) V_r^H (or V_r diag(f) V_r^H X), f = max(s - beta, 0)/s,
//      fused per block of rows (columns) so X is read once and written in place.
// SVST(X) = U max(S - beta, 0) V^H = X V diag(max(s-beta,0)/s) V^H, so the
// output equals the reference's U*Diagonal(s_thresh)*Vt with only the
// right (left) singular vectors formed.  A zero patch has no eigenvalue
// above tau > 0 and maps to zero (src/recon.jl:185).
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <type_traits>

#include "ozaki.hpp"
#include "reduce.hpp"
#include "sbr.hpp"
#include "svst.hpp"

namespace mslr {

constexpr int GT = 64;       // Gram tile edge
constexpr int GK = 16;       // Gram k-step
constexpr int KCH = 4096;    // contraction chunk per workgroup
constexpr int MAXM = 1024;   // largest Gram dimension
constexpr int NSLOT = 64;    // concurrent inverse-iteration solves per patch
constexpr int MAXQ = MAXM / 64;
constexpr int WS_PER = 256;  // doubles per Gram row of eigen workspace per patch (>= 4*NSLOT and >= threads)

// ------------------------------------------------------------------ Gram


// C[a,b] += sum_k conj(A(k,a)) A(k,b), A(k,a) = base[k*ks + a*cs], for the
// upper tile pair (ti <= tj) of patch ip and contraction chunk kc.
// AT is the product / accumulation type: float for the SVST path, double for
// the nuclear norm (whose small singular values need the extra precision).
template <class AT>
__global__ __launch_bounds__(256) void k_gram_tiles(const c64 *__restrict__ P, int64_t pstride,
                                                    int64_t ks, int64_t cs, int m, int64_t Kc,
                                                    int nkc, int kch, int ntile,
                                                    C2<AT> *__restrict__ part) {
    using CA = C2<AT>;
    __shared__ CA As[GK][GT + 1];
    __shared__ CA Bs[GK][GT + 1];
    const int ip = blockIdx.z;
    const int kc = blockIdx.y;
    // linear upper-triangular tile index -> (ti, tj), ti <= tj
    int rem = blockIdx.x, ti = 0;
    while (rem >= ntile - ti) {
        rem -= ntile - ti;
        ++ti;
    }
    const int tj = ti + rem;
    const int a0 = ti * GT, b0 = tj * GT;
    const c64 *A = P + ip * pstride;
    const int tid = threadIdx.x;
    const int tx = tid & 15, ty = tid >> 4;
    CA acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = CA{AT(0), AT(0)};
    const int64_t kbeg = (int64_t)kc * kch;
    const int64_t kend = std::min<int64_t>(Kc, kbeg + kch);
    const bool kfast = ks == 1;
    for (int64_t k0 = kbeg; k0 < kend; k0 += GK) {
#pragma unroll
        for (int r = 0; r < (GK * GT) / 256; ++r) {
            const int idx = tid + 256 * r;
            int kk, aa;
            if (kfast) {
                kk = idx % GK;
                aa = idx / GK;
            } else {
                aa = idx % GT;
                kk = idx / GT;
            }
            const int64_t k = k0 + kk;
            c64 va = {0.f, 0.f}, vb = {0.f, 0.f};
            if (k < kend) {
                if (a0 + aa < m) va = A[k * ks + (int64_t)(a0 + aa) * cs];
                if (b0 + aa < m) vb = A[k * ks + (int64_t)(b0 + aa) * cs];
            }
            As[kk][aa] = CA{(AT)va.re, (AT)va.im};
            Bs[kk][aa] = CA{(AT)vb.re, (AT)vb.im};
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < GK; ++kk) {
            CA av[4], bv[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) av[i] = As[kk][tx + 16 * i];
#pragma unroll
            for (int j = 0; j < 4; ++j) bv[j] = Bs[kk][ty + 16 * j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    // conj(a) * b
                    acc[i][j].re += av[i].re * bv[j].re + av[i].im * bv[j].im;
                    acc[i][j].im += av[i].re * bv[j].im - av[i].im * bv[j].re;
                }
        }
        __syncthreads();
    }
    CA *out = part + ((int64_t)ip * nkc + kc) * (int64_t)m * m;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int a = a0 + tx + 16 * i, b = b0 + ty + 16 * j;
            if (a < m && b < m) out[a + (int64_t)m * b] = acc[i][j];
        }
}

// Epilogue of the Gram tile kernels.  With one contraction chunk the sums are
// final: entries a <= b go straight into the Hermitian matrix with their mirror
// (as k_gram_reduce writes them: real diagonal, conjugate for conj_out, a
// non-finite sum flagged and zeroed); otherwise into the partial tiles.
__device__ __forceinline__ void gram_store(c128 *__restrict__ part, c128 *__restrict__ G, int ip,
                                           int nkc, int kc, int m, int a, int b, double sr,
                                           double si, int conj_out, int *err) {
    if (a >= m || b >= m) return;
    if (G == nullptr) {
        part[((int64_t)ip * nkc + kc) * (int64_t)m * m + a + (int64_t)m * b] = c128{sr, si};
        return;
    }
    if (a > b) return;
    if (conj_out) si = -si;
    if (!isfinite(sr) || !isfinite(si)) {
        if (err) *err = 1;
        sr = 0.0;
        si = 0.0;
    }
    c128 *g = G + (int64_t)ip * m * m;
    if (a == b) {
        g[a + (int64_t)m * b] = c128{sr, 0.0};
    } else {
        g[a + (int64_t)m * b] = c128{sr, si};
        g[b + (int64_t)m * a] = c128{sr, -si};
    }
}

// FP64 Gram tiles on the matrix cores (v_mfma_f64_16x16x4f64): same tiling,
// staging and output as k_gram_tiles<double>.  Per wave a 32x32 sub-tile as
// 2x2 blocks of 16x16; each complex block takes four real MFMAs per k-step of
// 4.  Operands are fed as (X(k, b), conj X(k, a)) so that the accumulator
// lanes run down the columns of the output.  Lane layout (D = P Q, P 16x4,
// Q 4x16): lane l holds P[l%16][l/16], Q[l/16][l%16] and D[l/16 + 4v][l%16].
typedef double md4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_gram_mfma(const c64 *__restrict__ P, int64_t pstride,
                                                   int64_t ks, int64_t cs, int m, int64_t Kc,
                                                   int nkc, int kch, int ntile,
                                                   c128 *__restrict__ part,
                                                   c128 *__restrict__ G, int conj_out,
                                                   int *__restrict__ err) {
    __shared__ double Ar[GK][GT + 2], Ai[GK][GT + 2];
    __shared__ double Br[GK][GT + 2], Bi[GK][GT + 2];
    // the tile pairs of one (patch, chunk) share column tiles: keep them on one XCD
    const int64_t nwg = (int64_t)gridDim.x * gridDim.y * gridDim.z;
    const int64_t wl = xcd_order(blockIdx.x + (int64_t)gridDim.x * (blockIdx.y + (int64_t)gridDim.y * blockIdx.z), nwg);
    const int ip = (int)(wl / ((int64_t)gridDim.x * gridDim.y));
    const int kc = (int)((wl / gridDim.x) % gridDim.y);
    int rem = (int)(wl % gridDim.x), ti = 0;
    while (rem >= ntile - ti) {
        rem -= ntile - ti;
        ++ti;
    }
    const int tj = ti + rem;
    const int a0 = ti * GT, b0 = tj * GT;
    const c64 *X = P + ip * pstride;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wa = (wave & 1) * 32, wb = (wave >> 1) * 32;
    const int r16 = lane & 15, k4 = lane >> 4;
    md4 accr[2][2], acci[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            accr[i][j] = md4{0.0, 0.0, 0.0, 0.0};
            acci[i][j] = md4{0.0, 0.0, 0.0, 0.0};
        }
    const int64_t kbeg = (int64_t)kc * kch;
    const int64_t kend = std::min<int64_t>(Kc, kbeg + kch);
    const bool kfast = ks == 1;
    for (int64_t k0 = kbeg; k0 < kend; k0 += GK) {
#pragma unroll
        for (int r = 0; r < (GK * GT) / 256; ++r) {
            const int idx = tid + 256 * r;
            const int kk = kfast ? idx % GK : idx / GT;
            const int aa = kfast ? idx / GK : idx % GT;
            const int64_t k = k0 + kk;
            float var = 0.f, vai = 0.f, vbr = 0.f, vbi = 0.f;
            if (k < kend) {
                if (a0 + aa < m) {
                    const c64 v = X[k * ks + (int64_t)(a0 + aa) * cs];
                    var = v.re;
                    vai = v.im;
                }
                if (b0 + aa < m) {
                    const c64 v = X[k * ks + (int64_t)(b0 + aa) * cs];
                    vbr = v.re;
                    vbi = v.im;
                }
            }
            Ar[kk][aa] = var;
            Ai[kk][aa] = vai;
            Br[kk][aa] = vbr;
            Bi[kk][aa] = vbi;
        }
        __syncthreads();
#pragma unroll
        for (int kq = 0; kq < GK; kq += 4) {
            const int kr = kq + k4;
            double pr[2], pi[2], pn[2], qr[2], qi[2];
#pragma unroll
            for (int j = 0; j < 2; ++j) {  // P = X(k, b) block rows b
                pr[j] = Br[kr][wb + 16 * j + r16];
                pi[j] = Bi[kr][wb + 16 * j + r16];
                pn[j] = -pr[j];
            }
#pragma unroll
            for (int i = 0; i < 2; ++i) {  // Q = conj X(k, a) block columns a
                qr[i] = Ar[kr][wa + 16 * i + r16];
                qi[i] = Ai[kr][wa + 16 * i + r16];
            }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    // x_b conj(x_a): re = br ar + bi ai, im = bi ar - br ai
                    md4 cr = accr[i][j], ci = acci[i][j];
                    cr = __builtin_amdgcn_mfma_f64_16x16x4f64(pr[j], qr[i], cr, 0, 0, 0);
                    cr = __builtin_amdgcn_mfma_f64_16x16x4f64(pi[j], qi[i], cr, 0, 0, 0);
                    ci = __builtin_amdgcn_mfma_f64_16x16x4f64(pi[j], qr[i], ci, 0, 0, 0);
                    ci = __builtin_amdgcn_mfma_f64_16x16x4f64(pn[j], qi[i], ci, 0, 0, 0);
                    accr[i][j] = cr;
                    acci[i][j] = ci;
                }
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int a = a0 + wa + 16 * i + r16;
                const int b = b0 + wb + 16 * j + k4 + 4 * v;
                gram_store(part, G, ip, nkc, kc, m, a, b, accr[i][j][v], acci[i][j][v], conj_out,
                           err);
            }
}

// G[a,b] = sum_kc part (FP64), Hermitian fill; conj_out flips to conj(C).
template <class AT>
__global__ void k_gram_reduce(const C2<AT> *__restrict__ part, int m, int nkc, int np, int conj_out,
                              c128 *__restrict__ G, int *err) {
    const int64_t mm = (int64_t)m * m;
    const int64_t n = mm * np;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t ip = e / mm;
        const int64_t r = e - ip * mm;
        const int a = (int)(r % m), b = (int)(r / m);
        if (a > b) continue;
        double sr = 0.0, si = 0.0;
        const C2<AT> *pp = part + ip * nkc * mm + r;
        for (int kc = 0; kc < nkc; ++kc) {
            const C2<AT> v = pp[kc * mm];
            sr += (double)v.re;
            si += (double)v.im;
        }
        if (conj_out) si = -si;
        if (!isfinite(sr) || !isfinite(si)) {
            if (err) *err = 1;
            sr = 0.0;
            si = 0.0;
        }
        c128 *g = G + ip * mm;
        if (a == b) {
            g[r] = c128{sr, 0.0};
        } else {
            g[r] = c128{sr, si};
            g[b + (int64_t)m * a] = c128{sr, -si};
        }
    }
}

// ------------------------------------------------------------------ tridiagonalisation

__device__ __forceinline__ c128 cdiv(c128 a, c128 b) {
    const double den = b.re * b.re + b.im * b.im;
    return c128{(a.re * b.re + a.im * b.im) / den, (a.im * b.re - a.re * b.im) / den};
}

// Lower Householder reduction to a real tridiagonal (the zhetrd sequence):
// Q^H G Q = T with Q = H(0) H(1) ... H(m-2), H(j) = I - tau_j v_j v_j^H,
// v_j(j+1) = 1, v_j(i) stored in G[i, j] for i > j+1; d, e real.
// Blocked (the latrd scheme): within a panel of TNB
// columns the trailing matrix is not updated; each column is first brought up
// to date with the panel's earlier reflectors, and its symmetric product uses
// the panel-start matrix corrected by the panel's V and W.  After the panel one
// rank-2*TNB update writes the trailing matrix once.  Per column the trailing
// matrix is read about once instead of read twice and written once.
constexpr int TNB = 16;
constexpr int TNT = 512;

__device__ __forceinline__ c128 wave_sum_c(c128 v) {
    for (int o = 32; o > 0; o >>= 1) {
        v.re += __shfl_xor(v.re, o, 64);
        v.im += __shfl_xor(v.im, o, 64);
    }
    return v;
}

__global__ __launch_bounds__(TNT) void k_tridiag_blk(c128 *__restrict__ Gall, int m,
                                                     double *__restrict__ dall,
                                                     double *__restrict__ eall,
                                                     c128 *__restrict__ tall,
                                                     c128 *__restrict__ VWall) {
    __shared__ c128 v[MAXM];
    __shared__ c128 w[MAXM];
    __shared__ c128 cv[TNB], cw[TNB];
    __shared__ double red[16];
    __shared__ c128 s_tau, s_scal, s_dot;
    __shared__ double s_beta;
    const int ip = blockIdx.x;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    constexpr int NW = TNT / 64;
    c128 *A = Gall + (int64_t)ip * m * m;
    c128 *V = VWall + (int64_t)ip * 2 * m * TNB;  // V[i + m*q]
    c128 *W = V + (int64_t)m * TNB;
    double *d = dall + (int64_t)ip * m;
    double *e = eall + (int64_t)ip * m;
    c128 *tv = tall + (int64_t)ip * m;
    for (int j0 = 0; j0 + 1 < m; j0 += TNB) {
        const int j1 = min(j0 + TNB, m - 1);  // panel columns j0 .. j1-1
        for (int j = j0; j < j1; ++j) {
            const int jj = j - j0;
            const int n = m - j - 1;
            c128 *col = A + (int64_t)m * j;
            // 1. bring column j (rows >= j) up to date with the panel's earlier reflectors
            if (jj > 0) {
                for (int i = j + tid; i < m; i += TNT) {
                    c128 a = col[i];
                    for (int q = 0; q < jj; ++q) {
                        const c128 vi = V[i + (int64_t)m * q], wi = W[i + (int64_t)m * q];
                        const c128 vj = V[j + (int64_t)m * q], wj = W[j + (int64_t)m * q];
                        a = a - vi * conj(wj) - wi * conj(vj);
                    }
                    col[i] = a;
                }
                __syncthreads();
            }
            // 2. Householder vector of column j below the diagonal
            double xn = 0.0;
            for (int i = j + 2 + tid; i < m; i += TNT) xn += abs2(col[i]);
            xn = block_sum(xn, red);
            if (tid == 0) {
                const c128 alpha = col[j + 1];
                if (xn == 0.0 && alpha.im == 0.0) {
                    s_tau = c128{0.0, 0.0};
                    s_beta = alpha.re;
                } else {
                    const double nrm = sqrt(alpha.re * alpha.re + alpha.im * alpha.im + xn);
                    const double beta = alpha.re >= 0.0 ? -nrm : nrm;
                    s_tau = c128{(beta - alpha.re) / beta, -alpha.im / beta};
                    s_scal = cdiv(c128{1.0, 0.0}, c128{alpha.re - beta, alpha.im});
                    s_beta = beta;
                }
                d[j] = col[j].re;
                e[j] = s_beta;
                tv[j] = s_tau;
            }
            __syncthreads();
            const c128 tau = s_tau;
            const bool reflect = !(tau.re == 0.0 && tau.im == 0.0);
            const c128 scal = s_scal;
            // v (rows j+1..m-1) -> LDS, column storage and panel V; zeros above
            for (int i = tid; i < m; i += TNT) {
                c128 vi = {0.0, 0.0};
                if (reflect && i >= j + 1) {
                    if (i == j + 1) {
                        vi = c128{1.0, 0.0};
                    } else {
                        vi = col[i] * scal;
                        col[i] = vi;
                    }
                }
                if (i >= j + 1) v[i - j - 1] = vi;
                V[i + (int64_t)m * jj] = vi;
            }
            __syncthreads();
            if (!reflect) {
                for (int i = tid; i < m; i += TNT) W[i + (int64_t)m * jj] = c128{0.0, 0.0};
                __syncthreads();
                continue;
            }
            // 3. w = tau (A22 v - V' (W'^H v) - W' (V'^H v)), A22 = panel-start trailing block
            for (int q = wave; q < 2 * jj; q += NW) {
                const c128 *P = (q < jj) ? W + (int64_t)m * q : V + (int64_t)m * (q - jj);
                c128 sacc = {0.0, 0.0};
                for (int i = lane; i < n; i += 64) sacc = sacc + cmulc(P[j + 1 + i], v[i]);
                sacc = wave_sum_c(sacc);
                if (lane == 0) {
                    if (q < jj) cv[q] = sacc;
                    else cw[q - jj] = sacc;
                }
            }
            const c128 *A22 = A + (j + 1) + (int64_t)m * (j + 1);
            for (int i = tid; i < n; i += TNT) {
                // 8 independent loads in flight per thread before their FMAs
                double ar = 0.0, ai = 0.0;
                const c128 *row = A22 + i;
                int k = 0;
                for (; k + 8 <= n; k += 8) {
                    c128 a[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) a[u] = row[(int64_t)m * (k + u)];
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        const c128 b = v[k + u];
                        ar += a[u].re * b.re - a[u].im * b.im;
                        ai += a[u].re * b.im + a[u].im * b.re;
                    }
                }
                for (; k < n; ++k) {
                    const c128 a = row[(int64_t)m * k];
                    const c128 b = v[k];
                    ar += a.re * b.re - a.im * b.im;
                    ai += a.re * b.im + a.im * b.re;
                }
                w[i] = c128{ar, ai};
            }
            __syncthreads();
            for (int i = tid; i < n; i += TNT) {
                c128 pi = w[i];
                for (int q = 0; q < jj; ++q)
                    pi = pi - V[j + 1 + i + (int64_t)m * q] * cv[q] - W[j + 1 + i + (int64_t)m * q] * cw[q];
                w[i] = tau * pi;
            }
            __syncthreads();
            // 4. alpha = -1/2 tau (w^H v); w += alpha v
            double dr = 0.0, di = 0.0;
            for (int i = tid; i < n; i += TNT) {
                const c128 t = cmulc(w[i], v[i]);
                dr += t.re;
                di += t.im;
            }
            dr = block_sum(dr, red);
            di = block_sum(di, red);
            if (tid == 0) s_dot = c128{dr, di};
            __syncthreads();
            const c128 alpha2 = (-0.5) * (tau * s_dot);
            for (int i = tid; i < m; i += TNT) {
                c128 wi = {0.0, 0.0};
                if (i >= j + 1) {
                    wi = w[i - j - 1] + alpha2 * v[i - j - 1];
                }
                W[i + (int64_t)m * jj] = wi;
            }
            __syncthreads();
        }
        // trailing update of rows/columns >= j1 with the panel's rank-2*nq correction
        const int nq = j1 - j0;
        if (j1 < m) {
            const int nt2 = m - j1;
            const int64_t nn = (int64_t)nt2 * nt2;
            for (int64_t idx = tid; idx < nn; idx += TNT) {
                const int i = j1 + (int)(idx % nt2), k = j1 + (int)(idx / nt2);
                c128 a = A[i + (int64_t)m * k];
                for (int q = 0; q < nq; ++q) {
                    const c128 vi = V[i + (int64_t)m * q], wi = W[i + (int64_t)m * q];
                    const c128 vk = V[k + (int64_t)m * q], wk = W[k + (int64_t)m * q];
                    a = a - vi * conj(wk) - wi * conj(vk);
                }
                A[i + (int64_t)m * k] = a;
            }
        }
        __syncthreads();
    }
    if (tid == 0) {
        d[m - 1] = A[(m - 1) + (int64_t)m * (m - 1)].re;
        e[m - 1] = 0.0;
        tv[m - 1] = c128{0.0, 0.0};
    }
}

// ------------------------------------------------------------------ tridiagonal eigenpairs

// Number of eigenvalues of T smaller than x.
__device__ int sturm_count(const double *d, const double *e, int m, double x, double pivmin) {
    int c = 0;
    double q = d[0] - x;
    if (fabs(q) <= pivmin) q = -pivmin;
    c += q < 0.0;
    for (int i = 1; i < m; ++i) {
        q = (d[i] - x) - e[i - 1] * e[i - 1] / q;
        if (fabs(q) <= pivmin) q = -pivmin;
        c += q < 0.0;
    }
    return c;
}

// 1/q from the hardware estimate and two Newton steps (correct to rounding for
// finite nonzero q; the callers keep |q| >= pivmin or tiny, so q is never 0 or inf).
__device__ __forceinline__ double fast_rcp(double q) {
    double r = __builtin_amdgcn_rcp(q);
    double t = __builtin_fma(-q, r, 1.0);
    r = __builtin_fma(r, t, r);
    t = __builtin_fma(-q, r, 1.0);
    return __builtin_fma(r, t, r);
}

// sturm_count with the division replaced by fast_rcp (the bisection steps;
// k_eig_count keeps the IEEE division for the count that fixes r).
// e2 = e .* e.
__device__ int sturm_count_fast(const double *__restrict__ d, const double *__restrict__ e2, int m,
                                double x, double pivmin) {
    int c = 0;
    double q = d[0] - x;
    if (fabs(q) <= pivmin) q = -pivmin;
    c += q < 0.0;
#pragma unroll 4
    for (int i = 1; i < m; ++i) {
        q = (d[i] - x) - e2[i - 1] * fast_rcp(q);
        if (fabs(q) <= pivmin) q = -pivmin;
        c += q < 0.0;
    }
    return c;
}

__device__ __forceinline__ double hash_unit(uint32_t a, uint32_t b) {
    uint32_t h = a * 0x9E3779B1u ^ (b + 0x7F4A7C15u) * 0x85EBCA77u;
    h ^= h >> 16;
    h *= 0x7feb352du;
    h ^= h >> 15;
    h *= 0x846ca68bu;
    h ^= h >> 16;
    return (double)(h & 0xFFFFFF) / 8388608.0 - 1.0;  // [-1, 1)
}

// Solve (T - lam I) x = x in place by Gaussian elimination with partial
// pivoting (row interchanges between i and i+1).  u0,u1,u2 hold the upper
// factor diagonals; arrays are strided by `st` (interleaved slots).
__device__ void tri_solve(const double *d, const double *e, int m, double lam, double tiny,
                          double *x, double *u0, double *u1, double *u2, int st) {
    // forward elimination applied to x
    double a = d[0] - lam;            // current pivot row diagonal
    double b = m > 1 ? e[0] : 0.0;    // current pivot row super-diagonal
    double c = 0.0;                   // current pivot row second super-diagonal
    double xi = x[0];
    for (int i = 0; i + 1 < m; ++i) {
        const double sub = e[i];          // T[i+1, i]
        const double dn = d[i + 1] - lam; // T[i+1, i+1]
        const double en = (i + 2 < m) ? e[i + 1] : 0.0;  // T[i+1, i+2]
        const double xn = x[(i + 1) * st];
        if (fabs(a) >= fabs(sub)) {
            const double piv = a != 0.0 ? a : tiny;
            const double l = sub / piv;
            u0[i * st] = piv;
            u1[i * st] = b;
            u2[i * st] = c;
            x[i * st] = xi;
            a = dn - l * b;
            b = en - l * c;
            c = 0.0;
            xi = xn - l * xi;
        } else {
            // swap rows i and i+1
            const double l = a / sub;
            u0[i * st] = sub;
            u1[i * st] = dn;
            u2[i * st] = en;
            x[i * st] = xn;
            const double na = b - l * dn;
            const double nb = c - l * en;
            a = na;
            b = nb;
            c = 0.0;
            xi = xi - l * xn;
        }
    }
    u0[(m - 1) * st] = a != 0.0 ? a : tiny;
    x[(m - 1) * st] = xi;
    // back substitution
    for (int i = m - 1; i >= 0; --i) {
        double s = x[i * st];
        if (i + 1 < m) s -= u1[i * st] * x[(i + 1) * st];
        if (i + 2 < m) s -= u2[i * st] * x[(i + 2) * st];
        x[i * st] = s / u0[i * st];
    }
}

// Eigenvector of the symmetric tridiagonal (d, e) for the eigenvalue lam by a
// twisted factorisation: dp = pivots of T - lam I = L D L^T (top-down, kept in
// z), dm = pivots of U D U^T (bottom-up, kept in dm with stride st), twist
// index r = argmin |dp_r + dm_r - (d_r - lam)|, then z_r = 1 and
// z_i = -(e_i/dp_i) z_{i+1} above r, z_{i+1} = -(e_i/dm_{i+1}) z_i below r.
// (z and dm do not alias d, e or each other: each value the recurrences carry stays
// in a register, and the loads of the unrolled loops are issued ahead of their use;
// e2 = e .* e)
__device__ void twisted_vector(const double *__restrict__ d, const double *__restrict__ e,
                               const double *__restrict__ e2, int m, double lam, double tiny,
                               double *__restrict__ z, double *__restrict__ dm, int st) {
    double p = d[0] - lam;
    if (p == 0.0) p = tiny;
    z[0] = p;
#pragma unroll 4
    for (int i = 0; i + 1 < m; ++i) {
        p = (d[i + 1] - lam) - e2[i] * fast_rcp(p);
        if (p == 0.0) p = tiny;
        z[i + 1] = p;
    }
    double q = d[m - 1] - lam;
    if (q == 0.0) q = tiny;
    dm[(int64_t)(m - 1) * st] = q;
    int best = m - 1;
    double gbest = fabs(z[m - 1]);
#pragma unroll 4
    for (int i = m - 2; i >= 0; --i) {
        const double di = d[i] - lam;
        q = di - e2[i] * fast_rcp(q);
        if (q == 0.0) q = tiny;
        dm[(int64_t)i * st] = q;
        const double g = fabs(z[i] + q - di);
        if (g < gbest) {
            gbest = g;
            best = i;
        }
    }
    z[best] = 1.0;
    double amax = 1.0;
    double zn = 1.0;  // z[i + 1]
#pragma unroll 4
    for (int i = best - 1; i >= 0; --i) {
        const double v = -(e[i] / z[i]) * zn;
        z[i] = v;
        zn = v;
        amax = fmax(amax, fabs(v));
    }
    double zp = 1.0;  // z[i]
#pragma unroll 4
    for (int i = best; i + 1 < m; ++i) {
        const double v = -(e[i] / dm[(int64_t)(i + 1) * st]) * zp;
        z[i + 1] = v;
        zp = v;
        amax = fmax(amax, fabs(v));
    }
    double nrm = 0.0;
#pragma unroll 4
    for (int i = 0; i < m; ++i) {
        const double t = z[i] / amax;
        nrm += t * t;
    }
    const double inv = 1.0 / (amax * sqrt(nrm));
#pragma unroll 4
    for (int i = 0; i < m; ++i) z[i] *= inv;
}

__device__ __forceinline__ c128 wave_sum(c128 v) {
    for (int o = 32; o > 0; o >>= 1) {
        v.re += __shfl_xor(v.re, o, 64);
        v.im += __shfl_xor(v.im, o, 64);
    }
    return v;
}

// Eigenpairs of the tridiagonal T of each patch above tau = beta^2, in three
// launches so that large problems spread over many workgroups:
//   k_eig_count   : rank r = #eigenvalues > tau (Sturm count), Gershgorin bounds;
//   k_eig_vectors : a group of lanes per eigenvalue: multisection, then its
//                   eigenvector by a twisted factorisation;
//   k_eig_clusters: tight clusters redone by inverse iteration, filter factors.
// r = -1 marks beta == 0 (identity).
constexpr int EIG_INFO = 4;  // per patch: tnorm, pivmin, upper and lower search bounds

// all != 0: every eigenpair is wanted (r = m), for the SVST derivative.
__global__ __launch_bounds__(64) void k_eig_count(const double *__restrict__ dall,
                                                  const double *__restrict__ eall, int m, int64_t nb,
                                                  const double *__restrict__ beta,
                                                  int beta_per_patch, double *__restrict__ info,
                                                  int *__restrict__ rall, int all) {
    const int64_t ip = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= nb) return;
    const double *d = dall + ip * m;
    const double *e = eall + ip * m;
    const float b32 = (float)beta[beta_per_patch ? ip : 0];
    const double bd = (double)b32;
    if (bd == 0.0 && !all) {
        rall[ip] = -1;
        return;
    }
    const double tau = bd * bd;
    double tn = 0.0, emax = 0.0, gl = INFINITY, gu = -INFINITY;
    for (int i = 0; i < m; ++i) {
        const double el = i > 0 ? fabs(e[i - 1]) : 0.0;
        const double er = i + 1 < m ? fabs(e[i]) : 0.0;
        tn = fmax(tn, fabs(d[i]) + el + er);
        gl = fmin(gl, d[i] - el - er);
        gu = fmax(gu, d[i] + el + er);
        if (i + 1 < m) emax = fmax(emax, e[i] * e[i]);
    }
    const double pivmin = DBL_MIN * fmax(1.0, emax);
    const double pad = 2.0 * DBL_EPSILON * fmax(fabs(gl), fabs(gu)) + pivmin;
    gu = gu + pad;
    int r = 0;
    if (all) r = m;
    else if (tau < gu) r = m - sturm_count(d, e, m, tau, pivmin);
    double *inf = info + ip * EIG_INFO;
    inf[0] = tn;
    inf[1] = pivmin;
    inf[2] = gu;
    inf[3] = all ? gl - pad : tau;  // every wanted eigenvalue lies above it
    rall[ip] = r;
}

// k_eig_count for small batches: one 64-thread workgroup per patch.  d and e are
// staged in LDS by the wave, the Gershgorin bounds are wave reductions (max / min:
// the same values in any order) and lane 0 takes the Sturm count from LDS, so a
// lone patch waits for no chain of dependent global loads; bit-identical results.
// hi: raise the waves' issue priority (the single matrix of a whole-image patch
// runs beside the batched kernels of the other scales, and its latency is the cost).
__device__ __forceinline__ double wave_max64(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}
__global__ __launch_bounds__(64) void k_eig_count_wave(const double *__restrict__ dall,
                                                       const double *__restrict__ eall, int m,
                                                       const double *__restrict__ beta,
                                                       int beta_per_patch,
                                                       double *__restrict__ info,
                                                       int *__restrict__ rall, int all, int hi) {
    if (hi) __builtin_amdgcn_s_setprio(3);
    extern __shared__ __align__(16) unsigned char eig_smem[];
    double *d = reinterpret_cast<double *>(eig_smem);
    double *e = d + m;
    const int64_t ip = blockIdx.x;
    const int lane = threadIdx.x;
    const float b32 = (float)beta[beta_per_patch ? ip : 0];
    const double bd = (double)b32;
    if (bd == 0.0 && !all) {
        if (lane == 0) rall[ip] = -1;
        return;
    }
    for (int i = lane; i < m; i += 64) {
        d[i] = dall[ip * m + i];
        e[i] = eall[ip * m + i];
    }
    __syncthreads();
    double tn = 0.0, emax = 0.0, gl = INFINITY, gu = -INFINITY;
    for (int i = lane; i < m; i += 64) {
        const double el = i > 0 ? fabs(e[i - 1]) : 0.0;
        const double er = i + 1 < m ? fabs(e[i]) : 0.0;
        tn = fmax(tn, fabs(d[i]) + el + er);
        gl = fmin(gl, d[i] - el - er);
        gu = fmax(gu, d[i] + el + er);
        if (i + 1 < m) emax = fmax(emax, e[i] * e[i]);
    }
    tn = wave_max64(tn);
    emax = wave_max64(emax);
    gl = -wave_max64(-gl);
    gu = wave_max64(gu);
    if (lane != 0) return;
    const double tau = bd * bd;
    const double pivmin = DBL_MIN * fmax(1.0, emax);
    const double pad = 2.0 * DBL_EPSILON * fmax(fabs(gl), fabs(gu)) + pivmin;
    gu = gu + pad;
    int r = 0;
    if (all) r = m;
    else if (tau < gu) r = m - sturm_count(d, e, m, tau, pivmin);
    double *inf = info + ip * EIG_INFO;
    inf[0] = tn;
    inf[1] = pivmin;
    inf[2] = gu;
    inf[3] = all ? gl - pad : tau;
    rall[ip] = r;
}

// grid (patches, cdiv(m * G, 64)); a group of G lanes per eigenvalue k < r.
// lam[k] is the k-th largest eigenvalue, found by multisection: each step the
// G lanes take Sturm counts at G interior points of [lo, hi] and the group keeps
// the sub-interval where the count passes the eigenvalue's index (log2(G+1)
// bits per step instead of one; the counts are monotone in the point).  Its
// eigenvector goes to Z column k, with the bottom-up pivots kept in U's column
// k (free until the back-transformation).
template <int G>
__global__ __launch_bounds__(64) void k_eig_vectors(const double *__restrict__ dall,
                                                    const double *__restrict__ eall, int m,
                                                    const double *__restrict__ beta,
                                                    int beta_per_patch,
                                                    const double *__restrict__ info,
                                                    const int *__restrict__ rall,
                                                    double *__restrict__ lamall,
                                                    double *__restrict__ Zall,
                                                    c64 *__restrict__ Uall, int k0, int hiprio) {
    static_assert(G >= 1 && G <= 64 && (64 % G) == 0, "group size");
    if (hiprio) __builtin_amdgcn_s_setprio(3);
    const int ip = blockIdx.x;
    const int r = rall[ip];
    if (k0 + (int)(blockIdx.y * 64) / G >= r) return;  // no wanted eigenvalue here
    // the patch's d, e and e .* e in LDS: the Sturm counts and the twisted
    // factorisation walk them sequentially, once per bisection step
    extern __shared__ __align__(16) unsigned char eig_smem[];
    double *d = reinterpret_cast<double *>(eig_smem);
    double *e = d + m;
    double *e2 = e + m;
#pragma unroll 4
    for (int i = threadIdx.x; i < m; i += 64) {  // (unrolled: the loads issue together)
        d[i] = dall[(int64_t)ip * m + i];
        const double ei = eall[(int64_t)ip * m + i];
        e[i] = ei;
        e2[i] = ei * ei;
    }
    __syncthreads();
    const int t = blockIdx.y * 64 + threadIdx.x;
    const int k = k0 + t / G, j = t % G;
    if (k >= r) return;  // whole groups leave together
    const int gbase = (threadIdx.x & 63) - j;
    const double *inf = info + (int64_t)ip * EIG_INFO;
    const double tnorm = inf[0], pivmin = inf[1], gu = inf[2];
    const int idx = m - 1 - k;  // ascending index: count(lo) <= idx < count(hi)
    double lo = inf[3], hi = gu;
    for (int it = 0; it < 200; ++it) {
        if (hi - lo <= 2.0 * DBL_EPSILON * fmax(fabs(lo), fabs(hi)) + pivmin) break;
        const double w = (hi - lo) / (double)(G + 1);
        if (!(lo + w > lo) || !(hi - w < hi)) break;
        const double xj = lo + (double)(j + 1) * w;
        const bool above = sturm_count_fast(d, e2, m, xj, pivmin) > idx;
        const unsigned long long bal = __ballot(above);
        const unsigned long long gm = (bal >> gbase) & (G == 64 ? ~0ull : ((1ull << G) - 1));
        const int f = gm ? __builtin_ctzll(gm) : G;  // first point above the index
        const double nlo = f == 0 ? lo : lo + (double)f * w;
        const double nhi = f == G ? hi : lo + (double)(f + 1) * w;
        if (!(nlo < nhi)) break;
        lo = nlo;
        hi = nhi;
    }
    if (j != 0) return;
    const double lk = 0.5 * (lo + hi);
    lamall[(int64_t)ip * m + k] = lk;
    const double tiny = DBL_EPSILON * fmax(tnorm, DBL_MIN);
    double *z = Zall + (int64_t)ip * m * m + (int64_t)k * m;
    double *dm = reinterpret_cast<double *>(Uall + (int64_t)ip * m * m + (int64_t)k * m);
    twisted_vector(d, e, e2, m, lk, tiny, z, dm, 1);
}

// One workgroup of NSLOT threads per patch.  Runs of eigenvalues closer than
// cltol (tight clusters) are redone: every member after the first by inverse
// iteration from a distinct start with re-orthogonalisation against the
// earlier members (one thread per cluster).  Then f_k = max(s_k - beta, 0)/s_k
// (src/recon.jl:195-196).
__global__ __launch_bounds__(NSLOT) void k_eig_clusters(
    const double *__restrict__ dall, const double *__restrict__ eall, int m,
    const double *__restrict__ beta, int beta_per_patch, const double *__restrict__ info,
    const int *__restrict__ rall, const double *__restrict__ lamall, double *__restrict__ Zall,
    double *__restrict__ wsall, float *__restrict__ fall) {
    const int ip = blockIdx.x;
    const int r = rall[ip];
    if (r <= 0) return;
    const int tid = threadIdx.x;
    const double *d = dall + (int64_t)ip * m;
    const double *e = eall + (int64_t)ip * m;
    const double *lam = lamall + (int64_t)ip * m;
    double *Z = Zall + (int64_t)ip * m * m;
    const double tnorm = info[(int64_t)ip * EIG_INFO];
    const double bd = (double)(float)beta[beta_per_patch ? ip : 0];
    const double cltol = 1e-12 * tnorm;
    const double pertol = 10.0 * DBL_EPSILON * fmax(tnorm, DBL_MIN);
    const double tiny = DBL_EPSILON * fmax(tnorm, DBL_MIN);
    double *ws = wsall + (int64_t)ip * WS_PER * m;
    double *xs = ws + tid;
    double *u0 = ws + (int64_t)NSLOT * m + tid;
    double *u1 = ws + (int64_t)2 * NSLOT * m + tid;
    double *u2 = ws + (int64_t)3 * NSLOT * m + tid;
    for (int k0 = tid; k0 < r; k0 += NSLOT) {
        if (k0 > 0 && lam[k0 - 1] - lam[k0] <= cltol) continue;  // not a cluster start
        int k1 = k0 + 1;
        while (k1 < r && lam[k1 - 1] - lam[k1] <= cltol) ++k1;
        double lprev = lam[k0];
        for (int k = k0 + 1; k < k1; ++k) {
            double lk = lam[k];
            if (lprev - lk < pertol) lk = lprev - pertol;
            lprev = lk;
            for (int i = 0; i < m; ++i) xs[i * NSLOT] = hash_unit((uint32_t)i, (uint32_t)k);
            for (int it = 0; it < 3; ++it) {
                tri_solve(d, e, m, lk, tiny, xs, u0, u1, u2, NSLOT);
                for (int q = k0; q < k; ++q) {
                    const double *zq = Z + (int64_t)q * m;
                    double dot = 0.0;
                    for (int i = 0; i < m; ++i) dot += zq[i] * xs[i * NSLOT];
                    for (int i = 0; i < m; ++i) xs[i * NSLOT] -= dot * zq[i];
                }
                double amax = 0.0;
                for (int i = 0; i < m; ++i) amax = fmax(amax, fabs(xs[i * NSLOT]));
                if (amax == 0.0) amax = 1.0;
                double nrm = 0.0;
                for (int i = 0; i < m; ++i) {
                    const double t = xs[i * NSLOT] / amax;
                    nrm += t * t;
                }
                nrm = amax * sqrt(nrm);
                const double inv = nrm > 0.0 ? 1.0 / nrm : 0.0;
                for (int i = 0; i < m; ++i) xs[i * NSLOT] *= inv;
            }
            double *zk = Z + (int64_t)k * m;
            for (int i = 0; i < m; ++i) zk[i] = xs[i * NSLOT];
        }
    }
    // cluster members may be rewritten by other threads: all are done before f
    __syncthreads();
    float *f = fall + (int64_t)ip * m;
    for (int k = tid; k < r; k += NSLOT) {
        const double sg = sqrt(fmax(lam[k], 0.0));
        f[k] = (float)(sg > bd ? (sg - bd) / sg : 0.0);
    }
}

// Back-transformation U = Q Z (Q = H(0) ... H(m-2)) for the r eigenvectors of
// each patch; grid (patches, cdiv(m, 256/G)), one chunk of eigenvectors per
// workgroup.  G lanes own one eigenvector (element i at lane i mod G, in
// registers); blocks of BTR reflectors are staged in LDS, and each reflector
// costs one G-lane shuffle reduction -- no workgroup barrier per reflector.
constexpr int BT_MAXE = 40;  // eigenvector elements per lane
constexpr int BTR = 8;       // reflectors per LDS block

template <int G>
__global__ __launch_bounds__(256) void k_backtransform(const c128 *__restrict__ Gall,
                                                       const c128 *__restrict__ tall,
                                                       const double *__restrict__ Zall,
                                                       const int *__restrict__ rall, int m,
                                                       c64 *__restrict__ Uall) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    c128 *vb = reinterpret_cast<c128 *>(smem_raw);  // BTR x m
    __shared__ c128 tb[BTR];
    constexpr int CH = 256 / G;
    const int ip = blockIdx.x;
    const int r = rall[ip];
    if (r <= 0) return;
    const int c0 = blockIdx.y * CH;  // this workgroup's eigenvectors c0 .. c0+CH-1
    if (c0 >= r) return;
    const int tid = threadIdx.x;
    const int kl = tid / G, t = tid % G;
    const c128 *A = Gall + (int64_t)ip * m * m;
    const c128 *tv = tall + (int64_t)ip * m;
    const double *Z = Zall + (int64_t)ip * m * m;
    c64 *U = Uall + (int64_t)ip * m * m;
    {
        const int k = c0 + kl;
        const bool active = k < r;
        c128 z[BT_MAXE];
#pragma unroll
        for (int q = 0; q < BT_MAXE; ++q) {
            const int i = t + G * q;
            z[q] = (active && i < m) ? c128{Z[(int64_t)k * m + i], 0.0} : c128{0.0, 0.0};
        }
        for (int jhi = m - 2; jhi >= 0; jhi -= BTR) {
            const int jlo = jhi - BTR + 1 > 0 ? jhi - BTR + 1 : 0;
            const int nb = jhi - jlo + 1;
            __syncthreads();
            for (int idx = tid; idx < nb * m; idx += 256) {
                const int jj = idx / m, i = idx - jj * m;
                const int j = jlo + jj;
                c128 v = {0.0, 0.0};
                if (i == j + 1) v = c128{1.0, 0.0};
                else if (i > j + 1) v = A[i + (int64_t)m * j];
                vb[jj * m + i] = v;
            }
            if (tid < nb) tb[tid] = tv[jlo + tid];
            __syncthreads();
            for (int jj = nb - 1; jj >= 0; --jj) {
                const c128 tau = tb[jj];
                if (tau.re == 0.0 && tau.im == 0.0) continue;
                const c128 *v = vb + jj * m;
                c128 s = {0.0, 0.0};
#pragma unroll
                for (int q = 0; q < BT_MAXE; ++q) {
                    const int i = t + G * q;
                    if (i < m) s = s + cmulc(v[i], z[q]);
                }
#pragma unroll
                for (int o = G / 2; o >= 1; o >>= 1) {
                    s.re += __shfl_xor(s.re, o, 64);
                    s.im += __shfl_xor(s.im, o, 64);
                }
                s = tau * s;
#pragma unroll
                for (int q = 0; q < BT_MAXE; ++q) {
                    const int i = t + G * q;
                    if (i < m) z[q] = z[q] - s * v[i];
                }
            }
        }
        if (active) {
#pragma unroll
            for (int q = 0; q < BT_MAXE; ++q) {
                const int i = t + G * q;
                if (i < m) U[(int64_t)k * m + i] = c64{(float)z[q].re, (float)z[q].im};
            }
        }
    }
}

// Batched complex GEMM C = A * B (FP32) on the matrix cores.
// A(i,k) = A[z*bA + i*ai + k*ak] * fs[z*bfs + k] (fs optional),
// B(k,j) = B[z*bB + k*bk + j*bj] (conjugated when conjB), C(i,j) = C[z*bC + i*ci + j*cj].
// K = kvar[z] when kvar is given (kvar[z] < 0 writes the identity), else Kfix.
// Grid (column tiles, row tiles, batch), renumbered onto XCDs (xcd_order).
// 64x64 output tile per workgroup of four waves, 32x32 per wave as 2x2 blocks of
// 16x16; each complex block is four v_mfma_f32_16x16x4f32 per k-step of 4 (real
// and imaginary parts as separate real products).  The MFMA operands are fed
// as (B^T, A^T) so that the accumulator lanes run down the columns of C.
// Lane layout of v_mfma_f32_16x16x4f32 (D = X Y, X 16x4, Y 4x16): lane l holds
// X[l%16][l/16], Y[l/16][l%16] and D[4*(l/16)+v][l%16], v = 0..3.
constexpr int MT = 64, MK = 16;
typedef float mf4 __attribute__((ext_vector_type(4)));

// Per-patch choice between the dense apply (Y = X W, W = U_r diag(f) U_r^H formed
// first) and the low-rank one (Y = (X U_r diag(f)) U_r^H, two skinny products):
// a patch with 0 <= r and 2r <= m takes the low-rank form.  mode 0: every patch
// (no selection); 1: dense patches only; 2: low-rank patches, N = r; 3: low-rank
// patches, K = r.  Decided per workgroup from r (uniform), so one launch of each
// form covers a batch with both kinds.
struct RankSel {
    const int *r = nullptr;
    int m = 0;
    int mode = 0;
    int lim = -1;  // modes 2 and 3: only patches keeping more than lim (when lim >= 0)
};

// Applies the selection for batch entry z: false when the workgroup has nothing
// to do; N or K narrowed to r for the low-rank products.
__device__ __forceinline__ bool rank_select(const RankSel &rs, int z, int &N, int &K) {
    if (rs.mode == 0) return true;
    const int r = rs.r[z];
    const bool low = r >= 0 && 2 * r <= rs.m;
    if (rs.mode == 1) return !low;
    if (!low || (rs.lim >= 0 && r <= rs.lim)) return false;
    if (rs.mode == 2) N = min(N, r);
    else K = r;
    return true;
}

__global__ __launch_bounds__(256) void k_cgemm(const c64 *__restrict__ A, int64_t bA, int64_t ai,
                                               int64_t ak, const c64 *__restrict__ B, int64_t bB,
                                               int64_t bk, int64_t bj, int conjB,
                                               const float *__restrict__ fs, int64_t bfs,
                                               c64 *__restrict__ C, int64_t bC, int64_t ci,
                                               int64_t cj, int M, int N, int Kfix,
                                               const int *__restrict__ kvar, RankSel rs,
                                               int accum) {
    __shared__ float Ar[MK][MT + 4], Ai[MK][MT + 4];
    __shared__ float Br[MK][MT + 4], Bi[MK][MT + 4];
    // the column tiles of one block of rows share A: keep them on one XCD
    const int64_t nwg = (int64_t)gridDim.x * gridDim.y * gridDim.z;
    const int64_t wl = xcd_order(blockIdx.x + (int64_t)gridDim.x * (blockIdx.y + (int64_t)gridDim.y * blockIdx.z), nwg);
    const int z = (int)(wl / ((int64_t)gridDim.x * gridDim.y));
    const int j0 = (int)(wl % gridDim.x) * MT, i0 = (int)((wl / gridDim.x) % gridDim.y) * MT;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    c64 *Cz = C + (int64_t)z * bC;
    int K = Kfix;
    if (!rank_select(rs, z, N, K) || j0 >= N) return;  // uniform per workgroup
    if (kvar) {
        K = kvar[z];
        if (K < 0) {  // identity
            for (int idx = tid; idx < MT * MT; idx += 256) {
                const int i = i0 + (idx & 63), j = j0 + (idx >> 6);
                if (i < M && j < N) Cz[i * ci + j * cj] = c64{i == j ? 1.f : 0.f, 0.f};
            }
            return;
        }
    }
    const c64 *Az = A + (int64_t)z * bA;
    const c64 *Bz = B + (int64_t)z * bB;
    const float *fz = fs ? fs + (int64_t)z * bfs : nullptr;
    const int wi = (wave & 1) * 32, wj = (wave >> 1) * 32;  // wave's sub-tile
    mf4 accr[2][2], acci[2][2];
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj2 = 0; bj2 < 2; ++bj2) {
            accr[bi][bj2] = mf4{0.f, 0.f, 0.f, 0.f};
            acci[bi][bj2] = mf4{0.f, 0.f, 0.f, 0.f};
        }
    const bool afast_i = ai == 1;
    const bool bfast_k = bk == 1;
    const int r16 = lane & 15, k4 = lane >> 4;
    constexpr int NR = (MK * MT) / 256;
    // global -> registers for one k-tile (the next tile is fetched while the
    // current one is multiplied)
    // (the loads of a tile are issued together; the scaling by fs and the
    // conjugation wait until the tile is staged)
    c64 ra[NR], rb[NR];
    float fa[NR];
    const float bsgn = conjB ? -1.f : 1.f;
    auto fetch = [&](int k0) {
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const int idx = tid + 256 * r;
            const int kk = afast_i ? idx / MT : idx % MK;
            const int ii = afast_i ? idx % MT : idx / MK;
            const int k = k0 + kk, i = i0 + ii;
            const bool ina = k < K && i < M;
            ra[r] = ina ? Az[i * ai + k * ak] : c64{0.f, 0.f};
            fa[r] = (fz && ina) ? fz[k] : 1.f;
            const int kb = bfast_k ? idx % MK : idx / MT;
            const int jb = bfast_k ? idx / MK : idx % MT;
            const int kq = k0 + kb, j = j0 + jb;
            rb[r] = (kq < K && j < N) ? Bz[kq * bk + j * bj] : c64{0.f, 0.f};
        }
    };
    if (K > 0) fetch(0);
    for (int k0 = 0; k0 < K; k0 += MK) {
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const int idx = tid + 256 * r;
            const int kk = afast_i ? idx / MT : idx % MK;
            const int ii = afast_i ? idx % MT : idx / MK;
            Ar[kk][ii] = ra[r].re * fa[r];
            Ai[kk][ii] = ra[r].im * fa[r];
            const int kb = bfast_k ? idx % MK : idx / MT;
            const int jb = bfast_k ? idx / MK : idx % MT;
            Br[kb][jb] = rb[r].re;
            Bi[kb][jb] = rb[r].im * bsgn;
        }
        __syncthreads();
        if (k0 + MK < K) fetch(k0 + MK);
#pragma unroll
        for (int ks = 0; ks < MK; ks += 4) {
            const int kr = ks + k4;
            float xr[2], xi[2], xn[2], yr[2], yi[2];
#pragma unroll
            for (int bj2 = 0; bj2 < 2; ++bj2) {  // X = B^T block rows j
                xr[bj2] = Br[kr][wj + 16 * bj2 + r16];
                xi[bj2] = Bi[kr][wj + 16 * bj2 + r16];
                xn[bj2] = -xi[bj2];
            }
#pragma unroll
            for (int bi = 0; bi < 2; ++bi) {  // Y = A^T block columns i
                yr[bi] = Ar[kr][wi + 16 * bi + r16];
                yi[bi] = Ai[kr][wi + 16 * bi + r16];
            }
#pragma unroll
            for (int bi = 0; bi < 2; ++bi)
#pragma unroll
                for (int bj2 = 0; bj2 < 2; ++bj2) {
                    mf4 cr = accr[bi][bj2], cim = acci[bi][bj2];
                    cr = __builtin_amdgcn_mfma_f32_16x16x4f32(xr[bj2], yr[bi], cr, 0, 0, 0);
                    cr = __builtin_amdgcn_mfma_f32_16x16x4f32(xn[bj2], yi[bi], cr, 0, 0, 0);
                    cim = __builtin_amdgcn_mfma_f32_16x16x4f32(xr[bj2], yi[bi], cim, 0, 0, 0);
                    cim = __builtin_amdgcn_mfma_f32_16x16x4f32(xi[bj2], yr[bi], cim, 0, 0, 0);
                    accr[bi][bj2] = cr;
                    acci[bi][bj2] = cim;
                }
        }
        __syncthreads();
    }
    // D[j][i] at lane (4*(l/16)+v, l%16): i = l%16 runs down the column of C
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj2 = 0; bj2 < 2; ++bj2)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = i0 + wi + 16 * bi + r16;
                const int j = j0 + wj + 16 * bj2 + 4 * k4 + v;
                if (i < M && j < N) {
                    c64 o = c64{accr[bi][bj2][v], acci[bi][bj2][v]};
                    if (accum) o = o + Cz[i * ci + j * cj];
                    Cz[i * ci + j * cj] = o;
                }
            }
}

// ------------------------------------------------------------------ split-BF16 products
// A FP32 value v is carried as hi + lo, hi = bf16(v), lo = bf16(v - hi), so
// |v - hi - lo| <= 2^-18 |v|; a product in the reconstruction GEMM is
// hi*hi' + hi*lo' + lo*hi' (the lo*lo' term is below the split error there; the
// Gram keeps it, below).  bf16 x bf16 products are exact in FP32.
// A complex value is one 32-bit word per part (bf16 re in the low half, bf16 im
// in the high half), so a run of K complex values is the run of 2K interleaved
// reals that v_mfma_f32_16x16x32_bf16 contracts: one instruction covers 16
// complex k.  Its lane layout (D = P Q, P 16x32, Q 32x16): lane l holds
// P[l%16][8(l/16) .. 8(l/16)+7], Q[8(l/16) .. +7][l%16] and D[4(l/16)+v][l%16].
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef unsigned u4 __attribute__((ext_vector_type(4)));

constexpr int BKS = 32;         // complex k per LDS stage
constexpr int BLD = BKS + 4;    // LDS row pitch in words: 16-byte reads of 16 rows hit distinct banks
constexpr int GB_T = 512;       // threads of the split-BF16 Gram (eight waves of 32 x 16)

__device__ __forceinline__ void split_c64(c64 v, unsigned &hi, unsigned &lo) {
    const bf2v h = {(__bf16)v.re, (__bf16)v.im};
    const bf2v l = {(__bf16)(v.re - (float)h[0]), (__bf16)(v.im - (float)h[1])};
    hi = __builtin_bit_cast(unsigned, h);
    lo = __builtin_bit_cast(unsigned, l);
}
// (re, im) -> (re, -im)
__device__ __forceinline__ u4 w_conj(u4 w) { return w ^ 0x80000000u; }
// (re, im) -> (im, re)
__device__ __forceinline__ u4 w_swap(u4 w) { return (w >> 16) | (w << 16); }
// (re, im) -> (im, -re)
__device__ __forceinline__ u4 w_swapneg(u4 w) { return w_swap(w) ^ 0x80000000u; }

__device__ __forceinline__ mf4 mfma_bf16(u4 p, u4 q, mf4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8, p),
                                                    __builtin_bit_cast(bf8, q), c, 0, 0, 0);
}
// c += P Q in split form
__device__ __forceinline__ mf4 mfma_split(u4 ph, u4 pl, u4 qh, u4 ql, mf4 c) {
    c = mfma_bf16(pl, qh, c);
    c = mfma_bf16(ph, ql, c);
    return mfma_bf16(ph, qh, c);
}

// Eight complex values of one row (column) of a tile into LDS as hi and lo words.
__device__ __forceinline__ void stage8(const c64 (&r)[8], unsigned *sh, unsigned *sl) {
    unsigned h[8], l[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) split_c64(r[e], h[e], l[e]);
    *(u4 *)sh = u4{h[0], h[1], h[2], h[3]};
    *(u4 *)(sh + 4) = u4{h[4], h[5], h[6], h[7]};
    *(u4 *)sl = u4{l[0], l[1], l[2], l[3]};
    *(u4 *)(sl + 4) = u4{l[4], l[5], l[6], l[7]};
}

// Ozaki slices (ozaki.hpp) of row-contiguous patch matrices (rows of each column
// contiguous, the columns of all patches one after another), a block of SPL_ROWS
// rows per exponent.  One wave per (column, block).

// One wave's block of up to SPL_ROWS values (v0: element lane, v1: lane + 64) as
// Ozaki words with the block's exponent, n valid values written at W[0 .. n).
__device__ __forceinline__ void split_block(c64 v0, c64 v1, int lane, int n, unsigned *Wh,
                                            unsigned *Wl) {
    float sg1, sg2;
    ozaki_block_quanta(
        fmaxf(fmaxf(fabsf(v0.re), fabsf(v0.im)), fmaxf(fabsf(v1.re), fabsf(v1.im))), sg1, sg2);
    unsigned h, l;
    if (lane < n) {
        ozaki_words(v0, sg1, sg2, h, l);
        Wh[lane] = h;
        Wl[lane] = l;
    }
    if (lane + 64 < n) {
        ozaki_words(v1, sg1, sg2, h, l);
        Wh[lane + 64] = h;
        Wl[lane + 64] = l;
    }
}

// Xc (nullable): also an FP32 copy of X (for products that overwrite X).
__global__ __launch_bounds__(256) void k_split_rows(const c64 *__restrict__ X, int64_t p,
                                                    int64_t ncol, unsigned *__restrict__ Wh,
                                                    unsigned *__restrict__ Wl,
                                                    c64 *__restrict__ Xc) {
    const int64_t nblk = (p + SPL_ROWS - 1) / SPL_ROWS;
    const int64_t wid = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (wid >= nblk * ncol) return;
    const int lane = threadIdx.x & 63;
    const int64_t col = wid / nblk, blk = wid - col * nblk;
    const int64_t off = col * p + blk * SPL_ROWS;
    const int n = (int)std::min<int64_t>(SPL_ROWS, p - blk * SPL_ROWS);
    const c64 v0 = lane < n ? X[off + lane] : c64{0.f, 0.f};
    const c64 v1 = lane + 64 < n ? X[off + lane + 64] : c64{0.f, 0.f};
    if (Xc) {
        if (lane < n) Xc[off + lane] = v0;
        if (lane + 64 < n) Xc[off + lane + 64] = v1;
    }
    split_block(v0, v1, lane, n, Wh + off, Wl + off);
}

// k_split_rows reading the patches from their image: patches ip0 .. ip0 + np of
// the geometry g, words and the FP32 copy Xc in patch-matrix order (columns of the
// np patches one after another).  One wave per (patch, block of SPL_ROWS rows,
// group of SG_T frames): the rows' voxel offsets once, then a frame per step.
// The patches' x runs share cache lines with their x neighbours' (a run of 10
// voxels is 80 bytes), so the waves are ordered patch before frame group and the
// workgroups mapped to XCDs so that consecutive ones share an XCD (workgroup b runs
// on XCD b mod 8): x-neighbouring patches then read those lines through one L2.
constexpr int SG_T = 16;
__global__ __launch_bounds__(256) void k_split_gather(const c64 *__restrict__ img, PatchGeom g,
                                                      int64_t ip0, int64_t np,
                                                      unsigned *__restrict__ Wh,
                                                      unsigned *__restrict__ Wl,
                                                      c64 *__restrict__ Xc) {
    const int64_t p = g.p;
    const int Nt = g.Nt;
    const int64_t nblk = (p + SPL_ROWS - 1) / SPL_ROWS;
    const int ntg = (Nt + SG_T - 1) / SG_T;
    const int64_t G = gridDim.x, b = blockIdx.x, G8 = G - G % 8;
    const int64_t L = b < G8 ? (b % 8) * (G8 / 8) + b / 8 : b;  // a bijection of [0, G)
    const int64_t wid = L * 4 + (threadIdx.x >> 6);
    if (wid >= np * nblk * ntg) return;
    const int lane = threadIdx.x & 63;
    const int64_t blk = wid % nblk, r = wid / nblk;
    const int64_t z = r % np;
    const int tg = (int)(r / np);
    const int n = (int)std::min<int64_t>(SPL_ROWS, p - blk * SPL_ROWS);
    const int r0 = (int)(blk * SPL_ROWS) + lane;
    const int64_t o0 = lane < n ? patch_voxel(g, ip0 + z, r0) : 0;
    const int64_t o1 = lane + 64 < n ? patch_voxel(g, ip0 + z, r0 + 64) : 0;
    const int64_t V = (int64_t)g.N[0] * g.N[1] * g.N[2];
    const int t1 = min(Nt, (tg + 1) * SG_T);
    // four frames' loads issued together (unconditional: o0 = o1 = 0 past the block,
    // voxel 0 of the frame, and the last frame repeated past t1), the lanes past the
    // block zeroed afterwards
    for (int t0 = tg * SG_T; t0 < t1; t0 += 4) {
        c64 a0[4], a1[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int t = min(t0 + q, t1 - 1);
            a0[q] = img[o0 + t * V];
            a1[q] = img[o1 + t * V];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int t = t0 + q;
            if (t >= t1) break;  // uniform per wave
            const c64 v0 = lane < n ? a0[q] : c64{0.f, 0.f};
            const c64 v1 = lane + 64 < n ? a1[q] : c64{0.f, 0.f};
            const int64_t off = (z * Nt + t) * p + blk * SPL_ROWS;
            if (Xc) {
                if (lane < n) Xc[off + lane] = v0;
                if (lane + 64 < n) Xc[off + lane + 64] = v1;
            }
            split_block(v0, v1, lane, n, Wh + off, Wl + off);
        }
    }
}

// Patches with fewer voxels than frames (the Gram X X^H contracts over time):
// row a of patch ip as time-contiguous words, W[(ip p + a) Nt + k], blocks of
// SPL_ROWS times sharing an exponent, so that k_gram_bf16 reads them as it reads
// row-contiguous words (and conjugates its result).  Consecutive waves take
// consecutive rows, which share the cache lines of the strided reads.
__global__ __launch_bounds__(256) void k_split_cols(const c64 *__restrict__ X, int64_t p,
                                                    int64_t Nt, int64_t nb,
                                                    unsigned *__restrict__ Wh,
                                                    unsigned *__restrict__ Wl) {
    const int64_t nblk = (Nt + SPL_ROWS - 1) / SPL_ROWS;
    const int64_t wid = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (wid >= nb * nblk * p) return;
    const int lane = threadIdx.x & 63;
    const int64_t a = wid % p, t = wid / p;
    const int64_t blk = t % nblk, ip = t / nblk;
    const int64_t k0 = blk * SPL_ROWS;
    const int n = (int)std::min<int64_t>(SPL_ROWS, Nt - k0);
    const c64 *src = X + ip * p * Nt + a;
    const c64 v0 = lane < n ? src[(k0 + lane) * p] : c64{0.f, 0.f};
    const c64 v1 = lane + 64 < n ? src[(k0 + lane + 64) * p] : c64{0.f, 0.f};
    const int64_t off = (ip * p + a) * Nt + k0;
    split_block(v0, v1, lane, n, Wh + off, Wl + off);
}

// Split-BF16 Gram tiles with error-free sums, from the Ozaki words of
// k_split_rows (row-contiguous patches): same tiling, contraction chunks and
// FP64 partial output as k_gram_mfma.  For an entry (a, b) all s1*s1 products of
// one SPL_ROWS-row block share one quantum and stay below 2^16 of it, so the
// MFMA sums over a block (256 products) are exact in FP32; so are those of the
// s1*s2 + s2*s1 terms (a quantum 2^8 smaller, 512 products below 2^15) and of
// the s2*s2 terms (2^16 smaller).  After each block the three sums go into FP64
// accumulators.  The result is the exact Gram of s1 + s2: the FP32 data
// perturbed by at most 2^-17 of each block's maximum, with no rounding of long
// sums (which would put eps32 of the dominant eigenvalue onto every small one).
// The s2*s2 terms are kept because dropping them is a one-signed bias of the
// diagonal, ~2^-20 of the dominant eigenvalue, which would bury eigenvalues 1e8
// below it.  Operands are (X(k, b), X(k, a)): G_re = [br, bi].[ar, ai],
// G_im = [bi, -br].[ar, ai] over interleaved k.  Eight waves of 32 x 16 sub-tiles (two
// 16 x 16 blocks, so 152 VGPRs and two waves per SIMD) and two LDS stages (one
// barrier per step).  The block sums being exact, any tiling or product order gives
// the same Gram bit for bit.
__global__ __launch_bounds__(GB_T) void k_gram_bf16(const unsigned *__restrict__ Wh,
                                                    const unsigned *__restrict__ Wl,
                                                    int64_t pstride, int64_t cs, int m, int64_t Kc,
                                                    int nkc, int kch, int ntile, int vec,
                                                    c128 *__restrict__ part, c128 *__restrict__ G,
                                                    int conj_out, int *__restrict__ err) {
    // two LDS stages: the next step's words are stored while the current one is read
    __shared__ __align__(16) unsigned Sh[2][2][GT][BLD];  // [stage][A cols a, B cols b][col][k]
    __shared__ __align__(16) unsigned Sl[2][2][GT][BLD];
    // the tile pairs of one (patch, chunk) share column tiles: keep them on one XCD
    const int64_t nwg = (int64_t)gridDim.x * gridDim.y * gridDim.z;
    const int64_t wl = xcd_order(blockIdx.x + (int64_t)gridDim.x * (blockIdx.y + (int64_t)gridDim.y * blockIdx.z), nwg);
    const int ip = (int)(wl / ((int64_t)gridDim.x * gridDim.y));
    const int kc = (int)((wl / gridDim.x) % gridDim.y);
    int rem = (int)(wl % gridDim.x), ti = 0;
    while (rem >= ntile - ti) {
        rem -= ntile - ti;
        ++ti;
    }
    const int tj = ti + rem;
    const int a0 = ti * GT, b0 = tj * GT;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    // eight waves, each a 32 (a) x 16 (b) sub-tile: two 16 x 16 blocks
    const int wa = (wave & 1) * 32, wb = (wave >> 1) * 16;
    const int r16 = lane & 15, g4 = lane >> 4;
    // load slot: one column, four consecutive k (one 16-byte word group); columns
    // past m load column m - 1 and are zeroed when staged, so that no loaded value
    // is consumed before its stage's put and the loads stay in flight across the
    // step (LDS-only barriers)
    const int lc = tid >> 3;
    const int lk = (tid & 7) * 4;
    const bool va = a0 + lc < m, vb = b0 + lc < m;
    const int64_t oa = ip * pstride + (int64_t)min(a0 + lc, m - 1) * cs;
    const int64_t ob = ip * pstride + (int64_t)min(b0 + lc, m - 1) * cs;
    mf4 hr[2], hi[2], xr[2], xi[2], lr[2], li[2];
    md4 dr[2], di[2];
    auto zero32 = [&]() {
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            hr[i] = mf4{0.f, 0.f, 0.f, 0.f};
            hi[i] = mf4{0.f, 0.f, 0.f, 0.f};
            xr[i] = mf4{0.f, 0.f, 0.f, 0.f};
            xi[i] = mf4{0.f, 0.f, 0.f, 0.f};
            lr[i] = mf4{0.f, 0.f, 0.f, 0.f};
            li[i] = mf4{0.f, 0.f, 0.f, 0.f};
        }
    };
    zero32();
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        dr[i] = md4{0.0, 0.0, 0.0, 0.0};
        di[i] = md4{0.0, 0.0, 0.0, 0.0};
    }
    const int64_t kbeg = (int64_t)kc * kch;
    const int64_t kend = std::min<int64_t>(Kc, kbeg + kch);
    u4 ah, al, bh, bl;
    auto fetch = [&](int64_t k0) {
        const int64_t k = k0 + lk;
        if (vec && k0 + BKS <= kend) {  // a whole step of 16-byte aligned rows (p % 4 == 0)
            ah = *(const u4 *)(Wh + oa + k);
            al = *(const u4 *)(Wl + oa + k);
            bh = *(const u4 *)(Wh + ob + k);
            bl = *(const u4 *)(Wl + ob + k);
        } else {
            unsigned t[4][4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const bool kin = k + e < kend;
                t[0][e] = (kin && va) ? Wh[oa + k + e] : 0u;
                t[1][e] = (kin && va) ? Wl[oa + k + e] : 0u;
                t[2][e] = (kin && vb) ? Wh[ob + k + e] : 0u;
                t[3][e] = (kin && vb) ? Wl[ob + k + e] : 0u;
            }
            ah = u4{t[0][0], t[0][1], t[0][2], t[0][3]};
            al = u4{t[1][0], t[1][1], t[1][2], t[1][3]};
            bh = u4{t[2][0], t[2][1], t[2][2], t[2][3]};
            bl = u4{t[3][0], t[3][1], t[3][2], t[3][3]};
        }
    };
    auto put = [&](int st) {
        const u4 z = {0u, 0u, 0u, 0u};
        *(u4 *)&Sh[st][0][lc][lk] = va ? ah : z;
        *(u4 *)&Sl[st][0][lc][lk] = va ? al : z;
        *(u4 *)&Sh[st][1][lc][lk] = vb ? bh : z;
        *(u4 *)&Sl[st][1][lc][lk] = vb ? bl : z;
    };
    fetch(kbeg);
    put(0);
    if (kbeg + BKS < kend) fetch(kbeg + BKS);
    __syncthreads();
    int st = 0;
    for (int64_t k0 = kbeg; k0 < kend; k0 += BKS) {
        // the next step's words into the other stage (its last readers passed the
        // barrier that ended the previous step), then the step after it in flight
        if (k0 + BKS < kend) {
            put(st ^ 1);
            if (k0 + 2 * BKS < kend) fetch(k0 + 2 * BKS);
        }
#pragma unroll
        for (int s = 0; s < BKS / 16; ++s) {
            const int kw = 16 * s + 4 * g4;
            u4 qh[2], ql[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                qh[i] = *(const u4 *)&Sh[st][0][wa + 16 * i + r16][kw];
                ql[i] = *(const u4 *)&Sl[st][0][wa + 16 * i + r16][kw];
            }
            const u4 ph = *(const u4 *)&Sh[st][1][wb + r16][kw];
            const u4 pl = *(const u4 *)&Sl[st][1][wb + r16][kw];
            const u4 ph2 = w_swapneg(ph), pl2 = w_swapneg(pl);
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                hr[i] = mfma_bf16(ph, qh[i], hr[i]);
                xr[i] = mfma_bf16(pl, qh[i], xr[i]);
                lr[i] = mfma_bf16(pl, ql[i], lr[i]);
                hi[i] = mfma_bf16(ph2, qh[i], hi[i]);
                xi[i] = mfma_bf16(pl2, qh[i], xi[i]);
                li[i] = mfma_bf16(pl2, ql[i], li[i]);
            }
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                xr[i] = mfma_bf16(ph, ql[i], xr[i]);
                xi[i] = mfma_bf16(ph2, ql[i], xi[i]);
            }
        }
        // end of a SPL_ROWS-row block (the chunk starts on a block boundary) or of
        // the chunk: the exact FP32 sums into FP64
        if (((k0 + BKS - kbeg) % SPL_ROWS) == 0 || k0 + BKS >= kend) {
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    dr[i][v] += ((double)hr[i][v] + (double)xr[i][v]) + (double)lr[i][v];
                    di[i][v] += ((double)hi[i][v] + (double)xi[i][v]) + (double)li[i][v];
                }
            zero32();
        }
        lds_barrier();  // (the next step's loads stay in flight)
        st ^= 1;
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int a = a0 + wa + 16 * i + r16;
            const int b = b0 + wb + 4 * g4 + v;
            gram_store(part, G, ip, nkc, kc, m, a, b, dr[i][v], di[i][v], conj_out, err);
        }
}

// Split-BF16 complex GEMM C = A B (the SVST apply), FP32 accumulation.  Same
// tiling, operand conventions and output layout as k_cgemm (64x64 per
// workgroup, 32x32 per wave; A(i,k) scaled by fs[k] when given, B conjugated when
// conjB), operands (B^T, A^T): C_re = [br, bi].[ar, -ai], C_im = [br, bi].[ai, ar].
// Each FP32 operand is split into hi + lo BF16 words when staged.  blockIdx.x runs
// over the column tiles of C so the tiles sharing a block of rows of A are
// dispatched together.
template <bool SCALED>
__global__ __launch_bounds__(256) void k_cgemm_bf16(const c64 *__restrict__ A, int64_t bA,
                                                    int64_t ai, int64_t ak,
                                                    const c64 *__restrict__ B, int64_t bB,
                                                    int64_t bk, int64_t bj, int conjB,
                                                    const float *__restrict__ fs, int64_t bfs,
                                                    c64 *__restrict__ C, int64_t bC, int64_t ci,
                                                    int64_t cj, int M, int N, int K, RankSel rs) {
    __shared__ __align__(16) unsigned Sh[2][MT][BLD];  // [A rows i, B cols j][row][k]
    __shared__ __align__(16) unsigned Sl[2][MT][BLD];
    // the column tiles of one block of rows share A: keep them on one XCD
    const int64_t nwg = (int64_t)gridDim.x * gridDim.y * gridDim.z;
    const int64_t wl = xcd_order(blockIdx.x + (int64_t)gridDim.x * (blockIdx.y + (int64_t)gridDim.y * blockIdx.z), nwg);
    const int z = (int)(wl / ((int64_t)gridDim.x * gridDim.y));
    const int j0 = (int)(wl % gridDim.x) * MT, i0 = (int)((wl / gridDim.x) % gridDim.y) * MT;
    if (!rank_select(rs, z, N, K) || j0 >= N) return;  // uniform per workgroup
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wi = (wave & 1) * 32, wj = (wave >> 1) * 32;
    const int r16 = lane & 15, g4 = lane >> 4;
    // load slots: one row of A (column of B), eight consecutive k
    const bool afast = ai == 1, bfast = bk == 1;
    const int ar_ = afast ? tid & 63 : tid >> 2;
    const int ak_ = afast ? (tid >> 6) * 8 : (tid & 3) * 8;
    const int bc_ = bfast ? tid >> 2 : tid & 63;
    const int bk_ = bfast ? (tid & 3) * 8 : (tid >> 6) * 8;
    const bool av = i0 + ar_ < M, bv = j0 + bc_ < N;
    const c64 *pa = A + (int64_t)z * bA + (int64_t)(i0 + ar_) * ai;
    const c64 *pb = B + (int64_t)z * bB + (int64_t)(j0 + bc_) * bj;
    const float *fz = fs ? fs + (int64_t)z * bfs : nullptr;
    mf4 accr[2][2], acci[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            accr[i][j] = mf4{0.f, 0.f, 0.f, 0.f};
            acci[i][j] = mf4{0.f, 0.f, 0.f, 0.f};
        }
    // (the loads of a tile are issued together; with SCALED the scaling by fs and
    // the conjugation of B wait until the tile is staged)
    c64 ra[8], rb[8];
    float fa[SCALED ? 8 : 1];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int ka = k0 + ak_ + e, kb = k0 + bk_ + e;
            const bool ina = av && ka < K;
            ra[e] = ina ? pa[(int64_t)ka * ak] : c64{0.f, 0.f};
            rb[e] = (bv && kb < K) ? pb[(int64_t)kb * bk] : c64{0.f, 0.f};
            if constexpr (SCALED) fa[e] = (fz && ina) ? fz[ka] : 1.f;
        }
    };
    if (K > 0) fetch(0);
    for (int k0 = 0; k0 < K; k0 += BKS) {
        if constexpr (SCALED) {
            const float bsgn = conjB ? -1.f : 1.f;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                ra[e] = c64{ra[e].re * fa[e], ra[e].im * fa[e]};
                rb[e].im *= bsgn;
            }
        }
        stage8(ra, &Sh[0][ar_][ak_], &Sl[0][ar_][ak_]);
        stage8(rb, &Sh[1][bc_][bk_], &Sl[1][bc_][bk_]);
        __syncthreads();
        if (k0 + BKS < K) fetch(k0 + BKS);
#pragma unroll
        for (int s = 0; s < BKS / 16; ++s) {
            const int kw = 16 * s + 4 * g4;
            u4 qh[2], ql[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                qh[i] = *(const u4 *)&Sh[0][wi + 16 * i + r16][kw];
                ql[i] = *(const u4 *)&Sl[0][wi + 16 * i + r16][kw];
            }
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const u4 ph = *(const u4 *)&Sh[1][wj + 16 * j + r16][kw];
                const u4 pl = *(const u4 *)&Sl[1][wj + 16 * j + r16][kw];
#pragma unroll
                for (int i = 0; i < 2; ++i) {
                    accr[i][j] = mfma_split(ph, pl, w_conj(qh[i]), w_conj(ql[i]), accr[i][j]);
                    acci[i][j] = mfma_split(ph, pl, w_swap(qh[i]), w_swap(ql[i]), acci[i][j]);
                }
            }
        }
        __syncthreads();
    }
    c64 *Cz = C + (int64_t)z * bC;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int ii = i0 + wi + 16 * i + r16;
                const int jj = j0 + wj + 16 * j + 4 * g4 + v;
                if (ii < M && jj < N) Cz[ii * ci + jj * cj] = c64{accr[i][j][v], acci[i][j][v]};
            }
}

// ---- split-BF16 reconstruction products C = A B' with row-contiguous A (the
// p >= Nt applies X W, X U_r and T f U_r^H).  The small right operand is split
// once per launch (k_apply_bsplit) into hi / lo words, B'(k, j) = fs[k] B(k, j)
// (conjugated when conjB), stored per batch entry as [column j][k], k padded to a
// multiple of BKS with zeros (the stages up to the selected K only);
// k_apply_bf16 then stages it into LDS with plain 16-byte copies and splits only
// A.  A workgroup takes a block of 64 rows of A and all the column tiles of
// 32 NB; four waves of 32 x 16 NB.  The grid strides over the (batch entry, row
// block) units, so entries that the rank selection skips cost no workgroup
// dispatch.
constexpr int AP_ROWS = 64;  // rows of A per unit

__global__ __launch_bounds__(256) void k_apply_bsplit(const c64 *__restrict__ B, int64_t bB,
                                                      int64_t bk, int64_t bj, int conjB,
                                                      const float *__restrict__ fs, int64_t bfs,
                                                      int N, int K, int Kp, RankSel rs,
                                                      unsigned *__restrict__ Wh,
                                                      unsigned *__restrict__ Wl, int64_t bW) {
    const int z = blockIdx.y;
    int Nz = N, Kz = K;
    if (!rank_select(rs, z, Nz, Kz)) return;
    // four consecutive k of one column j per thread (Kp is a multiple of BKS), the
    // grid striding over the entry (the host caps it: most of a batch's entries are
    // skipped by the rank selection, or narrowed to a few columns or k)
    const int64_t ne = (int64_t)Nz * Kp;
    const int kend = (Kz + BKS - 1) / BKS * BKS;  // stages k_apply_bf16 reads
    for (int64_t e = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4; e < ne;
         e += (int64_t)gridDim.x * blockDim.x * 4) {
        const int j = (int)(e / Kp), k0 = (int)(e - (int64_t)j * Kp);
        if (k0 >= kend) continue;
        unsigned h[4], l[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int k = k0 + v, kc = min(k, Kz - 1);  // Kz >= 1 past the test above
            c64 b = B[(int64_t)z * bB + (int64_t)kc * bk + (int64_t)j * bj];
            if (conjB) b.im = -b.im;
            if (fs) {
                const float f = fs[(int64_t)z * bfs + kc];
                b = c64{b.re * f, b.im * f};
            }
            split_c64(k < Kz ? b : c64{0.f, 0.f}, h[v], l[v]);
        }
        *(u4 *)(Wh + (int64_t)z * bW + e) = u4{h[0], h[1], h[2], h[3]};
        *(u4 *)(Wl + (int64_t)z * bW + e) = u4{l[0], l[1], l[2], l[3]};
    }
}

// Workgroups per batch entry of k_apply_bsplit: enough for the whole entry alone, at
// most about 16384 over the batch.
static unsigned bsplit_groups(int64_t bW, int64_t nz) {
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(cdiv(bW, 1024), cdiv(16384, nz)));
}

// AW: A arrives as the split pass's Ozaki words (wa: s1, then s2 at wa + wlo, same
// element offsets as A), staged into LDS as they are; else A is FP32 and split per
// element when staged.
template <int NB, bool AW>
__global__ __launch_bounds__(256, 2) void k_apply_bf16(const c64 *__restrict__ A,
                                                       const unsigned *__restrict__ wa,
                                                       int64_t wlo, int64_t bA,
                                                       int64_t ak, const unsigned *__restrict__ Wh,
                                                       const unsigned *__restrict__ Wl, int64_t bW,
                                                       int Kp, c64 *__restrict__ C, int64_t bC,
                                                       int64_t cj, int M, int N, int K, int nb,
                                                       RankSel rs, PatchGeom sg, int64_t sip0,
                                                       int scat) {
    constexpr int TN = 32 * NB;               // columns per tile
    constexpr int BQ = TN * BKS / 4 / 256;    // 16-byte B words per thread and part
    static_assert(TN * BKS / 4 % 256 == 0, "B tile must split evenly over the threads");
    __shared__ __align__(16) unsigned Ah[AP_ROWS][BLD], Al[AP_ROWS][BLD];
    __shared__ __align__(16) unsigned Bh[TN][BLD], Bl[TN][BLD];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wi = (wave & 1) * 32, wj = (wave >> 1) * (16 * NB);
    const int r16 = lane & 15, g4 = lane >> 4;
    const int ar = tid & 63, akq = (tid >> 6) * 8;  // A slot: one row, eight consecutive k
    const int nrt = (M + AP_ROWS - 1) / AP_ROWS;
    const int64_t nunits = (int64_t)nb * nrt;
    for (int64_t u = blockIdx.x; u < nunits; u += gridDim.x) {
        const int z = (int)(u / nrt);
        const int i0 = (int)(u - (int64_t)z * nrt) * AP_ROWS;
        int Nz = N, Kz = K;
        if (!rank_select(rs, z, Nz, Kz)) continue;  // uniform per workgroup
        const bool av = i0 + ar < M;
        const c64 *pa = A + (int64_t)z * bA + i0 + ar;
        const unsigned *pw = wa + (int64_t)z * bA + i0 + ar;
        const unsigned *wh = Wh + (int64_t)z * bW, *wl = Wl + (int64_t)z * bW;
        c64 *Cz = scat ? C : C + (int64_t)z * bC;
        // one k stage: the A tile stays in LDS for all the column tiles
        const bool one = Kz <= BKS;
        c64 ra[AW ? 1 : 8];
        unsigned rwh[AW ? 8 : 1], rwl[AW ? 8 : 1];
        auto fetch = [&](int k0) {
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int k = k0 + akq + e;
                const bool ok = av && k < Kz;
                if constexpr (AW) {
                    rwh[e] = ok ? pw[(int64_t)k * ak] : 0u;
                    rwl[e] = ok ? pw[wlo + (int64_t)k * ak] : 0u;
                } else {
                    ra[e] = ok ? pa[(int64_t)k * ak] : c64{0.f, 0.f};
                }
            }
        };
        for (int j0 = 0; j0 < Nz; j0 += TN) {
            mf4 accr[2][NB], acci[2][NB];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < NB; ++j) {
                    accr[i][j] = mf4{0.f, 0.f, 0.f, 0.f};
                    acci[i][j] = mf4{0.f, 0.f, 0.f, 0.f};
                }
            u4 rbh[BQ], rbl[BQ];
            auto fetch_b = [&](int k0) {
#pragma unroll
                for (int q = 0; q < BQ; ++q) {
                    const int f = tid + 256 * q;
                    const int jj = f / (BKS / 4), kq = (f % (BKS / 4)) * 4;
                    const bool ok = j0 + jj < Nz;
                    const int64_t o = (int64_t)(ok ? j0 + jj : 0) * Kp + k0 + kq;
                    const u4 h = *(const u4 *)(wh + o), l = *(const u4 *)(wl + o);
                    rbh[q] = ok ? h : u4{0u, 0u, 0u, 0u};
                    rbl[q] = ok ? l : u4{0u, 0u, 0u, 0u};
                }
            };
            const bool sa = !one || j0 == 0;
            if (sa) fetch(0);
            for (int k0 = 0; k0 < Kz; k0 += BKS) {
                fetch_b(k0);
                if (sa) {
                    if constexpr (AW) {
                        *(u4 *)&Ah[ar][akq] = u4{rwh[0], rwh[1], rwh[2], rwh[3]};
                        *(u4 *)&Ah[ar][akq + 4] = u4{rwh[4], rwh[5], rwh[6], rwh[7]};
                        *(u4 *)&Al[ar][akq] = u4{rwl[0], rwl[1], rwl[2], rwl[3]};
                        *(u4 *)&Al[ar][akq + 4] = u4{rwl[4], rwl[5], rwl[6], rwl[7]};
                    } else {
                        stage8(ra, &Ah[ar][akq], &Al[ar][akq]);
                    }
                }
#pragma unroll
                for (int q = 0; q < BQ; ++q) {
                    const int f = tid + 256 * q;
                    const int jj = f / (BKS / 4), kq = (f % (BKS / 4)) * 4;
                    *(u4 *)&Bh[jj][kq] = rbh[q];
                    *(u4 *)&Bl[jj][kq] = rbl[q];
                }
                __syncthreads();
                if (k0 + BKS < Kz) fetch(k0 + BKS);
#pragma unroll
                for (int s = 0; s < BKS / 16; ++s) {
                    const int kw = 16 * s + 4 * g4;
                    u4 ch[2], cl[2], sh[2], sl[2];
#pragma unroll
                    for (int i = 0; i < 2; ++i) {
                        const u4 qh = *(const u4 *)&Ah[wi + 16 * i + r16][kw];
                        const u4 ql = *(const u4 *)&Al[wi + 16 * i + r16][kw];
                        ch[i] = w_conj(qh);
                        cl[i] = w_conj(ql);
                        sh[i] = w_swap(qh);
                        sl[i] = w_swap(ql);
                    }
#pragma unroll
                    for (int j = 0; j < NB; ++j) {
                        if (j0 + wj + 16 * j >= Nz) break;  // uniform per wave
                        const u4 ph = *(const u4 *)&Bh[wj + 16 * j + r16][kw];
                        const u4 pl = *(const u4 *)&Bl[wj + 16 * j + r16][kw];
#pragma unroll
                        for (int i = 0; i < 2; ++i) {
                            accr[i][j] = mfma_split(ph, pl, ch[i], cl[i], accr[i][j]);
                            acci[i][j] = mfma_split(ph, pl, sh[i], sl[i], acci[i][j]);
                        }
                    }
                }
                __syncthreads();
            }
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const int ii = i0 + wi + 16 * i + r16;
                if (ii >= M) continue;
                // row ii of the patch: its voxel when scattering into the image
                const int64_t ro = scat ? patch_voxel(sg, sip0 + z, ii) : ii;
#pragma unroll
                for (int j = 0; j < NB; ++j)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int jj = j0 + wj + 16 * j + 4 * g4 + v;
                        if (jj < Nz) Cz[ro + (int64_t)jj * cj] = c64{accr[i][j][v], acci[i][j][v]};
                    }
            }
        }
    }
}

// The low-rank reconstruction Y = (X U_r) diag(f) U_r^H of the patches keeping
// r <= LR_MAX singular values (and 2r <= m), one pass per 64-row block: T = X U_r
// from the split pass's words (as k_apply_bf16<2, true> computes it), kept on chip as
// the split words of T, then Y = T B2 with B2 = diag(f) U_r^H split (as
// k_apply_bf16<2, false> computes it from T in memory: the same products in the same
// order, so the same bits).  Saves T's round trip through memory and a launch; the
// read of X and the (scattered) write of Y overlap in one kernel.
// W1 (U_r, [column j][k], k < Kp1) and W2 (f U_r^H, [column j][k], k < Kp2) as
// k_apply_bsplit writes them.
constexpr int LR_MAX = 64;
constexpr int LR_LD = LR_MAX + 4;  // LDS row pitch of T's words
__global__ __launch_bounds__(256, 2) void k_apply_lr_bf16(
    const unsigned *__restrict__ wa, int64_t wlo, int64_t bA, const unsigned *__restrict__ W1h,
    const unsigned *__restrict__ W1l, int64_t bW1, int Kp1, const unsigned *__restrict__ W2h,
    const unsigned *__restrict__ W2l, int64_t bW2, int Kp2, c64 *__restrict__ C, int64_t bC,
    int64_t cj, int M, int N2, int K1, int nb, const int *__restrict__ rall, int m,
    PatchGeom sg, int64_t sip0, int scat) {
    constexpr int NB = 2, TN = 32 * NB;
    constexpr int BQ = TN * BKS / 4 / 256;
    __shared__ __align__(16) unsigned Ah[AP_ROWS][BLD], Al[AP_ROWS][BLD];
    __shared__ __align__(16) unsigned Bh[TN][BLD], Bl[TN][BLD];
    __shared__ __align__(16) unsigned Th[AP_ROWS][LR_LD], Tl[AP_ROWS][LR_LD];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wi = (wave & 1) * 32, wj = (wave >> 1) * (16 * NB);
    const int r16 = lane & 15, g4 = lane >> 4;
    const int ar = tid & 63, akq = (tid >> 6) * 8;
    const int nrt = (M + AP_ROWS - 1) / AP_ROWS;
    const int64_t nunits = (int64_t)nb * nrt;
    for (int64_t u = blockIdx.x; u < nunits; u += gridDim.x) {
        const int z = (int)(u / nrt);
        const int i0 = (int)(u - (int64_t)z * nrt) * AP_ROWS;
        const int r = rall[z];
        if (!(r >= 0 && 2 * r <= m && r <= LR_MAX)) continue;  // uniform per workgroup
        const bool av = i0 + ar < M;
        const unsigned *pw = wa + (int64_t)z * bA + i0 + ar;
        c64 *Cz = scat ? C : C + (int64_t)z * bC;
        mf4 accr[2][NB], acci[2][NB];
        auto zero_acc = [&]() {
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < NB; ++j) {
                    accr[i][j] = mf4{0.f, 0.f, 0.f, 0.f};
                    acci[i][j] = mf4{0.f, 0.f, 0.f, 0.f};
                }
        };
        // ---- phase 1: T = X U_r (N = r, K = K1)
        zero_acc();
        {
            unsigned rwh[8], rwl[8];
            auto fetch = [&](int k0) {
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int k = k0 + akq + e;
                    const bool ok = av && k < K1;
                    rwh[e] = ok ? pw[(int64_t)k * M] : 0u;
                    rwl[e] = ok ? pw[wlo + (int64_t)k * M] : 0u;
                }
            };
            const unsigned *wh = W1h + (int64_t)z * bW1, *wl = W1l + (int64_t)z * bW1;
            fetch(0);
            for (int k0 = 0; k0 < K1; k0 += BKS) {
#pragma unroll
                for (int q = 0; q < BQ; ++q) {
                    const int f = tid + 256 * q;
                    const int jj = f / (BKS / 4), kq = (f % (BKS / 4)) * 4;
                    const bool ok = jj < r;
                    const int64_t o = (int64_t)(ok ? jj : 0) * Kp1 + k0 + kq;
                    const u4 h = *(const u4 *)(wh + o), l = *(const u4 *)(wl + o);
                    *(u4 *)&Bh[jj][kq] = ok ? h : u4{0u, 0u, 0u, 0u};
                    *(u4 *)&Bl[jj][kq] = ok ? l : u4{0u, 0u, 0u, 0u};
                }
                *(u4 *)&Ah[ar][akq] = u4{rwh[0], rwh[1], rwh[2], rwh[3]};
                *(u4 *)&Ah[ar][akq + 4] = u4{rwh[4], rwh[5], rwh[6], rwh[7]};
                *(u4 *)&Al[ar][akq] = u4{rwl[0], rwl[1], rwl[2], rwl[3]};
                *(u4 *)&Al[ar][akq + 4] = u4{rwl[4], rwl[5], rwl[6], rwl[7]};
                __syncthreads();
                if (k0 + BKS < K1) fetch(k0 + BKS);
#pragma unroll
                for (int s = 0; s < BKS / 16; ++s) {
                    const int kw = 16 * s + 4 * g4;
                    u4 ch[2], cl[2], sh[2], sl[2];
#pragma unroll
                    for (int i = 0; i < 2; ++i) {
                        const u4 qh = *(const u4 *)&Ah[wi + 16 * i + r16][kw];
                        const u4 ql = *(const u4 *)&Al[wi + 16 * i + r16][kw];
                        ch[i] = w_conj(qh);
                        cl[i] = w_conj(ql);
                        sh[i] = w_swap(qh);
                        sl[i] = w_swap(ql);
                    }
#pragma unroll
                    for (int j = 0; j < NB; ++j) {
                        if (wj + 16 * j >= r) break;  // uniform per wave
                        const u4 ph = *(const u4 *)&Bh[wj + 16 * j + r16][kw];
                        const u4 pl = *(const u4 *)&Bl[wj + 16 * j + r16][kw];
#pragma unroll
                        for (int i = 0; i < 2; ++i) {
                            accr[i][j] = mfma_split(ph, pl, ch[i], cl[i], accr[i][j]);
                            acci[i][j] = mfma_split(ph, pl, sh[i], sl[i], acci[i][j]);
                        }
                    }
                }
                __syncthreads();
            }
        }
        // T's split words into LDS (T[row][k], k < r; zero past r up to the padded
        // stage end), as the second product's A operand
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < NB; ++j)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int row = wi + 16 * i + r16;
                    const int col = wj + 16 * j + 4 * g4 + v;
                    const bool keep = col < r && i0 + row < M;
                    unsigned h, l;
                    split_c64(keep ? c64{accr[i][j][v], acci[i][j][v]} : c64{0.f, 0.f}, h, l);
                    Th[row][col] = h;
                    Tl[row][col] = l;
                }
        __syncthreads();
        // ---- phase 2: Y = T B2 (N = N2, K = r), column tiles of TN
        const unsigned *w2h = W2h + (int64_t)z * bW2, *w2l = W2l + (int64_t)z * bW2;
        for (int j0 = 0; j0 < N2; j0 += TN) {
            zero_acc();
            for (int k0 = 0; k0 < r; k0 += BKS) {
#pragma unroll
                for (int q = 0; q < BQ; ++q) {
                    const int f = tid + 256 * q;
                    const int jj = f / (BKS / 4), kq = (f % (BKS / 4)) * 4;
                    const bool ok = j0 + jj < N2;
                    const int64_t o = (int64_t)(ok ? j0 + jj : 0) * Kp2 + k0 + kq;
                    const u4 h = *(const u4 *)(w2h + o), l = *(const u4 *)(w2l + o);
                    *(u4 *)&Bh[jj][kq] = ok ? h : u4{0u, 0u, 0u, 0u};
                    *(u4 *)&Bl[jj][kq] = ok ? l : u4{0u, 0u, 0u, 0u};
                }
                __syncthreads();
#pragma unroll
                for (int s = 0; s < BKS / 16; ++s) {
                    const int kw = 16 * s + 4 * g4;
                    u4 ch[2], cl[2], sh[2], sl[2];
#pragma unroll
                    for (int i = 0; i < 2; ++i) {
                        const u4 qh = *(const u4 *)&Th[wi + 16 * i + r16][k0 + kw];
                        const u4 ql = *(const u4 *)&Tl[wi + 16 * i + r16][k0 + kw];
                        ch[i] = w_conj(qh);
                        cl[i] = w_conj(ql);
                        sh[i] = w_swap(qh);
                        sl[i] = w_swap(ql);
                    }
#pragma unroll
                    for (int j = 0; j < NB; ++j) {
                        if (j0 + wj + 16 * j >= N2) break;  // uniform per wave
                        const u4 ph = *(const u4 *)&Bh[wj + 16 * j + r16][kw];
                        const u4 pl = *(const u4 *)&Bl[wj + 16 * j + r16][kw];
#pragma unroll
                        for (int i = 0; i < 2; ++i) {
                            accr[i][j] = mfma_split(ph, pl, ch[i], cl[i], accr[i][j]);
                            acci[i][j] = mfma_split(ph, pl, sh[i], sl[i], acci[i][j]);
                        }
                    }
                }
                __syncthreads();
            }
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const int ii = i0 + wi + 16 * i + r16;
                if (ii >= M) continue;
                const int64_t ro = scat ? patch_voxel(sg, sip0 + z, ii) : ii;
#pragma unroll
                for (int j = 0; j < NB; ++j)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int jj = j0 + wj + 16 * j + 4 * g4 + v;
                        if (jj < N2) Cz[ro + (int64_t)jj * cj] = c64{accr[i][j][v], acci[i][j][v]};
                    }
            }
        }
        __syncthreads();  // T and the B tile are rewritten by the next unit
    }
}

// m == 1: each matrix is a vector v (len = p*Nt); SVST = v * max(|v| - beta, 0)/|v|.
// nuc (nullable): the output's singular value max(|v| - beta, 0) per matrix.
__global__ __launch_bounds__(256) void k_svst_m1(c64 *__restrict__ P, int64_t len, int64_t Np,
                                                 const double *__restrict__ beta,
                                                 int beta_per_patch, double *__restrict__ nuc) {
    const int lane = threadIdx.x & 63;
    const int64_t ip = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (ip >= Np) return;
    c64 *v = P + ip * len;
    double s = 0.0;
    for (int64_t i = lane; i < len; i += 64) {
        const c64 a = v[i];
        s += (double)a.re * a.re + (double)a.im * a.im;
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    const double bd = (double)(float)beta[beta_per_patch ? ip : 0];
    const double sg = sqrt(s);
    if (nuc && lane == 0) nuc[ip] = fmax(sg - bd, 0.0);
    if (bd == 0.0) return;
    const float fk = sg > bd ? (float)((sg - bd) / sg) : 0.f;
    for (int64_t i = lane; i < len; i += 64) {
        const c64 a = v[i];
        v[i] = c64{a.re * fk, a.im * fk};
    }
}

// Sum of sqrt(max(lambda,0)) over all eigenvalues of the symmetric tridiagonal
// (d, e) of order m (the nuclear norm of a matrix whose Gram matrix it is), by
// bisection on Sturm counts; every thread of the workgroup calls it, the sum is
// valid in thread 0.
__device__ double tri_sqrt_eig_sum(const double *__restrict__ d, const double *__restrict__ e,
                                   int m) {
    __shared__ double red[16];
    __shared__ double s_gl, s_gu, s_pivmin;
    if (threadIdx.x == 0) {
        double emax = 0.0, gl = INFINITY, gu = -INFINITY;
        for (int i = 0; i < m; ++i) {
            const double el = i > 0 ? fabs(e[i - 1]) : 0.0;
            const double er = i + 1 < m ? fabs(e[i]) : 0.0;
            gl = fmin(gl, d[i] - el - er);
            gu = fmax(gu, d[i] + el + er);
            if (i + 1 < m) emax = fmax(emax, e[i] * e[i]);
        }
        s_pivmin = DBL_MIN * fmax(1.0, emax);
        const double pad = 2.0 * DBL_EPSILON * fmax(fabs(gl), fabs(gu)) + s_pivmin;
        s_gl = gl - pad;
        s_gu = gu + pad;
    }
    __syncthreads();
    double acc = 0.0;
    for (int k = threadIdx.x; k < m; k += blockDim.x) {
        double lo = s_gl, hi = s_gu;
        for (int it = 0; it < 200; ++it) {
            const double mid = 0.5 * (lo + hi);
            if (hi - lo <= 2.0 * DBL_EPSILON * fmax(fabs(lo), fabs(hi)) + s_pivmin) break;
            if (mid <= lo || mid >= hi) break;
            if (sturm_count(d, e, m, mid, s_pivmin) > k)
                hi = mid;
            else
                lo = mid;
        }
        acc += sqrt(fmax(0.5 * (lo + hi), 0.0));
    }
    return block_sum(acc, red);
}

__global__ __launch_bounds__(256) void k_nucnorm(const double *__restrict__ dall,
                                                 const double *__restrict__ eall, int m,
                                                 double *__restrict__ out) {
    const int ip = blockIdx.x;
    const double s = tri_sqrt_eig_sum(dall + (int64_t)ip * m, eall + (int64_t)ip * m, m);
    if (threadIdx.x == 0) out[ip] = s;
}

// Nuclear norm of each SVST output, sum_k max(sigma_k - beta, 0), from the
// eigenvalues the thresholding found: lam[k] (k < r) are the r Gram eigenvalues
// above beta^2.  r < 0 (beta = 0: the output is the input) takes all
// eigenvalues of the tridiagonal by bisection.  One workgroup per patch.
__global__ __launch_bounds__(64) void k_out_nucnorm(const double *__restrict__ dall,
                                                    const double *__restrict__ eall, int m,
                                                    const double *__restrict__ beta,
                                                    int beta_per_patch,
                                                    const int *__restrict__ rall,
                                                    const double *__restrict__ lamall,
                                                    double *__restrict__ out) {
    __shared__ double red[16];
    const int ip = blockIdx.x;
    const int r = rall[ip];
    if (r < 0) {
        const double s = tri_sqrt_eig_sum(dall + (int64_t)ip * m, eall + (int64_t)ip * m, m);
        if (threadIdx.x == 0) out[ip] = s;
        return;
    }
    const double bd = (double)(float)beta[beta_per_patch ? ip : 0];
    const double *lam = lamall + (int64_t)ip * m;
    double acc = 0.0;
    for (int k = threadIdx.x; k < r; k += blockDim.x) acc += fmax(sqrt(fmax(lam[k], 0.0)) - bd, 0.0);
    const double s = block_sum(acc, red);
    if (threadIdx.x == 0) out[ip] = s;
}

__global__ void k_m1_nucnorm(const c64 *__restrict__ P, int64_t len, int64_t Np,
                             double *__restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int64_t ip = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (ip >= Np) return;
    const c64 *v = P + ip * len;
    double s = 0.0;
    for (int64_t i = lane; i < len; i += 64) {
        const c64 a = v[i];
        s += (double)a.re * a.re + (double)a.im * a.im;
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) out[ip] = sqrt(s);
}

// ------------------------------------------------------------------ host driver

struct SvstWork {
    int64_t m, nkc, chunk;
    int kch;    // contraction rows per Gram workgroup
    c128 *VW;   // panel V, W of the blocked tridiagonalisation
    void *part;  // partial Gram tiles, C2<AT>
    c128 *G;
    double *d, *e, *lam, *Z, *ws;
    c128 *tau;
    c64 *U;
    float *f;
    int *r;
    double *info;  // EIG_INFO doubles per patch
    bool sbr;      // two-stage tridiagonalisation (sbr.hip)
    SbrWork sw;
};

// Tridiagonalisation: the two-stage path (sbr.hip) where the size allows;
// MSLR_SVST_REDUCTION = "one" selects the one-stage path everywhere.
static bool use_sbr(int64_t m) {
    const char *e = getenv("MSLR_SVST_REDUCTION");
    return sbr_eligible(m) && !(e != nullptr && strcmp(e, "one") == 0);
}

static size_t svst_bytes_per_patch(int64_t m, int64_t nkc) {
    return (size_t)(2 * TNB * m * 16 + nkc * m * m * 16 + m * m * 16 + m * 8 * 3 + m * 16 + m * m * 8 +
                    (int64_t)WS_PER * m * 8 + m * m * 8 + m * 4 + 16 + EIG_INFO * 8) +
           (use_sbr(m) ? sbr_bytes_per_matrix((int)m) : 0);
}

static SvstWork svst_work(mslr_ctx_s *ctx, int64_t m, int64_t Kc, int64_t Np) {
    SvstWork w{};
    w.m = m;
    w.kch = ctx->tuning.gram_chunk > 0 ? ctx->tuning.gram_chunk : KCH;
    w.nkc = cdiv(Kc, w.kch);
    // patches per chunk from a 16 GB workspace (per stream): the 3179 m = 387
    // patches of a 10^3 scale at Nt = 387 in one chunk instead of nine (each chunk's
    // launches end in a partly idle round)
    size_t budget = (size_t)16 << 30;
    if (const char *s = getenv("MSLR_SVST_WORK_MB")) {
        long mb = atol(s);
        if (mb > 0) budget = (size_t)mb << 20;
    }
    const size_t per = svst_bytes_per_patch(m, w.nkc);
    int64_t chunk = (int64_t)(budget / per);
    chunk = std::max<int64_t>(1, std::min<int64_t>(chunk, Np));
    chunk = std::min<int64_t>(chunk, 65535);
    w.chunk = chunk;
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const int64_t mm = m * m;
    size_t off = 0;
    size_t o_part = off; off += al((size_t)chunk * w.nkc * mm * 16);
    size_t o_G = off; off += al((size_t)chunk * mm * 16);
    size_t o_d = off; off += al((size_t)chunk * m * 8);
    size_t o_e = off; off += al((size_t)chunk * m * 8);
    size_t o_lam = off; off += al((size_t)chunk * m * 8);
    size_t o_tau = off; off += al((size_t)chunk * m * 16);
    size_t o_Z = off; off += al((size_t)chunk * mm * 8);
    size_t o_ws = off; off += al((size_t)chunk * WS_PER * m * 8);
    size_t o_U = off; off += al((size_t)chunk * mm * 8);
    size_t o_f = off; off += al((size_t)chunk * m * 4);
    size_t o_r = off; off += al((size_t)chunk * 4);
    size_t o_VW = off; off += al((size_t)chunk * 2 * TNB * m * 16);
    size_t o_info = off; off += al((size_t)chunk * EIG_INFO * 8);
    w.sbr = use_sbr(m);
    size_t o_T1 = off, o_V2 = off, o_tau2 = off, o_W1 = off, o_band = off;
    if (w.sbr) {
        const int mi = (int)m;
        o_T1 = off; off += al((size_t)chunk * sbr_panels(mi) * SBR_B * SBR_B * 16);
        o_V2 = off; off += al((size_t)chunk * m * sbr_steps(mi) * SBR_B * 16);
        o_tau2 = off; off += al((size_t)chunk * m * sbr_steps(mi) * 16);
        o_W1 = off; off += al((size_t)chunk * sbr_w1_elems(mi) * 16);
        if (sbr_global_band(m)) {
            o_band = off; off += al((size_t)chunk * 2 * SBR_B * m * 16);
        }
    }
    char *base = (char *)ctx->scratch_get(SLOT_GRAM, off);
    w.part = (void *)(base + o_part);
    w.G = (c128 *)(base + o_G);
    w.d = (double *)(base + o_d);
    w.e = (double *)(base + o_e);
    w.lam = (double *)(base + o_lam);
    w.tau = (c128 *)(base + o_tau);
    w.Z = (double *)(base + o_Z);
    w.ws = (double *)(base + o_ws);
    w.U = (c64 *)(base + o_U);
    w.f = (float *)(base + o_f);
    w.r = (int *)(base + o_r);
    w.VW = (c128 *)(base + o_VW);
    w.info = (double *)(base + o_info);
    if (w.sbr) {
        w.sw.T1 = (c128 *)(base + o_T1);
        w.sw.V2 = (c128 *)(base + o_V2);
        w.sw.tau2 = (c128 *)(base + o_tau2);
        w.sw.W1 = (c128 *)(base + o_W1);
        if (sbr_global_band(m)) w.sw.Band = (c128 *)(base + o_band);
    }
    return w;
}

template <class AT>
static void gram_and_tridiag(mslr_ctx_s *ctx, const SvstWork &w, const c64 *P, int64_t p,
                             int64_t Nt, int64_t nb, int *err_dev, const unsigned *Wh = nullptr,
                             const unsigned *Wl = nullptr) {
    const int m = (int)w.m;
    const bool left = p < Nt;
    const int64_t Kc = left ? Nt : p;
    const int64_t ks = left ? p : 1;
    const int64_t cs = left ? 1 : p;
    const int ntile = (int)cdiv(m, GT);
    const int npair = ntile * (ntile + 1) / 2;
    dim3 gg((unsigned)npair, (unsigned)w.nkc, (unsigned)nb);
    {
        MSLR_PROF(ctx, "gram_tiles");
        if constexpr (std::is_same<AT, double>::value) {
            // one contraction chunk: the tiles write the Hermitian matrices directly
            c128 *Gd = w.nkc == 1 ? w.G : nullptr;
            if (Wh)  // split words, contraction-contiguous (k_split_rows / k_split_cols)
                k_gram_bf16<<<gg, GB_T, 0, ctx->stream>>>(Wh, Wl, p * Nt, Kc, m, Kc, (int)w.nkc, w.kch,
                                                         ntile, (Kc % 4) == 0 ? 1 : 0,
                                                         (c128 *)w.part, Gd, left ? 1 : 0, err_dev);
            else
                k_gram_mfma<<<gg, 256, 0, ctx->stream>>>(P, p * Nt, ks, cs, m, Kc, (int)w.nkc, w.kch,
                                                         ntile, (c128 *)w.part, Gd, left ? 1 : 0,
                                                         err_dev);
        }
        else
            k_gram_tiles<AT><<<gg, 256, 0, ctx->stream>>>(P, p * Nt, ks, cs, m, Kc, (int)w.nkc, w.kch,
                                                          ntile, (C2<AT> *)w.part);
        HIP_LAUNCH_CHECK();
    }
    const int64_t ne = (int64_t)m * m * nb;
    const unsigned gr = (unsigned)std::min<int64_t>(cdiv(ne, 256), 1 << 20);
    if (!(std::is_same<AT, double>::value && w.nkc == 1)) {
        MSLR_PROF(ctx, "gram_reduce");
        k_gram_reduce<AT><<<gr, 256, 0, ctx->stream>>>((const C2<AT> *)w.part, m, (int)w.nkc, (int)nb,
                                                       left ? 1 : 0, w.G, err_dev);
        HIP_LAUNCH_CHECK();
    }
    if (w.sbr) {
        sbr_reduce(ctx, w.G, m, nb, w.sw, w.d, w.e);
    } else {
        MSLR_PROF(ctx, "tridiag");
        k_tridiag_blk<<<(unsigned)nb, TNT, 0, ctx->stream>>>(w.G, m, w.d, w.e, w.tau, w.VW);
        HIP_LAUNCH_CHECK();
    }
}

static void launch_backtransform(mslr_ctx_s *ctx, const SvstWork &w, int m, int64_t nb) {
    const size_t lds = (size_t)BTR * m * sizeof(c128);
    MSLR_PROF(ctx, "backtransform");
#define BT_LAUNCH(G_)                                                                          \
    {                                                                                          \
        static bool attr = false;                                                              \
        if (!attr) {                                                                           \
            HIP_CHECK(hipFuncSetAttribute((const void *)k_backtransform<G_>,                   \
                                          hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024)); \
            attr = true;                                                                       \
        }                                                                                      \
        dim3 g_((unsigned)nb, (unsigned)cdiv(m, 256 / G_));                                    \
        k_backtransform<G_><<<g_, 256, lds, ctx->stream>>>(w.G, w.tau, w.Z, w.r, m, w.U);       \
    }
    if (m <= 8 * BT_MAXE) BT_LAUNCH(8)
    else if (m <= 16 * BT_MAXE) BT_LAUNCH(16)
    else BT_LAUNCH(32)
#undef BT_LAUNCH
    HIP_LAUNCH_CHECK();
}

static RankSel rank_sel_at(RankSel rs, int64_t z0) {
    if (rs.r) rs.r += z0;
    return rs;
}

static void cgemm(mslr_ctx_s *ctx, const char *tag, const c64 *A, int64_t bA, int64_t ai,
                  int64_t ak, const c64 *B, int64_t bB, int64_t bk, int64_t bj, int conjB,
                  const float *fs, int64_t bfs, c64 *C, int64_t bC, int64_t ci, int64_t cj, int M,
                  int N, int Kfix, const int *kvar, int64_t nbatch, RankSel rs = RankSel{},
                  int accum = 0) {
    for (int64_t z0 = 0; z0 < nbatch; z0 += 65535) {
        const int64_t nz = std::min<int64_t>(65535, nbatch - z0);
        dim3 g((unsigned)cdiv(N, MT), (unsigned)cdiv(M, MT), (unsigned)nz);
        MSLR_PROF(ctx, tag);
        k_cgemm<<<g, 256, 0, ctx->stream>>>(A + z0 * bA, bA, ai, ak, B + z0 * bB, bB, bk, bj, conjB,
                                            fs ? fs + z0 * bfs : nullptr, bfs, C + z0 * bC, bC, ci,
                                            cj, M, N, Kfix, kvar ? kvar + z0 : nullptr,
                                            rank_sel_at(rs, z0), accum);
        HIP_LAUNCH_CHECK();
    }
}

static void cgemm_bf16(mslr_ctx_s *ctx, const char *tag, const c64 *A, int64_t bA, int64_t ai,
                       int64_t ak, const c64 *B, int64_t bB, int64_t bk, int64_t bj, int conjB,
                       const float *fs, int64_t bfs, c64 *C, int64_t bC, int64_t ci, int64_t cj,
                       int M, int N, int K, int64_t nbatch, RankSel rs = RankSel{}) {
    for (int64_t z0 = 0; z0 < nbatch; z0 += 65535) {
        const int64_t nz = std::min<int64_t>(65535, nbatch - z0);
        dim3 g((unsigned)cdiv(N, MT), (unsigned)cdiv(M, MT), (unsigned)nz);
        MSLR_PROF(ctx, tag);
        if (fs || conjB)
            k_cgemm_bf16<true><<<g, 256, 0, ctx->stream>>>(
                A + z0 * bA, bA, ai, ak, B + z0 * bB, bB, bk, bj, conjB,
                fs ? fs + z0 * bfs : nullptr, bfs, C + z0 * bC, bC, ci, cj, M, N, K,
                rank_sel_at(rs, z0));
        else
            k_cgemm_bf16<false><<<g, 256, 0, ctx->stream>>>(A + z0 * bA, bA, ai, ak, B + z0 * bB,
                                                            bB, bk, bj, 0, nullptr, 0, C + z0 * bC,
                                                            bC, ci, cj, M, N, K,
                                                            rank_sel_at(rs, z0));
        HIP_LAUNCH_CHECK();
    }
}

// C = A B' in split BF16 for row-contiguous A and C (ai = ci = 1): B' = (fs[k] B(k, j),
// conjugated when conjB) split once into hi / lo words, then k_apply_bf16.
// scatter (nullable): C is the image of an exact tiling's patches, batch entry z
// being patch sip0 + z (cj = V, the frame stride).
// Aw (nullable): A as Ozaki words (s1 at Aw, s2 at Aw + Awlo, A's element offsets)
// instead of the FP32 A.
static void apply_bf16(mslr_ctx_s *ctx, const char *tag, const c64 *A, int64_t bA, int64_t ak,
                       const c64 *B, int64_t bB, int64_t bk, int64_t bj, int conjB, const float *fs,
                       int64_t bfs, c64 *C, int64_t bC, int64_t cj, int M, int N, int K,
                       int64_t nbatch, RankSel rs, const PatchGeom *scatter = nullptr,
                       int64_t sip0 = 0, const unsigned *Aw = nullptr, int64_t Awlo = 0) {
    if (M <= 0 || N <= 0 || K <= 0 || nbatch <= 0) return;
    const int Kp = (int)cdiv(K, BKS) * BKS;
    const int64_t bW = (int64_t)N * Kp;
    const int64_t zc = std::min<int64_t>(nbatch, 65535);
    unsigned *Wh = (unsigned *)ctx->scratch_get(SLOT_BSPLIT, (size_t)(2 * zc * bW) * sizeof(unsigned));
    unsigned *Wl = Wh + zc * bW;
    MSLR_PROF(ctx, tag);
    for (int64_t z0 = 0; z0 < nbatch; z0 += zc) {
        const int nz = (int)std::min<int64_t>(zc, nbatch - z0);
        const RankSel rz = rank_sel_at(rs, z0);
        dim3 gs(bsplit_groups(bW, nz), (unsigned)nz);
        k_apply_bsplit<<<gs, 256, 0, ctx->stream>>>(B + z0 * bB, bB, bk, bj, conjB,
                                                    fs ? fs + z0 * bfs : nullptr, bfs, N, K, Kp, rz,
                                                    Wh, Wl, bW);
        HIP_LAUNCH_CHECK();
        const int64_t units = (int64_t)nz * cdiv(M, AP_ROWS);
        const unsigned g = (unsigned)std::min<int64_t>(units, 4 * (int64_t)std::max(ctx->num_cus, 1));
        const unsigned *Awz = Aw ? Aw + z0 * bA : nullptr;
        // (with Aw the kernel reads no FP32 A; any valid address stands in)
        const c64 *Az = A ? A + z0 * bA : reinterpret_cast<const c64 *>(Awz);
        c64 *Cz = scatter ? C : C + z0 * bC;
        const PatchGeom sg = scatter ? *scatter : PatchGeom{};
        const int64_t s0 = sip0 + z0;
        const int sc = scatter ? 1 : 0;
        auto go = [&](auto kern) {
            kern<<<g, 256, 0, ctx->stream>>>(Az, Awz, Awlo, bA, ak, Wh, Wl, bW, Kp, Cz, bC, cj, M, N,
                                             K, nz, rz, sg, s0, sc);
        };
        // column tiles per pass: five for the dense products (K = m: the A tile is
        // staged once per k step for 160 columns), at most two for the low-rank ones
        // (N or K narrowed to r per patch: the smaller tile keeps more waves resident;
        // c1 scale 2: X U_r 0.42 -> 0.35 ms, T f U_r^H 0.77 -> 0.67 ms, while the
        // whole-image X W takes 1.70 ms with two against 1.27 with five)
        const int nbt = rs.mode >= 2 ? (N > 32 ? 2 : 1) : (N > 64 ? 5 : (N > 32 ? 2 : 1));
        if (Aw) {
            if (nbt == 5) go(k_apply_bf16<5, true>);
            else if (nbt == 2) go(k_apply_bf16<2, true>);
            else go(k_apply_bf16<1, true>);
        } else {
            if (nbt == 5) go(k_apply_bf16<5, false>);
            else if (nbt == 2) go(k_apply_bf16<2, false>);
            else go(k_apply_bf16<1, false>);
        }
        HIP_LAUNCH_CHECK();
    }
}

// Y = (X U_r) diag(f) U_r^H for the patches keeping r <= LR_MAX (k_apply_lr_bf16),
// X as the split pass's words (Aw, s2 at Aw + Awlo), U (m x m, batch stride bU), f
// (batch stride bf); Y as apply_bf16 writes it (scatter: into the image).  The
// patches keeping more go through the two-product path (RankSel lim = LR_MAX).
static void apply_lowrank_fused(mslr_ctx_s *ctx, const unsigned *Aw, int64_t Awlo, int64_t bA,
                                const c64 *U, int64_t bU, const float *f, int64_t bf, c64 *C,
                                int64_t bC, int64_t cj, int M, int m, int nbatch, const int *r,
                                const PatchGeom *scatter, int64_t sip0) {
    const int half = m / 2;
    if (M <= 0 || half < 1 || nbatch <= 0) return;
    const int Kp1 = (int)cdiv(m, BKS) * BKS, Kp2 = (int)cdiv(half, BKS) * BKS;
    const int64_t bW1 = (int64_t)half * Kp1, bW2 = (int64_t)m * Kp2;
    const int64_t zc = std::min<int64_t>(nbatch, 65535);
    unsigned *W = (unsigned *)ctx->scratch_get(SLOT_BSPLIT_LR,
                                               (size_t)(2 * zc * (bW1 + bW2)) * sizeof(unsigned));
    unsigned *W1h = W, *W1l = W + zc * bW1, *W2h = W + 2 * zc * bW1, *W2l = W2h + zc * bW2;
    MSLR_PROF(ctx, "apply_lr");
    for (int64_t z0 = 0; z0 < nbatch; z0 += zc) {
        const int nz = (int)std::min<int64_t>(zc, nbatch - z0);
        // the same words as the two-product path's splits (modes 2 and 3 of its RankSel)
        const RankSel s1{r + z0, m, 2}, s2{r + z0, m, 3};
        k_apply_bsplit<<<dim3(bsplit_groups(bW1, nz), (unsigned)nz), 256, 0, ctx->stream>>>(
            U + z0 * bU, bU, 1, m, 0, nullptr, 0, half, m, Kp1, s1, W1h, W1l, bW1);
        HIP_LAUNCH_CHECK();
        k_apply_bsplit<<<dim3(bsplit_groups(bW2, nz), (unsigned)nz), 256, 0, ctx->stream>>>(
            U + z0 * bU, bU, m, 1, 1, f + z0 * bf, bf, m, half, Kp2, s2, W2h, W2l, bW2);
        HIP_LAUNCH_CHECK();
        const int64_t units = (int64_t)nz * cdiv(M, AP_ROWS);
        const unsigned g = (unsigned)std::min<int64_t>(units, 4 * (int64_t)std::max(ctx->num_cus, 1));
        const PatchGeom sg = scatter ? *scatter : PatchGeom{};
        k_apply_lr_bf16<<<g, 256, 0, ctx->stream>>>(
            Aw + z0 * bA, Awlo, bA, W1h, W1l, bW1, Kp1, W2h, W2l, bW2, Kp2,
            scatter ? C : C + z0 * bC, bC, cj, M, m, m, nz, r + z0, m, sg, sip0 + z0,
            scatter ? 1 : 0);
        HIP_LAUNCH_CHECK();
    }
}

// Reconstruction products that the split-BF16 form runs in FP32 on the matrix
// cores (bit 0 the dense X W, bit 1 X U_r, bit 2 T f U_r^H).
constexpr int SVST_APPLY_FP32 = 0;

// MSLR_SVST_APPLY=dense turns the low-rank apply off (every patch forms W).
static bool svst_lowrank_apply() {
    static const bool on = [] {
        const char *e = std::getenv("MSLR_SVST_APPLY");
        return !(e && std::string(e) == "dense");
    }();
    return on;
}

// The products writing the output (bits 0 and 2 of apply_fp32) stay split BF16.
static int apply_fp32_mask(const mslr_ctx_s *ctx) {
    return ctx->tuning.apply_fp32 > 0 ? ctx->tuning.apply_fp32 - 1 : SVST_APPLY_FP32;
}

bool svst_writes_direct(const mslr_ctx_s *ctx, int64_t p, int64_t Nt) {
    return ctx->products == MSLR_PRODUCTS_SPLIT_BF16 && p >= Nt && Nt > 1 &&
           (apply_fp32_mask(ctx) & 5) == 0;
}

bool svst_reads_words(const mslr_ctx_s *ctx, int64_t p, int64_t Nt) {
    return ctx->products == MSLR_PRODUCTS_SPLIT_BF16 && p >= Nt && Nt > 1 &&
           (apply_fp32_mask(ctx) & 3) == 0;
}

void svst_batched(mslr_ctx_s *ctx, const c64 *Pin, c64 *Pout, int64_t p, int64_t Nt, int64_t Np,
                  const double *beta_dev, int per_patch, int *err_dev, double *nuc_out,
                  const PatchGeom *scatter, const PatchGeom *gather, const unsigned *words) {
    if (Np == 0 || p == 0 || Nt == 0) return;
    const int64_t m = std::min(p, Nt);
    MSLR_REQUIRE(m <= MAXM, MSLR_ERR_UNSUPPORTED,
                 "min(patch voxels, Nt) must be <= " + std::to_string(MAXM));
    const bool direct = svst_writes_direct(ctx, p, Nt);
    const bool in_place = Pin == Pout;
    MSLR_REQUIRE(!in_place || direct, MSLR_ERR_ARG, "svst_batched needs distinct input and output");
    MSLR_REQUIRE(!scatter || (direct && scatter->p == p && scatter->Nt == Nt &&
                              scatter->exact_tiling() && scatter->Np == Np),
                 MSLR_ERR_ARG, "svst_batched scatters only an exact tiling's patches");
    MSLR_REQUIRE(!gather || (direct && gather->p == p && gather->Nt == Nt && gather->Np == Np),
                 MSLR_ERR_ARG, "svst_batched gathers patches only for split products with p >= Nt");
    // With split products and p >= Nt, the products that read X (the dense X W and
    // X U_r) take the split pass's Ozaki words as their left operand (x = s1 + s2 to
    // 2^-17 of the block maximum) when both run split; else they read an FP32 copy
    // of the patches, written by the split pass, where Y overwrites X.
    const bool words_a = svst_reads_words(ctx, p, Nt);
    MSLR_REQUIRE(!words || (words_a && !gather), MSLR_ERR_ARG,
                 "svst_batched takes split words only where every product reads them");
    const bool copy = (in_place || gather) && !words_a;
    const int64_t len = p * Nt;
    if (m == 1) {
        HIP_CHECK(hipMemcpyAsync(Pout, Pin, (size_t)(len * Np) * sizeof(c64),
                                 hipMemcpyDeviceToDevice, ctx->stream));
        const unsigned nb = (unsigned)cdiv(Np, 4);
        MSLR_PROF(ctx, "svst_m1");
        k_svst_m1<<<nb, 256, 0, ctx->stream>>>(Pout, len, Np, beta_dev, per_patch, nuc_out);
        HIP_LAUNCH_CHECK();
        return;
    }
    const bool left = p < Nt;
    const SvstWork w = svst_work(ctx, m, left ? Nt : p, Np);
    MSLR_REQUIRE(!words || w.chunk >= Np, MSLR_ERR_UNSUPPORTED,
                 "split words handed in need the batch in one workspace chunk");
    // split-BF16 Gram and apply products (the mixed-precision form)
    const bool split = ctx->products == MSLR_PRODUCTS_SPLIT_BF16;
    // W = V_r diag(f) V_r^H per patch (m x m, FP32), reuses the Gram partial space
    c64 *Wm = (c64 *)w.part;
    for (int64_t ip0 = 0; ip0 < Np; ip0 += w.chunk) {
        const int64_t nb = std::min<int64_t>(w.chunk, Np - ip0);
        const c64 *Pc = gather ? nullptr : Pin + ip0 * len;  // (gathered: the copy below)
        c64 *Yc = scatter ? Pout : Pout + ip0 * len;
        // in place: the reconstruction products read the split pass's FP32 copy
        c64 *Pcopy = copy ? (c64 *)ctx->scratch_get(SLOT_SVST_COPY, (size_t)(len * nb) * sizeof(c64))
                          : nullptr;
        // split products: the chunk's patches once as split-BF16 words for the Gram
        // (and, with words_a, for the products that read X)
        const unsigned *Wh = nullptr, *Wl = nullptr;
        if (words) {
            Wh = words;
            Wl = words + len * nb;
        } else if (split) {
            unsigned *wb = (unsigned *)ctx->scratch_get(SLOT_SPLIT, (size_t)len * nb * 8);
            MSLR_PROF(ctx, "split");
            if (!left) {
                const int64_t nw = cdiv(p, SPL_ROWS) * Nt * nb;
                if (gather) {
                    const int64_t ng = cdiv(p, SPL_ROWS) * cdiv(Nt, SG_T) * nb;
                    k_split_gather<<<(unsigned)cdiv(ng, 4), 256, 0, ctx->stream>>>(
                        Pin, *gather, ip0, nb, wb, wb + len * nb, Pcopy);
                } else {
                    k_split_rows<<<(unsigned)cdiv(nw, 4), 256, 0, ctx->stream>>>(
                        Pc, p, Nt * nb, wb, wb + len * nb, Pcopy);
                }
            } else {
                const int64_t nw = cdiv(Nt, SPL_ROWS) * p * nb;
                k_split_cols<<<(unsigned)cdiv(nw, 4), 256, 0, ctx->stream>>>(Pc, p, Nt, nb, wb,
                                                                             wb + len * nb);
            }
            HIP_LAUNCH_CHECK();
            Wh = wb;
            Wl = wb + len * nb;
        }
        // FP64 products: the Gram squares the dynamic range of X, and at production
        // shapes (global patch s1/s_noise ~ 1e4) FP32 products would bury the
        // singular values near the threshold under rounding error.
        gram_and_tridiag<double>(ctx, w, Pc, p, Nt, nb, err_dev, Wh, Wl);
        {
            const double *bp = beta_dev + (per_patch ? ip0 : 0);
            MSLR_PROF(ctx, "eigvec");
            const int hi = svst_small_batch(nb);
            if (hi)
                k_eig_count_wave<<<(unsigned)nb, 64, 2 * m * sizeof(double), ctx->stream>>>(
                    w.d, w.e, (int)m, bp, per_patch, w.info, w.r, 0, hi);
            else
                k_eig_count<<<(unsigned)cdiv(nb, 64), 64, 0, ctx->stream>>>(w.d, w.e, (int)m, nb, bp,
                                                                             per_patch, w.info, w.r, 0);
            HIP_LAUNCH_CHECK();
            // lanes per eigenvalue: many when the eigenvalues are few (the bisection
            // latency is the cost), few when the batch already fills the device.
            // (Splitting a patch's eigenvalues over a 16-lane and a 4-lane launch
            // serialises the two latencies for patches keeping more than the first
            // launch covers; measured slower beside the other scales' kernels.)
            const int64_t lim = ctx->tuning.eig_lanes_limit == 0 ? 64 * 1024
                                                                   : ctx->tuning.eig_lanes_limit;
            if (nb * m <= lim) {
                dim3 ge((unsigned)nb, (unsigned)cdiv(m * 16, 64));
                k_eig_vectors<16><<<ge, 64, 3 * m * sizeof(double), ctx->stream>>>(w.d, w.e, (int)m, bp, per_patch,
                                                              w.info, w.r, w.lam, w.Z, w.U, 0, hi);
            } else {
                dim3 ge((unsigned)nb, (unsigned)cdiv(m * 4, 64));
                k_eig_vectors<4><<<ge, 64, 3 * m * sizeof(double), ctx->stream>>>(w.d, w.e, (int)m, bp, per_patch,
                                                             w.info, w.r, w.lam, w.Z, w.U, 0, hi);
            }
            HIP_LAUNCH_CHECK();
            k_eig_clusters<<<(unsigned)nb, NSLOT, 0, ctx->stream>>>(w.d, w.e, (int)m, bp, per_patch,
                                                                    w.info, w.r, w.lam, w.Z, w.ws,
                                                                    w.f);
            HIP_LAUNCH_CHECK();
            if (nuc_out) {
                MSLR_PROF(ctx, "out_nucnorm");
                k_out_nucnorm<<<(unsigned)nb, 64, 0, ctx->stream>>>(w.d, w.e, (int)m, bp, per_patch,
                                                                    w.r, w.lam, nuc_out + ip0);
                HIP_LAUNCH_CHECK();
            }
        }
        if (w.sbr)
            sbr_backtransform(ctx, w.G, (int)m, nb, w.sw, w.Z, w.r, w.U);
        else
            launch_backtransform(ctx, w, (int)m, nb);
        const int mi = (int)m;
        // p >= Nt: patches keeping at most half their singular values take the
        // low-rank form Y = (X U_r) diag(f) U_r^H (4 p m r flops instead of
        // 2 p m^2 + m^2 r); the others W = U_r diag(f) U_r^H, then Y = X W.  The
        // choice is per patch, on the device, from r.
        const int half = mi / 2;
        const bool lowrank = !left && half >= 1 && ctx->tuning.svst_apply != 1 && svst_lowrank_apply();
        const RankSel dense = lowrank ? RankSel{w.r, mi, 1} : RankSel{};
        // W = (U diag f) U^H ; r < 0 (beta = 0) gives the identity, r = 0 gives 0
        cgemm(ctx, "form_w", w.U, m * m, 1, m, w.U, m * m, m, 1, 1, w.f, m, Wm, m * m, 1, m, mi, mi,
              0, w.r, nb, dense);
        // split-BF16 or FP32 matrix-core products per reconstruction product
        // (apply_fp32: 1 + mask of bit 0 dense, bit 1 X U_r, bit 2 T f U_r^H).  With
        // p >= Nt the right operands (W, U_r, f U_r^H) are m x m at most and are split
        // once (apply_bf16); W X (p < Nt) would split the patches themselves, so it
        // splits per tile (cgemm_bf16).
        const int amask = apply_fp32_mask(ctx);
        // out: a product writing Y (scattered into the image with scatter)
        // xw: the left operand is X, taken as the split words when words_a
        auto prod = [&](int bit, bool out, const char *tag, const c64 *A, int64_t bA, int64_t ai,
                        int64_t ak, const c64 *B, int64_t bB, int64_t bk, int64_t bj, int conjB,
                        const float *fs, int64_t bfs, c64 *C, int64_t bC, int64_t ci, int64_t cj,
                        int M, int N, int K, RankSel rs, bool xw = false) {
            if (split && !(amask & bit) && !left) {
                const bool sc = out && scatter;
                const bool aw = xw && words_a;
                apply_bf16(ctx, tag, A, bA, ak, B, bB, bk, bj, conjB, fs, bfs, C, bC,
                           sc ? (int64_t)scatter->N[0] * scatter->N[1] * scatter->N[2] : cj, M, N,
                           K, nb, rs, sc ? scatter : nullptr, ip0, aw ? Wh : nullptr,
                           aw ? len * nb : 0);
            } else if (split && !(amask & bit)) {
                cgemm_bf16(ctx, tag, A, bA, ai, ak, B, bB, bk, bj, conjB, fs, bfs, C, bC, ci, cj, M,
                           N, K, nb, rs);
            } else {
                cgemm(ctx, tag, A, bA, ai, ak, B, bB, bk, bj, conjB, fs, bfs, C, bC, ci, cj, M, N, K,
                      nullptr, nb, rs);
            }
        };
        const c64 *Pa = copy ? Pcopy : Pc;  // the products' left operand
        if (!left) {
            // Y (p x m) = X (p x m) W
            prod(1, true, "apply", Pa, len, 1, p, Wm, m * m, 1, m, 0, nullptr, 0, Yc, len, 1, p,
                 (int)p, mi, mi, dense, true);
        } else {
            // Y (p x Nt) = W (p x p) X
            prod(1, true, "apply", Wm, m * m, 1, m, Pa, len, 1, p, 0, nullptr, 0, Yc, len, 1, p,
                 (int)p, (int)Nt, mi, RankSel{});
        }
        if (lowrank) {
            // T (p x r) = X U_r, then Y = (T diag f) U_r^H; T holds at most m/2 columns
            const int64_t bT = p * half;
            c64 *T = (c64 *)ctx->scratch_get(SLOT_LOWRANK, (size_t)(bT * nb) * sizeof(c64));
            RankSel s1{w.r, mi, 2}, s2{w.r, mi, 3};
            if (split && !(amask & 6) && words_a && ctx->tuning.svst_apply == 0) {
                // r <= LR_MAX in one kernel, T kept on chip; the rest below
                apply_lowrank_fused(ctx, Wh, len * nb, len, w.U, m * m, w.f, m, Yc, len,
                                    scatter ? (int64_t)scatter->N[0] * scatter->N[1] * scatter->N[2]
                                            : (int64_t)p,
                                    (int)p, mi, (int)nb, w.r, scatter, ip0);
                s1.lim = s2.lim = LR_MAX;
            }
            prod(2, false, "apply_xu", Pa, len, 1, p, w.U, m * m, 1, m, 0, nullptr, 0, T, bT, 1, p,
                 (int)p, half, mi, s1, true);
            prod(4, true, "apply_tu", T, bT, 1, p, w.U, m * m, m, 1, 1, w.f, m, Yc, len, 1, p,
                 (int)p, mi, half, s2);
        }
    }
}

// ------------------------------------------------------------------ derivative
// Vector-Jacobian product of Y = SVST(Z, beta) = Z g(G), G = Z^H Z = V diag(l) V^H,
// g(l) = max(1 - beta / sqrt(l), 0), for p >= Nt.  With Wb = Z^H Ybar,
// S = (Wb + Wb^H) / 2, M = V^H S V and the divided differences D of g:
//   Zbar = Ybar W + 2 Z V (D o M) V^H,   betabar = -sum_{s_k > beta} M_kk / s_k.

// S = (Wb + Wb^H) / 2 per m x m matrix.
__global__ __launch_bounds__(256) void k_vjp_sym(const c64 *__restrict__ Wb, c64 *__restrict__ S,
                                                 int m) {
    const int64_t mm = (int64_t)m * m;
    const c64 *w = Wb + (int64_t)blockIdx.y * mm;
    c64 *o = S + (int64_t)blockIdx.y * mm;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < mm;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(t % m), j = (int)(t / m);
        const c64 a = w[t], b = w[(int64_t)i * m + j];
        o[t] = c64{0.5f * (a.re + b.re), 0.5f * (a.im - b.im)};
    }
}

// M <- 2 D o M (in place) and betabar of each patch; one workgroup per patch.
// lam: the patch's Gram eigenvalues (V's column order), beta cast to Float32 as in
// the forward.  D in forms without the eigenvalue difference in a denominator:
// both above beta^2: beta / (s_i s_j (s_i + s_j)); one above: g(l_i) / (l_i - l_j),
// where l_i - l_j >= (s_i - beta)(s_i + beta) > 0; both below: 0.
__global__ __launch_bounds__(256) void k_vjp_divdiff(c64 *__restrict__ Mall,
                                                     const double *__restrict__ lamall, int m,
                                                     const double *__restrict__ beta,
                                                     int beta_per_patch,
                                                     double *__restrict__ betabar) {
    __shared__ double red[4];
    const int ip = blockIdx.x;
    c64 *M = Mall + (int64_t)ip * m * m;
    const double *lam = lamall + (int64_t)ip * m;
    const double bd = (double)(float)beta[beta_per_patch ? ip : 0];
    double bb = 0.0;
    for (int t = threadIdx.x; t < m * m; t += blockDim.x) {
        const int i = t % m, j = t / m;
        const double li = fmax(lam[i], 0.0), lj = fmax(lam[j], 0.0);
        const double si = sqrt(li), sj = sqrt(lj);
        const bool ui = si > bd, uj = sj > bd;
        double dd = 0.0;
        if (ui && uj) dd = bd / (si * sj * (si + sj));
        else if (ui) dd = (1.0 - bd / si) / (li - lj);
        else if (uj) dd = (1.0 - bd / sj) / (lj - li);
        const c64 v = M[t];
        if (i == j && ui) bb -= (double)v.re / si;
        const float f = (float)(2.0 * dd);
        M[t] = c64{v.re * f, v.im * f};
    }
    bb = block_sum(bb, red);
    if (threadIdx.x == 0) betabar[ip] = bb;
}

void svst_vjp_batched(mslr_ctx_s *ctx, const c64 *Z, const c64 *Yb, c64 *Zb, int64_t p, int64_t Nt,
                      int64_t Np, const double *beta_dev, int per_patch, double *betabar_dev,
                      int *err_dev) {
    if (Np == 0) return;
    MSLR_REQUIRE(p >= Nt && Nt >= 2, MSLR_ERR_UNSUPPORTED,
                 "the SVST derivative needs patch voxels >= Nt >= 2");
    const int64_t m = Nt;
    MSLR_REQUIRE(m <= MAXM, MSLR_ERR_UNSUPPORTED, "Nt must be <= " + std::to_string(MAXM));
    const int64_t len = p * Nt, mm = m * m;
    const int mi = (int)m;
    const SvstWork w = svst_work(ctx, m, p, Np);
    c64 *Wm = (c64 *)w.part;
    c64 *buf = (c64 *)ctx->scratch_get(SLOT_VJP, (size_t)(5 * mm * w.chunk) * sizeof(c64));
    c64 *Wb = buf, *Sx = buf + mm * w.chunk, *T1 = buf + 2 * mm * w.chunk;
    c64 *M1 = buf + 3 * mm * w.chunk, *T2 = buf + 4 * mm * w.chunk;
    c64 *Gb = Wb;  // Wb is dead once S is formed
    for (int64_t ip0 = 0; ip0 < Np; ip0 += w.chunk) {
        const int64_t nb = std::min<int64_t>(w.chunk, Np - ip0);
        const c64 *Zc = Z + ip0 * len, *Ybc = Yb + ip0 * len;
        c64 *Zbc = Zb + ip0 * len;
        const double *bp = beta_dev + (per_patch ? ip0 : 0);
        // every eigenpair of G
        gram_and_tridiag<double>(ctx, w, Zc, p, Nt, nb, err_dev);
        {
            MSLR_PROF(ctx, "vjp_eig");
            k_eig_count<<<(unsigned)cdiv(nb, 64), 64, 0, ctx->stream>>>(w.d, w.e, mi, nb, bp,
                                                                         per_patch, w.info, w.r, 1);
            HIP_LAUNCH_CHECK();
            dim3 ge((unsigned)nb, (unsigned)cdiv(m * 4, 64));
            k_eig_vectors<4><<<ge, 64, 3 * m * sizeof(double), ctx->stream>>>(w.d, w.e, mi, bp, per_patch, w.info, w.r,
                                                         w.lam, w.Z, w.U, 0, 0);
            HIP_LAUNCH_CHECK();
            k_eig_clusters<<<(unsigned)nb, NSLOT, 0, ctx->stream>>>(w.d, w.e, mi, bp, per_patch,
                                                                    w.info, w.r, w.lam, w.Z, w.ws,
                                                                    w.f);
            HIP_LAUNCH_CHECK();
        }
        if (w.sbr)
            sbr_backtransform(ctx, w.G, mi, nb, w.sw, w.Z, w.r, w.U);
        else
            launch_backtransform(ctx, w, mi, nb);
        // W = V diag(g) V^H
        cgemm(ctx, "vjp_w", w.U, mm, 1, m, w.U, mm, m, 1, 1, w.f, m, Wm, mm, 1, m, mi, mi, 0, w.r,
              nb);
        // Wb = Z^H Ybar, formed as its transpose Ybar^T conj(Z) written column-major
        cgemm(ctx, "vjp_zy", Ybc, len, p, 1, Zc, len, 1, p, 1, nullptr, 0, Wb, mm, m, 1, mi, mi,
              (int)p, nullptr, nb);
        {
            MSLR_PROF(ctx, "vjp_sym");
            dim3 gs((unsigned)cdiv(mm, 256), (unsigned)nb);
            k_vjp_sym<<<gs, 256, 0, ctx->stream>>>(Wb, Sx, mi);
            HIP_LAUNCH_CHECK();
        }
        // M = V^H S V (V^H T1 as the transpose T1^T conj(V))
        cgemm(ctx, "vjp_rot", Sx, mm, 1, m, w.U, mm, 1, m, 0, nullptr, 0, T1, mm, 1, m, mi, mi, mi,
              nullptr, nb);
        cgemm(ctx, "vjp_rot", T1, mm, m, 1, w.U, mm, 1, m, 1, nullptr, 0, M1, mm, m, 1, mi, mi, mi,
              nullptr, nb);
        {
            MSLR_PROF(ctx, "vjp_divdiff");
            k_vjp_divdiff<<<(unsigned)nb, 256, 0, ctx->stream>>>(M1, w.lam, mi, bp, per_patch,
                                                                 betabar_dev + ip0);
            HIP_LAUNCH_CHECK();
        }
        // 2 Gbar = V (2 D o M) V^H
        cgemm(ctx, "vjp_rot", M1, mm, 1, m, w.U, mm, m, 1, 1, nullptr, 0, T2, mm, 1, m, mi, mi, mi,
              nullptr, nb);
        cgemm(ctx, "vjp_rot", w.U, mm, 1, m, T2, mm, 1, m, 0, nullptr, 0, Gb, mm, 1, m, mi, mi, mi,
              nullptr, nb);
        // Zbar = Ybar W + Z (2 Gbar)
        cgemm(ctx, "vjp_out", Ybc, len, 1, p, Wm, mm, 1, m, 0, nullptr, 0, Zbc, len, 1, p, (int)p,
              mi, mi, nullptr, nb);
        cgemm(ctx, "vjp_out", Zc, len, 1, p, Gb, mm, 1, m, 0, nullptr, 0, Zbc, len, 1, p, (int)p, mi,
              mi, nullptr, nb, RankSel{}, 1);
    }
}

// Pullback of the SVST of Np matrices with one row or one column: the group shrink
// y = v max(1 - beta/|v|, 0) of each vector v (the p x Nt values of the matrix, stored
// contiguously).  With <a, b> = Re sum conj(a) b, for |v| > beta
//   vbar = (1 - beta/|v|) ybar + (beta / |v|^3) <ybar, v> v,   betabar = -<ybar, v> / |v|
// and vbar = 0, betabar = 0 otherwise.  One wave per vector, FP64 sums; beta rounded
// to Float32 as in the forward (src/recon.jl:195).
__global__ __launch_bounds__(256) void k_vjp_vector(const c64 *__restrict__ Z,
                                                    const c64 *__restrict__ Yb,
                                                    c64 *__restrict__ Zb, int64_t len, int64_t Np,
                                                    const double *__restrict__ beta, int per_patch,
                                                    double *__restrict__ betabar) {
    const int lane = threadIdx.x & 63;
    const int64_t ip = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (ip >= Np) return;  // the whole wave
    const c64 *z = Z + ip * len;
    const c64 *yb = Yb + ip * len;
    c64 *zb = Zb + ip * len;
    double s2 = 0.0, d = 0.0;
    for (int64_t i = lane; i < len; i += 64) {
        const c64 a = z[i], b = yb[i];
        s2 += (double)a.re * a.re + (double)a.im * a.im;
        d += (double)b.re * a.re + (double)b.im * a.im;
    }
    for (int o = 32; o > 0; o >>= 1) {
        s2 += __shfl_xor(s2, o, 64);
        d += __shfl_xor(d, o, 64);
    }
    const double be = (double)(float)beta[per_patch ? ip : 0];
    const double s = sqrt(s2);
    if (s > be) {
        const double g = 1.0 - be / s, c = be * d / (s * s2);
        for (int64_t i = lane; i < len; i += 64) {
            const c64 a = z[i], b = yb[i];
            zb[i] = c64{(float)(g * b.re + c * a.re), (float)(g * b.im + c * a.im)};
        }
        if (lane == 0) betabar[ip] = -d / s;
    } else {
        for (int64_t i = lane; i < len; i += 64) zb[i] = c64{0.0f, 0.0f};
        if (lane == 0) betabar[ip] = 0.0;
    }
}

// dst[:, :, i] = src[:, :, i]^H for Np r x c matrices (32 x 32 tiles through LDS).
__global__ __launch_bounds__(256) void k_conj_transpose(const c64 *__restrict__ src,
                                                        c64 *__restrict__ dst, int64_t r, int64_t c) {
    __shared__ c64 tile[32][33];
    const int64_t ip = blockIdx.z;
    const int64_t i0 = (int64_t)blockIdx.x * 32, j0 = (int64_t)blockIdx.y * 32;
    const c64 *S = src + ip * r * c;
    c64 *D = dst + ip * r * c;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    for (int k = ty; k < 32; k += 8) {
        const int64_t i = i0 + tx, j = j0 + k;
        if (i < r && j < c) tile[k][tx] = S[i + r * j];
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const int64_t j = j0 + tx, i = i0 + k;
        if (i < r && j < c) {
            const c64 v = tile[tx][k];
            D[j + c * i] = c64{v.re, -v.im};
        }
    }
}

static void conj_transpose(mslr_ctx_s *ctx, const c64 *src, c64 *dst, int64_t r, int64_t c,
                           int64_t Np) {
    MSLR_PROF(ctx, "vjp_transpose");
    dim3 g((unsigned)cdiv(r, 32), (unsigned)cdiv(c, 32), (unsigned)Np);
    k_conj_transpose<<<g, 256, 0, ctx->stream>>>(src, dst, r, c);
    HIP_LAUNCH_CHECK();
}

void svst_vjp_any(mslr_ctx_s *ctx, const c64 *Z, const c64 *Yb, c64 *Zb, int64_t p, int64_t Nt,
                  int64_t Np, const double *beta_dev, int per_patch, double *betabar_dev,
                  int *err_dev) {
    if (Np == 0 || p == 0 || Nt == 0) return;
    if (std::min(p, Nt) == 1) {
        MSLR_PROF(ctx, "vjp_vector");
        k_vjp_vector<<<(unsigned)cdiv(Np, 4), 256, 0, ctx->stream>>>(Z, Yb, Zb, p * Nt, Np, beta_dev,
                                                                      per_patch, betabar_dev);
        HIP_LAUNCH_CHECK();
        return;
    }
    if (p >= Nt) {
        svst_vjp_batched(ctx, Z, Yb, Zb, p, Nt, Np, beta_dev, per_patch, betabar_dev, err_dev);
        return;
    }
    // fewer rows than columns: Y^H = SVST(Z^H), the pullback of the transposed problem
    const size_t len = (size_t)(p * Nt * Np);
    c64 *T = (c64 *)ctx->scratch_get(SLOT_VJP_T, 3 * len * sizeof(c64));
    c64 *ZT = T, *YbT = T + len, *ZbT = T + 2 * len;
    conj_transpose(ctx, Z, ZT, p, Nt, Np);
    conj_transpose(ctx, Yb, YbT, p, Nt, Np);
    svst_vjp_batched(ctx, ZT, YbT, ZbT, Nt, p, Np, beta_dev, per_patch, betabar_dev, err_dev);
    conj_transpose(ctx, ZbT, Zb, Nt, p, Np);
}

// Nuclear norm of every patch into per (device, Np doubles).
static void nucnorm_per_patch(mslr_ctx_s *ctx, const c64 *P, int64_t p, int64_t Nt, int64_t Np,
                              double *per) {
    const int64_t m = std::min(p, Nt);
    MSLR_REQUIRE(m <= MAXM, MSLR_ERR_UNSUPPORTED, "min(patch voxels, Nt) too large");
    if (m == 1) {
        k_m1_nucnorm<<<(unsigned)cdiv(Np, 4), 256, 0, ctx->stream>>>(P, p * Nt, Np, per);
        HIP_LAUNCH_CHECK();
    } else {
        const bool left = p < Nt;
        const SvstWork w = svst_work(ctx, m, left ? Nt : p, Np);
        for (int64_t ip0 = 0; ip0 < Np; ip0 += w.chunk) {
            const int64_t nb = std::min<int64_t>(w.chunk, Np - ip0);
            gram_and_tridiag<double>(ctx, w, P + ip0 * p * Nt, p, Nt, nb, nullptr);
            k_nucnorm<<<(unsigned)nb, 256, 0, ctx->stream>>>(w.d, w.e, (int)m, per + ip0);
            HIP_LAUNCH_CHECK();
        }
    }
}

void nucnorm_batched(mslr_ctx_s *ctx, const c64 *P, int64_t p, int64_t Nt, int64_t Np,
                     double *out_dev) {
    if (Np == 0) {
        HIP_CHECK(hipMemsetAsync(out_dev, 0, sizeof(double), ctx->stream));
        return;
    }
    double *per = (double *)ctx->scratch_get(SLOT_EIG, (size_t)(Np + 1) * sizeof(double));
    nucnorm_per_patch(ctx, P, p, Nt, Np, per);
    sum_partials(ctx, per, Np, out_dev);
}

}  // namespace mslr

using namespace mslr;

extern "C" {

mslr_status_t mslr_svst_batched(mslr_ctx ctx, void *P_dev, int64_t p, int64_t Nt, int64_t Np,
                                const double *beta_host, int per_patch) {
    return guarded([&] {
        mslr_ctx_s *c = check_ctx(ctx);
        MSLR_REQUIRE(p >= 0 && Nt >= 0 && Np >= 0, MSLR_ERR_ARG, "negative size");
        MSLR_REQUIRE(beta_host != nullptr, MSLR_ERR_ARG, "beta is null");
        const int64_t nbeta = per_patch ? Np : 1;
        for (int64_t i = 0; i < nbeta; ++i)
            MSLR_REQUIRE(beta_host[i] >= 0.0 && std::isfinite(beta_host[i]), MSLR_ERR_ARG,
                         "beta must be finite and >= 0");
        if (Np == 0 || p == 0 || Nt == 0) return;
        double *bd = (double *)c->scratch_get(SLOT_APPLY, (size_t)(nbeta + 1) * sizeof(double) + 64);
        int *err = (int *)(bd + nbeta + 1);
        HIP_CHECK(hipMemcpyAsync(bd, beta_host, (size_t)nbeta * sizeof(double),
                                 hipMemcpyHostToDevice, c->stream));
        HIP_CHECK(hipMemsetAsync(err, 0, sizeof(int), c->stream));
        c64 *Y = (c64 *)c->scratch_get(SLOT_PATCH_OUT, (size_t)(p * Nt * Np) * sizeof(c64));
        svst_batched(c, (const c64 *)P_dev, Y, p, Nt, Np, bd, per_patch, err);
        HIP_CHECK(hipMemcpyAsync(P_dev, Y, (size_t)(p * Nt * Np) * sizeof(c64),
                                 hipMemcpyDeviceToDevice, c->stream));
        int herr = 0;
        HIP_CHECK(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIP_CHECK(hipStreamSynchronize(c->stream));
        MSLR_REQUIRE(herr == 0, MSLR_ERR_NUMERIC, "non-finite values in a patch matrix");
    });
}

mslr_status_t mslr_svst_vjp(mslr_ctx ctx, const void *Z_dev, const void *Ybar_dev, void *Zbar_dev,
                            int64_t p, int64_t Nt, int64_t Np, const double *beta_host,
                            int per_patch, double *betabar_host) {
    return guarded([&] {
        mslr_ctx_s *c = check_ctx(ctx);
        MSLR_REQUIRE(Z_dev && Ybar_dev && Zbar_dev && beta_host && betabar_host, MSLR_ERR_ARG,
                     "null argument");
        MSLR_REQUIRE(p >= 0 && Nt >= 0 && Np >= 0, MSLR_ERR_ARG, "negative size");
        MSLR_REQUIRE(Zbar_dev != Z_dev && Zbar_dev != Ybar_dev, MSLR_ERR_ARG,
                     "Zbar must not alias the inputs");
        const int64_t nbeta = per_patch ? Np : 1;
        for (int64_t i = 0; i < nbeta; ++i)
            MSLR_REQUIRE(beta_host[i] >= 0.0 && std::isfinite(beta_host[i]), MSLR_ERR_ARG,
                         "beta must be finite and >= 0");
        for (int64_t i = 0; i < nbeta; ++i) betabar_host[i] = 0.0;
        if (Np == 0 || p == 0 || Nt == 0) return;
        double *bd = (double *)c->scratch_get(SLOT_APPLY, (size_t)(nbeta + Np + 2) * sizeof(double) + 64);
        double *bb = bd + nbeta;
        int *err = (int *)(bb + Np + 1);
        HIP_CHECK(hipMemcpyAsync(bd, beta_host, (size_t)nbeta * sizeof(double),
                                 hipMemcpyHostToDevice, c->stream));
        HIP_CHECK(hipMemsetAsync(err, 0, sizeof(int), c->stream));
        svst_vjp_any(c, (const c64 *)Z_dev, (const c64 *)Ybar_dev, (c64 *)Zbar_dev, p, Nt, Np,
                     bd, per_patch, bb, err);
        std::vector<double> h((size_t)Np);
        int herr = 0;
        HIP_CHECK(hipMemcpyAsync(h.data(), bb, (size_t)Np * sizeof(double), hipMemcpyDeviceToHost,
                                 c->stream));
        HIP_CHECK(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIP_CHECK(hipStreamSynchronize(c->stream));
        MSLR_REQUIRE(herr == 0, MSLR_ERR_NUMERIC, "non-finite values in a patch matrix");
        for (int64_t i = 0; i < Np; ++i) betabar_host[per_patch ? i : 0] += h[(size_t)i];
    });
}

mslr_status_t mslr_patch_nucnorm(mslr_ctx ctx, const void *P_dev, int64_t p, int64_t Nt,
                                 int64_t Np, double *out) {
    return guarded([&] {
        mslr_ctx_s *c = check_ctx(ctx);
        MSLR_REQUIRE(out != nullptr, MSLR_ERR_ARG, "out is null");
        MSLR_REQUIRE(p >= 1 && Nt >= 1 && Np >= 0, MSLR_ERR_ARG, "bad sizes");
        double *od = (double *)c->scratch_get(SLOT_APPLY, 64);
        nucnorm_batched(c, (const c64 *)P_dev, p, Nt, Np, od);
        HIP_CHECK(hipMemcpyAsync(out, od, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIP_CHECK(hipStreamSynchronize(c->stream));
    });
}

mslr_status_t mslr_patch_nucnorm_weighted(mslr_ctx ctx, const void *P_dev, int64_t p, int64_t Nt,
                                          int64_t Np, const double *lambdas_host, double *out) {
    return guarded([&] {
        mslr_ctx_s *c = check_ctx(ctx);
        MSLR_REQUIRE(out != nullptr && (lambdas_host != nullptr || Np == 0), MSLR_ERR_ARG,
                     "out or lambdas is null");
        MSLR_REQUIRE(p >= 1 && Nt >= 1 && Np >= 1, MSLR_ERR_ARG, "bad sizes");
        double *per = (double *)c->scratch_get(SLOT_EIG, (size_t)(Np + 1) * sizeof(double));
        nucnorm_per_patch(c, (const c64 *)P_dev, p, Nt, Np, per);
        std::vector<double> h((size_t)Np);
        HIP_CHECK(hipMemcpyAsync(h.data(), per, (size_t)Np * sizeof(double), hipMemcpyDeviceToHost,
                                 c->stream));
        HIP_CHECK(hipStreamSynchronize(c->stream));
        // a patch whose largest singular value is 0 (nuclear norm 0) adds nothing;
        // the mean over all patches, in patch order
        double tot = 0.0;
        for (int64_t i = 0; i < Np; ++i) tot += h[(size_t)i] > 0.0 ? lambdas_host[i] * h[(size_t)i] : 0.0;
        *out = tot / (double)Np;
    });
}

}  // extern "C"
