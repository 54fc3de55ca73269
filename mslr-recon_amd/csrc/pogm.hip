// POGM iteration on the device (src/mirt_mod.jl:111-185 with the driver
// closures of scripts/reconstruct.jl:174-202).
//
// With the reference's in-place prox (g_prox overwrites and returns its
// argument), xnew and znew are the same array, so per iteration
//   g   = A'(A sum_s X_s - y)                      (same g for every scale)
//   u   = x - alpha g
//   z   = u + beta (u - u_old) + gamma (u - x)     (the zeta_old term is 0)
//   x   = patchSVST_s(z_s, zeta lambda_s)  for each scale s
//   restart iff sum Re(-g .* (u - u_old)) <= cutoff ||g|| ||u - u_old||
// The scalar recurrence (t, beta, gamma, zeta, sigma) lives in device memory
// and is advanced by single-workgroup kernels, so an iteration has no host
// synchronisation and can be captured in a hipGraph.
//
// Cost log (the run_recon logger, scripts/reconstruct.jl:176-191,205 with
// src/mirt_mod.jl:106,181): for every iterate x_k, k = 0..niter,
//   dc_cost(x_k)  = 0.5 ||A sum_s X_s - y||^2
//   reg_cost(x_k) = sum_s lambda_s sum_patches ||patch of X_s||_*
// dc_cost(x_{k-1}) is the squared norm of the residual iteration k forms for its
// gradient anyway (partials from the x-pass of A); the last iterate's takes one
// more A when the log is read.  reg_cost(x_k): with disjoint patches the prox
// output's patches are the SVST outputs, whose singular values are
// max(s - beta, 0) of the eigenvalues the thresholding already found; with
// overlapping patches the new X_s is cut into patches again and their
// singular values taken from Gram eigenvalues (no vectors).
#include <algorithm>
#include <cmath>
#include <cstdio>

#include "patches.hpp"
#include "ozaki.hpp"
#include "reduce.hpp"
#include "sense.hpp"
#include "svst.hpp"

namespace mslr {

struct PogmDev {
    double alpha;
    double told, sig, tnew;
    double beta, gamma, zeta;
    double restart_cutoff;
    int iter;    // iterations completed
    int niter;   // iterations of the run
    int restart; // MSLR_RESTART_*
    int err;     // set when a Gram matrix held non-finite values
};

// frame groups of the gradient + update: at c1 (90x90x60x300, 18 coils) 1, 2, 4, 6 and
// 8 groups measured 47.7, 47.7, 48.2, 48.1, 47.9 ms per mixed iteration: the FFT passes
// fill every wave slot, so the other stream's memory-bound kernels find no room
constexpr int SENSE_CHUNKS_DEFAULT = 1;

// scal_log (nullable): alpha, beta, gamma, zeta of each iteration, for the
// derivative's backward pass.
__global__ void k_pogm_pre(PogmDev *s, const double *__restrict__ lambdas, int nscales,
                           double *__restrict__ thresholds, double *__restrict__ scal_log) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int k = s->iter + 1;
    const double told = s->told;
    const double tnew = (k == s->niter) ? 0.5 * (1.0 + sqrt(1.0 + 8.0 * told * told))
                                        : 0.5 * (1.0 + sqrt(1.0 + 4.0 * told * told));
    const double beta = (told - 1.0) / tnew;
    const double gamma = s->sig * told / tnew;
    const double zeta = s->alpha * (1.0 + beta + gamma);
    s->tnew = tnew;
    s->beta = beta;
    s->gamma = gamma;
    s->zeta = zeta;
    for (int i = 0; i < nscales; ++i) thresholds[i] = zeta * lambdas[i];
    if (scal_log) {
        double *o = scal_log + 4 * (k - 1);
        o[0] = s->alpha;
        o[1] = beta;
        o[2] = gamma;
        o[3] = zeta;
    }
}

// dst[iter] <- X (the prox input z of the running iteration), iter from the device
// counter so that graph replays fill their own slots.
__global__ __launch_bounds__(256) void k_ckpt_copy(const float4 *__restrict__ X,
                                                   float4 *__restrict__ dst, int64_t n4,
                                                   const PogmDev *s) {
    float4 *d = dst + (int64_t)s->iter * n4;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4;
         i += (int64_t)gridDim.x * blockDim.x)
        d[i] = X[i];
}

// Backward of z = (1 + b + c) u - b u_old - c x and u = x - alpha g for one
// iteration, per element over all scales (zb holds the prox input cotangent):
//   ut = ub + (1 + b + c) zb ;  ub <- -b zb ;  xb <- ut - c zb ;  gs = -alpha sum_s ut
__global__ __launch_bounds__(256) void k_vjp_state(const c64 *__restrict__ zb, c64 *__restrict__ ub,
                                                   c64 *__restrict__ xb, c64 *__restrict__ gs,
                                                   int64_t n, int nscales, float alpha, float b,
                                                   float c) {
    const float c1 = 1.f + b + c;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        float sr = 0.f, si = 0.f;
        for (int k = 0; k < nscales; ++k) {
            const int64_t o = i + k * n;
            const c64 z = zb[o], u = ub[o];
            const float tr = u.re + c1 * z.re, ti = u.im + c1 * z.im;
            ub[o] = c64{-b * z.re, -b * z.im};
            xb[o] = c64{tr - c * z.re, ti - c * z.im};
            sr += tr;
            si += ti;
        }
        gs[i] = c64{-alpha * sr, -alpha * si};
    }
}

// xb_s += h for every scale.
__global__ __launch_bounds__(256) void k_vjp_add(c64 *__restrict__ xb, const c64 *__restrict__ h,
                                                 int64_t n, int nscales) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const c64 v = h[i];
        for (int k = 0; k < nscales; ++k) xb[i + k * n] = xb[i + k * n] + v;
    }
}

// lam_bar[s] += zeta * sum(betabar[0..np)) (fixed order, one workgroup).
__global__ __launch_bounds__(256) void k_vjp_lambda(const double *__restrict__ betabar, int64_t np,
                                                    double zeta, double *__restrict__ lam_bar) {
    __shared__ double red[4];
    double acc = 0.0;
    for (int64_t i = threadIdx.x; i < np; i += blockDim.x) acc += betabar[i];
    acc = block_sum(acc, red);
    if (threadIdx.x == 0) *lam_bar += zeta * acc;
}

// X <- z, U <- u; block partials of (restart dot, ||g||^2, ||u - u_old||^2).
// One element of one scale (FP64 arithmetic on the FP32 state).
__device__ __forceinline__ void pogm_elem(float xr, float xi, float uor, float uoi, double gr,
                                          double gi, double alpha, double beta, double gamma,
                                          float &zr_out, float &zi_out, float &ur_out,
                                          float &ui_out, double &dot, double &gg, double &dd) {
    const double ur = xr - alpha * gr, ui = xi - alpha * gi;
    const double dr = ur - uor, di = ui - uoi;
    const double zr = ur + beta * dr + gamma * (ur - xr);
    const double zi = ui + beta * di + gamma * (ui - xi);
    zr_out = (float)zr;
    zi_out = (float)zi;
    ur_out = (float)ur;
    ui_out = (float)ui;
    dot -= gr * dr - gi * di;  // Re(-g .* d), no conjugate
    gg += gr * gr + gi * gi;
    dd += dr * dr + di * di;
}

// The update fused with the last transpose of the gradient: one workgroup per (block
// of TB yz-lines, frame f of the chunk).  g of the block is read from the plane-major
// gradient Pm[f][x][yz] and transposed through LDS; then, in image order, u and z of
// every scale (pogm_elem) and the block's partials of the restart test, stored at
// (frame ta + f, block) so that the order of the FP64 sums does not depend on how
// the frames were grouped.  X and U start at frame ta.
// wk >= 0 (even Nx, TB Nx a multiple of SPL_ROWS): z of scale wk, the whole-image
// patch, goes out as its split words (Wh, Wl, frame-major like X from frame ta)
// instead of into X: each wave's 64 pairs of x-neighbours are one SPL_ROWS-row block
// of a frame's column, split as the SVST's split pass splits it.
__global__ __launch_bounds__(256) void k_planes_update(const c64 *__restrict__ Pm, int Nx, int NYZ,
                                                       int TB, int64_t ta, c64 *__restrict__ X,
                                                       c64 *__restrict__ U, int64_t sstride,
                                                       int nscales, const PogmDev *__restrict__ s,
                                                       double *__restrict__ partials,
                                                       unsigned *__restrict__ Wh,
                                                       unsigned *__restrict__ Wl, int wk) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    c64 *tile = reinterpret_cast<c64 *>(smem_raw);
    __shared__ double red[16];
    const int f = blockIdx.y;
    const int l0 = blockIdx.x * TB;
    const int nl = min(TB, NYZ - l0);
    const int64_t V = (int64_t)Nx * NYZ;
    const c64 *src = Pm + (int64_t)f * V + l0;
    for (int e = threadIdx.x; e < nl * Nx; e += blockDim.x) {
        const int x = e / nl;
        const int l = e - x * nl;
        tile[x * (TB + 1) + l] = src[(int64_t)x * NYZ + l];
    }
    __syncthreads();
    const double alpha = s->alpha, beta = s->beta, gamma = s->gamma;
    double dot = 0.0, gg = 0.0, dd = 0.0;
    const int64_t base = (int64_t)f * V + (int64_t)l0 * Nx;
    if ((Nx & 1) == 0) {
        // two x-neighbours per thread: 16-byte accesses (base, the frame offset and
        // the scale stride are even with Nx), the X and U of up to four scales
        // loaded before any store
        const int64_t s2 = sstride / 2;
        float4 *X4 = reinterpret_cast<float4 *>(X + base);
        float4 *U4 = reinterpret_cast<float4 *>(U + base);
        uint2 *Wh2 = reinterpret_cast<uint2 *>(Wh + (wk >= 0 ? base : 0));
        uint2 *Wl2 = reinterpret_cast<uint2 *>(Wl + (wk >= 0 ? base : 0));
        const int half = Nx / 2, npair = nl * half;
        const int lane = threadIdx.x & 63;
        // whole waves step together (the split's block maximum is a wave reduction)
        for (int e0 = threadIdx.x - lane; e0 < npair; e0 += blockDim.x) {
            const int e = e0 + lane;
            const bool act = e < npair;
            const int l = act ? e / half : 0;
            const int x = act ? 2 * (e - l * half) : 0;
            const c64 g0 = tile[x * (TB + 1) + l], g1 = tile[(x + 1) * (TB + 1) + l];
            for (int k0 = 0; k0 < nscales; k0 += 4) {
                float4 xv[4], uv[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    if (act && k0 + q < nscales) {
                        xv[q] = X4[(int64_t)(k0 + q) * s2 + e];
                        uv[q] = U4[(int64_t)(k0 + q) * s2 + e];
                    }
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    if (k0 + q < nscales) {
                        float4 zo = make_float4(0.f, 0.f, 0.f, 0.f), uo;
                        if (act) {
                            pogm_elem(xv[q].x, xv[q].y, uv[q].x, uv[q].y, g0.re, g0.im, alpha,
                                      beta, gamma, zo.x, zo.y, uo.x, uo.y, dot, gg, dd);
                            pogm_elem(xv[q].z, xv[q].w, uv[q].z, uv[q].w, g1.re, g1.im, alpha,
                                      beta, gamma, zo.z, zo.w, uo.z, uo.w, dot, gg, dd);
                            U4[(int64_t)(k0 + q) * s2 + e] = uo;
                        }
                        if (k0 + q == wk) {  // uniform over the wave
                            float sg1, sg2;
                            ozaki_block_quanta(fmaxf(fmaxf(fabsf(zo.x), fabsf(zo.y)),
                                                     fmaxf(fabsf(zo.z), fabsf(zo.w))),
                                               sg1, sg2);
                            unsigned h0, l0w, h1, l1w;
                            ozaki_words(c64{zo.x, zo.y}, sg1, sg2, h0, l0w);
                            ozaki_words(c64{zo.z, zo.w}, sg1, sg2, h1, l1w);
                            if (act) {
                                Wh2[e] = make_uint2(h0, h1);
                                Wl2[e] = make_uint2(l0w, l1w);
                            }
                        } else if (act) {
                            X4[(int64_t)(k0 + q) * s2 + e] = zo;
                        }
                    }
                }
            }
        }
    } else {
        for (int e = threadIdx.x; e < nl * Nx; e += blockDim.x) {
            const int l = e / Nx;
            const int x = e - l * Nx;
            const c64 gv = tile[x * (TB + 1) + l];
            const int64_t i = base + e;
            for (int k = 0; k < nscales; ++k) {
                const c64 xv = X[(int64_t)k * sstride + i], uv = U[(int64_t)k * sstride + i];
                float zr, zi, ur, ui;
                pogm_elem(xv.re, xv.im, uv.re, uv.im, gv.re, gv.im, alpha, beta, gamma, zr, zi, ur,
                          ui, dot, gg, dd);
                X[(int64_t)k * sstride + i] = c64{zr, zi};
                U[(int64_t)k * sstride + i] = c64{ur, ui};
            }
        }
    }
    const double a = block_sum(dot, red);
    const double b = block_sum(gg, red);
    const double c = block_sum(dd, red);
    if (threadIdx.x == 0) {
        const int64_t o = 3 * ((ta + f) * (int64_t)gridDim.x + blockIdx.x);
        partials[o + 0] = a;
        partials[o + 1] = b;
        partials[o + 2] = c;
    }
}

__global__ void k_pogm_post(PogmDev *s, const double *__restrict__ partials, int nblk,
                            int *__restrict__ restart_log) {
    __shared__ double red[16];
    double a = 0.0, b = 0.0, c = 0.0;
    const int per = (nblk + blockDim.x - 1) / blockDim.x;
    const int i0 = threadIdx.x * per;
    const int i1 = min(nblk, i0 + per);
    for (int i = i0; i < i1; ++i) {
        a += partials[3 * i];
        b += partials[3 * i + 1];
        c += partials[3 * i + 2];
    }
    a = block_sum(a, red);
    b = block_sum(b, red);
    c = block_sum(c, red);
    if (threadIdx.x == 0) {
        const int k = s->iter + 1;
        int rs = 0;
        if (s->restart == MSLR_RESTART_GR) rs = a <= s->restart_cutoff * sqrt(b) * sqrt(c);
        if (rs) {
            s->tnew = 1.0;
            s->sig = 1.0;
        }
        restart_log[k - 1] = rs;
        s->told = s->tnew;
        s->iter = k;
    }
}

// costs[2 iter] = 0.5 sum(part): dc_cost of the iterate the gradient was taken at.
__global__ __launch_bounds__(1024) void k_cost_dc(const double *__restrict__ part, int64_t n,
                                                  const PogmDev *s, double *__restrict__ costs) {
    __shared__ double red[16];
    double acc = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) acc += part[i];  // fixed order
    acc = block_sum(acc, red);
    if (threadIdx.x == 0) costs[2 * s->iter] = 0.5 * acc;
}

// costs[2 iter + 1] = sum_s lambda_s nuc_s: reg_cost of the newest iterate.
__global__ void k_cost_reg(const double *__restrict__ nuc, const double *__restrict__ lambdas,
                           int nscales, const PogmDev *s, double *__restrict__ costs) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double r = 0.0;
    for (int k = 0; k < nscales; ++k) r += lambdas[k] * nuc[k];
    costs[2 * s->iter + 1] = r;
}

__global__ void k_fill_zero(c64 *__restrict__ x, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        x[i] = c64{0.f, 0.f};
}

static void fill_zero(mslr_ctx_s *ctx, c64 *x, int64_t n) {
    if (n <= 0) return;
    const unsigned nb = (unsigned)std::min<int64_t>(cdiv(n, 256), 1 << 16);
    k_fill_zero<<<nb, 256, 0, ctx->stream>>>(x, n);
    HIP_LAUNCH_CHECK();
}

// patchSVST of 1x1x1 patches tiling the image (stride 1): each voxel's row over
// time x_v (1 x Nt) is one patch, its only singular value is ||x_v||, so
// out_v = x_v max(||x_v|| - beta, 0) / ||x_v|| (beta rounded to Float32 as in
// src/recon.jl:195).  Equals img2patches -> SVST -> patches2img without the two
// transposing copies; in place allowed.  nuc_part (nullable): per-workgroup sums
// of the outputs' singular values max(||x_v|| - beta, 0).
__global__ __launch_bounds__(256) void k_prox_voxel(const c64 *img, c64 *out, int64_t V, int Nt,
                                                    const double *__restrict__ beta,
                                                    double *__restrict__ nuc_part) {
    __shared__ double red[4];
    const double bd = (double)(float)beta[0];
    double nacc = 0.0;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
         v += (int64_t)gridDim.x * blockDim.x) {
        float fk = 1.f;
        if (bd != 0.0 || nuc_part) {
            double s = 0.0;
            for (int t = 0; t < Nt; ++t) {
                const c64 a = img[v + V * t];
                s += (double)a.re * a.re + (double)a.im * a.im;
            }
            const double sg = sqrt(s);
            if (bd != 0.0) fk = sg > bd ? (float)((sg - bd) / sg) : 0.f;
            nacc += fmax(sg - bd, 0.0);
        }
        for (int t = 0; t < Nt; ++t) {
            const c64 a = img[v + V * t];
            out[v + V * t] = c64{a.re * fk, a.im * fk};
        }
    }
    if (nuc_part) {
        nacc = block_sum(nacc, red);
        if (threadIdx.x == 0) nuc_part[blockIdx.x] = nacc;
    }
}

// The same with one read of the data: a workgroup of 256 threads takes 16 voxels,
// 16 threads per voxel each holding every 16th frame of its time course in registers
// (at most PV_T frames per thread); the squared norm is summed per thread over its
// frames and then over the 16 threads in a fixed order.
constexpr int PV_V = 16, PV_G = 16, PV_T = 32;
__global__ __launch_bounds__(256) void k_prox_voxel_reg(const c64 *img, c64 *out, int64_t V,
                                                        int Nt, const double *__restrict__ beta,
                                                        double *__restrict__ nuc_part) {
    __shared__ double part[PV_G][PV_V];
    __shared__ float fks[PV_V];
    __shared__ double red[4];
    const double bd = (double)(float)beta[0];
    const int vl = threadIdx.x % PV_V, g = threadIdx.x / PV_V;
    double nacc = 0.0;
    for (int64_t v0 = (int64_t)blockIdx.x * PV_V; v0 < V; v0 += (int64_t)gridDim.x * PV_V) {
        const int64_t v = v0 + vl;
        const bool ok = v < V;
        c64 a[PV_T];
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < PV_T; ++q) {
            const int t = g + PV_G * q;
            a[q] = c64{0.f, 0.f};
            if (ok && t < Nt) a[q] = img[v + V * t];
            s += (double)a[q].re * a[q].re + (double)a[q].im * a[q].im;
        }
        part[g][vl] = s;
        __syncthreads();
        if (g == 0) {
            double tot = 0.0;
            for (int i = 0; i < PV_G; ++i) tot += part[i][vl];
            const double sg = sqrt(tot);
            float fk = 1.f;
            if (bd != 0.0) fk = sg > bd ? (float)((sg - bd) / sg) : 0.f;
            fks[vl] = fk;
            if (ok) nacc += fmax(sg - bd, 0.0);
        }
        __syncthreads();
        const float fk = fks[vl];
#pragma unroll
        for (int q = 0; q < PV_T; ++q) {
            const int t = g + PV_G * q;
            if (ok && t < Nt) out[v + V * t] = c64{a[q].re * fk, a[q].im * fk};
        }
    }
    if (nuc_part) {
        nacc = block_sum(nacc, red);
        if (threadIdx.x == 0) nuc_part[blockIdx.x] = nacc;
    }
}

static bool prox_voxel_reg(int Nt) { return Nt <= PV_G * PV_T; }

static unsigned prox_voxel_blocks(int64_t V, int Nt) {
    return prox_voxel_reg(Nt) ? (unsigned)std::min<int64_t>(cdiv(V, PV_V), 1 << 16)
                              : (unsigned)std::min<int64_t>(cdiv(V, 256), 1 << 16);
}

// Per-unit partial count prox_scale writes for the nuclear norm of its output.
static int64_t nuc_units(const PatchGeom &g) {
    const int64_t V = (int64_t)g.N[0] * g.N[1] * g.N[2];
    if (g.p == 1 && g.Np == V && g.exact_tiling()) return prox_voxel_blocks(V, g.Nt);
    return std::max<int64_t>(1, g.Np);
}

// *nuc_tot = sum over patches of the nuclear norm of img's patches (eigenvalues
// of the Gram matrices only).
static void scale_nucnorm(mslr_ctx_s *ctx, const PatchGeom &g, const c64 *img, double *nuc_tot) {
    const int64_t V = (int64_t)g.N[0] * g.N[1] * g.N[2];
    MSLR_PROF(ctx, "nucnorm");
    if (g.Np == 1 && g.p == V) {
        nucnorm_batched(ctx, img, g.p, g.Nt, 1, nuc_tot);
        return;
    }
    const size_t bytes = (size_t)std::max<int64_t>(1, g.p * g.Nt * g.Np) * sizeof(c64);
    c64 *P = (c64 *)ctx->scratch_get(SLOT_PATCH, bytes);
    launch_img2patches(ctx, g, img, P, 0, g.Np);
    nucnorm_batched(ctx, P, g.p, g.Nt, g.Np, nuc_tot);
}

// patchSVST for one scale: out <- patches2img(SVST(img2patches(img))).
// Two layouts need no patch copies: 1x1x1 patches with stride 1 (one voxel per
// patch), and a single patch covering the image (the patch matrix is the image).
// With split-BF16 products (p >= Nt) the SVST's split pass gathers the patches from
// the image (no img2patches) and the SVST writes out itself: the single patch in
// place, an exact tiling's patches scattered to their voxels (no patches2img).
// nuc_part/nuc_tot (device, nullable together): *nuc_tot <- sum over patches of
// the nuclear norm of out's patches; nuc_part holds nuc_units(g) partials.
// words (nullable): for the single patch with split-BF16 products, its split words,
// written by the caller (the update) in place of img (fused_words_scale).
static void prox_scale(mslr_ctx_s *ctx, const PatchGeom &g, const c64 *img, c64 *out,
                       const double *beta_dev, int *err, double *nuc_part, double *nuc_tot,
                       const unsigned *words = nullptr) {
    const int64_t V = (int64_t)g.N[0] * g.N[1] * g.N[2];
    if (g.p == 1 && g.Np == V && g.exact_tiling()) {
        const unsigned nb = prox_voxel_blocks(V, g.Nt);
        {
            MSLR_PROF(ctx, "prox_voxel");
            if (prox_voxel_reg(g.Nt))
                k_prox_voxel_reg<<<nb, 256, 0, ctx->stream>>>(img, out, V, g.Nt, beta_dev, nuc_part);
            else
                k_prox_voxel<<<nb, 256, 0, ctx->stream>>>(img, out, V, g.Nt, beta_dev, nuc_part);
            HIP_LAUNCH_CHECK();
        }
        if (nuc_part) sum_partials(ctx, nuc_part, nb, nuc_tot);
        return;
    }
    const size_t bytes = (size_t)std::max<int64_t>(1, g.p * g.Nt * g.Np) * sizeof(c64);
    // split-BF16 products write out directly: in place for the single patch,
    // scattered to the voxels for an exact tiling
    const bool direct = svst_writes_direct(ctx, g.p, g.Nt);
    MSLR_REQUIRE(!words || (g.Np == 1 && g.p == V && direct), MSLR_ERR_ARG,
                 "split words are handed in only for the single patch");
    if (g.Np == 1 && g.p == V && direct) {
        svst_batched(ctx, img, out, g.p, g.Nt, 1, beta_dev, 0, err, nuc_part, nullptr, nullptr,
                     words);
        if (nuc_part) sum_partials(ctx, nuc_part, 1, nuc_tot);
        return;
    }
    if (direct && g.Np > 1 && g.exact_tiling()) {
        svst_batched(ctx, img, out, g.p, g.Nt, g.Np, beta_dev, 0, err, nuc_part, &g, &g);
        if (nuc_part) sum_partials(ctx, nuc_part, g.Np, nuc_tot);
        return;
    }
    c64 *Y = (c64 *)ctx->scratch_get(SLOT_PATCH_OUT, bytes);
    if (direct && g.Np > 1) {
        // overlapping patches: gathered by the split pass, averaged by patches2img
        const bool disjoint = g.disjoint();
        svst_batched(ctx, img, Y, g.p, g.Nt, g.Np, beta_dev, 0, err, disjoint ? nuc_part : nullptr,
                     nullptr, &g);
        launch_patches2img(ctx, g, Y, out);
        if (!nuc_part) return;
        if (disjoint)
            sum_partials(ctx, nuc_part, g.Np, nuc_tot);
        else
            scale_nucnorm(ctx, g, out, nuc_tot);
        return;
    }
    if (g.Np == 1 && g.p == V) {
        svst_batched(ctx, img, Y, g.p, g.Nt, 1, beta_dev, 0, err, nuc_part);
        {
            MSLR_PROF(ctx, "patch_copy");
            HIP_CHECK(hipMemcpyAsync(out, Y, bytes, hipMemcpyDeviceToDevice, ctx->stream));
        }
        if (nuc_part) sum_partials(ctx, nuc_part, 1, nuc_tot);
        return;
    }
    // disjoint patches come back unchanged from X_s: their singular values are
    // the SVST's outputs
    const bool disjoint = g.disjoint();
    c64 *P = (c64 *)ctx->scratch_get(SLOT_PATCH, bytes);
    launch_img2patches(ctx, g, img, P, 0, g.Np);
    svst_batched(ctx, P, Y, g.p, g.Nt, g.Np, beta_dev, 0, err, disjoint ? nuc_part : nullptr);
    launch_patches2img(ctx, g, Y, out);
    if (!nuc_part) return;
    if (disjoint)
        sum_partials(ctx, nuc_part, g.Np, nuc_tot);
    else
        scale_nucnorm(ctx, g, out, nuc_tot);
}

// Backward of x_s = patchSVST(z_s, beta) = p2i(SVST(i2p(z_s))) for one scale:
// zb <- i2p^T(SVST'^T(p2i^T(xb))), betabar per patch into bb (device), where
// p2i^T(xb) = i2p(xb / count) and i2p^T sums the covering patches.  tmp: an image
// of scratch.  Patches need at least Nt voxels (1x1x1 patches are not covered).
static void prox_vjp_scale(mslr_ctx_s *ctx, const PatchGeom &g, const c64 *z, const c64 *xb,
                           c64 *zb, const double *beta_dev, double *bb, int *err, c64 *tmp) {
    const int64_t V = (int64_t)g.N[0] * g.N[1] * g.N[2];
    if (g.Np == 1 && g.p == V) {
        svst_vjp_any(ctx, z, xb, zb, g.p, g.Nt, 1, beta_dev, 0, bb, err);
        return;
    }
    const bool disjoint = g.disjoint();
    const size_t bytes = (size_t)(g.p * g.Nt * g.Np) * sizeof(c64);
    c64 *P = (c64 *)ctx->scratch_get(SLOT_PATCH, bytes);
    c64 *Pb = (c64 *)ctx->scratch_get(SLOT_PATCH_OUT, bytes);
    c64 *Qb = (c64 *)ctx->scratch_get(SLOT_VJP_P, bytes);
    launch_img2patches(ctx, g, z, P, 0, g.Np);
    if (disjoint) {
        launch_img2patches(ctx, g, xb, Pb, 0, g.Np);
    } else {
        launch_div_coverage(ctx, g, xb, tmp);
        launch_img2patches(ctx, g, tmp, Pb, 0, g.Np);
    }
    svst_vjp_any(ctx, P, Pb, Qb, g.p, g.Nt, g.Np, beta_dev, 0, bb, err);
    if (disjoint)
        launch_patches2img(ctx, g, Qb, zb);  // every covered voxel once: a copy
    else
        launch_patches2img_sum(ctx, g, Qb, zb);
}

}  // namespace mslr

using namespace mslr;

struct mslr_pogm_s {
    mslr_sense_s *sense = nullptr;
    mslr_ctx_s *ctx = nullptr;
    int64_t V = 0, Nt = 0, S = 0, K = 0, Nc = 0;
    int niter = 0, form = 0;
    std::vector<PatchGeom> geoms;
    std::vector<double> lambdas;
    double L = 0;
    // device
    c64 *X = nullptr, *U = nullptr, *y = nullptr;
    PogmDev *dev = nullptr;
    double *lam_dev = nullptr, *thr_dev = nullptr, *partials = nullptr;
    int *log_dev = nullptr;
    int iters_done = 0;
    bool initialised = false;
    // cost log: costs[2k], costs[2k+1] = dc_cost, reg_cost of x_k (k = 0..niter);
    // dc_part: residual partials; nuc_tot[s] and nuc_part[s]: per-scale nuclear norms
    bool log_costs = false;
    double *costs = nullptr, *dc_part = nullptr, *nuc_tot = nullptr;
    int64_t n_dc_part = 0;
    std::vector<double *> nuc_part;
    int dc_final_at = -1;  // iterate whose dc_cost was last completed by an extra A
    // derivative form: the prox inputs z_k of every iteration and the scalars
    bool checkpoint = false;
    c64 *ckpt = nullptr;
    double *scal_log = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // scratch generation the graph was captured at / of the last eager iteration
    uint64_t graph_gen = 0, warm_gen = 0;
    bool warm = false;
    // the gradient's frame groups, each followed by its part of the update; the
    // batched forms alternate them between the context stream and gpipe
    std::vector<std::pair<int64_t, int64_t>> gchunks;
    int upd_lines = 0;       // yz-lines per workgroup of the fused transpose + update
    int64_t n_upd_part = 0;  // partials of the restart test (frames x line blocks)
    hipStream_t gpipe = nullptr;
    hipEvent_t ev_gfork = nullptr, ev_gjoin = nullptr;
    // split words of the whole-image patch's prox input, written by the update
    // (fused_words_scale); allocated at the first iteration that uses them
    unsigned *words = nullptr;
    // batched forms: one stream per scale for the prox (fork / join with events)
    std::vector<hipStream_t> aux;
    std::vector<hipEvent_t> ev_join;
    hipEvent_t ev_fork = nullptr;
    std::vector<void *> allocs;

    void *alloc(size_t b) {
        void *p = nullptr;
        HIP_CHECK(hipMalloc(&p, std::max<size_t>(b, 256)));
        allocs.push_back(p);
        return p;
    }
};

static void drop_graph(mslr_pogm_s *st) {
    if (st->exec) hipGraphExecDestroy(st->exec);
    if (st->graph) hipGraphDestroy(st->graph);
    st->exec = nullptr;
    st->graph = nullptr;
}

static void pogm_free(mslr_pogm_s *st) {
    if (!st) return;
    drop_graph(st);
    for (auto s2 : st->aux) hipStreamDestroy(s2);
    if (st->gpipe) hipStreamDestroy(st->gpipe);
    if (st->ev_gfork) hipEventDestroy(st->ev_gfork);
    if (st->ev_gjoin) hipEventDestroy(st->ev_gjoin);
    for (auto e : st->ev_join) hipEventDestroy(e);
    if (st->ev_fork) hipEventDestroy(st->ev_fork);
    for (void *p : st->allocs) hipFree(p);
    delete st;
}

// reg_cost of the newest iterate from the per-scale nuclear norms.
static void log_reg_cost(mslr_pogm_s *st) {
    if (!st->log_costs) return;
    mslr_ctx_s *ctx = st->ctx;
    MSLR_PROF(ctx, "cost_reg");
    k_cost_reg<<<1, 64, 0, ctx->stream>>>(st->nuc_tot, st->lam_dev, (int)st->S, st->dev,
                                          st->costs);
    HIP_LAUNCH_CHECK();
}

// The scale whose prox input the update writes as split words instead of into X,
// or -1: the single patch covering the image when its SVST reads X only through the
// split words (split-BF16 products, p >= Nt) and writes its output in place, and
// the update's waves cover whole SPL_ROWS-row blocks (even Nx, TB Nx a multiple of
// SPL_ROWS).  Not with checkpoints, which keep each prox input.  ctx->products must
// hold the iteration's setting.
static int fused_words_scale(const mslr_pogm_s *st, const mslr_ctx_s *ctx, int TB, int Nx) {
    if (st->checkpoint || Nx % 2 != 0 || ((int64_t)TB * Nx) % SPL_ROWS != 0) return -1;
    for (int64_t k = 0; k < st->S; ++k) {
        const PatchGeom &g = st->geoms[k];
        if (g.Np == 1 && g.p == st->V && svst_writes_direct(ctx, g.p, g.Nt) &&
            svst_reads_words(ctx, g.p, g.Nt))
            return (int)k;
    }
    return -1;
}

// Enqueue one iteration on the context stream.
static void pogm_iteration(mslr_pogm_s *st) {
    mslr_ctx_s *ctx = st->ctx;
    const int64_t n = st->V * st->Nt;
    // the mixed form runs the SVST products in split BF16; the context setting
    // is restored after the prox
    struct ProductsScope {
        mslr_ctx_s *c;
        int saved;
        ProductsScope(mslr_ctx_s *c_, int form) : c(c_), saved(c_->products) {
            if (form == MSLR_FORM_MIXED) c->products = MSLR_PRODUCTS_SPLIT_BF16;
            else c->products = MSLR_PRODUCTS_FP32;
        }
        ~ProductsScope() { c->products = saved; }
    } products_scope(ctx, st->form);
    {   // beta, gamma, zeta of this iteration (they do not depend on the gradient)
        MSLR_PROF(ctx, "pogm_pre");
        k_pogm_pre<<<1, 64, 0, ctx->stream>>>(st->dev, st->lam_dev, (int)st->S, st->thr_dev,
                                              st->scal_log);
        HIP_LAUNCH_CHECK();
    }
    // gradient frame group by frame group, each group's update right behind its
    // adjoint (on the stream the group ran on)
    SensePipe sp;
    if (st->gpipe && !ctx->prof) {
        sp.stream = st->gpipe;
        sp.slot_base = (size_t)SLOT_COUNT * (size_t)(st->S + 1);
        sp.fork = st->ev_gfork;
        sp.join = st->ev_gjoin;
    }
    // the update reads each frame group's gradient where the adjoint leaves it
    // (plane-major), fused with the transpose back to image order
    const int TB = st->upd_lines;
    const int Nx = st->sense->N[0], NYZ = st->sense->NYZ;
    const size_t tile_lds = (size_t)Nx * (TB + 1) * sizeof(c64);
    const int wk = fused_words_scale(st, ctx, TB, Nx);
    if (wk >= 0 && !st->words) st->words = (unsigned *)st->alloc((size_t)n * 2 * sizeof(unsigned));
    sense_grad_pipelined(st->sense, st->X, (int)st->S, n, st->y, nullptr,
                         st->log_costs ? st->dc_part : nullptr, st->gchunks, sp,
                         [&](size_t, int64_t ta, int64_t nf, const c64 *planes) {
                             MSLR_PROF(ctx, "pogm_update");
                             const int64_t o = ta * st->V;
                             dim3 grid((unsigned)cdiv(NYZ, TB), (unsigned)nf);
                             k_planes_update<<<grid, 256, tile_lds, ctx->stream>>>(
                                 planes, Nx, NYZ, TB, ta, st->X + o, st->U + o, n, (int)st->S,
                                 st->dev, st->partials, wk >= 0 ? st->words + o : nullptr,
                                 wk >= 0 ? st->words + n + o : nullptr, wk);
                             HIP_LAUNCH_CHECK();
                         });
    if (st->log_costs) {
        MSLR_PROF(ctx, "cost_dc");
        k_cost_dc<<<1, 1024, 0, ctx->stream>>>(st->dc_part, st->n_dc_part, st->dev, st->costs);
        HIP_LAUNCH_CHECK();
    }
    if (st->checkpoint) {
        MSLR_PROF(ctx, "checkpoint");
        const int64_t n4 = n * st->S / 2;
        k_ckpt_copy<<<2048, 256, 0, ctx->stream>>>((const float4 *)st->X, (float4 *)st->ckpt, n4,
                                                   st->dev);
        HIP_LAUNCH_CHECK();
    }
    {
        MSLR_PROF(ctx, "pogm_post");
        k_pogm_post<<<1, 256, 0, ctx->stream>>>(st->dev, st->partials, (int)st->n_upd_part,
                                                st->log_dev);
        HIP_LAUNCH_CHECK();
    }
    int *err = (int *)((char *)st->dev + offsetof(PogmDev, err));
    static const char *sfx[] = {"@s1", "@s2", "@s3", "@s4", "@s5", "@s6", "@s7", "@s8"};
    auto words = [&](int64_t k) { return k == wk ? st->words : nullptr; };
    // serial scales without aux streams, and whenever kernels are being timed (so
    // each kernel's events bracket it alone)
    auto nuc_part = [&](int64_t k) { return st->log_costs ? st->nuc_part[k] : nullptr; };
    auto nuc_tot = [&](int64_t k) { return st->log_costs ? st->nuc_tot + k : nullptr; };
    if (st->aux.empty() || ctx->prof) {
        for (int64_t k = 0; k < st->S; ++k) {
            c64 *Xk = st->X + k * n;
            ctx->prof_suffix = k < 8 ? sfx[k] : "@s9+";
            prox_scale(ctx, st->geoms[k], Xk, Xk, st->thr_dev + k, err, nuc_part(k), nuc_tot(k),
                       words(k));
        }
        ctx->prof_suffix = "";
        log_reg_cost(st);
        return;
    }
    // Scales are independent: each runs on its own stream with its own scratch.
    // The 1x1x1 scales (one HBM pass, no scratch) go first on the main stream, so
    // their traffic does not overlap the other scales' gathers and Gram passes.
    hipStream_t main = ctx->stream;
    auto voxel_scale = [&](int64_t k) {
        const PatchGeom &g = st->geoms[k];
        return g.p == 1 && g.Np == (int64_t)g.N[0] * g.N[1] * g.N[2] && g.exact_tiling();
    };
    for (int64_t k = 0; k < st->S; ++k) {
        if (!voxel_scale(k)) continue;
        ctx->prof_suffix = k < 8 ? sfx[k] : "@s9+";
        prox_scale(ctx, st->geoms[k], st->X + k * n, st->X + k * n, st->thr_dev + k, err,
                   nuc_part(k), nuc_tot(k));
    }
    ctx->prof_suffix = "";
    HIP_CHECK(hipEventRecord(st->ev_fork, main));
    try {
        for (int64_t k = 0; k < st->S; ++k) {
            if (voxel_scale(k)) continue;
            c64 *Xk = st->X + k * n;
            HIP_CHECK(hipStreamWaitEvent(st->aux[k], st->ev_fork, 0));
            ctx->stream = st->aux[k];
            ctx->slot_base = (size_t)SLOT_COUNT * (size_t)(k + 1);
            ctx->prof_suffix = k < 8 ? sfx[k] : "@s9+";
            prox_scale(ctx, st->geoms[k], Xk, Xk, st->thr_dev + k, err, nuc_part(k), nuc_tot(k),
                       words(k));
            HIP_CHECK(hipEventRecord(st->ev_join[k], st->aux[k]));
        }
    } catch (...) {
        ctx->stream = main;
        ctx->slot_base = 0;
        ctx->prof_suffix = "";
        throw;
    }
    ctx->stream = main;
    ctx->slot_base = 0;
    ctx->prof_suffix = "";
    for (int64_t k = 0; k < st->S; ++k)
        if (!voxel_scale(k)) HIP_CHECK(hipStreamWaitEvent(main, st->ev_join[k], 0));
    log_reg_cost(st);
}

extern "C" {

mslr_status_t mslr_pogm_create(mslr_sense sense, const mslr_pogm_desc *desc, mslr_pogm *out) {
    return guarded([&] {
        MSLR_REQUIRE(out && sense && desc, MSLR_ERR_ARG, "null argument");
        *out = nullptr;
        sense->ctx->activate();
        const mslr_pogm_desc &d = *desc;
        MSLR_REQUIRE(d.mom == MSLR_MOM_POGM || d.mom == MSLR_MOM_FPGM || d.mom == MSLR_MOM_PGM,
                     MSLR_ERR_ARG, "mom must be pogm, fpgm or pgm");
        MSLR_REQUIRE(d.restart == MSLR_RESTART_GR || d.restart == MSLR_RESTART_NONE,
                     MSLR_ERR_ARG, "restart must be gr or none");
        MSLR_REQUIRE(d.L >= 0, MSLR_ERR_ARG, "f_L < 0");
        MSLR_REQUIRE(d.bsig >= 0, MSLR_ERR_ARG, "bsig < 0");
        MSLR_REQUIRE(std::fabs(d.restart_cutoff) < 1, MSLR_ERR_ARG, "restart_cutoff");
        MSLR_REQUIRE(d.mom == MSLR_MOM_POGM, MSLR_ERR_UNSUPPORTED,
                     "the device loop implements mom = pogm");
        MSLR_REQUIRE(d.bsig == 1.0, MSLR_ERR_UNSUPPORTED, "the device loop implements bsig = 1");
        MSLR_REQUIRE(d.L > 0, MSLR_ERR_ARG, "f_L must be > 0");
        MSLR_REQUIRE(d.niter >= 0, MSLR_ERR_ARG, "niter < 0");
        MSLR_REQUIRE(d.nscales >= 1 && d.patch_sizes && d.strides && d.lambdas, MSLR_ERR_ARG,
                     "scales");
        MSLR_REQUIRE(d.N[0] == sense->N[0] && d.N[1] == sense->N[1] && d.N[2] == sense->N[2] &&
                         d.Nt == sense->Nt,
                     MSLR_ERR_SHAPE, "image size differs from the SENSE plan");
        MSLR_REQUIRE(d.form == MSLR_FORM_PLAIN || d.form == MSLR_FORM_BATCHED ||
                         d.form == MSLR_FORM_MIXED,
                     MSLR_ERR_ARG, "form");
        auto *st = new mslr_pogm_s();
        try {
            st->sense = sense;
            st->ctx = sense->ctx;
            st->V = sense->V;
            st->Nt = sense->Nt;
            st->S = d.nscales;
            st->K = sense->K;
            st->Nc = sense->Nc;
            st->niter = d.niter;
            st->form = d.form;
            st->L = d.L;
            for (int64_t k = 0; k < d.nscales; ++k) {
                st->geoms.push_back(make_geom(d.N, d.Nt, d.patch_sizes + 3 * k, d.strides + 3 * k));
                st->lambdas.push_back(d.lambdas[k]);
            }
            const int64_t n = st->V * st->Nt;
            st->X = (c64 *)st->alloc((size_t)(n * st->S) * sizeof(c64));
            st->U = (c64 *)st->alloc((size_t)(n * st->S) * sizeof(c64));
            st->y = (c64 *)st->alloc((size_t)std::max<int64_t>(1, st->K * st->Nc * st->Nt) *
                                     sizeof(c64));
            st->dev = (PogmDev *)st->alloc(sizeof(PogmDev));
            st->lam_dev = (double *)st->alloc(sizeof(double) * st->S);
            st->thr_dev = (double *)st->alloc(sizeof(double) * st->S);
            {
                const int nch = sense->ctx->tuning.sense_chunks > 0 ? sense->ctx->tuning.sense_chunks
                                                                    : SENSE_CHUNKS_DEFAULT;
                st->gchunks = sense_grad_chunks(sense, nch);
            }
            // restart-test partials of the fused update: one triple per (frame, line block)
            st->upd_lines = sense->ctx->tuning.update_lines > 0
                                ? std::min(sense->ctx->tuning.update_lines, sense->NYZ)
                                : sense_transpose_lines(sense);
            MSLR_REQUIRE((size_t)sense->N[0] * (st->upd_lines + 1) * sizeof(c64) <= 160 * 1024 - 256,
                         MSLR_ERR_UNSUPPORTED, "update_lines: the Nx x lines tile exceeds the LDS");
            st->n_upd_part = cdiv(sense->NYZ, st->upd_lines) * (int64_t)st->Nt;
            st->partials = (double *)st->alloc(sizeof(double) * 3 * (size_t)st->n_upd_part);
            HIP_CHECK(hipFuncSetAttribute((const void *)k_planes_update,
                                          hipFuncAttributeMaxDynamicSharedMemorySize,
                                          160 * 1024 - 256));
            st->log_dev = (int *)st->alloc(sizeof(int) * std::max(1, d.niter));
            HIP_CHECK(hipMemcpy(st->lam_dev, st->lambdas.data(), sizeof(double) * st->S,
                                hipMemcpyHostToDevice));
            PogmDev h{};
            h.alpha = 1.0 / d.L;
            h.told = 1.0;
            h.sig = 1.0;
            h.tnew = 1.0;
            h.restart_cutoff = d.restart_cutoff;
            h.iter = 0;
            h.niter = d.niter;
            h.restart = d.restart;
            HIP_CHECK(hipMemcpy(st->dev, &h, sizeof(h), hipMemcpyHostToDevice));
            HIP_CHECK(hipMemset(st->log_dev, 0, sizeof(int) * std::max(1, d.niter)));
            MSLR_REQUIRE(d.log_costs == 0 || d.log_costs == 1, MSLR_ERR_ARG, "log_costs must be 0 or 1");
            st->log_costs = d.log_costs == 1;
            MSLR_REQUIRE(d.checkpoint == 0 || d.checkpoint == 1, MSLR_ERR_ARG,
                         "checkpoint must be 0 or 1");
            st->checkpoint = d.checkpoint == 1;
            if (st->checkpoint) {
                MSLR_REQUIRE((n * st->S) % 2 == 0, MSLR_ERR_UNSUPPORTED,
                             "checkpointing needs an even element count");
                // one 5D prox input per iteration: refuse what cannot fit rather than
                // failing inside hipMalloc
                const double need = (double)std::max(1, d.niter) * (double)(n * st->S) * sizeof(c64);
                size_t free_b = 0, total_b = 0;
                HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
                if (need > 0.9 * (double)free_b) {
                    char msg[256];
                    std::snprintf(msg, sizeof msg,
                                  "checkpoint=1 keeps niter = %d prox inputs of %.2f GB each "
                                  "(%.1f GB), more than the %.1f GB free on the device",
                                  d.niter, (double)(n * st->S) * sizeof(c64) / 1e9, need / 1e9,
                                  (double)free_b / 1e9);
                    MSLR_FAIL(MSLR_ERR_UNSUPPORTED, msg);
                }
                st->ckpt = (c64 *)st->alloc((size_t)std::max(1, d.niter) * (size_t)(n * st->S) *
                                            sizeof(c64));
                st->scal_log = (double *)st->alloc(sizeof(double) * 4 * (size_t)std::max(1, d.niter));
            }
            if (st->log_costs) {
                st->costs = (double *)st->alloc(sizeof(double) * 2 * (size_t)(d.niter + 1));
                st->n_dc_part = st->Nt * st->Nc * (int64_t)sense->nlb;
                st->dc_part = (double *)st->alloc(sizeof(double) * (size_t)st->n_dc_part);
                st->nuc_tot = (double *)st->alloc(sizeof(double) * (size_t)st->S);
                for (const PatchGeom &g : st->geoms)
                    st->nuc_part.push_back((double *)st->alloc(sizeof(double) * (size_t)nuc_units(g)));
            }
            if (d.form != MSLR_FORM_PLAIN && st->gchunks.size() > 1) {
                HIP_CHECK(hipStreamCreateWithFlags(&st->gpipe, hipStreamNonBlocking));
                HIP_CHECK(hipEventCreateWithFlags(&st->ev_gfork, hipEventDisableTiming));
                HIP_CHECK(hipEventCreateWithFlags(&st->ev_gjoin, hipEventDisableTiming));
            }
            if (d.form != MSLR_FORM_PLAIN && d.nscales > 1) {
                for (int64_t k = 0; k < d.nscales; ++k) {
                    hipStream_t s2 = nullptr;
                    HIP_CHECK(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking));
                    st->aux.push_back(s2);
                    hipEvent_t e = nullptr;
                    HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
                    st->ev_join.push_back(e);
                }
                HIP_CHECK(hipEventCreateWithFlags(&st->ev_fork, hipEventDisableTiming));
            }
            *out = st;
        } catch (...) {
            pogm_free(st);
            throw;
        }
    });
}

mslr_status_t mslr_pogm_destroy(mslr_pogm st) {
    return guarded([&] {
        if (!st) return;
        st->ctx->activate();
        HIP_CHECK(hipStreamSynchronize(st->ctx->stream));
        pogm_free(st);
    });
}

mslr_status_t mslr_pogm_init(mslr_pogm st, const void *y_dev, const void *X0_dev) {
    return guarded([&] {
        MSLR_REQUIRE(st && y_dev, MSLR_ERR_ARG, "null argument");
        mslr_ctx_s *ctx = st->ctx;
        ctx->activate();
        const int64_t n = st->V * st->Nt;
        sense_ref_to_internal(st->sense, (const c64 *)y_dev, st->y);
        if (X0_dev) {
            HIP_CHECK(hipMemcpyAsync(st->X, X0_dev, (size_t)(n * st->S) * sizeof(c64),
                                     hipMemcpyDeviceToDevice, ctx->stream));
        } else {
            sense_adjoint_internal(st->sense, st->y, st->X);
            fill_zero(ctx, st->X + n, n * (st->S - 1));
        }
        HIP_CHECK(hipMemcpyAsync(st->U, st->X, (size_t)(n * st->S) * sizeof(c64),
                                 hipMemcpyDeviceToDevice, ctx->stream));
        // Reset the scalar recurrence (told = sig = 1, src/mirt_mod.jl:100).
        PogmDev cur{};
        HIP_CHECK(hipMemcpyAsync(&cur, st->dev, sizeof(cur), hipMemcpyDeviceToHost, ctx->stream));
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
        cur.told = 1.0;
        cur.sig = 1.0;
        cur.tnew = 1.0;
        cur.iter = 0;
        cur.err = 0;
        HIP_CHECK(hipMemcpyAsync(st->dev, &cur, sizeof(cur), hipMemcpyHostToDevice, ctx->stream));
        if (st->log_costs) {
            // reg_cost(x0) (the logger's call at iteration 0, src/mirt_mod.jl:106)
            for (int64_t k = 0; k < st->S; ++k)
                scale_nucnorm(ctx, st->geoms[k], st->X + k * n, st->nuc_tot + k);
            log_reg_cost(st);
        }
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
        st->iters_done = 0;
        st->dc_final_at = -1;
        st->initialised = true;
    });
}

mslr_status_t mslr_pogm_run(mslr_pogm st, int n) {
    return guarded([&] {
        MSLR_REQUIRE(st, MSLR_ERR_ARG, "null state");
        MSLR_REQUIRE(st->initialised, MSLR_ERR_ARG, "mslr_pogm_init has not been called");
        MSLR_REQUIRE(n >= 0 && st->iters_done + n <= st->niter, MSLR_ERR_ARG,
                     "more iterations than niter");
        mslr_ctx_s *ctx = st->ctx;
        ctx->activate();
        for (int i = 0; i < n; ++i) {
            // kernel timing needs eager launches
            const bool use_graph = st->form != MSLR_FORM_PLAIN && !ctx->prof;
            // a scratch slot reallocated since the capture (by any call on the
            // context) leaves the graph pointing at freed memory
            if (st->exec && st->graph_gen != ctx->scratch_gen) drop_graph(st);
            if (use_graph && st->exec) {
                HIP_CHECK(hipGraphLaunch(st->exec, ctx->stream));
            } else if (use_graph && st->warm && st->warm_gen == ctx->scratch_gen) {
                // Scratch is sized by an eager iteration; capture the next.
                HIP_CHECK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
                try {
                    pogm_iteration(st);
                } catch (...) {
                    hipGraph_t gtmp = nullptr;
                    hipStreamEndCapture(ctx->stream, &gtmp);
                    if (gtmp) hipGraphDestroy(gtmp);
                    throw;
                }
                HIP_CHECK(hipStreamEndCapture(ctx->stream, &st->graph));
                HIP_CHECK(hipGraphInstantiate(&st->exec, st->graph, nullptr, nullptr, 0));
                st->graph_gen = ctx->scratch_gen;
                HIP_CHECK(hipGraphLaunch(st->exec, ctx->stream));
            } else {
                pogm_iteration(st);
                st->warm = true;
                st->warm_gen = ctx->scratch_gen;
            }
            st->iters_done++;
        }
    });
}

void *mslr_pogm_x(mslr_pogm st) { return st ? (void *)st->X : nullptr; }

mslr_status_t mslr_pogm_log(mslr_pogm st, int *iters_done, int *restarts) {
    return guarded([&] {
        MSLR_REQUIRE(st, MSLR_ERR_ARG, "null state");
        mslr_ctx_s *ctx = st->ctx;
        ctx->activate();
        PogmDev cur{};
        HIP_CHECK(hipMemcpyAsync(&cur, st->dev, sizeof(cur), hipMemcpyDeviceToHost, ctx->stream));
        if (restarts && st->iters_done > 0)
            HIP_CHECK(hipMemcpyAsync(restarts, st->log_dev, sizeof(int) * st->iters_done,
                                     hipMemcpyDeviceToHost, ctx->stream));
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
        if (iters_done) *iters_done = cur.iter;
        MSLR_REQUIRE(cur.err == 0, MSLR_ERR_NUMERIC, "non-finite values in a patch matrix");
    });
}

mslr_status_t mslr_pogm_costs(mslr_pogm st, int *n_out, double *dc_costs, double *reg_costs) {
    return guarded([&] {
        MSLR_REQUIRE(st && n_out, MSLR_ERR_ARG, "null argument");
        MSLR_REQUIRE(st->log_costs, MSLR_ERR_ARG, "the state was created with log_costs = 0");
        MSLR_REQUIRE(st->initialised, MSLR_ERR_ARG, "mslr_pogm_init has not been called");
        mslr_ctx_s *ctx = st->ctx;
        ctx->activate();
        const int k = st->iters_done;
        if (st->dc_final_at != k) {
            // dc_cost of the newest iterate: its residual has not been formed yet
            const int64_t n = st->V * st->Nt;
            sense_resid_internal(st->sense, st->X, (int)st->S, n, st->y, st->dc_part);
            MSLR_PROF(ctx, "cost_dc");
            k_cost_dc<<<1, 1024, 0, ctx->stream>>>(st->dc_part, st->n_dc_part, st->dev, st->costs);
            HIP_LAUNCH_CHECK();
            st->dc_final_at = k;
        }
        std::vector<double> h(2 * (size_t)(k + 1));
        HIP_CHECK(hipMemcpyAsync(h.data(), st->costs, sizeof(double) * h.size(),
                                 hipMemcpyDeviceToHost, ctx->stream));
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
        *n_out = k + 1;
        for (int i = 0; i <= k; ++i) {
            if (dc_costs) dc_costs[i] = h[2 * i];
            if (reg_costs) reg_costs[i] = h[2 * i + 1];
        }
    });
}

mslr_status_t mslr_pogm_lambda_vjp(mslr_pogm st, const void *xbar_dev, double *lambda_bar) {
    return guarded([&] {
        MSLR_REQUIRE(st && xbar_dev && lambda_bar, MSLR_ERR_ARG, "null argument");
        MSLR_REQUIRE(st->checkpoint, MSLR_ERR_ARG, "the state was created with checkpoint = 0");
        MSLR_REQUIRE(st->initialised, MSLR_ERR_ARG, "mslr_pogm_init has not been called");
        mslr_ctx_s *ctx = st->ctx;
        ctx->activate();
        const int T = st->iters_done;
        const int64_t S = st->S, n = st->V * st->Nt;
        for (int64_t k = 0; k < S; ++k) lambda_bar[k] = 0.0;
        if (T == 0) return;
        std::vector<double> sc(4 * (size_t)T);
        HIP_CHECK(hipMemcpyAsync(sc.data(), st->scal_log, sizeof(double) * sc.size(),
                                 hipMemcpyDeviceToHost, ctx->stream));
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
        int64_t npmax = 1;
        for (const PatchGeom &g : st->geoms) npmax = std::max<int64_t>(npmax, g.Np);
        // backward state (allocated once per call; the forward's buffers stay intact)
        std::vector<void *> tmp;
        auto dalloc = [&](size_t b) {
            void *p = nullptr;
            HIP_CHECK(hipMalloc(&p, std::max<size_t>(b, 256)));
            tmp.push_back(p);
            return p;
        };
        try {
            c64 *xb = (c64 *)dalloc((size_t)(n * S) * sizeof(c64));
            c64 *ub = (c64 *)dalloc((size_t)(n * S) * sizeof(c64));
            c64 *zb = (c64 *)dalloc((size_t)(n * S) * sizeof(c64));
            c64 *gs = (c64 *)dalloc((size_t)n * sizeof(c64));
            c64 *h = (c64 *)dalloc((size_t)n * sizeof(c64));
            c64 *timg = (c64 *)dalloc((size_t)n * sizeof(c64));
            c64 *y0 = (c64 *)dalloc((size_t)std::max<int64_t>(1, st->K * st->Nc * st->Nt) *
                                    sizeof(c64));
            double *bb = (double *)dalloc(sizeof(double) * (size_t)npmax);
            double *lb = (double *)dalloc(sizeof(double) * (size_t)S);
            double *thr = (double *)dalloc(sizeof(double) * (size_t)S);
            int *err = (int *)dalloc(sizeof(int));
            fill_zero(ctx, y0, st->K * st->Nc * st->Nt);
            fill_zero(ctx, ub, n * S);
            HIP_CHECK(hipMemsetAsync(lb, 0, sizeof(double) * S, ctx->stream));
            HIP_CHECK(hipMemsetAsync(err, 0, sizeof(int), ctx->stream));
            for (int64_t k = 0; k < S; ++k)
                HIP_CHECK(hipMemcpyAsync(xb + k * n, xbar_dev, (size_t)n * sizeof(c64),
                                         hipMemcpyDeviceToDevice, ctx->stream));
            const unsigned nbl = (unsigned)std::min<int64_t>(cdiv(n, 256), 1 << 16);
            std::vector<double> thr_h((size_t)S);
            for (int it = T; it >= 1; --it) {
                const double alpha = sc[4 * (it - 1)], beta = sc[4 * (it - 1) + 1];
                const double gamma = sc[4 * (it - 1) + 2], zeta = sc[4 * (it - 1) + 3];
                for (int64_t k = 0; k < S; ++k) thr_h[k] = zeta * st->lambdas[k];
                HIP_CHECK(hipMemcpyAsync(thr, thr_h.data(), sizeof(double) * S,
                                         hipMemcpyHostToDevice, ctx->stream));
                const c64 *zk = st->ckpt + (int64_t)(it - 1) * n * S;
                for (int64_t k = 0; k < S; ++k) {
                    prox_vjp_scale(ctx, st->geoms[k], zk + k * n, xb + k * n, zb + k * n, thr + k, bb,
                                   err, timg);
                    MSLR_PROF(ctx, "vjp_lambda");
                    k_vjp_lambda<<<1, 256, 0, ctx->stream>>>(bb, st->geoms[k].Np, zeta, lb + k);
                    HIP_LAUNCH_CHECK();
                }
                {
                    MSLR_PROF(ctx, "vjp_state");
                    k_vjp_state<<<nbl, 256, 0, ctx->stream>>>(zb, ub, xb, gs, n, (int)S,
                                                              (float)alpha, (float)beta,
                                                              (float)gamma);
                    HIP_LAUNCH_CHECK();
                }
                // h = A'A gs (the gradient map's derivative, the same for every scale)
                sense_grad_internal(st->sense, gs, 1, n, y0, h, nullptr);
                {
                    MSLR_PROF(ctx, "vjp_state");
                    k_vjp_add<<<nbl, 256, 0, ctx->stream>>>(xb, h, n, (int)S);
                    HIP_LAUNCH_CHECK();
                }
            }
            int herr = 0;
            HIP_CHECK(hipMemcpyAsync(lambda_bar, lb, sizeof(double) * S, hipMemcpyDeviceToHost,
                                     ctx->stream));
            HIP_CHECK(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
            HIP_CHECK(hipStreamSynchronize(ctx->stream));
            for (void *p : tmp) hipFree(p);
            tmp.clear();
            MSLR_REQUIRE(herr == 0, MSLR_ERR_NUMERIC, "non-finite values in a patch matrix");
        } catch (...) {
            hipStreamSynchronize(ctx->stream);
            for (void *p : tmp) hipFree(p);
            throw;
        }
    });
}

mslr_status_t mslr_pogm_thresholds(mslr_pogm st, double *beta) {
    return guarded([&] {
        MSLR_REQUIRE(st && beta, MSLR_ERR_ARG, "null argument");
        st->ctx->activate();
        HIP_CHECK(hipMemcpyAsync(beta, st->thr_dev, sizeof(double) * st->S, hipMemcpyDeviceToHost,
                                 st->ctx->stream));
        HIP_CHECK(hipStreamSynchronize(st->ctx->stream));
    });
}

mslr_status_t mslr_prox_scales(mslr_ctx ctx, void *X_dev, const int64_t N[3], int64_t Nt,
                               int64_t nscales, const int64_t *patch_sizes, const int64_t *strides,
                               const double *beta_host) {
    return guarded([&] {
        mslr_ctx_s *c = check_ctx(ctx);
        MSLR_REQUIRE(X_dev && N && patch_sizes && strides && beta_host, MSLR_ERR_ARG,
                     "null argument");
        MSLR_REQUIRE(nscales >= 1, MSLR_ERR_ARG, "nscales must be >= 1");
        std::vector<PatchGeom> gs;
        for (int64_t k = 0; k < nscales; ++k) {
            MSLR_REQUIRE(beta_host[k] >= 0 && std::isfinite(beta_host[k]), MSLR_ERR_ARG,
                         "beta must be finite and >= 0");
            gs.push_back(make_geom(N, Nt, patch_sizes + 3 * k, strides + 3 * k));
        }
        const int64_t n = N[0] * N[1] * N[2] * Nt;
        double *bd = (double *)c->scratch_get(SLOT_APPLY, (size_t)(nscales + 1) * sizeof(double) + 64);
        int *err = (int *)(bd + nscales);
        HIP_CHECK(hipMemcpyAsync(bd, beta_host, sizeof(double) * nscales, hipMemcpyHostToDevice,
                                 c->stream));
        HIP_CHECK(hipMemsetAsync(err, 0, sizeof(int), c->stream));
        for (int64_t k = 0; k < nscales; ++k) {
            c64 *Xk = (c64 *)X_dev + k * n;
            prox_scale(c, gs[k], Xk, Xk, bd + k, err, nullptr, nullptr);
        }
        int herr = 0;
        HIP_CHECK(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIP_CHECK(hipStreamSynchronize(c->stream));
        MSLR_REQUIRE(herr == 0, MSLR_ERR_NUMERIC, "non-finite values in a patch matrix");
    });
}

mslr_status_t mslr_patch_svst(mslr_ctx ctx, const void *img_dev, void *out_dev,
                              const int64_t N[3], int64_t Nt, const int64_t ps[3],
                              const int64_t ss[3], double beta) {
    return guarded([&] {
        mslr_ctx_s *c = check_ctx(ctx);
        MSLR_REQUIRE(beta >= 0 && std::isfinite(beta), MSLR_ERR_ARG, "beta must be >= 0");
        PatchGeom g = make_geom(N, Nt, ps, ss);
        double *bd = (double *)c->scratch_get(SLOT_APPLY, 64);
        int *err = (int *)(bd + 1);
        HIP_CHECK(hipMemcpyAsync(bd, &beta, sizeof(double), hipMemcpyHostToDevice, c->stream));
        HIP_CHECK(hipMemsetAsync(err, 0, sizeof(int), c->stream));
        prox_scale(c, g, (const c64 *)img_dev, (c64 *)out_dev, bd, err, nullptr, nullptr);
        int herr = 0;
        HIP_CHECK(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIP_CHECK(hipStreamSynchronize(c->stream));
        MSLR_REQUIRE(herr == 0, MSLR_ERR_NUMERIC, "non-finite values in a patch matrix");
    });
}

}  // extern "C"
