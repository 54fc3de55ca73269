/*
 * mslr.h - C ABI of the MI355X (gfx950) MSLR reconstruction library.
 *
 * The library accelerates one POGM iteration of multi-scale low-rank (MSLR)
 * 3D+time MRI reconstruction: SENSE forward/adjoint, per-scale patch singular
 * value soft-thresholding (SVST) and the POGM momentum/restart update.
 *
 * Conventions (all entry points):
 *   - Arrays are column-major (Julia order, x fastest).  Complex values are
 *     interleaved (re, im) pairs: ComplexF32 = 2 x float, ComplexF64 = 2 x double.
 *   - Pointers named *_dev are device pointers obtained from mslr_malloc (or any
 *     hipMalloc'd memory on the context's device); *_host are host pointers.
 *   - Scalars are double.  Sizes are int64_t.
 *   - Every call returns an mslr_status_t; on failure mslr_last_error() returns
 *     a message for the calling thread.  Status codes map to the exception kinds
 *     the reference throws: MSLR_ERR_ARG ~ ArgumentError, MSLR_ERR_SHAPE ~
 *     AssertionError/DimensionMismatch, MSLR_ERR_DEVICE ~ error("use_gpu=true
 *     requires ...").
 *   - A context owns one HIP stream.  All work of a context is enqueued on that
 *     stream in call order; calls on one context must be serialised by the
 *     caller (like the reference GPU path, which is single-stream).  Entry
 *     points that return host-visible results synchronise the stream.
 *   - The library never frees memory it does not own.
 *
 * Reference interfaces replaced (rextlfung/mslr-recon, file:line):
 *   Asense_gpu / MIRT.Asense          src/sense_gpu.jl:56-108, scripts/reconstruct.jl:135-142
 *   block_diag over frames + flatten   scripts/reconstruct.jl:142,145-147
 *   img2patches / patches2img          src/recon.jl:53-78, 96-126
 *   SVST / patchSVST / patchSVST_gpu   src/recon.jl:183-269, src/sense_gpu.jl:122-160
 *   patch_nucnorm                      src/recon.jl:138-146
 *   pogm_mod (+ _gr_restart)           src/mirt_mod.jl:26-29, 74-188
 *   poweriter_mod                      src/mirt_mod.jl:215-238
 *   dc_cost_grad / g_prox / lambdas    scripts/reconstruct.jl:159-202
 */
#ifndef MSLR_H
#define MSLR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    MSLR_OK = 0,
    MSLR_ERR_ARG = 1,         /* invalid argument value (ArgumentError) */
    MSLR_ERR_SHAPE = 2,       /* shape / mask / sample-count mismatch (AssertionError) */
    MSLR_ERR_DEVICE = 3,      /* HIP runtime error or no usable GPU */
    MSLR_ERR_NOMEM = 4,       /* device allocation failed */
    MSLR_ERR_UNSUPPORTED = 5, /* size outside what the kernels support (e.g. FFT length with a prime factor > 5) */
    MSLR_ERR_NUMERIC = 6      /* non-finite values or eigensolver failure */
} mslr_status_t;

typedef struct mslr_ctx_s *mslr_ctx;
typedef struct mslr_sense_s *mslr_sense;
typedef struct mslr_pogm_s *mslr_pogm;

/* ---------------------------------------------------------------- library */

/* Message of the last failed call on this thread ("" if none). */
const char *mslr_last_error(void);
/* ABI version (major*10000 + minor*100 + patch). */
int mslr_version(void);

/* --------------------------------------------------------- instrumentation */
/* No counterpart in the reference.  While enabled, every kernel launch of the
 * context is bracketed by HIP events on the context stream (the POGM loop then
 * runs without hipGraph capture).  Enabling or disabling clears the records. */
mslr_status_t mslr_prof_enable(mslr_ctx ctx, int enable);
/* Per-kernel lines "name launches total_ms\n" of the records so far (synchronises). */
mslr_status_t mslr_prof_report(mslr_ctx ctx, char *buf, size_t len);

/* ---------------------------------------------------------------- context */

/* Replaces the reference's CUDA.functional() probe and device selection
 * (scripts/reconstruct.jl:45-55, 73-81).  device = HIP device ordinal. */
mslr_status_t mslr_ctx_create(int device, mslr_ctx *out);
mslr_status_t mslr_ctx_destroy(mslr_ctx ctx);
mslr_status_t mslr_ctx_sync(mslr_ctx ctx);
/* Device name and total / free memory in bytes (println at scripts/reconstruct.jl:78-80). */
mslr_status_t mslr_ctx_info(mslr_ctx ctx, char *name, size_t name_len,
                            size_t *total_bytes, size_t *free_bytes);

/* Products of the SVST Gram matrix and reconstruction GEMM on this context
 * (mslr_svst_batched, mslr_patch_svst).  No counterpart in the reference, whose
 * patch SVD runs in FP32 LAPACK (src/recon.jl:183-198): this selects the
 * mixed-precision form.  FP32 (default): FP64 Gram products of the FP32 data,
 * FP32 reconstruction.  SPLIT_BF16: each FP32 operand as hi + lo BF16 on the
 * matrix cores (hi*hi + hi*lo + lo*hi, FP32 accumulation; Gram sums in FP64);
 * the eigensolve stays FP64.  The POGM MSLR_FORM_MIXED form uses SPLIT_BF16 for
 * its prox whatever this setting is. */
#define MSLR_PRODUCTS_FP32       0
#define MSLR_PRODUCTS_SPLIT_BF16 1
mslr_status_t mslr_ctx_set_products(mslr_ctx ctx, int products);

/* Launch shapes on this context.  No counterpart in the reference (its SVD and
 * FFTs are library calls); set by the autotuner
 * (mslr_hip.autotune) or by hand.  0 in a field keeps the built-in choice.
 * Read when work is enqueued: a POGM state's captured graph keeps the shapes it
 * was captured with, so create the state after setting them.
 *   gram_chunk:      contraction rows per Gram workgroup (built-in 4096); a
 *                    multiple of 256 in [256, 65536].  Partial Grams of the chunks
 *                    are summed in FP64, so the choice changes rounding only.
 *   eig_lanes_limit: batches with patches x m at or below this use 16 lanes per
 *                    eigenvector, larger ones 4 (built-in 65536); -1 = always 4.
 *   svst_apply:      1 = form W for every patch (dense apply); 0 = low-rank
 *                    products X U_r for patches with 2 r <= m (built-in; split BF16:
 *                    those keeping r <= 64 in one kernel holding X U_r on chip);
 *                    2 = the low-rank products always as two launches (X U_r through
 *                    memory; the same bits as 0).
 *   sense_chunks:    frame groups of the POGM gradient (built-in 1, at most 64):
 *                    A and A' run group by group and each group's update follows its
 *                    adjoint; the batched forms alternate groups between two streams
 *                    so the groups' memory-bound and FFT-bound kernels overlap.
 *                    Changes the summation order of the restart test's FP64 sums only.
 *   yz_threads_fwd / yz_threads_adj: threads of the compiled y-z FFT passes of
 *                    the 90 x 60 plane, 512, 768 or 1024 (built-in 768 both ways).
 *                    Results are bit-identical.
 *   update_lines:    yz-lines per workgroup of the POGM update, which is fused with
 *                    the gradient's last transpose (plane-major -> image order): 0 =
 *                    built-in 64 (fewer where an Nx x 64 tile exceeds the LDS), else a
 *                    multiple of 8 in [8, 64]; read at mslr_pogm_create.  Changes the
 *                    grouping of the restart test's FP64 sums only.
 *   sb2st_schedule:  band -> tridiagonal kernel for Gram matrices above 300 (band in
 *                    global memory): 0 = 12 waves starting sweeps 3 time steps apart
 *                    (built-in), 1 = 16 waves 2 apart, 2 = 6 waves 5 apart, 3 = 4 waves
 *                    8 apart.  Results are bit-identical.
 *   apply_fp32:      with split-BF16 products, which SVST reconstruction products run
 *                    in FP32 on the matrix cores instead: 0 = built-in, else 1 + a
 *                    mask of bit 0 (the dense Y = X W), bit 1 (T = X U_r) and bit 2
 *                    (Y = T f U_r^H); 1 = all split BF16, 8 = all FP32.  Changes
 *                    rounding only; the FP32 products form ignores it.
 * Invalid values -> MSLR_ERR_ARG. */
typedef struct mslr_tuning {
    int gram_chunk;
    int eig_lanes_limit;
    int svst_apply;
    int sense_chunks;
    int yz_threads_fwd;
    int yz_threads_adj;
    int update_lines;
    int sb2st_schedule;
    int apply_fp32;
    int reserved[3];
} mslr_tuning;
mslr_status_t mslr_ctx_set_tuning(mslr_ctx ctx, const mslr_tuning *t);
mslr_status_t mslr_ctx_get_tuning(mslr_ctx ctx, mslr_tuning *t);

/* ------------------------------------------------------- device buffers */
/* Replaces cu(...) / Array(...) transfers (scripts/reconstruct.jl:122-130, 233-238). */
mslr_status_t mslr_malloc(mslr_ctx ctx, size_t bytes, void **dev_ptr);
mslr_status_t mslr_free(mslr_ctx ctx, void *dev_ptr);
mslr_status_t mslr_upload(mslr_ctx ctx, void *dst_dev, const void *src_host, size_t bytes);
mslr_status_t mslr_download(mslr_ctx ctx, void *dst_host, const void *src_dev, size_t bytes);
mslr_status_t mslr_copy_d2d(mslr_ctx ctx, void *dst_dev, const void *src_dev, size_t bytes);
mslr_status_t mslr_zero(mslr_ctx ctx, void *dev_ptr, size_t bytes);

/* ------------------------------------------------------ SENSE operator */
/*
 * All-frames SENSE plan: the block-diagonal stack of per-frame Asense operators
 *   A = block_diag([Asense(mask[:,:,:,t], smaps) for t in 1:Nt]...)
 * (scripts/reconstruct.jl:135-142; per-frame operator src/sense_gpu.jl:56-108).
 *
 *   forward : y[:, c, t] = scale * (fftshift o FFT3 o ifftshift)(smaps[...,c] .* x[...,t])[mask_t]
 *   adjoint : the exact conjugate transpose of forward (unnormalised inverse DFT
 *             times the same scale).  src/sense_gpu.jl:93 applies a normalised
 *             ifft instead, which is 1/prod(N) times this adjoint.
 *   scale   = 1/sqrt(Nx*Ny*Nz) if unitary else 1.
 *   fft_forward = 0 uses the positive-exponent DFT normalised by 1/prod(N) in the
 *             forward (fftshift o ifft o ifftshift, src/sense_gpu.jl:77); the adjoint
 *             stays the exact conjugate transpose.
 *
 * smaps_host : ComplexF32 (Nx,Ny,Nz,Nc), used as given (normalise before, as
 *              scripts/reconstruct.jl:89 does).
 * mask_host  : uint8 (Nx,Ny,Nz,Nt), nonzero = sampled.  Every frame must have
 *              the same sample count K (scripts/reconstruct.jl:116-118), else
 *              MSLR_ERR_SHAPE.
 * k-space layout (K, Nc, Nt): sample k of frame t is the k-th sampled voxel of
 * mask_t in column-major order (findall(vec(samp)), src/sense_gpu.jl:68).
 * Each of Nx, Ny, Nz must factor into 2, 3 and 5 and Ny*Nz <= 8192.
 */
mslr_status_t mslr_sense_create(mslr_ctx ctx, const int64_t N[3], int64_t Nc, int64_t Nt,
                                const float *smaps_host, const uint8_t *mask_host,
                                int fft_forward, int unitary, mslr_sense *out);
mslr_status_t mslr_sense_destroy(mslr_sense plan);
/* Samples per frame per coil (K = sum(samp), src/sense_gpu.jl:63). */
int64_t mslr_sense_nsamples(mslr_sense plan);

/* x_dev: ComplexF32 (Nx,Ny,Nz,nframes) holding frames t0..t0+nframes-1;
 * y_dev: ComplexF32 (K,Nc,nframes).  Frames are 0-based. */
mslr_status_t mslr_sense_forward(mslr_sense plan, const void *x_dev, void *y_dev,
                                 int64_t t0, int64_t nframes);
mslr_status_t mslr_sense_adjoint(mslr_sense plan, const void *y_dev, void *x_dev,
                                 int64_t t0, int64_t nframes);
/* mslr_sense_forward without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments and
 * result as mslr_sense_forward (the same kernels, so bit-identical output); the
 * intermediates live in a 64 MiB array of the library's, one per device, set
 * aside when the library is loaded on that device and held until the process
 * ends.  Frames run in chunks that fit it; a frame needing more than 64 MiB
 * (its V ComplexF32 image plus Nc*Nx*lines sampled-line values) gives
 * MSLR_ERR_UNSUPPORTED before any work is queued.  All calls on one device share
 * the array: calls that overlap on one device (other contexts or threads) race. */
mslr_status_t mslr_sense_forward_noalloc(mslr_sense plan, const void *x_dev, void *y_dev,
                                         int64_t t0, int64_t nframes);
/* The same in ComplexF64 (used by the FP64 power iteration). */
mslr_status_t mslr_sense_forward_f64(mslr_sense plan, const void *x_dev, void *y_dev,
                                     int64_t t0, int64_t nframes);
mslr_status_t mslr_sense_adjoint_f64(mslr_sense plan, const void *y_dev, void *x_dev,
                                     int64_t t0, int64_t nframes);
/* Forward residual r = A sum_k X_k - y (reference sample order) and rnorm2 = ||r||^2
 * (FP64 sum of FP32 values), the data term of dc_cost (scripts/reconstruct.jl:176-178).
 * r_dev: device (K, Nc, Nt), or NULL to keep only the norm. */
mslr_status_t mslr_sense_residual(mslr_sense plan, const void *X_dev, int64_t nscales,
                                  const void *y_dev, void *r_dev, double *rnorm2);
/*
 * Gradient of the data-consistency term on the scale sum
 * (dc_cost_grad, scripts/reconstruct.jl:174,176-183):
 *   s = sum_k X[:,:,:,:,k];  r = A s - y;  g = A' r;  dc = 0.5*||r||^2
 * X_dev: ComplexF32 (Nx,Ny,Nz,Nt,nscales); y_dev: (K,Nc,Nt); g_dev: (Nx,Ny,Nz,Nt).
 * dc_cost (host, nullable) receives 0.5*||r||^2 accumulated in FP64.
 */
mslr_status_t mslr_sense_grad(mslr_sense plan, const void *X_dev, int64_t nscales,
                              const void *y_dev, void *g_dev, double *dc_cost);
/* sigma_1(A) by power iteration in FP64 (poweriter_mod, src/mirt_mod.jl:215-238):
 * x0 = ones, at most niter steps, stop when |ratio-ratio_old|/ratio < tol.
 * Returns sigma1 = ||A x||/||x|| and the number of steps taken. */
mslr_status_t mslr_sense_poweriter(mslr_sense plan, int niter, double tol,
                                   double *sigma1, int *iters);

/* ------------------------------------------------------- patch operators */
/*
 * Patch geometry (src/recon.jl:53-78): patch size clamped to the image, stride s,
 * Nsteps = cld(N - ps, s), start_i = min(i*s, N - ps) (0-based), patches ordered
 * x fastest, voxels within a patch column-major.  Np = prod(Nsteps+1).
 */
mslr_status_t mslr_patch_geometry(const int64_t N[3], const int64_t ps[3], const int64_t ss[3],
                                  int64_t ps_clamped[3], int64_t nsteps[3], int64_t *Np);
/* img_dev: ComplexF32 (Nx,Ny,Nz,Nt) -> P_dev: ComplexF32 (p,Nt,Np). */
mslr_status_t mslr_img2patches(mslr_ctx ctx, const void *img_dev, const int64_t N[3], int64_t Nt,
                               const int64_t ps[3], const int64_t ss[3], void *P_dev);
/* P_dev (p,Nt,Np) -> img_dev (Nx,Ny,Nz,Nt): patch sum / max(coverage,1), summed in
 * patch order (src/recon.jl:96-126). */
mslr_status_t mslr_patches2img(mslr_ctx ctx, const void *P_dev, const int64_t N[3], int64_t Nt,
                               const int64_t ps[3], const int64_t ss[3], void *img_dev);

/* ------------------------------------------------------------------ SVST */
/*
 * Batched SVST (src/recon.jl:183-198, 217-241): for each of the Np matrices
 * X_i (p x Nt, ComplexF32, stored consecutively) in place
 *   X_i <- U diag(max(s - beta_i, 0)) V^H,     X_i == 0 -> 0.
 * beta_host: one value (per_patch = 0) or Np values (per_patch = 1).  Each
 * beta is rounded to Float32 first (src/recon.jl:195).  beta < 0 -> MSLR_ERR_ARG.
 * Method: Gram matrix (FP64 products of the FP32 data, or split-BF16 products
 * with FP64 sums, see mslr_ctx_set_products), FP64 Hermitian eigensolve,
 * reconstruction X_i V_r diag(f) V_r^H.
 */
mslr_status_t mslr_svst_batched(mslr_ctx ctx, void *P_dev, int64_t p, int64_t Nt, int64_t Np,
                                const double *beta_host, int per_patch);
/* Vector-Jacobian product of the batched SVST above (no counterpart in the
 * reference; the SVST pullback an rrule of SVST, src/recon.jl:183-205, would call):
 * for Y_i = SVST(Z_i, beta_i) and cotangents Ybar_i, writes Zbar_i (device, distinct
 * from Z and Ybar) and betabar = d<Ybar, Y>/d beta (host: Np values with per_patch,
 * else their sum).  Z, Ybar, Zbar: ComplexF32 (p, Nt, Np), any p, Nt >= 1.  Uses
 * every eigenpair of the Gram matrix Z_i^H Z_i (FP64; for p < Nt that of Z_i Z_i^H,
 * through Y^H = SVST(Z^H)) and safeguarded divided differences of
 * g(l) = max(1 - beta / sqrt(l), 0) (no 1/(s_i^2 - s_j^2) is formed); a single row or
 * column takes the group shrink's closed-form derivative. */
mslr_status_t mslr_svst_vjp(mslr_ctx ctx, const void *Z_dev, const void *Ybar_dev, void *Zbar_dev,
                            int64_t p, int64_t Nt, int64_t Np, const double *beta_host,
                            int per_patch, double *betabar_host);
/* Sum over patches of the nuclear norms (patch_nucnorm, src/recon.jl:138-146). */
mslr_status_t mslr_patch_nucnorm(mslr_ctx ctx, const void *P_dev, int64_t p, int64_t Nt,
                                 int64_t Np, double *out);
/* Weighted form (patch_nucnorm(P, lambdas), src/recon.jl:157-166): the mean over the
 * Np patches of lambdas_host[i] * ||P_i||_*, a patch with no nonzero singular value
 * adding 0; all patches' nuclear norms in one batched pass.  Np >= 1. */
mslr_status_t mslr_patch_nucnorm_weighted(mslr_ctx ctx, const void *P_dev, int64_t p,
                                          int64_t Nt, int64_t Np, const double *lambdas_host,
                                          double *out);
/* The in-place prox of run_recon over all scales (g_prox, scripts/reconstruct.jl:196-202):
 * X[..., k] <- patchSVST(X[..., k], beta[k], patch_sizes[k], strides[k]) for the 5D
 * ComplexF32 array X (N[0], N[1], N[2], Nt, nscales); beta_host[k] = c * lambda_k.
 * Scales one after another on the context stream. */
mslr_status_t mslr_prox_scales(mslr_ctx ctx, void *X_dev, const int64_t N[3], int64_t Nt,
                               int64_t nscales, const int64_t *patch_sizes, const int64_t *strides,
                               const double *beta_host);
/* patchSVST (src/recon.jl:251-256): img2patches -> SVST -> patches2img. */
mslr_status_t mslr_patch_svst(mslr_ctx ctx, const void *img_dev, void *out_dev,
                              const int64_t N[3], int64_t Nt, const int64_t ps[3],
                              const int64_t ss[3], double beta);

/* ------------------------------------------------------------------ setup */
/* The setup of run_recon (scripts/reconstruct.jl:88-118,145-147) on the device.
 *
 * smaps <- smaps ./ (sqrt.(sum(abs2.(smaps); dims=4)) .+ eps(Float32)) in place
 * (scripts/reconstruct.jl:89): ComplexF32 (Nx,Ny,Nz,Nc), FP32 throughout, the coils
 * summed in order.  Asynchronous. */
mslr_status_t mslr_normalise_smaps(mslr_ctx ctx, void *smaps_dev, const int64_t N[3], int64_t Nc);
/* Mask inference, validation and flatten (scripts/reconstruct.jl:109-118,145-147) of the
 * zero-filled k-space ksp0_dev, ComplexF32 (Nx,Ny,Nz,Nvc,Nt):
 *   mask_dev (uint8 (Nx,Ny,Nz,Nt), device, written) = ksp0[:,:,:,1,:] .!= 0;
 *   MSLR_ERR_SHAPE "Coil c has a different sampling pattern" for the first coil whose
 *   nonzeros differ from coil 1's, then "Frame t has a different sample count" for the
 *   first frame whose count differs from the previous frame's (1-based, as the
 *   reference's @assert messages);
 *   *K = samples per frame; y_dev (nullable; y_capacity values) receives (K, Nvc, Nt),
 *   ksp[:, :, t] = reshape(ksp0, :, Nvc, Nt)[vec(mask_t), :, t].
 * Synchronises (the assertions are decided on the host). */
mslr_status_t mslr_kspace_flatten(mslr_ctx ctx, const void *ksp0_dev, const int64_t N[3],
                                  int64_t Nvc, int64_t Nt, uint8_t *mask_dev, void *y_dev,
                                  int64_t y_capacity, int64_t *K);
/* Scale factor of run_recon: quantile(vec(abs.(sense_comb(ksp0, smaps))), q), type 7
 * (scripts/reconstruct.jl:102-103; sense_comb src/recon.jl:312-323), from the flattened
 * k-space y_dev (K, Nc, Nt) and the plan's maps: sense_comb = A'y / (sqrt(V) (sum_c |S_c|^2
 * + eps)) with A' the plan's exact adjoint (any 2/3/5 size, odd ones included); the order
 * statistics by a radix select on the device.  Synchronises. */
mslr_status_t mslr_sense_scale_factor(mslr_sense plan, const void *y_dev, double q, double *sf);
/* y <- y ./ max(sf, eps(Float32)) in place for n ComplexF32 values (the division in FP64,
 * as ComplexF32 ./= Float64; scripts/reconstruct.jl:104).  Asynchronous. */
mslr_status_t mslr_scale_kspace(mslr_ctx ctx, void *y_dev, int64_t n, double sf);

/* ------------------------------------------------------------------ analysis */
/* tSNR(img) (src/analysis.jl:39-44): out[v] = mean_t |img[v,t]| / (std_t |img[v,t]| +
 * eps(Float32)), std with n - 1, for img ComplexF32 (V, Nt) column-major (V = Nx Ny Nz);
 * out: Float32 (V), device.  FP64 sums.  Asynchronous on the context stream. */
mslr_status_t mslr_tsnr(mslr_ctx ctx, const void *img_dev, int64_t V, int64_t Nt, void *out_dev);
/* detrend!(img) (src/analysis.jl:122-132) in place: img[v,t] -= t * b_v, t = 1..Nt,
 * b_v the least-squares slope of img[v,:] against [1 t].  Asynchronous. */
mslr_status_t mslr_detrend(mslr_ctx ctx, void *img_dev, int64_t V, int64_t Nt);

/* ------------------------------------------------------------------ POGM */
/* Forms of the POGM iteration. */
#define MSLR_FORM_PLAIN   0  /* per-scale launches, FP32 Gram products */
#define MSLR_FORM_BATCHED 1  /* all scales per launch + hipGraph-captured iteration */
#define MSLR_FORM_MIXED   2  /* BATCHED with split-BF16 Gram products (FP32 accumulate) */

/* Momentum / restart options of pogm_mod (src/mirt_mod.jl:74-93). */
#define MSLR_MOM_POGM 0
#define MSLR_MOM_FPGM 1
#define MSLR_MOM_PGM  2
#define MSLR_RESTART_GR   0
#define MSLR_RESTART_NONE 1

typedef struct {
    int64_t N[3];
    int64_t Nt;
    int64_t nscales;
    const int64_t *patch_sizes; /* nscales x 3 */
    const int64_t *strides;     /* nscales x 3 */
    const double *lambdas;      /* nscales: regularisation weights lambda_k */
    double L;                   /* Lipschitz constant f_L (nscales * sigma1^2) */
    int niter;                  /* total iterations of the run (the last one uses the 8*t^2 rule) */
    int mom;                    /* MSLR_MOM_* (only POGM is implemented on the device) */
    int restart;                /* MSLR_RESTART_* */
    double restart_cutoff;      /* must satisfy |restart_cutoff| < 1 */
    double bsig;                /* gamma decrease factor, >= 0 */
    int form;                   /* MSLR_FORM_* */
    int log_costs;              /* 1: keep dc_cost and reg_cost of every iterate (the logger) */
    int checkpoint;             /* 1: keep each iteration's prox input for mslr_pogm_lambda_vjp */
} mslr_pogm_desc;

/*
 * POGM state on the device for the aliased in-place prox of the reference
 * (scripts/reconstruct.jl:196-202 with src/mirt_mod.jl:131-184).  Holds
 * X (Nx,Ny,Nz,Nt,nscales) and U_old, the per-iteration scalars (t, sigma,
 * zeta) and a per-iteration restart log.
 */
mslr_status_t mslr_pogm_create(mslr_sense sense, const mslr_pogm_desc *desc, mslr_pogm *out);
mslr_status_t mslr_pogm_destroy(mslr_pogm st);
/* Set y (ComplexF32 (K,Nc,Nt), reference sample order) and X0.  X0_dev == NULL
 * means X0[...,1] = A' y and zeros elsewhere (scripts/reconstruct.jl:209-214). */
mslr_status_t mslr_pogm_init(mslr_pogm st, const void *y_dev, const void *X0_dev);
/* Run n further iterations (iter counts continue across calls). */
mslr_status_t mslr_pogm_run(mslr_pogm st, int n);
/* Device pointer to the current X (ComplexF32 (Nx,Ny,Nz,Nt,nscales)); valid
 * until mslr_pogm_destroy.  After the last iteration it is the returned xnew. */
void *mslr_pogm_x(mslr_pogm st);
/* Iterations done so far and the restart flags (one per iteration, host array
 * of length >= iterations done). */
mslr_status_t mslr_pogm_log(mslr_pogm st, int *iters_done, int *restarts);
/* The logger's costs of x_0 .. x_k after k iterations (k + 1 values each, host
 * arrays, either may be NULL), as run_recon collects them
 * (scripts/reconstruct.jl:176-191,205,228-229 with src/mirt_mod.jl:106,181):
 *   dc_costs[i]  = 0.5 ||A sum_s X_s - y||^2           (FP32 residual, FP64 sum)
 *   reg_costs[i] = sum_s lambda_s patch_nucnorm(img2patches(X_s))   (FP64)
 * The restart flags are mslr_pogm_log's (x_0's is false).  Needs log_costs = 1
 * at create; reading the log after the last iteration costs one forward A. */
mslr_status_t mslr_pogm_costs(mslr_pogm st, int *n, double *dc_costs, double *reg_costs);
/* Derivative form (no counterpart in the reference, which has no autodiff; the
 * pullback a ChainRulesCore rrule of run_recon's POGM loop,
 * scripts/reconstruct.jl:220-226, would call).  After k iterations of a state
 * created with checkpoint = 1: lambda_bar[s] = d<xbar, sum_s X_s>/d lambda_s through
 * the k unrolled iterations, xbar (device, ComplexF32 (Nx,Ny,Nz,Nt)) the cotangent
 * of the final image sum (for the MSE to x_ref: 2 (sum_s X_s - x_ref) / numel).
 * The restart decisions are held fixed.  Every patch geometry: overlapping,
 * edge-clipped, fewer voxels than frames, 1x1x1.  Allocates its backward state for
 * the call. */
mslr_status_t mslr_pogm_lambda_vjp(mslr_pogm st, const void *xbar_dev, double *lambda_bar);
/* Per-scale SVST thresholds zeta*lambda_k used in the last iteration (host, nscales). */
mslr_status_t mslr_pogm_thresholds(mslr_pogm st, double *beta);

#ifdef __cplusplus
}
#endif

#endif /* MSLR_H */
