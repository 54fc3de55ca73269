// Batched singular value soft-thresholding of patch matrices.
#pragma once

#include "common.hpp"
#include "patches.hpp"

namespace mslr {

// SVST of Np matrices X_i (p x Nt, ComplexF32, consecutive), Pin -> Pout (distinct):
//   X_i <- U diag(max(s - beta_i, 0)) V^H.
// beta_dev: device doubles, beta_i = beta_dev[per_patch ? i : 0] (rounded to
// Float32 inside, src/recon.jl:195).  err_dev (device int, nullable) is set to
// 1 when a Gram matrix holds non-finite values.  nuc_out (device, Np doubles,
// nullable) receives each output's nuclear norm sum_k max(s_k - beta_i, 0).
//
// Options of the split-BF16 form with p >= Nt (svst_writes_direct):
//   gather (nullable): Pin is the image (V x Nt) of the geometry's patches, which
//     the split pass reads through the geometry (no img2patches pass);
//   scatter (nullable): the patches are those of an exact tiling (every voxel in
//     one patch) and Pout is their image: each output patch is written straight to
//     its voxels, so no patches2img pass follows;
//   Pin == Pout (in place, also with gather and scatter on one image).
// The products that read X (X W, X U_r) take the split pass's words as their left
// operand when svst_reads_words; otherwise, with gather or in place, the split pass
// keeps an FP32 copy of the patch matrices for them.
//   words (nullable; svst_reads_words, no gather, one chunk): the split words of
//     Pin (ozaki.hpp: wh of all Np p x Nt matrices, then wl) already written by
//     the caller; no split pass runs and Pin is not read.
void svst_batched(mslr_ctx_s *ctx, const c64 *Pin, c64 *Pout, int64_t p, int64_t Nt, int64_t Np,
                  const double *beta_dev, int per_patch, int *err_dev, double *nuc_out = nullptr,
                  const PatchGeom *scatter = nullptr, const PatchGeom *gather = nullptr,
                  const unsigned *words = nullptr);

// Whether svst_batched takes the gather, scatter and in-place options above for
// these shapes on ctx.
bool svst_writes_direct(const mslr_ctx_s *ctx, int64_t p, int64_t Nt);

// Whether every product of svst_batched reads the input only through its split
// words (split-BF16 products, p >= Nt), so that a caller may hand them in.
bool svst_reads_words(const mslr_ctx_s *ctx, int64_t p, int64_t Nt);

// Vector-Jacobian product of the batched SVST (p >= Nt >= 2): Zbar (distinct from
// the inputs) <- the cotangent of Z, betabar_dev[i] <- d<Ybar, Y>/d beta_i per
// matrix (device, Np doubles).  Takes every eigenpair of each Gram matrix.
void svst_vjp_batched(mslr_ctx_s *ctx, const c64 *Z, const c64 *Yb, c64 *Zb, int64_t p, int64_t Nt,
                      int64_t Np, const double *beta_dev, int per_patch, double *betabar_dev,
                      int *err_dev);

// The same for any shape: min(p, Nt) = 1 by the group shrink's closed form, p < Nt
// through the conjugate transposes (Y^H = SVST(Z^H)), p >= Nt >= 2 as above.
void svst_vjp_any(mslr_ctx_s *ctx, const c64 *Z, const c64 *Yb, c64 *Zb, int64_t p, int64_t Nt,
                  int64_t Np, const double *beta_dev, int per_patch, double *betabar_dev,
                  int *err_dev);

// Sum of singular values of each matrix, accumulated into *out_dev (device double).
void nucnorm_batched(mslr_ctx_s *ctx, const c64 *P, int64_t p, int64_t Nt, int64_t Np,
                     double *out_dev);

}  // namespace mslr
