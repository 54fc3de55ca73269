// Two-stage reduction of batched Hermitian matrices (FP64) to real tridiagonal
// form: dense -> band of width SBR_B (block Householder, GEMM-shaped updates),
// then band -> tridiagonal (bulge chasing; the band in LDS up to SBR_MAXM, in
// global memory above), and the back-transformation
// of tridiagonal eigenvectors through both stages.
#pragma once

#include "common.hpp"

namespace mslr {

constexpr int SBR_B = 16;       // bandwidth of the intermediate band matrix
constexpr int SBR_MINM = 64;    // smaller matrices use the one-stage reduction
constexpr int SBR_MAXM = 300;   // the band (2*SBR_B diagonals) fits in LDS up to here
constexpr int SBR_MAXG = 480;   // band in global memory up to here (sweep schedule and
                                // the back-transformation's reflector staging)

inline bool sbr_eligible(int64_t m) { return m >= SBR_MINM && m <= SBR_MAXG; }
inline bool sbr_global_band(int64_t m) { return m > SBR_MAXM; }
// panels of the dense -> band stage
inline int sbr_panels(int m) {
    int n = 0;
    for (int c0 = 0; m - c0 - SBR_B >= 2; c0 += SBR_B) ++n;
    return n;
}
// reflector slots per sweep of the band -> tridiagonal stage
inline int sbr_steps(int m) { return (m - 1 + SBR_B - 1) / SBR_B; }

constexpr int SBR_S1_WAVES = 8;  // waves of the dense -> band kernel

// per matrix workspace of the dense -> band stage in global memory: the panel's
// Y = A V and W = Y T - 1/2 V M ((m - B) x B each) and the waves' B x B partials of
// V^H Y
__host__ __device__ inline size_t sbr_w1_elems(int m) {
    return (size_t)2 * (m - SBR_B) * SBR_B + (size_t)SBR_S1_WAVES * SBR_B * SBR_B;
}

struct SbrWork {
    c128 *T1 = nullptr;    // (sbr_panels(m), B, B) per matrix: compact-WY T of each panel
    c128 *V2 = nullptr;    // (m, sbr_steps(m), B) per matrix: bulge-chasing reflectors
    c128 *tau2 = nullptr;  // (m, sbr_steps(m)) per matrix
    c128 *W1 = nullptr;    // sbr_w1_elems(m) per matrix
    c128 *Band = nullptr;  // (2B, m) per matrix when sbr_global_band(m)
};

inline size_t sbr_bytes_per_matrix(int m) {
    return (size_t)sbr_panels(m) * SBR_B * SBR_B * 16 + (size_t)m * sbr_steps(m) * SBR_B * 16 +
           (size_t)m * sbr_steps(m) * 16 + sbr_w1_elems(m) * 16 +
           (sbr_global_band(m) ? (size_t)2 * SBR_B * m * 16 : 0);
}

// G (nb matrices, m x m, full Hermitian storage) -> d, e (real tridiagonal).
// G keeps the panel reflectors below the band.
// Batches of at most this many matrices (the whole-image patch) run their eigensolver
// kernels at raised wave priority: their latency is the cost, and the batched kernels
// of the other scales share the CUs with them.
constexpr int64_t SVST_SMALL_BATCH = 8;
inline int svst_small_batch(int64_t nb) { return nb <= SVST_SMALL_BATCH ? 1 : 0; }

void sbr_reduce(mslr_ctx_s *ctx, c128 *G, int m, int64_t nb, const SbrWork &w, double *d,
                double *e);

// U[:, k] = Q1 Q2 Z[:, k] for k < r (per matrix), U as ComplexF32 (m x m per matrix).
void sbr_backtransform(mslr_ctx_s *ctx, const c128 *G, int m, int64_t nb, const SbrWork &w,
                       const double *Z, const int *r, c64 *U);

}  // namespace mslr
