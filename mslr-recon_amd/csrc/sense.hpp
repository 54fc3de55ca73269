// All-frames SENSE plan shared by the SENSE entry points and the POGM driver.
#pragma once

#include "common.hpp"
#include "fft.hpp"

struct mslr_sense_s {
    mslr_ctx_s *ctx = nullptr;
    int N[3] = {0, 0, 0};
    int64_t V = 0;    // Nx*Ny*Nz
    int NYZ = 0;      // Ny*Nz
    int Nc = 0;
    int Nt = 0;
    int64_t K = 0;    // samples per frame per coil
    int sgn = -1;     // exponent sign of the forward DFT
    double scale = 1.0;
    mslr::FftDesc d[3];
    int LBx = 1;      // yz-lines per workgroup in the x-axis passes
    int nt_yz = 64;   // threads per workgroup in the yz-plane passes
    size_t chunk_bytes = 0;
    bool use_fixed = true;  // compile-time FFT lengths where available

    // device arrays
    mslr::c64 *smaps32 = nullptr;   // smaps * pre-phase, (V, Nc)
    mslr::c128 *smaps64 = nullptr;
    mslr::c64 *tw32[3] = {nullptr, nullptr, nullptr};   // exp(+2 pi i m / n)
    mslr::c128 *tw64[3] = {nullptr, nullptr, nullptr};
    mslr::c64 *ph32[3] = {nullptr, nullptr, nullptr};   // post-phase per axis
    mslr::c128 *ph64[3] = {nullptr, nullptr, nullptr};
    int *ptr = nullptr;    // (Nx+1, Nt): first internal sample of plane kx in frame t
    int *off = nullptr;    // (K, Nt): ky + Ny*kz of internal sample j
    int *perm = nullptr;   // (K, Nt): reference index of internal sample j
    std::vector<void *> allocs;
};

namespace mslr {

// Internal sample order: frame-major, then kx plane, then ky + Ny*kz.
// y_int = y_ref permuted, both (K, Nc, Nt).
void sense_ref_to_internal(mslr_sense_s *P, const c64 *y_ref, c64 *y_int);

// g = A'(A sum_k X_k - y_int), frames 0..Nt-1, everything in internal sample
// order.  dc_partial (device, nullable) receives Nt*Nc*Nx partial ||r||^2.
void sense_grad_internal(mslr_sense_s *P, const c64 *X, int nscales, int64_t sstride,
                         const c64 *y_int, c64 *g, double *dc_partial);

// x = A' y for all frames, y in internal order.
void sense_adjoint_internal(mslr_sense_s *P, const c64 *y_int, c64 *x);

}  // namespace mslr
