// All-frames SENSE plan shared by the SENSE entry points and the POGM driver.
#pragma once

#include <functional>
#include <utility>
#include <vector>

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
    int LBx = 1;      // sampled lines per workgroup in the x-axis passes
    int nlb = 1;      // line blocks per (coil, frame): cdiv(max lines per frame, LBx)
    int nt_yz = 64;   // threads per workgroup in the yz-plane passes
    size_t chunk_bytes = 0;
    bool use_fixed = true;  // compile-time FFT lengths where available

    // device arrays
    mslr::c64 *smaps32 = nullptr;   // smaps * pre-phase, plane-major (Nyz, Nx, Nc)
    mslr::c128 *smaps64 = nullptr;
    mslr::c64 *tw32[3] = {nullptr, nullptr, nullptr};   // exp(+2 pi i m / n)
    mslr::c128 *tw64[3] = {nullptr, nullptr, nullptr};
    mslr::c64 *ph32[3] = {nullptr, nullptr, nullptr};   // post-phase per axis
    mslr::c128 *ph64[3] = {nullptr, nullptr, nullptr};
    // Sampled lines: the distinct (ky, kz) of a frame having at least one sample.
    int *lptr = nullptr;     // (Nt+1): first global line of frame t
    int *line_yz = nullptr;  // (lines): ky + Ny*kz, ascending within a frame
    int *kbeg = nullptr;     // (lines): first internal sample of the line in its frame
    int *lk = nullptr;       // (K, Nt): (l << 16) | kx of internal sample j (l: line of the frame)
    // two-stage y-z passes (compiled plane sizes, FP32, default sign)
    int yz2 = 0;                       // 1: tables below built
    mslr::c64 *smaps_yz2 = nullptr;    // coil maps in the y-stage-A lane order
    int *line_pos = nullptr;           // (lines): LDS offset of the line's (ky, kz)
    // z-stage-B butterflies holding at least one sampled line, per frame: zb_list[lptr[t] ..
    // lptr[t] + zb_cnt[t]) (ascending; at most one per line)
    int *zb_list = nullptr;
    int *zb_cnt = nullptr;
    mslr::c64 *tw_yz2[2] = {nullptr, nullptr};  // y twiddles, forward / inverse sign
    // two-stage fused x passes (compiled line lengths, FP32, default sign)
    int x2 = 0;                                 // 1: tables below built
    mslr::c64 *tw_x2[2] = {nullptr, nullptr};   // x twiddles (Cooley-Tukey split)
    bool full_lines = false;                    // every sampled line holds all Nx samples
    std::vector<int> lptr_h; // host copy of lptr
    std::vector<void *> allocs;
};

namespace mslr {

// Sample order: frame-major, then sampled line (ascending ky + Ny*kz), then kx,
// which is the reference's column-major order; kept as an explicit step
// (a copy when the buffers differ) for callers holding reference-order data.
void sense_ref_to_internal(mslr_sense_s *P, const c64 *y_ref, c64 *y_int);

// g = A'(A sum_k X_k - y_int), frames 0..Nt-1, everything in internal sample
// order.  dc_partial (device, nullable) receives Nt*Nc*nlb partial ||r||^2.
void sense_grad_internal(mslr_sense_s *P, const c64 *X, int nscales, int64_t sstride,
                         const c64 *y_int, c64 *g, double *dc_partial);

// Frame chunks of the pipelined gradient: nch groups of consecutive frames (as
// equal as the frame count allows), each split further where it exceeds the
// plan's scratch budget.
std::vector<std::pair<int64_t, int64_t>> sense_grad_chunks(mslr_sense_s *P, int nch);

// Second stream of the pipelined gradient: chunks alternate between the context
// stream and `stream` (scratch slots offset by slot_base there); fork/join events.
struct SensePipe {
    hipStream_t stream = nullptr;  // nullptr: every chunk on the context stream, in order
    size_t slot_base = 0;
    hipEvent_t fork = nullptr, join = nullptr;
};

// g = A'(A sum_k X_k - y_int) chunk by chunk; after(i, ta, nf, planes) enqueues work
// on the chunk's frames [ta, ta+nf) behind chunk i's adjoint on the stream the chunk
// ran on (ctx->stream while it is called).  With g == nullptr the chunk's gradient is
// left plane-major in scratch, planes[f][x][y + Ny z] (f = 0..nf-1), valid until the
// stream's next chunk; else it is written to g and planes holds the same values.  The
// context stream waits for all chunks.
void sense_grad_pipelined(mslr_sense_s *P, const c64 *X, int nscales, int64_t sstride,
                          const c64 *y_int, c64 *g, double *dc_partial,
                          const std::vector<std::pair<int64_t, int64_t>> &chunks,
                          const SensePipe &pipe,
                          const std::function<void(size_t, int64_t, int64_t, const c64 *)> &after);

// yz-lines per workgroup of the plane-major <-> image transposes (an LDS tile of
// Nx x (lines + 1) values).
int sense_transpose_lines(mslr_sense_s *P);

// Partial ||A sum_k X_k - y_int||^2 (Nt*Nc*nlb doubles into dc_partial), the
// forward half of sense_grad_internal.
void sense_resid_internal(mslr_sense_s *P, const c64 *X, int nscales, int64_t sstride,
                          const c64 *y_int, double *dc_partial);

// x = A' y for all frames, y in internal order.
void sense_adjoint_internal(mslr_sense_s *P, const c64 *y_int, c64 *x);

// Bytes of the file-scope scratch array of mslr_sense_forward_noalloc (per device).
#ifndef MSLR_NOALLOC_SCRATCH_BYTES
#define MSLR_NOALLOC_SCRATCH_BYTES (64u << 20)
#endif
// Looks up this device's address of that array (the device must be current);
// once per device, at context creation.
void sense_noalloc_init(int device);

}  // namespace mslr
