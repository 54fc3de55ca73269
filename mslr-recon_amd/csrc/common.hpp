// Shared definitions of the MSLR HIP library: complex arithmetic, error
// reporting, the context (device + stream + workspace) and launch helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mslr.h"

namespace mslr {

// ------------------------------------------------------------------ complex
template <class T>
struct C2 {
    T re, im;
};

template <class T>
__host__ __device__ __forceinline__ C2<T> mk(T r, T i) { return C2<T>{r, i}; }
template <class T>
__host__ __device__ __forceinline__ C2<T> operator+(C2<T> a, C2<T> b) { return {a.re + b.re, a.im + b.im}; }
template <class T>
__host__ __device__ __forceinline__ C2<T> operator-(C2<T> a, C2<T> b) { return {a.re - b.re, a.im - b.im}; }
template <class T>
__host__ __device__ __forceinline__ C2<T> operator*(C2<T> a, C2<T> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <class T>
__host__ __device__ __forceinline__ C2<T> operator*(T s, C2<T> a) { return {s * a.re, s * a.im}; }
template <class T>
__host__ __device__ __forceinline__ C2<T> conj(C2<T> a) { return {a.re, -a.im}; }
// conj(a) * b
template <class T>
__host__ __device__ __forceinline__ C2<T> cmulc(C2<T> a, C2<T> b) {
    return {a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re};
}
template <class T>
__host__ __device__ __forceinline__ T abs2(C2<T> a) { return a.re * a.re + a.im * a.im; }

using c64 = C2<float>;
using c128 = C2<double>;

// ------------------------------------------------------------------ errors
struct Error : std::runtime_error {
    mslr_status_t code;
    Error(mslr_status_t c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string &msg);
void clear_last_error();

#define MSLR_FAIL(code, msg)                                               \
    do {                                                                   \
        throw ::mslr::Error((code), std::string(__func__) + ": " + (msg)); \
    } while (0)

#define MSLR_REQUIRE(cond, code, msg) \
    do {                              \
        if (!(cond)) MSLR_FAIL(code, msg); \
    } while (0)

#define HIP_CHECK(expr)                                                              \
    do {                                                                             \
        hipError_t e_ = (expr);                                                      \
        if (e_ != hipSuccess)                                                        \
            throw ::mslr::Error(MSLR_ERR_DEVICE, std::string(#expr) + " failed: " +  \
                                                     hipGetErrorString(e_));          \
    } while (0)

#define HIP_LAUNCH_CHECK() HIP_CHECK(hipGetLastError())

// Wrap an entry point body: converts exceptions into status codes.
template <class F>
mslr_status_t guarded(F &&f) {
    try {
        clear_last_error();
        f();
        return MSLR_OK;
    } catch (const Error &e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc &) {
        set_last_error("host allocation failed");
        return MSLR_ERR_NOMEM;
    } catch (const std::exception &e) {
        set_last_error(e.what());
        return MSLR_ERR_DEVICE;
    }
}

// ------------------------------------------------------------------ context
// Grow-only device scratch arena, reused across calls on one context.
struct Scratch {
    void *ptr = nullptr;
    size_t bytes = 0;
};

}  // namespace mslr

struct mslr_ctx_s {
    int device = 0;
    hipStream_t stream = nullptr;
    int num_cus = 0;
    std::vector<mslr::Scratch> scratch;  // indexed slots, see ScratchSlot
    size_t slot_base = 0;                // added to every slot (per-stream scratch sets)
    uint64_t scratch_gen = 0;            // bumped whenever a slot is reallocated
    int products = 0;                    // MSLR_PRODUCTS_*: SVST Gram / apply products
    mslr_tuning tuning{};                // SVST launch shapes (0 = built-in)

    // Kernel timing with HIP events on `stream` (enabled by mslr_prof_enable).
    struct ProfRec {
        int tag;
        hipEvent_t a, b;
    };
    bool prof = false;
    const char *prof_suffix = "";  // appended to kernel names (e.g. the scale being processed)
    std::vector<std::string> prof_tags;
    std::vector<ProfRec> prof_recs;
    std::vector<hipEvent_t> prof_pool;

    void *scratch_get(size_t slot, size_t bytes);
    void activate() const;
    int prof_tag(const char *name);
    hipEvent_t prof_event();
};

namespace mslr {
// Records HIP events around the launches in its scope when profiling is on.
struct ProfScope {
    mslr_ctx_s *ctx;
    int tag = -1;
    hipEvent_t a = nullptr, b = nullptr;
    ProfScope(mslr_ctx_s *c, const char *name) : ctx(c) {
        if (!ctx->prof) return;
        tag = ctx->prof_tag((std::string(name) + ctx->prof_suffix).c_str());
        a = ctx->prof_event();
        b = ctx->prof_event();
        hipEventRecord(a, ctx->stream);
    }
    ~ProfScope() {
        if (tag < 0) return;
        hipEventRecord(b, ctx->stream);
        ctx->prof_recs.push_back({tag, a, b});
    }
};
}  // namespace mslr
#define MSLR_PROF(ctx, name) ::mslr::ProfScope prof_scope_##__LINE__((ctx), (name))

namespace mslr {

inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Work-group index renumbered so that consecutive numbers run on one XCD: the
// dispatcher hands consecutive work-groups to the 8 XCDs in turn, each with its
// own L2, so work-groups that share operands and are numbered consecutively would
// otherwise fetch the shared data once per XCD.  The last n % 8 keep their order.
__device__ __forceinline__ int64_t xcd_order(int64_t w, int64_t n) {
    constexpr int NXCD = 8;
    const int64_t per = n / NXCD;
    if (w >= per * NXCD) return w;
    return (w % NXCD) * per + w / NXCD;
}

// Work-group barrier that orders LDS accesses only.  __syncthreads is a fence over
// all memory as well, so it also waits for the wave's global loads in flight; this
// one leaves them in flight, for loops that prefetch global data across their
// barriers while the waves exchange data through LDS alone (global memory written
// before it is not visible to the other waves after it).
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Lookup of the context behind a handle with validation.
inline mslr_ctx_s *check_ctx(mslr_ctx c) {
    if (!c) throw Error(MSLR_ERR_ARG, "null context");
    c->activate();
    return c;
}

// Scratch slots used by the library (one arena per slot per context).
enum ScratchSlot : size_t {
    SLOT_FFT = 0,       // SENSE intermediate (sampled k-space lines per coil)
    SLOT_RESID = 1,     // SENSE residual / staging
    SLOT_REDUCE = 2,    // FP64 partial sums
    SLOT_GRAM_PART = 3, // split-K partial Gram tiles (FP32)
    SLOT_GRAM = 4,      // FP64 Gram matrices / eigen workspace
    SLOT_EIG = 5,       // eigenvectors + filter factors
    SLOT_APPLY = 6,     // apply intermediate X V_r
    SLOT_PATCH = 7,     // materialised patches
    SLOT_PATCH_OUT = 8, // image staging for patch operators
    SLOT_PWR = 9,       // power-iteration vectors
    SLOT_PLANES = 10,   // SENSE plane-major image frames
    SLOT_SPLIT = 11,    // patch data as split-BF16 words (mixed-precision products)
    SLOT_LOWRANK = 12,  // X U_r of the low-rank SVST apply
    SLOT_VJP = 13,      // m x m intermediates of the SVST derivative
    SLOT_VJP_P = 14,    // patch cotangents of the prox derivative
    SLOT_SETUP = 15,    // run_recon setup: chunk counts, magnitudes, selection state
    SLOT_VJP_T = 16,    // conjugate-transposed patches of the SVST derivative (p < Nt)
    SLOT_BSPLIT = 17,   // right operand of the split-BF16 apply as hi / lo words
    SLOT_SVST_COPY = 18,// FP32 copy of an SVST input that is overwritten in place
    SLOT_BSPLIT_LR = 19,// U_r and f U_r^H as hi / lo words for the fused low-rank apply
    SLOT_COUNT = 20
};

}  // namespace mslr
