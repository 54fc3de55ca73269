// Context, device buffers and error reporting of the C ABI.
#include <mutex>

#include "common.hpp"
#include "sense.hpp"

namespace mslr {

static thread_local std::string g_last_error;

void set_last_error(const std::string &msg) { g_last_error = msg; }
void clear_last_error() { g_last_error.clear(); }

}  // namespace mslr

using namespace mslr;

void mslr_ctx_s::activate() const {
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess)
        throw Error(MSLR_ERR_DEVICE, std::string("hipSetDevice failed: ") + hipGetErrorString(e));
}

void *mslr_ctx_s::scratch_get(size_t slot, size_t bytes) {
    slot += slot_base;
    if (slot >= scratch.size()) scratch.resize(slot + 1);
    Scratch &s = scratch[slot];
    if (bytes == 0) bytes = 256;
    if (s.bytes >= bytes) return s.ptr;
    // Growing: work queued on the old buffer (on any stream) must finish first.
    HIP_CHECK(hipDeviceSynchronize());
    if (s.ptr) HIP_CHECK(hipFree(s.ptr));
    s.ptr = nullptr;
    s.bytes = 0;
    const size_t want = (bytes + 4095) & ~(size_t)4095;
    hipError_t e = hipMalloc(&s.ptr, want);
    if (e != hipSuccess) {
        s.ptr = nullptr;
        throw Error(MSLR_ERR_NOMEM, "scratch allocation of " + std::to_string(want) +
                                        " bytes failed: " + hipGetErrorString(e));
    }
    s.bytes = want;
    ++scratch_gen;  // captured graphs holding the old address are stale now
    return s.ptr;
}

int mslr_ctx_s::prof_tag(const char *name) {
    for (size_t i = 0; i < prof_tags.size(); ++i)
        if (prof_tags[i] == name) return (int)i;
    prof_tags.emplace_back(name);
    return (int)prof_tags.size() - 1;
}

hipEvent_t mslr_ctx_s::prof_event() {
    if (!prof_pool.empty()) {
        hipEvent_t e = prof_pool.back();
        prof_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    HIP_CHECK(hipEventCreate(&e));
    return e;
}

extern "C" {

const char *mslr_last_error(void) { return g_last_error.c_str(); }

mslr_status_t mslr_prof_enable(mslr_ctx ctx, int enable) {
    return guarded([&] {
        mslr_ctx_s *c = check_ctx(ctx);
        HIP_CHECK(hipStreamSynchronize(c->stream));
        for (auto &r : c->prof_recs) {
            c->prof_pool.push_back(r.a);
            c->prof_pool.push_back(r.b);
        }
        c->prof_recs.clear();
        c->prof = enable != 0;
    });
}

mslr_status_t mslr_prof_report(mslr_ctx ctx, char *buf, size_t len) {
    return guarded([&] {
        mslr_ctx_s *c = check_ctx(ctx);
        MSLR_REQUIRE(buf && len > 0, MSLR_ERR_ARG, "null buffer");
        HIP_CHECK(hipStreamSynchronize(c->stream));
        std::vector<double> tot(c->prof_tags.size(), 0.0);
        std::vector<long> cnt(c->prof_tags.size(), 0);
        for (auto &r : c->prof_recs) {
            float ms = 0.f;
            HIP_CHECK(hipEventElapsedTime(&ms, r.a, r.b));
            tot[r.tag] += ms;
            cnt[r.tag] += 1;
        }
        std::string out;
        for (size_t i = 0; i < tot.size(); ++i) {
            if (!cnt[i]) continue;
            char line[256];
            std::snprintf(line, sizeof line, "%s %ld %.6f\n", c->prof_tags[i].c_str(), cnt[i], tot[i]);
            out += line;
        }
        std::snprintf(buf, len, "%s", out.c_str());
    });
}

int mslr_version(void) { return 100; }

mslr_status_t mslr_ctx_create(int device, mslr_ctx *out) {
    return guarded([&] {
        MSLR_REQUIRE(out != nullptr, MSLR_ERR_ARG, "out is null");
        *out = nullptr;
        int n = 0;
        hipError_t e = hipGetDeviceCount(&n);
        if (e != hipSuccess || n == 0)
            throw Error(MSLR_ERR_DEVICE, "no HIP device is available");
        MSLR_REQUIRE(device >= 0 && device < n, MSLR_ERR_ARG, "device ordinal out of range");
        HIP_CHECK(hipSetDevice(device));
        hipDeviceProp_t prop;
        HIP_CHECK(hipGetDeviceProperties(&prop, device));
        auto *c = new mslr_ctx_s();
        c->device = device;
        c->num_cus = prop.multiProcessorCount;
        hipError_t se = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (se != hipSuccess) {
            delete c;
            throw Error(MSLR_ERR_DEVICE, std::string("hipStreamCreate failed: ") +
                                             hipGetErrorString(se));
        }
        c->scratch.resize(SLOT_COUNT);
        try {
            sense_noalloc_init(device);
        } catch (...) {
            hipStreamDestroy(c->stream);
            delete c;
            throw;
        }
        *out = c;
    });
}

mslr_status_t mslr_ctx_destroy(mslr_ctx ctx) {
    return guarded([&] {
        if (!ctx) return;
        ctx->activate();
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
        for (auto &s : ctx->scratch)
            if (s.ptr) hipFree(s.ptr);
        for (auto &r : ctx->prof_recs) {
            hipEventDestroy(r.a);
            hipEventDestroy(r.b);
        }
        for (auto e : ctx->prof_pool) hipEventDestroy(e);
        hipStreamDestroy(ctx->stream);
        delete ctx;
    });
}

mslr_status_t mslr_ctx_set_products(mslr_ctx ctx, int products) {
    return guarded([&] {
        mslr_ctx_s *c = check_ctx(ctx);
        MSLR_REQUIRE(products == MSLR_PRODUCTS_FP32 || products == MSLR_PRODUCTS_SPLIT_BF16,
                     MSLR_ERR_ARG, "products must be MSLR_PRODUCTS_FP32 or MSLR_PRODUCTS_SPLIT_BF16");
        c->products = products;
    });
}

mslr_status_t mslr_ctx_set_tuning(mslr_ctx ctx, const mslr_tuning *t) {
    return guarded([&] {
        mslr_ctx_s *c = check_ctx(ctx);
        MSLR_REQUIRE(t, MSLR_ERR_ARG, "null tuning");
        MSLR_REQUIRE(t->gram_chunk == 0 ||
                         (t->gram_chunk >= 256 && t->gram_chunk <= 65536 && t->gram_chunk % 256 == 0),
                     MSLR_ERR_ARG, "gram_chunk must be 0 or a multiple of 256 in [256, 65536]");
        MSLR_REQUIRE(t->eig_lanes_limit >= -1, MSLR_ERR_ARG, "eig_lanes_limit must be >= -1");
        MSLR_REQUIRE(t->svst_apply >= 0 && t->svst_apply <= 2, MSLR_ERR_ARG,
                     "svst_apply must be 0, 1 or 2");
        MSLR_REQUIRE(t->sense_chunks >= 0 && t->sense_chunks <= 64, MSLR_ERR_ARG,
                     "sense_chunks must be in [0, 64]");
        for (int v : {t->yz_threads_fwd, t->yz_threads_adj})
            MSLR_REQUIRE(v == 0 || v == 512 || v == 768 || v == 1024, MSLR_ERR_ARG,
                         "yz_threads_fwd / yz_threads_adj must be 0, 512, 768 or 1024");
        MSLR_REQUIRE(t->update_lines == 0 ||
                         (t->update_lines >= 8 && t->update_lines <= 64 && t->update_lines % 8 == 0),
                     MSLR_ERR_ARG, "update_lines must be 0 or a multiple of 8 in [8, 64]");
        MSLR_REQUIRE(t->sb2st_schedule >= 0 && t->sb2st_schedule <= 3, MSLR_ERR_ARG,
                     "sb2st_schedule must be in [0, 3]");
        MSLR_REQUIRE(t->apply_fp32 >= 0 && t->apply_fp32 <= 8, MSLR_ERR_ARG,
                     "apply_fp32 must be in [0, 8]");
        for (int r : t->reserved) MSLR_REQUIRE(r == 0, MSLR_ERR_ARG, "reserved fields must be 0");
        c->tuning = *t;
    });
}

mslr_status_t mslr_ctx_get_tuning(mslr_ctx ctx, mslr_tuning *t) {
    return guarded([&] {
        mslr_ctx_s *c = check_ctx(ctx);
        MSLR_REQUIRE(t, MSLR_ERR_ARG, "null tuning");
        *t = c->tuning;
    });
}

mslr_status_t mslr_ctx_sync(mslr_ctx ctx) {
    return guarded([&] {
        mslr_ctx_s *c = check_ctx(ctx);
        HIP_CHECK(hipStreamSynchronize(c->stream));
    });
}

mslr_status_t mslr_ctx_info(mslr_ctx ctx, char *name, size_t name_len, size_t *total_bytes,
                            size_t *free_bytes) {
    return guarded([&] {
        mslr_ctx_s *c = check_ctx(ctx);
        hipDeviceProp_t prop;
        HIP_CHECK(hipGetDeviceProperties(&prop, c->device));
        if (name && name_len > 0) {
            std::snprintf(name, name_len, "%s", prop.name);
        }
        size_t fr = 0, tot = 0;
        HIP_CHECK(hipMemGetInfo(&fr, &tot));
        if (total_bytes) *total_bytes = tot;
        if (free_bytes) *free_bytes = fr;
    });
}

mslr_status_t mslr_malloc(mslr_ctx ctx, size_t bytes, void **dev_ptr) {
    return guarded([&] {
        mslr_ctx_s *c = check_ctx(ctx);
        (void)c;
        MSLR_REQUIRE(dev_ptr != nullptr, MSLR_ERR_ARG, "dev_ptr is null");
        *dev_ptr = nullptr;
        hipError_t e = hipMalloc(dev_ptr, bytes == 0 ? 256 : bytes);
        if (e != hipSuccess) {
            *dev_ptr = nullptr;
            throw Error(MSLR_ERR_NOMEM, "hipMalloc of " + std::to_string(bytes) +
                                            " bytes failed: " + hipGetErrorString(e));
        }
    });
}

mslr_status_t mslr_free(mslr_ctx ctx, void *dev_ptr) {
    return guarded([&] {
        mslr_ctx_s *c = check_ctx(ctx);
        if (!dev_ptr) return;
        HIP_CHECK(hipStreamSynchronize(c->stream));
        HIP_CHECK(hipFree(dev_ptr));
    });
}

mslr_status_t mslr_upload(mslr_ctx ctx, void *dst_dev, const void *src_host, size_t bytes) {
    return guarded([&] {
        mslr_ctx_s *c = check_ctx(ctx);
        if (bytes == 0) return;
        MSLR_REQUIRE(dst_dev && src_host, MSLR_ERR_ARG, "null pointer");
        HIP_CHECK(hipMemcpyAsync(dst_dev, src_host, bytes, hipMemcpyHostToDevice, c->stream));
        HIP_CHECK(hipStreamSynchronize(c->stream));
    });
}

mslr_status_t mslr_download(mslr_ctx ctx, void *dst_host, const void *src_dev, size_t bytes) {
    return guarded([&] {
        mslr_ctx_s *c = check_ctx(ctx);
        if (bytes == 0) return;
        MSLR_REQUIRE(dst_host && src_dev, MSLR_ERR_ARG, "null pointer");
        HIP_CHECK(hipMemcpyAsync(dst_host, src_dev, bytes, hipMemcpyDeviceToHost, c->stream));
        HIP_CHECK(hipStreamSynchronize(c->stream));
    });
}

mslr_status_t mslr_copy_d2d(mslr_ctx ctx, void *dst_dev, const void *src_dev, size_t bytes) {
    return guarded([&] {
        mslr_ctx_s *c = check_ctx(ctx);
        if (bytes == 0) return;
        MSLR_REQUIRE(dst_dev && src_dev, MSLR_ERR_ARG, "null pointer");
        HIP_CHECK(hipMemcpyAsync(dst_dev, src_dev, bytes, hipMemcpyDeviceToDevice, c->stream));
    });
}

mslr_status_t mslr_zero(mslr_ctx ctx, void *dev_ptr, size_t bytes) {
    return guarded([&] {
        mslr_ctx_s *c = check_ctx(ctx);
        if (bytes == 0) return;
        MSLR_REQUIRE(dev_ptr, MSLR_ERR_ARG, "null pointer");
        HIP_CHECK(hipMemsetAsync(dev_ptr, 0, bytes, c->stream));
    });
}

}  // extern "C"
