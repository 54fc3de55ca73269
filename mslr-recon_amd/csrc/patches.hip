// Patch extraction / recombination (src/recon.jl:53-126).
//
// img2patches is a pure gather; patches2img is written gather-style: every
// output voxel sums the patches covering it in increasing patch index (the
// order in which src/recon.jl:111-120 accumulates them), then divides by
// max(count, 1) in FP32, so it is bit-identical to the sequential loop and
// needs no atomics.
#include <algorithm>

#include "patches.hpp"

namespace mslr {

bool PatchGeom::disjoint() const {
    for (int a = 0; a < 3; ++a) {
        // consecutive starts at least ps apart
        for (int i = 1; i < ns[a]; ++i)
            if (start(a, i) - start(a, i - 1) < ps[a]) return false;
    }
    return true;
}

bool PatchGeom::exact_tiling() const {
    if (!disjoint()) return false;
    for (int a = 0; a < 3; ++a)
        if (ns[a] * ps[a] != N[a]) return false;
    return true;
}

PatchGeom make_geom(const int64_t N[3], int64_t Nt, const int64_t ps[3], const int64_t ss[3]) {
    PatchGeom g{};
    MSLR_REQUIRE(Nt >= 1, MSLR_ERR_ARG, "Nt must be >= 1");
    g.Nt = (int)Nt;
    g.Np = 1;
    g.p = 1;
    for (int a = 0; a < 3; ++a) {
        MSLR_REQUIRE(N[a] >= 1, MSLR_ERR_ARG, "image size must be positive");
        MSLR_REQUIRE(ps[a] >= 1, MSLR_ERR_ARG, "patch size must be positive");
        MSLR_REQUIRE(ss[a] >= 1, MSLR_ERR_ARG, "stride must be positive");
        g.N[a] = (int)N[a];
        g.ps[a] = (int)std::min<int64_t>(ps[a], N[a]);
        g.ss[a] = (int)ss[a];
        const int64_t nsteps = cdiv(N[a] - g.ps[a], ss[a]);  // cld(N - ps, s)
        g.ns[a] = (int)(nsteps + 1);
        g.Np *= g.ns[a];
        g.p *= g.ps[a];
    }
    return g;
}

// One thread per (i, t, patch) element of P.
__global__ void k_img2patches(const c64 *__restrict__ img, c64 *__restrict__ P, PatchGeom g,
                              int64_t ip0, int64_t np) {
    const int64_t V = (int64_t)g.N[0] * g.N[1] * g.N[2];
    const int64_t n = g.p * g.Nt * np;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e % g.p;
        const int64_t r = e / g.p;
        const int64_t t = r % g.Nt;
        const int64_t ip = ip0 + r / g.Nt;
        const int ix = (int)(ip % g.ns[0]);
        const int iy = (int)((ip / g.ns[0]) % g.ns[1]);
        const int iz = (int)(ip / ((int64_t)g.ns[0] * g.ns[1]));
        const int dx = (int)(i % g.ps[0]);
        const int dy = (int)((i / g.ps[0]) % g.ps[1]);
        const int dz = (int)(i / ((int64_t)g.ps[0] * g.ps[1]));
        const int x = g.start(0, ix) + dx;
        const int y = g.start(1, iy) + dy;
        const int z = g.start(2, iz) + dz;
        P[e] = img[x + (int64_t)g.N[0] * (y + (int64_t)g.N[1] * z) + V * t];
    }
}

// One thread per (patch element i, patch), looping over the frames: the index
// arithmetic once per thread, then strided copies (P rows coalesced, image
// reads in runs of ps[0] voxels).
__global__ __launch_bounds__(256) void k_img2patches_fr(const c64 *__restrict__ img,
                                                        c64 *__restrict__ P, PatchGeom g,
                                                        int64_t ip0, int64_t np) {
    const int64_t nrc = (g.p + 255) / 256;
    const int64_t ipl = (int64_t)blockIdx.x / nrc;
    const int64_t i = ((int64_t)blockIdx.x - ipl * nrc) * 256 + threadIdx.x;
    if (i >= g.p || ipl >= np) return;
    const int64_t V = (int64_t)g.N[0] * g.N[1] * g.N[2];
    const c64 *src = img + patch_voxel(g, ip0 + ipl, (int)i);
    c64 *dst = P + i + g.p * g.Nt * ipl;
#pragma unroll 4
    for (int t = 0; t < g.Nt; ++t) dst[g.p * t] = src[V * t];
}

// The inverse for an exact tiling (every voxel in exactly one patch): a copy,
// one thread per image voxel looping over the frames (image writes coalesced,
// patch reads in runs of ps[0]).
__global__ __launch_bounds__(256) void k_patches2img_tiled(const c64 *__restrict__ P,
                                                           c64 *__restrict__ img, PatchGeom g) {
    const int64_t V = (int64_t)g.N[0] * g.N[1] * g.N[2];
    const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (v >= V) return;
    const int x = (int)(v % g.N[0]);
    const int y = (int)((v / g.N[0]) % g.N[1]);
    const int z = (int)(v / ((int64_t)g.N[0] * g.N[1]));
    // exact tiling: patch starts are multiples of the patch size
    const int ix = x / g.ps[0], iy = y / g.ps[1], iz = z / g.ps[2];
    const int dx = x - ix * g.ps[0], dy = y - iy * g.ps[1], dz = z - iz * g.ps[2];
    const int64_t ip = ix + (int64_t)g.ns[0] * (iy + (int64_t)g.ns[1] * iz);
    const int64_t i = dx + (int64_t)g.ps[0] * (dy + (int64_t)g.ps[1] * dz);
    const c64 *src = P + i + g.p * g.Nt * ip;
    c64 *dst = img + v;
#pragma unroll 4
    for (int t = 0; t < g.Nt; ++t) dst[V * t] = src[g.p * t];
}

// Range of patch indices along axis a covering coordinate x: [lo, hi].
__device__ __forceinline__ void cover_range(const PatchGeom &g, int a, int x, int &lo, int &hi) {
    int l = x - g.ps[a] + 1;
    l = l <= 0 ? 0 : (l + g.ss[a] - 1) / g.ss[a];
    lo = g.ns[a];
    hi = -1;
    for (int i = l; i < g.ns[a]; ++i) {
        const int s = g.start(a, i);
        if (s > x) break;
        if (s + g.ps[a] > x) {
            if (i < lo) lo = i;
            hi = i;
        }
    }
}

// One thread per output element (voxel, t).  divide = false leaves the sum of the
// covering patches (the adjoint of img2patches).
__global__ void k_patches2img(const c64 *__restrict__ P, c64 *__restrict__ img, PatchGeom g,
                              int divide) {
    const int64_t V = (int64_t)g.N[0] * g.N[1] * g.N[2];
    const int64_t n = V * g.Nt;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t v = e % V;
        const int64_t t = e / V;
        const int x = (int)(v % g.N[0]);
        const int y = (int)((v / g.N[0]) % g.N[1]);
        const int z = (int)(v / ((int64_t)g.N[0] * g.N[1]));
        int xl, xh, yl, yh, zl, zh;
        cover_range(g, 0, x, xl, xh);
        cover_range(g, 1, y, yl, yh);
        cover_range(g, 2, z, zl, zh);
        c64 acc = {0.f, 0.f};
        float cnt = 0.f;
        for (int iz = zl; iz <= zh; ++iz) {
            const int dz = z - g.start(2, iz);
            for (int iy = yl; iy <= yh; ++iy) {
                const int dy = y - g.start(1, iy);
                for (int ix = xl; ix <= xh; ++ix) {
                    const int dx = x - g.start(0, ix);
                    const int64_t ip = ix + (int64_t)g.ns[0] * (iy + (int64_t)g.ns[1] * iz);
                    const int64_t i = dx + (int64_t)g.ps[0] * (dy + (int64_t)g.ps[1] * dz);
                    acc = acc + P[i + g.p * (t + (int64_t)g.Nt * ip)];
                    cnt += 1.f;
                }
            }
        }
        const float d = (divide && cnt > 1.f) ? cnt : 1.f;
        img[e] = c64{acc.re / d, acc.im / d};
    }
}

// out = img / max(coverage count, 1) per voxel (the adjoint of the division in
// patches2img); in place allowed.
__global__ void k_div_coverage(const c64 *img, c64 *out, PatchGeom g) {
    const int64_t V = (int64_t)g.N[0] * g.N[1] * g.N[2];
    const int64_t n = V * g.Nt;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t v = e % V;
        const int x = (int)(v % g.N[0]);
        const int y = (int)((v / g.N[0]) % g.N[1]);
        const int z = (int)(v / ((int64_t)g.N[0] * g.N[1]));
        int xl, xh, yl, yh, zl, zh;
        cover_range(g, 0, x, xl, xh);
        cover_range(g, 1, y, yl, yh);
        cover_range(g, 2, z, zl, zh);
        const float cnt = (float)(max(xh - xl + 1, 0) * max(yh - yl + 1, 0) * max(zh - zl + 1, 0));
        const float d = cnt > 1.f ? cnt : 1.f;
        const c64 a = img[e];
        out[e] = c64{a.re / d, a.im / d};
    }
}

static unsigned grid_for(int64_t n) {
    int64_t nb = cdiv(n, 256);
    if (nb > 1 << 20) nb = 1 << 20;
    if (nb < 1) nb = 1;
    return (unsigned)nb;
}

void launch_img2patches(mslr_ctx_s *ctx, const PatchGeom &g, const c64 *img, c64 *P, int64_t ip0,
                        int64_t np) {
    const int64_t n = g.p * g.Nt * np;
    if (n == 0) return;
    {
        MSLR_PROF(ctx, "img2patches");
        const int64_t nb = cdiv(g.p, 256) * np;
        if (nb < (int64_t)1 << 31)
            k_img2patches_fr<<<(unsigned)nb, 256, 0, ctx->stream>>>(img, P, g, ip0, np);
        else
            k_img2patches<<<grid_for(n), 256, 0, ctx->stream>>>(img, P, g, ip0, np);
        HIP_LAUNCH_CHECK();
    }
}

void launch_patches2img(mslr_ctx_s *ctx, const PatchGeom &g, const c64 *P, c64 *img) {
    const int64_t n = (int64_t)g.N[0] * g.N[1] * g.N[2] * g.Nt;
    {
        MSLR_PROF(ctx, "patches2img");
        const int64_t nb = cdiv((int64_t)g.N[0] * g.N[1] * g.N[2], 256);
        if (g.exact_tiling() && nb < (int64_t)1 << 31)
            k_patches2img_tiled<<<(unsigned)nb, 256, 0, ctx->stream>>>(P, img, g);
        else
            k_patches2img<<<grid_for(n), 256, 0, ctx->stream>>>(P, img, g, 1);
        HIP_LAUNCH_CHECK();
    }
}

void launch_patches2img_sum(mslr_ctx_s *ctx, const PatchGeom &g, const c64 *P, c64 *img) {
    const int64_t n = (int64_t)g.N[0] * g.N[1] * g.N[2] * g.Nt;
    if (n == 0) return;
    MSLR_PROF(ctx, "patches2img");
    k_patches2img<<<grid_for(n), 256, 0, ctx->stream>>>(P, img, g, 0);
    HIP_LAUNCH_CHECK();
}

void launch_div_coverage(mslr_ctx_s *ctx, const PatchGeom &g, const c64 *img, c64 *out) {
    const int64_t n = (int64_t)g.N[0] * g.N[1] * g.N[2] * g.Nt;
    if (n == 0) return;
    MSLR_PROF(ctx, "div_coverage");
    k_div_coverage<<<grid_for(n), 256, 0, ctx->stream>>>(img, out, g);
    HIP_LAUNCH_CHECK();
}

}  // namespace mslr

using namespace mslr;

extern "C" {

mslr_status_t mslr_patch_geometry(const int64_t N[3], const int64_t ps[3], const int64_t ss[3],
                                  int64_t ps_clamped[3], int64_t nsteps[3], int64_t *Np) {
    return guarded([&] {
        MSLR_REQUIRE(N && ps && ss, MSLR_ERR_ARG, "null argument");
        PatchGeom g = make_geom(N, 1, ps, ss);
        for (int a = 0; a < 3; ++a) {
            if (ps_clamped) ps_clamped[a] = g.ps[a];
            if (nsteps) nsteps[a] = g.ns[a] - 1;
        }
        if (Np) *Np = g.Np;
    });
}

mslr_status_t mslr_img2patches(mslr_ctx ctx, const void *img_dev, const int64_t N[3], int64_t Nt,
                               const int64_t ps[3], const int64_t ss[3], void *P_dev) {
    return guarded([&] {
        mslr_ctx_s *c = check_ctx(ctx);
        PatchGeom g = make_geom(N, Nt, ps, ss);
        launch_img2patches(c, g, (const c64 *)img_dev, (c64 *)P_dev, 0, g.Np);
    });
}

mslr_status_t mslr_patches2img(mslr_ctx ctx, const void *P_dev, const int64_t N[3], int64_t Nt,
                               const int64_t ps[3], const int64_t ss[3], void *img_dev) {
    return guarded([&] {
        mslr_ctx_s *c = check_ctx(ctx);
        PatchGeom g = make_geom(N, Nt, ps, ss);
        launch_patches2img(c, g, (const c64 *)P_dev, (c64 *)img_dev);
    });
}

}  // extern "C"
