// Patch geometry (src/recon.jl:53-78) and the patch gather / scatter kernels.
#pragma once

#include "common.hpp"

namespace mslr {

struct PatchGeom {
    int N[3];     // image size
    int ps[3];    // patch size, clamped to the image
    int ss[3];    // stride
    int ns[3];    // patches per axis (Nsteps + 1)
    int64_t Np;   // number of patches
    int64_t p;    // voxels per patch
    int Nt;

    // 0-based start of patch index i along axis a: min(i*s, N - ps).
    __host__ __device__ __forceinline__ int start(int a, int i) const {
        const int s = i * ss[a];
        const int lim = N[a] - ps[a];
        return s < lim ? s : lim;
    }
    // true when no voxel is covered by more than one patch
    bool disjoint() const;
    // true when every voxel is covered exactly once (non-overlapping tiling)
    bool exact_tiling() const;
};

PatchGeom make_geom(const int64_t N[3], int64_t Nt, const int64_t ps[3], const int64_t ss[3]);

// Image offset (frame 0) of element i of patch ip.
__device__ __forceinline__ int64_t patch_voxel(const PatchGeom &g, int64_t ip, int i) {
    const int ix = (int)(ip % g.ns[0]);
    const int iy = (int)((ip / g.ns[0]) % g.ns[1]);
    const int iz = (int)(ip / ((int64_t)g.ns[0] * g.ns[1]));
    const int dx = i % g.ps[0];
    const int dy = (i / g.ps[0]) % g.ps[1];
    const int dz = i / (g.ps[0] * g.ps[1]);
    const int x = g.start(0, ix) + dx;
    const int y = g.start(1, iy) + dy;
    const int z = g.start(2, iz) + dz;
    return x + (int64_t)g.N[0] * (y + (int64_t)g.N[1] * z);
}

// P (p, Nt, Np) <- img (Nx,Ny,Nz,Nt), patches [ip0, ip0+np) only (P holds those np patches).
void launch_img2patches(mslr_ctx_s *ctx, const PatchGeom &g, const c64 *img, c64 *P,
                        int64_t ip0, int64_t np);
// img <- sum of covering patches / max(count, 1), summed in patch order.
void launch_patches2img(mslr_ctx_s *ctx, const PatchGeom &g, const c64 *P, c64 *img);
// img <- sum of covering patches (the adjoint of launch_img2patches).
void launch_patches2img_sum(mslr_ctx_s *ctx, const PatchGeom &g, const c64 *P, c64 *img);
// out <- img / max(count, 1) per voxel (the adjoint of the division); in place allowed.
void launch_div_coverage(mslr_ctx_s *ctx, const PatchGeom &g, const c64 *img, c64 *out);

}  // namespace mslr
