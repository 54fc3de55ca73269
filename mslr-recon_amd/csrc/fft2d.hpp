// Two-stage in-place FFTs of the (y, z) planes for the SENSE y-z passes.
//
// Each axis of length N = N1 * N2 is held in LDS as an N1 x N2 array: position
// p = a * N2 + b.  Stage A transforms over a (N1 points, one butterfly per b),
// stage B over b (N2 points, one butterfly per a); each butterfly reads its
// points into registers, transforms them there and writes them back to the same
// positions, so a stage needs no second buffer and no barrier between its reads
// and writes.  Two index maps make this a DFT:
//   Cooley-Tukey (CT): position (a, b) holds spatial index b + N2 a and, after
//     both stages, frequency a + N1 b; stage B multiplies by W_N^(a b) (before its
//     DFT in the forward direction, after it in the inverse one).
//   Good-Thomas prime factor (PFA, gcd(N1, N2) = 1): position (a, b) holds spatial
//     index (N2 a + N1 b) mod N and frequency k with k = a (mod N1), k = b (mod N2);
//     no twiddles at all.
// The inverse transform with the same maps takes the frequency layout back to the
// spatial one, so forward and adjoint share one layout.  All LDS offsets of a
// butterfly are compile-time constants from its base, and the in-register DFTs of
// composite coprime sizes are prime-factor codelets (no internal twiddles either).
#pragma once

#include "fft_fixed.hpp"

namespace mslr {

// k in [0, R1 R2) with k = k1 (mod R1) and k = k2 (mod R2)
constexpr int crt_index(int k1, int k2, int R1, int R2) {
    for (int k = 0; k < R1 * R2; ++k)
        if (k % R1 == k1 && k % R2 == k2) return k;
    return -1;
}

// In-register prime-factor DFT of R1*R2 points (gcd(R1, R2) = 1), natural order
// in and out: x'[n1][n2] = v[(R2 n1 + R1 n2) mod R], a 2-D DFT without twiddles,
// X[k] = X'[k mod R1][k mod R2].
template <class T, int R1, int R2, int SGN>
struct DftPfa {
    static __device__ __forceinline__ void run(C2<T> *v) {
        constexpr int R = R1 * R2;
        C2<T> t[R];
#pragma unroll
        for (int n2 = 0; n2 < R2; ++n2) {
            C2<T> u[R1];
#pragma unroll
            for (int n1 = 0; n1 < R1; ++n1) u[n1] = v[(R2 * n1 + R1 * n2) % R];
            DftR<T, R1, SGN>::run(u);
#pragma unroll
            for (int k1 = 0; k1 < R1; ++k1) t[k1 * R2 + n2] = u[k1];
        }
#pragma unroll
        for (int k1 = 0; k1 < R1; ++k1) {
            C2<T> u[R2];
#pragma unroll
            for (int n2 = 0; n2 < R2; ++n2) u[n2] = t[k1 * R2 + n2];
            DftR<T, R2, SGN>::run(u);
#pragma unroll
            for (int k2 = 0; k2 < R2; ++k2) v[crt_index(k1, k2, R1, R2)] = u[k2];
        }
    }
};

// Codelets of the plane stages: prime-factor forms where the size splits into
// coprime factors, the Cooley-Tukey register DFTs otherwise.
template <class T, int R, int SGN>
struct DftP : DftR<T, R, SGN> {};
template <class T, int SGN>
struct DftP<T, 6, SGN> : DftPfa<T, 2, 3, SGN> {};
template <class T, int SGN>
struct DftP<T, 10, SGN> : DftPfa<T, 2, 5, SGN> {};
template <class T, int SGN>
struct DftP<T, 12, SGN> : DftPfa<T, 4, 3, SGN> {};
template <class T, int SGN>
struct DftP<T, 15, SGN> : DftPfa<T, 3, 5, SGN> {};
template <class T, int SGN>
struct DftP<T, 20, SGN> : DftPfa<T, 4, 5, SGN> {};

// One axis: N = N1 * N2 with the CT or PFA index maps.
template <int N1_, int N2_, bool PFA_>
struct PlaneAxis {
    static constexpr int N1 = N1_, N2 = N2_, N = N1_ * N2_;
    static constexpr bool PFA = PFA_;
    // spatial index held by position (a, b)
    static constexpr __host__ __device__ int spatial(int a, int b) {
        return PFA ? (N2 * a + N1 * b) % N : N2 * a + b;
    }
    // position a * N2 + b holding frequency k after the forward transform
    static constexpr __host__ __device__ int freq_position(int k) {
        return PFA ? (k % N1) * N2 + (k % N2) : (k % N1) * N2 + k / N1;
    }
};

// The plane plan: y axis CT or PFA, z axis PFA; LDS row stride LY (complex values)
// >= NY; NT threads per workgroup.  Stage butterfly counts:
//   yA: N2y * NZ (lane u: b = u % N2y, z position pz = u / N2y)
//   yB: N1y * NZ (lane u: a = u % N1y, pz = u / N1y)
//   zA: N2z * NY (lane u: y position py = u % NY, b = u / NY)
//   zB: N1z * NY (lane u: py = u % NY, a = u / NY)
template <class AY, class AZ, int NT_>
struct PlanePlan {
    static_assert(AZ::PFA, "the z axis uses the prime-factor map");
    using Y = AY;
    using Z = AZ;
    static constexpr int NY = AY::N, NZ = AZ::N, NYZ = NY * NZ;
    static constexpr int LY = NY;
    static constexpr int NT = NT_;
    static constexpr int UYA = AY::N2 * NZ, UYB = AY::N1 * NZ;
    static constexpr int UZA = AZ::N2 * NY, UZB = AZ::N1 * NY;
    static constexpr int rounds(int u) { return (u + NT - 1) / NT; }
    static constexpr int RYA = rounds(UYA), RYB = rounds(UYB), RZA = rounds(UZA),
                         RZB = rounds(UZB);
    // LDS offset of y position py and z position pz
    static constexpr __host__ __device__ int lds(int py, int pz) { return pz * LY + py; }
    // spatial z of z position pz
    static constexpr __host__ __device__ int zspat(int pz) {
        return AZ::spatial(pz / AZ::N2, pz % AZ::N2);
    }
};

}  // namespace mslr
