// Fixed-length Stockham FFT on lines in LDS, specialised at compile time.
//
// Each length is split into two (or one) composite radices whose butterflies
// run entirely in registers (radix 6, 8, 9, 10 are two-level Cooley-Tukey
// DFTs with constant twiddles), so a length-90 line takes 2 LDS round trips
// instead of the 4 of a radix-2/3/5 decomposition, and every index
// computation divides by compile-time constants.
#pragma once

#include "fft.hpp"

namespace mslr {

// ---------------------------------------------------------------- constexpr trig
constexpr double kPi = 3.14159265358979323846264338327950288;

constexpr double c_angle(int num, int den) {
    // 2 pi num / den reduced to (-pi, pi]
    const int m = ((num % den) + den) % den;
    double a = 2.0 * kPi * (double)m / (double)den;
    if (a > kPi) a -= 2.0 * kPi;
    return a;
}
constexpr double c_cos(double x) {
    double x2 = x * x, term = 1.0, sum = 1.0;
    for (int i = 1; i < 28; ++i) {
        term *= -x2 / ((2.0 * i - 1.0) * (2.0 * i));
        sum += term;
    }
    return sum;
}
constexpr double c_sin(double x) {
    double x2 = x * x, term = x, sum = x;
    for (int i = 1; i < 28; ++i) {
        term *= -x2 / ((2.0 * i) * (2.0 * i + 1.0));
        sum += term;
    }
    return sum;
}

// ---------------------------------------------------------------- register DFTs
// dftR<T, R, SGN>(v): in-place DFT of R values with exponent sign SGN.
template <class T, int R, int SGN>
struct DftR;

template <class T, int SGN>
struct DftR<T, 1, SGN> {
    static __device__ __forceinline__ void run(C2<T> *) {}
};
template <class T, int SGN>
struct DftR<T, 2, SGN> {
    static __device__ __forceinline__ void run(C2<T> *v) {
        C2<T> a = v[0], b = v[1];
        v[0] = a + b;
        v[1] = a - b;
    }
};
template <class T, int SGN>
struct DftR<T, 3, SGN> {
    static __device__ __forceinline__ void run(C2<T> *v) { dft3(v, SGN); }
};
template <class T, int SGN>
struct DftR<T, 4, SGN> {
    static __device__ __forceinline__ void run(C2<T> *v) { dft4(v, SGN); }
};
template <class T, int SGN>
struct DftR<T, 5, SGN> {
    static __device__ __forceinline__ void run(C2<T> *v) { dft5(v, SGN); }
};

// Two-level DFT of R1*R2 values: X[k1 + R1 k2] =
//   sum_j2 W_R^(j2 k1) W_R2^(j2 k2) sum_j1 x[j2 + R2 j1] W_R1^(j1 k1)
template <class T, int R1, int R2, int SGN>
struct DftCT {
    static __device__ __forceinline__ void run(C2<T> *v) {
        constexpr int R = R1 * R2;
        C2<T> a[R];
#pragma unroll
        for (int j2 = 0; j2 < R2; ++j2) {
            C2<T> t[R1];
#pragma unroll
            for (int j1 = 0; j1 < R1; ++j1) t[j1] = v[j2 + R2 * j1];
            DftR<T, R1, SGN>::run(t);
#pragma unroll
            for (int k1 = 0; k1 < R1; ++k1) {
                const double ang = c_angle(SGN * j2 * k1, R);
                if (j2 * k1 == 0) {
                    a[j2 + R2 * k1] = t[k1];
                } else {
                    const C2<T> w = {(T)c_cos(ang), (T)c_sin(ang)};
                    a[j2 + R2 * k1] = t[k1] * w;
                }
            }
        }
#pragma unroll
        for (int k1 = 0; k1 < R1; ++k1) {
            C2<T> t[R2];
#pragma unroll
            for (int j2 = 0; j2 < R2; ++j2) t[j2] = a[j2 + R2 * k1];
            DftR<T, R2, SGN>::run(t);
#pragma unroll
            for (int k2 = 0; k2 < R2; ++k2) v[k1 + R1 * k2] = t[k2];
        }
    }
};
template <class T, int SGN>
struct DftR<T, 6, SGN> : DftCT<T, 2, 3, SGN> {};
template <class T, int SGN>
struct DftR<T, 8, SGN> : DftCT<T, 2, 4, SGN> {};
template <class T, int SGN>
struct DftR<T, 9, SGN> : DftCT<T, 3, 3, SGN> {};
template <class T, int SGN>
struct DftR<T, 10, SGN> : DftCT<T, 2, 5, SGN> {};

// ---------------------------------------------------------------- plans
// Radix sequence per supported length (first pass radix R0 with span 1,
// second R1 with span R0).
template <int N>
struct FixPlan {
    static constexpr bool ok = false;
    static constexpr int R0 = 1, R1 = 1;
};
#define MSLR_FIXPLAN(N_, A_, B_)                 \
    template <>                                  \
    struct FixPlan<N_> {                         \
        static constexpr bool ok = true;         \
        static constexpr int R0 = A_, R1 = B_;   \
    };
MSLR_FIXPLAN(90, 10, 9)
MSLR_FIXPLAN(60, 10, 6)
MSLR_FIXPLAN(64, 8, 8)
MSLR_FIXPLAN(40, 10, 4)
MSLR_FIXPLAN(32, 8, 4)
MSLR_FIXPLAN(24, 8, 3)
MSLR_FIXPLAN(12, 4, 3)
MSLR_FIXPLAN(10, 10, 1)
#undef MSLR_FIXPLAN

inline bool fixed_length(int n) {
    switch (n) {
        case 90: case 60: case 64: case 40: case 32: case 24: case 12: case 10:
            return true;
        default:
            return false;
    }
}

// One Stockham pass of radix R and span NS over `nlines` lines of length N.
template <class T, int N, int R, int NS, int SGN, int MAXB>
__device__ __forceinline__ void fix_pass(C2<T> *buf, int nlines, int lstride, int estride,
                                         const C2<T> *__restrict__ tw, int tid, int nthr) {
    constexpr int NB = N / R;
    constexpr int TSTEP = N / (NS * R);
    const int total = NB * nlines;
    C2<T> v[MAXB][R];
    int ob[MAXB];
#pragma unroll
    for (int q = 0; q < MAXB; ++q) {
        const int b = tid + q * nthr;
        if (b < total) {
            const int line = b / NB;
            const int j = b - line * NB;
            const int k = j % NS;
            const C2<T> *src = buf + line * lstride + j * estride;
#pragma unroll
            for (int i = 0; i < R; ++i) v[q][i] = src[i * NB * estride];
            if (NS > 1) {
#pragma unroll
                for (int i = 1; i < R; ++i) {
                    C2<T> w = tw[i * k * TSTEP];
                    if (SGN < 0) w.im = -w.im;
                    v[q][i] = v[q][i] * w;
                }
            }
            DftR<T, R, SGN>::run(v[q]);
            ob[q] = line * lstride + ((j - k) * R + k) * estride;
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MAXB; ++q) {
        const int b = tid + q * nthr;
        if (b < total) {
#pragma unroll
            for (int i = 0; i < R; ++i) buf[ob[q] + i * NS * estride] = v[q][i];
        }
    }
    __syncthreads();
}

// Transform of nlines <= MAXLINES lines of length N with NT threads.
template <class T, int N, int SGN, int MAXLINES, int NT>
__device__ __forceinline__ void fix_fft(C2<T> *buf, int nlines, int lstride, int estride,
                                        const C2<T> *__restrict__ tw) {
    using P = FixPlan<N>;
    static_assert(P::ok, "no fixed plan for this length");
    static_assert(P::R0 * P::R1 == N, "bad plan");
    const int tid = threadIdx.x;
    constexpr int MB0 = (MAXLINES * (N / P::R0) + NT - 1) / NT;
    fix_pass<T, N, P::R0, 1, SGN, MB0>(buf, nlines, lstride, estride, tw, tid, NT);
    if (P::R1 > 1) {
        constexpr int MB1 = (MAXLINES * (N / P::R1) + NT - 1) / NT;
        fix_pass<T, N, P::R1, P::R0, SGN, MB1>(buf, nlines, lstride, estride, tw, tid, NT);
    }
}

}  // namespace mslr
