// In-register DFT codelets of compile-time size: radix 2, 3, 4, 5 and the two-level
// Cooley-Tukey forms 6, 8, 9, 10 (constant twiddles, computed at compile time).  The
// prime-factor codelets and the two-stage LDS transforms built on them are in
// fft2d.hpp.
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

}  // namespace mslr
