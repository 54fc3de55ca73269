// Mixed-radix (2/3/4/5) Stockham FFT on lines held in LDS.
//
// A workgroup transforms `nlines` lines of length n that live in shared memory
// (element i of line l at buf[l*lstride + i*estride]).  Each radix pass reads
// its butterfly inputs into registers, synchronises, writes the outputs to
// their autosorted positions and synchronises again, so a transform needs no
// second LDS buffer.  Twiddles come from a per-length table tw[m] = exp(+2*pi*i*m/n)
// (conjugated on the fly for the negative-exponent direction).
#pragma once

#include "common.hpp"

namespace mslr {

constexpr int FFT_MAX_PASSES = 8;

struct FftDesc {
    int n;
    int npass;
    int radix[FFT_MAX_PASSES];
};

// Factor n into radices 4, 2, 3, 5 (radix-4 first).  Returns false when n has
// a prime factor larger than 5 or needs more passes than supported.
inline bool fft_make_desc(int n, FftDesc &d) {
    d.n = n;
    d.npass = 0;
    if (n < 1) return false;
    int m = n;
    auto push = [&](int r) {
        if (d.npass >= FFT_MAX_PASSES) return false;
        d.radix[d.npass++] = r;
        return true;
    };
    while (m % 4 == 0) {
        if (!push(4)) return false;
        m /= 4;
    }
    while (m % 2 == 0) {
        if (!push(2)) return false;
        m /= 2;
    }
    while (m % 3 == 0) {
        if (!push(3)) return false;
        m /= 3;
    }
    while (m % 5 == 0) {
        if (!push(5)) return false;
        m /= 5;
    }
    return m == 1;
}

// In-register DFT of size R with exponent sign sgn (-1 forward, +1 inverse).
template <class T, int R>
__device__ __forceinline__ void dft_small(C2<T> *v, int sgn);

template <class T>
__device__ __forceinline__ C2<T> mul_i(C2<T> a, int sgn) {  // sgn * i * a
    return sgn > 0 ? C2<T>{-a.im, a.re} : C2<T>{a.im, -a.re};
}

template <>
__device__ __forceinline__ void dft_small<float, 2>(c64 *v, int) {
    c64 a = v[0], b = v[1];
    v[0] = a + b;
    v[1] = a - b;
}
template <>
__device__ __forceinline__ void dft_small<double, 2>(c128 *v, int) {
    c128 a = v[0], b = v[1];
    v[0] = a + b;
    v[1] = a - b;
}

template <class T>
__device__ __forceinline__ void dft3(C2<T> *v, int sgn) {
    const T h = T(0.5);
    const T s3 = T(0.86602540378443864676372317075294);
    C2<T> t1 = v[1] + v[2];
    C2<T> t2 = v[0] - h * t1;
    C2<T> t3 = s3 * mul_i(v[1] - v[2], sgn);
    v[0] = v[0] + t1;
    v[1] = t2 + t3;
    v[2] = t2 - t3;
}
template <>
__device__ __forceinline__ void dft_small<float, 3>(c64 *v, int sgn) { dft3(v, sgn); }
template <>
__device__ __forceinline__ void dft_small<double, 3>(c128 *v, int sgn) { dft3(v, sgn); }

template <class T>
__device__ __forceinline__ void dft4(C2<T> *v, int sgn) {
    C2<T> a = v[0] + v[2], b = v[0] - v[2];
    C2<T> c = v[1] + v[3], d = mul_i(v[1] - v[3], sgn);
    v[0] = a + c;
    v[1] = b + d;
    v[2] = a - c;
    v[3] = b - d;
}
template <>
__device__ __forceinline__ void dft_small<float, 4>(c64 *v, int sgn) { dft4(v, sgn); }
template <>
__device__ __forceinline__ void dft_small<double, 4>(c128 *v, int sgn) { dft4(v, sgn); }

template <class T>
__device__ __forceinline__ void dft5(C2<T> *v, int sgn) {
    const T c1 = T(0.30901699437494742410229341718282);   // cos(2pi/5)
    const T c2 = T(-0.80901699437494742410229341718282);  // cos(4pi/5)
    const T s1 = T(0.95105651629515357211643933337938);   // sin(2pi/5)
    const T s2 = T(0.58778525229247312916870595463907);   // sin(4pi/5)
    C2<T> t1 = v[1] + v[4], t2 = v[2] + v[3];
    C2<T> t3 = v[1] - v[4], t4 = v[2] - v[3];
    C2<T> a1 = v[0] + c1 * t1 + c2 * t2;
    C2<T> a2 = v[0] + c2 * t1 + c1 * t2;
    C2<T> b1 = mul_i(s1 * t3 + s2 * t4, sgn);
    C2<T> b2 = mul_i(s2 * t3 - s1 * t4, sgn);
    v[0] = v[0] + t1 + t2;
    v[1] = a1 + b1;
    v[4] = a1 - b1;
    v[2] = a2 + b2;
    v[3] = a2 - b2;
}
template <>
__device__ __forceinline__ void dft_small<float, 5>(c64 *v, int sgn) { dft5(v, sgn); }
template <>
__device__ __forceinline__ void dft_small<double, 5>(c128 *v, int sgn) { dft5(v, sgn); }

// One Stockham pass of radix R over all lines.  MAXE bounds the elements one
// thread holds: nlines*n <= MAXE*nthr must hold.
template <class T, int R, int MAXE>
__device__ __forceinline__ void fft_pass(C2<T> *buf, int n, int Ns, int nlines, int lstride,
                                         int estride, const C2<T> *__restrict__ tw, int sgn,
                                         int tid, int nthr) {
    constexpr int MAXB = (MAXE + R - 1) / R;
    const int nb = n / R;
    const int total = nb * nlines;
    const int tstep = n / (Ns * R);
    C2<T> v[MAXB][R];
    int ob[MAXB];
#pragma unroll
    for (int q = 0; q < MAXB; ++q) {
        const int b = tid + q * nthr;
        if (b < total) {
            const int line = b / nb;
            const int j = b - line * nb;
            const int k = j % Ns;
            const C2<T> *src = buf + line * lstride + j * estride;
#pragma unroll
            for (int i = 0; i < R; ++i) v[q][i] = src[i * nb * estride];
            if (Ns > 1) {
#pragma unroll
                for (int i = 1; i < R; ++i) {
                    C2<T> w = tw[i * k * tstep];
                    if (sgn < 0) w.im = -w.im;
                    v[q][i] = v[q][i] * w;
                }
            }
            dft_small<T, R>(v[q], sgn);
            ob[q] = line * lstride + ((j - k) * R + k) * estride;
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MAXB; ++q) {
        const int b = tid + q * nthr;
        if (b < total) {
#pragma unroll
            for (int i = 0; i < R; ++i) buf[ob[q] + i * Ns * estride] = v[q][i];
        }
    }
    __syncthreads();
}

// Full transform of `nlines` lines (the caller synchronises before the call;
// the call ends synchronised).
template <class T, int MAXE>
__device__ __forceinline__ void fft_lines(C2<T> *buf, const FftDesc &d, int nlines, int lstride,
                                          int estride, const C2<T> *__restrict__ tw, int sgn) {
    const int tid = threadIdx.x;
    const int nthr = blockDim.x;
    int Ns = 1;
    for (int p = 0; p < d.npass; ++p) {
        const int R = d.radix[p];
        switch (R) {
            case 2:
                fft_pass<T, 2, MAXE>(buf, d.n, Ns, nlines, lstride, estride, tw, sgn, tid, nthr);
                break;
            case 3:
                fft_pass<T, 3, MAXE>(buf, d.n, Ns, nlines, lstride, estride, tw, sgn, tid, nthr);
                break;
            case 4:
                fft_pass<T, 4, MAXE>(buf, d.n, Ns, nlines, lstride, estride, tw, sgn, tid, nthr);
                break;
            default:
                fft_pass<T, 5, MAXE>(buf, d.n, Ns, nlines, lstride, estride, tw, sgn, tid, nthr);
                break;
        }
        Ns *= R;
    }
}

}  // namespace mslr
