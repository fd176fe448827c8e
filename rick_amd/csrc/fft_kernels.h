// Complex arithmetic and in-register DFT butterflies shared by the FFT
// kernels (fft.hip: power-of-two Stockham engine; fft_gen.hip: mixed-radix
// and Bluestein transforms of any length).  All transforms use the sign of
// FFTW_BACKWARD, exp(+2*pi*i*n*k/N).
#pragma once

#include <hip/hip_runtime.h>

namespace rick {
namespace fftk {

template <typename T>
struct C2 {
    T x, y;
};

template <typename T>
__device__ __forceinline__ C2<T> cadd(C2<T> a, C2<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ C2<T> csub(C2<T> a, C2<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ C2<T> cmul(C2<T> a, C2<T> b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// Multiplication by +i (backward transforms use exp(+2 pi i ...)).
template <typename T>
__device__ __forceinline__ C2<T> mul_i(C2<T> a) { return {-a.y, a.x}; }

// exp(+2*pi*i*m/16) with m known at compile time.
template <typename T, int M>
__device__ __forceinline__ C2<T> rot16(C2<T> a)
{
    constexpr int m = M & 15;
    constexpr double c1 = 0.92387953251128675613, s1 = 0.38268343236508977173,
                     h = 0.70710678118654752440;
    if constexpr (m == 0) return a;
    else if constexpr (m == 4) return mul_i(a);
    else if constexpr (m == 8) return {-a.x, -a.y};
    else if constexpr (m == 12) return {a.y, -a.x};
    else {
        constexpr double tr[16] = {1, c1, h, s1, 0, -s1, -h, -c1, -1, -c1, -h, -s1, 0, s1, h, c1};
        constexpr double re = tr[m], im = tr[(m + 12) & 15];
        return cmul(a, C2<T>{(T)re, (T)im});
    }
}

template <typename T>
__device__ __forceinline__ void dft4(C2<T> &x0, C2<T> &x1, C2<T> &x2, C2<T> &x3)
{
    const C2<T> a = cadd(x0, x2), b = csub(x0, x2), c = cadd(x1, x3), d = mul_i(csub(x1, x3));
    x0 = cadd(a, c);
    x2 = csub(a, c);
    x1 = cadd(b, d);
    x3 = csub(b, d);
}

// In-register DFT with sign +1: X[k] = sum_n x[n] exp(+2 pi i n k / R).
template <typename T, int R>
__device__ __forceinline__ void dftR(C2<T> *v)
{
    if constexpr (R == 2) {
        const C2<T> a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (R == 4) {
        dft4(v[0], v[1], v[2], v[3]);
    } else if constexpr (R == 8) {
        // n = 2*n1 + n2, k = k1 + 4*k2.
        C2<T> e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
        C2<T> o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
        dft4(e0, e1, e2, e3);
        dft4(o0, o1, o2, o3);
        o1 = rot16<T, 2>(o1);
        o2 = rot16<T, 4>(o2);
        o3 = rot16<T, 6>(o3);
        v[0] = cadd(e0, o0); v[4] = csub(e0, o0);
        v[1] = cadd(e1, o1); v[5] = csub(e1, o1);
        v[2] = cadd(e2, o2); v[6] = csub(e2, o2);
        v[3] = cadd(e3, o3); v[7] = csub(e3, o3);
    } else {
        // n = 4*n1 + n2, k = k1 + 4*k2.
        C2<T> y[4][4];
#pragma unroll
        for (int n2 = 0; n2 < 4; ++n2) {
            y[n2][0] = v[n2];
            y[n2][1] = v[4 + n2];
            y[n2][2] = v[8 + n2];
            y[n2][3] = v[12 + n2];
            dft4(y[n2][0], y[n2][1], y[n2][2], y[n2][3]);
        }
        y[1][1] = rot16<T, 1>(y[1][1]); y[1][2] = rot16<T, 2>(y[1][2]); y[1][3] = rot16<T, 3>(y[1][3]);
        y[2][1] = rot16<T, 2>(y[2][1]); y[2][2] = rot16<T, 4>(y[2][2]); y[2][3] = rot16<T, 6>(y[2][3]);
        y[3][1] = rot16<T, 3>(y[3][1]); y[3][2] = rot16<T, 6>(y[3][2]); y[3][3] = rot16<T, 9>(y[3][3]);
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) {
            dft4(y[0][k1], y[1][k1], y[2][k1], y[3][k1]);
            v[k1] = y[0][k1];
            v[k1 + 4] = y[1][k1];
            v[k1 + 8] = y[2][k1];
            v[k1 + 12] = y[3][k1];
        }
    }
}

// Radix 3: X1,2 = x0 - (x1+x2)/2 +- i*(sqrt(3)/2)*(x1-x2).
template <typename T>
__device__ __forceinline__ void dft3(C2<T> *v)
{
    constexpr double h = 0.86602540378443864676;  // sin(2 pi/3)
    const C2<T> t1 = cadd(v[1], v[2]);
    const C2<T> t2 = {v[0].x - T(0.5) * t1.x, v[0].y - T(0.5) * t1.y};
    const C2<T> d = csub(v[1], v[2]);
    const C2<T> e = mul_i(C2<T>{(T)h * d.x, (T)h * d.y});
    v[0] = cadd(v[0], t1);
    v[1] = cadd(t2, e);
    v[2] = csub(t2, e);
}

// Radix 5 from the symmetric sums a = x1+x4, x2+x3 and differences b = x1-x4,
// x2-x3 (c_k = cos(2 pi k/5), s_k = sin(2 pi k/5)).
template <typename T>
__device__ __forceinline__ void dft5(C2<T> *v)
{
    constexpr double c1 = 0.30901699437494742410, c2 = -0.80901699437494742410,
                     s1 = 0.95105651629515357212, s2 = 0.58778525229247312917;
    const C2<T> a1 = cadd(v[1], v[4]), a2 = cadd(v[2], v[3]);
    const C2<T> b1 = csub(v[1], v[4]), b2 = csub(v[2], v[3]);
    const C2<T> x0 = v[0];
    const C2<T> r1 = {x0.x + (T)c1 * a1.x + (T)c2 * a2.x, x0.y + (T)c1 * a1.y + (T)c2 * a2.y};
    const C2<T> r2 = {x0.x + (T)c2 * a1.x + (T)c1 * a2.x, x0.y + (T)c2 * a1.y + (T)c1 * a2.y};
    const C2<T> i1 = mul_i(C2<T>{(T)s1 * b1.x + (T)s2 * b2.x, (T)s1 * b1.y + (T)s2 * b2.y});
    const C2<T> i2 = mul_i(C2<T>{(T)s2 * b1.x - (T)s1 * b2.x, (T)s2 * b1.y - (T)s1 * b2.y});
    v[0] = cadd(x0, cadd(a1, a2));
    v[1] = cadd(r1, i1);
    v[4] = csub(r1, i1);
    v[2] = cadd(r2, i2);
    v[3] = csub(r2, i2);
}

// Any radix R by direct summation, with the R-th roots of unity w[m] =
// exp(+2 pi i m/R) supplied by the caller (R = 7, 11, 13 here).
template <typename T, int R>
__device__ __forceinline__ void dft_direct(C2<T> *v, const C2<T> *w)
{
    C2<T> out[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        C2<T> acc = v[0];
#pragma unroll
        for (int n = 1; n < R; ++n) acc = cadd(acc, cmul(v[n], w[(n * k) % R]));
        out[k] = acc;
    }
#pragma unroll
    for (int k = 0; k < R; ++k) v[k] = out[k];
}

}  // namespace fftk
}  // namespace rick
