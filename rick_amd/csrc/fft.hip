// Batched backward C2C FFT of w-planes for gfx950 (the fftw_data stage,
// fourier_transform.c:13-93: per plane a 2-D transform with sign +1, n0 = Ny,
// n1 = Nx, scaled by 1/(Nx*Ny)), operating directly on the interleaved
// plane-major grid layout.
//
// Structure:
//   * a Stockham autosort engine: each thread holds 16 complex values in
//     registers; a pass of radix R (16, 8, 4 or 2) reads R values, applies the
//     inter-pass twiddles, does an in-register DFT and writes back through LDS;
//     the first pass reads global memory and the last writes global memory.
//   * k_rows: lines are grid rows (contiguous u), G rows per workgroup.
//   * k_cols: lines are columns of a W-column strip; a line of length L has
//     element n at row  row_base + n*row_stride  and is written to row
//     out_base + n*out_stride, optionally multiplied by W_N^(n*tw_mult).  With
//     L = Ny this is the whole column transform; for long columns the
//     transform is split into two such passes (N = N1*N2, four-step):
//       A: N1-point DFTs over rows n2 + N2*n1, twiddle W_N^(k1*n2),
//          written to rows n2*N1 + k1 (out of place);
//       B: N2-point DFTs over rows k1 + N1*n2, written back in place, which
//          leaves X[k1 + N1*k2] at row k1 + N1*k2 (natural order).
// Twiddle tables hold exp(+2*pi*i*m/N) for the axis length N; an L-point
// sub-transform uses them with stride N/L.
#include "rick_internal.h"

#include <algorithm>

namespace rick {
namespace {

constexpr int kE = 16;  // complex values per thread

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

// exp(+2*pi*i*m/16), m = 0..15, as literals.
__device__ __forceinline__ double tw16_re(int m)
{
    constexpr double c[16] = {1.0, 0.92387953251128675613, 0.70710678118654752440,
                              0.38268343236508977173, 0.0, -0.38268343236508977173,
                              -0.70710678118654752440, -0.92387953251128675613, -1.0,
                              -0.92387953251128675613, -0.70710678118654752440,
                              -0.38268343236508977173, 0.0, 0.38268343236508977173,
                              0.70710678118654752440, 0.92387953251128675613};
    return c[m & 15];
}
__device__ __forceinline__ double tw16_im(int m) { return tw16_re(m - 4); }

// Multiply by exp(+2*pi*i*m/16) with m a compile-time constant after
// unrolling; the trivial cases avoid multiplications.
template <typename T>
__device__ __forceinline__ C2<T> rot16(C2<T> a, int m)
{
    m &= 15;
    if (m == 0) return a;
    if (m == 4) return mul_i(a);
    if (m == 8) return {-a.x, -a.y};
    if (m == 12) return {a.y, -a.x};
    const C2<T> w = {(T)tw16_re(m), (T)tw16_im(m)};
    return cmul(a, w);
}

// In-register DFT with sign +1: X[k] = sum_n x[n] exp(+2 pi i n k / R).
template <typename T>
__device__ __forceinline__ void dft2(C2<T> *v)
{
    const C2<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
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
template <typename T>
__device__ __forceinline__ void dft4(C2<T> *v) { dft4(v[0], v[1], v[2], v[3]); }
template <typename T>
__device__ __forceinline__ void dft8(C2<T> *v)
{
    // n = 2*n1 + n2, k = k1 + 4*k2.
    C2<T> y[2][4];
#pragma unroll
    for (int n2 = 0; n2 < 2; ++n2) {
#pragma unroll
        for (int n1 = 0; n1 < 4; ++n1) y[n2][n1] = v[2 * n1 + n2];
        dft4(y[n2]);
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) y[n2][k1] = rot16(y[n2][k1], 2 * n2 * k1);
    }
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        v[k1] = cadd(y[0][k1], y[1][k1]);
        v[k1 + 4] = csub(y[0][k1], y[1][k1]);
    }
}
template <typename T>
__device__ __forceinline__ void dft16(C2<T> *v)
{
    // n = 4*n1 + n2, k = k1 + 4*k2.
    C2<T> y[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
#pragma unroll
        for (int n1 = 0; n1 < 4; ++n1) y[n2][n1] = v[4 * n1 + n2];
        dft4(y[n2]);
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) y[n2][k1] = rot16(y[n2][k1], n2 * k1);
    }
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        C2<T> z0 = y[0][k1], z1 = y[1][k1], z2 = y[2][k1], z3 = y[3][k1];
        dft4(z0, z1, z2, z3);
        v[k1] = z0;
        v[k1 + 4] = z1;
        v[k1 + 8] = z2;
        v[k1 + 12] = z3;
    }
}
template <typename T, int R>
__device__ __forceinline__ void dftR(C2<T> *v)
{
    if constexpr (R == 2) dft2(v);
    else if constexpr (R == 4) dft4(v);
    else if constexpr (R == 8) dft8(v);
    else dft16(v);
}

__host__ __device__ __forceinline__ int lpad(int n) { return n + (n >> 4); }

// Line layouts.  Rows: line g element n at LDS g*LP + lpad(n), butterfly b ->
// (g = b / (L/R), j = b % (L/R)).  Columns: element (n, c) at n*W + c,
// butterfly b -> (c = b % W, j = b / W).
struct RowMap {
    int L, LP;
    __device__ int lds(int g, int n) const { return g * LP + lpad(n); }
};
struct ColMap {
    int L, W;
    __device__ int lds(int c, int n) const { return n * W + c; }
};

// One radix-R pass over the data held by this workgroup.  Values come from
// `src` (registers already loaded) and leave through LDS or registers; this
// helper only does twiddle + DFT on one butterfly.
template <typename T, int R>
__device__ __forceinline__ void butterfly(C2<T> *v, int j, int Ns, int L, int ts,
                                          const C2<T> *__restrict__ tw)
{
    if (Ns > 1) {
        const int k = j % Ns;
        const int step = (L / (Ns * R)) * ts;
#pragma unroll
        for (int r = 1; r < R; ++r) v[r] = cmul(v[r], tw[(r * k) * step]);
    }
    dftR<T, R>(v);
}

struct Passes {
    int np;
    int r[6];
};

__host__ Passes plan_passes(int L)
{
    Passes p{};
    int rest = L;
    while (rest >= 16 && (rest % 16) == 0) { p.r[p.np++] = 16; rest /= 16; }
    if (rest > 1) p.r[p.np++] = rest;  // 2, 4 or 8
    return p;
}

// Generic driver: Layout supplies global addressing for (line, n).
template <typename T, int R, bool ROWS>
__device__ __forceinline__ void do_pass(C2<T> *vals, int tid, int nthreads, int L, int G, int Ns,
                                        int ts, const C2<T> *tw, bool from_lds, bool to_lds,
                                        C2<T> *lds, int LP, int W,
                                        const C2<T> *gin, C2<T> *gout,
                                        size_t line_base0, size_t line_step, size_t estride,
                                        size_t out_base0, size_t out_step, size_t ostride,
                                        int tw_mult, int twN_stride, T scale)
{
    // butterflies per thread
    constexpr int BPT = kE / R;
    const int LR = L / R;
#pragma unroll
    for (int i = 0; i < BPT; ++i) {
        const int b = tid + i * nthreads;
        int g, j;
        if (ROWS) { g = b / LR; j = b % LR; }
        else { g = b % W; j = b / W; }
        C2<T> *v = vals + i * R;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int n = j + r * LR;
            if (from_lds) {
                v[r] = ROWS ? lds[g * LP + lpad(n)] : lds[n * W + g];
            } else {
                v[r] = gin[line_base0 + (size_t)g * line_step + (size_t)n * estride];
            }
        }
    }
    if (from_lds) __syncthreads();
#pragma unroll
    for (int i = 0; i < BPT; ++i) {
        const int b = tid + i * nthreads;
        const int j = ROWS ? b % LR : b / W;
        butterfly<T, R>(vals + i * R, j, Ns, L, ts, tw);
    }
#pragma unroll
    for (int i = 0; i < BPT; ++i) {
        const int b = tid + i * nthreads;
        int g, j;
        if (ROWS) { g = b / LR; j = b % LR; }
        else { g = b % W; j = b / W; }
        C2<T> *v = vals + i * R;
        const int k = j % Ns;
        const int dbase = (j / Ns) * Ns * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int n = dbase + r * Ns;
            if (to_lds) {
                if (ROWS) lds[g * LP + lpad(n)] = v[r];
                else lds[n * W + g] = v[r];
            } else {
                C2<T> o = v[r];
                if (tw_mult) o = cmul(o, tw[(size_t)((n * tw_mult) * twN_stride)]);
                o.x *= scale;
                o.y *= scale;
                gout[out_base0 + (size_t)g * out_step + (size_t)n * ostride] = o;
            }
        }
    }
    if (to_lds) __syncthreads();
}

template <typename T, bool ROWS>
__device__ void run_passes(const Passes &P, int tid, int nthreads, int L, int G, int ts,
                           const C2<T> *tw, C2<T> *lds, int LP, int W, const C2<T> *gin,
                           C2<T> *gout, size_t line_base0, size_t line_step, size_t estride,
                           size_t out_base0, size_t out_step, size_t ostride, int tw_mult,
                           int twN_stride, T scale)
{
    C2<T> vals[kE];
    int Ns = 1;
    for (int p = 0; p < P.np; ++p) {
        const bool from_lds = p > 0, to_lds = p + 1 < P.np;
        const int R = P.r[p];
#define RICK_PASS(RR)                                                                        \
    do_pass<T, RR, ROWS>(vals, tid, nthreads, L, G, Ns, ts, tw, from_lds, to_lds, lds, LP, W, \
                         gin, gout, line_base0, line_step, estride, out_base0, out_step,    \
                         ostride, tw_mult, twN_stride, scale)
        if (R == 16) RICK_PASS(16);
        else if (R == 8) RICK_PASS(8);
        else if (R == 4) RICK_PASS(4);
        else RICK_PASS(2);
#undef RICK_PASS
        Ns *= R;
    }
}

// Rows: grid of (planes * ny / G) workgroups of L*G/16 threads.
template <typename T>
__global__ void __launch_bounds__(512)
k_rows(Passes P, int L, int G, int rows_total, int ts, const C2<T> *__restrict__ tw,
       const C2<T> *gin, C2<T> *gout, T scale)
{
    extern __shared__ unsigned char smem_raw[];
    C2<T> *lds = reinterpret_cast<C2<T> *>(smem_raw);
    const int LP = lpad(L) + 1;
    const size_t row0 = (size_t)blockIdx.x * G;
    const size_t base = row0 * (size_t)L;
    (void)rows_total;
    run_passes<T, true>(P, threadIdx.x, blockDim.x, L, G, ts, tw, lds, LP, 0, gin, gout, base,
                        (size_t)L, 1, base, (size_t)L, 1, 0, 0, scale);
}

// Columns: workgroup (x: strip of W columns, y: row-set index s, z: plane).
// Line element n of column c sits at row  rb + n*rs  with rb = s*sb_in + s2,
// see the file comment; output row  ob + n*os.
struct ColArgs {
    int L, W, nx, ny;
    int in_set_mul, in_stride;     // input row  = s*in_set_mul + n*in_stride
    int out_set_mul, out_stride;   // output row = s*out_set_mul + n*out_stride
    int tw_by_set;                 // twiddle W_N^(n*s) after the DFT (four-step A)
    int ts;                        // table stride for the L-point DFT
    int twN_stride;                // table stride for W_N (N = ny)
};

template <typename T>
__global__ void __launch_bounds__(512)
k_cols(Passes P, ColArgs a, const C2<T> *__restrict__ tw, const C2<T> *gin, C2<T> *gout, T scale)
{
    extern __shared__ unsigned char smem_raw[];
    C2<T> *lds = reinterpret_cast<C2<T> *>(smem_raw);
    const int c0 = blockIdx.x * a.W;
    const int s = blockIdx.y;
    const size_t plane = (size_t)blockIdx.z * (size_t)a.nx * (size_t)a.ny;
    const size_t in_base = plane + (size_t)s * a.in_set_mul * (size_t)a.nx + c0;
    const size_t out_base = plane + (size_t)s * a.out_set_mul * (size_t)a.nx + c0;
    run_passes<T, false>(P, threadIdx.x, blockDim.x, a.L, a.W, a.ts, tw, lds, 0, a.W, gin, gout,
                         in_base, 1, (size_t)a.in_stride * a.nx, out_base, 1,
                         (size_t)a.out_stride * a.nx, a.tw_by_set ? s : 0, a.twN_stride, scale);
}

// Kernels using more than 64 KB of dynamic LDS must say so before launch.
template <typename K>
hipError_t allow_lds(K kernel, size_t bytes)
{
    if (bytes <= 65536) return hipSuccess;
    return hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

template <typename T>
hipError_t fft_impl(const FftPlan &p, const void *in_v, void *out_v, void *tmp_v,
                    double scale_d, hipStream_t s, Timing *tm)
{
    using CT = C2<T>;
    const CT *in = static_cast<const CT *>(in_v);
    CT *out = static_cast<CT *>(out_v);
    const CT *twx = static_cast<const CT *>(p.tw_x);
    const CT *twy = static_cast<const CT *>(p.tw_y);
    const int nx = p.nx, ny = p.ny, planes = p.planes;
    const T one = T(1), scale = (T)scale_d;
    // Elements per workgroup: 4096, or the line length if longer.
    const int chunk = 4096;

    // ---- columns (v axis, length ny) ----
    const CT *src = in;
    if (ny > 1) {
        if (!fft_needs_tmp(sizeof(T) == 8 ? RICK_FP64 : RICK_FP32, ny)) {
            ColArgs a{};
            a.L = ny;
            a.W = min(nx, max(1, chunk / ny));
            a.nx = nx;
            a.ny = ny;
            a.in_set_mul = 0;
            a.in_stride = 1;
            a.out_set_mul = 0;
            a.out_stride = 1;
            a.tw_by_set = 0;
            a.ts = 1;
            a.twN_stride = 1;
            if (nx % a.W) return hipErrorInvalidValue;
            const int threads = a.L * a.W / kE;
            const size_t lds = sizeof(CT) * (size_t)a.L * a.W;
            if (hipError_t e = allow_lds(k_cols<T>, lds); e != hipSuccess) return e;
            tm->begin(RICK_STAGE_FFT_COLS, s);
            hipLaunchKernelGGL(k_cols<T>, dim3(nx / a.W, 1, planes), dim3(threads), lds, s,
                               plan_passes(a.L), a, twy, src, out, one);
            tm->end(RICK_STAGE_FFT_COLS, s);
        } else {
            // Four-step: ny = N1 * N2 with N1 = 64.
            const int N1 = 64, N2 = ny / 64;
            ColArgs A{};
            A.L = N1;
            A.W = min(nx, chunk / N1);
            A.nx = nx;
            A.ny = ny;
            A.in_set_mul = 1;        // row = n2 + n1*N2
            A.in_stride = N2;
            A.out_set_mul = N1;      // row = n2*N1 + k1
            A.out_stride = 1;
            A.tw_by_set = 1;         // W_N^(k1*n2)
            A.ts = N2;
            A.twN_stride = 1;
            ColArgs B{};
            B.L = N2;
            B.W = min(nx, max(1, chunk / N2));
            B.nx = nx;
            B.ny = ny;
            B.in_set_mul = 1;        // row = k1 + n2*N1
            B.in_stride = N1;
            B.out_set_mul = 1;       // row = k1 + k2*N1
            B.out_stride = N1;
            B.tw_by_set = 0;
            B.ts = N1;
            B.twN_stride = 1;
            if (nx % A.W || nx % B.W) return hipErrorInvalidValue;
            const size_t ldsA = sizeof(CT) * (size_t)A.L * A.W;
            const size_t ldsB = sizeof(CT) * (size_t)B.L * B.W;
            const Passes PA = plan_passes(A.L), PB = plan_passes(B.L);
            if (hipError_t e = allow_lds(k_cols<T>, std::max(ldsA, ldsB)); e != hipSuccess)
                return e;
            if (in_v != out_v) {
                // A: in -> out (rows permuted), B: in place on out.
                tm->begin(RICK_STAGE_FFT_COLS, s);
                hipLaunchKernelGGL(k_cols<T>, dim3(nx / A.W, N2, planes), dim3(A.L * A.W / kE),
                                   ldsA, s, PA, A, twy, src, out, one);
                tm->end(RICK_STAGE_FFT_COLS, s);
                tm->begin(RICK_STAGE_FFT_COLS_B, s);
                hipLaunchKernelGGL(k_cols<T>, dim3(nx / B.W, N1, planes), dim3(B.L * B.W / kE),
                                   ldsB, s, PB, B, twy, out, out, one);
                tm->end(RICK_STAGE_FFT_COLS_B, s);
            } else {
                // In place: per plane, A: plane -> tmp, B: tmp -> plane.
                if (tmp_v == nullptr) return hipErrorInvalidValue;
                CT *tmp = static_cast<CT *>(tmp_v);
                const size_t pl = (size_t)nx * ny;
                for (int iw = 0; iw < planes; ++iw) {
                    tm->begin(RICK_STAGE_FFT_COLS, s);
                    hipLaunchKernelGGL(k_cols<T>, dim3(nx / A.W, N2, 1), dim3(A.L * A.W / kE),
                                       ldsA, s, PA, A, twy, out + iw * pl, tmp, one);
                    tm->end(RICK_STAGE_FFT_COLS, s);
                    tm->begin(RICK_STAGE_FFT_COLS_B, s);
                    hipLaunchKernelGGL(k_cols<T>, dim3(nx / B.W, N1, 1), dim3(B.L * B.W / kE),
                                       ldsB, s, PB, B, twy, tmp, out + iw * pl, one);
                    tm->end(RICK_STAGE_FFT_COLS_B, s);
                }
            }
        }
        src = out;
    }
    // ---- rows (u axis, length nx), in place on out, scaled ----
    if (nx > 1) {
        const int G = min(ny, max(1, chunk / nx));
        const int rows_total = planes * ny;
        if (rows_total % G) return hipErrorInvalidValue;
        const int threads = nx * G / kE;
        const int LP = lpad(nx) + 1;
        const size_t lds = sizeof(CT) * (size_t)LP * G;
        if (hipError_t e = allow_lds(k_rows<T>, lds); e != hipSuccess) return e;
        tm->begin(RICK_STAGE_FFT_ROWS, s);
        hipLaunchKernelGGL(k_rows<T>, dim3(rows_total / G), dim3(threads), lds, s,
                           plan_passes(nx), nx, G, rows_total, 1, twx, src, out, scale);
        tm->end(RICK_STAGE_FFT_ROWS, s);
    } else if (src != out || scale != one) {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace

bool fft_needs_tmp(int precision, int ny)
{
    (void)precision;
    return ny > 512;  // keeps >= 8 columns (>= 64 B) per row access in every pass
}

hipError_t launch_fft(const FftPlan &p, int precision, const void *in, void *out, void *tmp,
                      double scale, hipStream_t s, Timing *tm)
{
    Timing off;
    if (tm == nullptr) tm = &off;
    if (precision == RICK_FP64) return fft_impl<double>(p, in, out, tmp, scale, s, tm);
    return fft_impl<float>(p, in, out, tmp, scale, s, tm);
}

}  // namespace rick
