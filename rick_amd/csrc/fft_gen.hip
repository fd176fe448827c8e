// Backward C2C FFT of w-planes for axis lengths the power-of-two engine
// (fft.hip) does not take: FFTW plans any Ny x Nx (fourier_transform.c:
// 34-36), so the fftw_data stage must too.
//
// Each axis of the 2-D transform is planned on its own:
//   * direct     -- one pass over lines of length L <= 4096 whose prime
//                   factors are <= 13: a mixed-radix Stockham transform held
//                   in LDS (radices 16, 8, 4, 2, 3, 5, 7, 11, 13), one
//                   workgroup per G = 4096 / L lines;
//   * four-step  -- N = N1 * N2 (both direct lengths, N1 closest to sqrt(N)):
//                   pass A does the N1-point DFTs with the twiddle W_N^(k1 n2)
//                   into a one-plane scratch, pass B the N2-point DFTs back,
//                   in natural order.  Used for long rows (N > 4096) and for
//                   columns longer than 256 (G >= 16 adjacent columns per
//                   workgroup keeps the strided reads in whole 256-byte runs);
//   * Bluestein  -- lengths with a prime factor above 13: X[k] = b[k] *
//                   sum_n (x[n] b[n]) conj(b[k-n]), b[n] = exp(+i pi n^2/N),
//                   as a cyclic convolution of power-of-two length
//                   M >= 2N-1 done with two M-point transforms of the
//                   engine above and the precomputed transform of the chirp.
// Twiddles are tables exp(+2 pi i m/N) of the axis length (computed on the
// host in long double); a pass of length L | N uses them with stride N/L.
#include <algorithm>
#include <cmath>
#include <vector>

#include "fft_kernels.h"
#include "rick_internal.h"
#include "rick_tables.h"

namespace rick {
namespace {

using namespace fftk;

constexpr int kGenThreads = 256;
constexpr int kGenElems = 4096;       // complex values per workgroup (G * L)
constexpr int kGenDirectCols = 256;   // longest column transform done in one pass
constexpr size_t kBlueChunk = size_t(1) << 25;   // Bluestein scratch: elements per chunk

// Line addressing of one launch.  Workgroup (group, set, plane) transforms
// lines g0 .. g0+G-1 of that set; element n of line g is read at
//   z*pz + s*ps_in  + g*pg_in  + n*pe_in
// and written at
//   z*pz + s*ps_out + g*pg_out + n*pe_out,
// after the optional four-step twiddle W_N^(n*s) (tw_by 1) or W_N^(n*g)
// (tw_by 2), N the table length, and the scale.
struct GenAddr {
    int nl, G, ngroups, sets;
    long long pz, ps_in, pg_in, pe_in, ps_out, pg_out, pe_out;
    int tw_by;
    int line_major;   // butterflies along a line are consecutive threads (contiguous lines)
};

// Butterfly b of a pass with Lr butterflies per line.
__device__ __forceinline__ void gen_map(int b, int Lr, int G, int line_major, int &g, int &j)
{
    if (line_major) { g = b / Lr; j = b % Lr; }
    else { g = b % G; j = b / G; }
}

__device__ __forceinline__ int gen_lds(int g, int n, int L, int G, int line_major)
{
    return line_major ? g * L + n : n * G + g;
}

template <typename T>
struct GenCtx {
    const GenPassDesc *gp;
    const GenAddr *a;
    const C2<T> *tw;
    long long ntab;
    const C2<T> *in;
    C2<T> *out;
    C2<T> *lds;
    size_t ibase, obase;
    int s, g0, nvalid;
    T scale;
};

// One radix-R pass (Stockham: Ns = product of the earlier radices).  The
// first pass reads global memory, later ones LDS; the last writes global
// memory.  Reads of a pass all complete (barrier) before its writes, so the
// LDS buffer is used in place.
template <typename T, int R>
__device__ __forceinline__ void gen_pass(const GenCtx<T> &c, int Ns, bool first, bool last)
{
    constexpr int BPT = (kGenElems / R + kGenThreads - 1) / kGenThreads;
    const GenAddr &a = *c.a;
    const int L = c.gp->L, G = a.G, Lr = L / R, nb = G * Lr;
    const int tid = threadIdx.x;
    C2<T> v[BPT][R];
    int gj[BPT][2];
#pragma unroll
    for (int i = 0; i < BPT; ++i) {
        const int b = tid + i * kGenThreads;
        gj[i][0] = -1;
        if (b < nb) {
            int g, j;
            gen_map(b, Lr, G, a.line_major, g, j);
            gj[i][0] = g;
            gj[i][1] = j;
#pragma unroll
            for (int q = 0; q < R; ++q) {
                const int n = j + q * Lr;
                if (first)
                    v[i][q] = g < c.nvalid ? c.in[c.ibase + (size_t)g * a.pg_in + (size_t)n * a.pe_in]
                                           : C2<T>{T(0), T(0)};
                else
                    v[i][q] = c.lds[gen_lds(g, n, L, G, a.line_major)];
            }
        }
    }
    if (!first) __syncthreads();
    C2<T> w[R];
    if constexpr (R == 7 || R == 11 || R == 13) {
        w[0] = C2<T>{T(1), T(0)};
#pragma unroll
        for (int m = 1; m < R; ++m) w[m] = c.tw[(size_t)m * (size_t)(c.ntab / R)];
    }
    const long long tstep = c.ntab / ((long long)Ns * R);
#pragma unroll
    for (int i = 0; i < BPT; ++i) {
        const int g = gj[i][0];
        if (g < 0) continue;
        const int j = gj[i][1], k = j % Ns;
        if (Ns > 1) {
#pragma unroll
            for (int q = 1; q < R; ++q)
                v[i][q] = cmul(v[i][q], c.tw[(size_t)((long long)k * q * tstep)]);
        }
        if constexpr (R == 2 || R == 4 || R == 8 || R == 16) dftR<T, R>(v[i]);
        else if constexpr (R == 3) dft3(v[i]);
        else if constexpr (R == 5) dft5(v[i]);
        else dft_direct<T, R>(v[i], w);
        const int dbase = (j / Ns) * Ns * R + k;
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const int n = dbase + q * Ns;
            if (last) {
                if (g >= c.nvalid) continue;   // (its twiddle index would pass the table)
                C2<T> o = v[i][q];
                if (a.tw_by) {
                    const long long e = a.tw_by == 1 ? c.s : c.g0 + g;
                    o = cmul(o, c.tw[(size_t)((long long)n * e)]);
                }
                o.x *= c.scale;
                o.y *= c.scale;
                c.out[c.obase + (size_t)g * a.pg_out + (size_t)n * a.pe_out] = o;
            } else {
                c.lds[gen_lds(g, n, L, G, a.line_major)] = v[i][q];
            }
        }
    }
    if (!last) __syncthreads();
}

template <typename T>
__global__ void __launch_bounds__(kGenThreads)
k_fft_gen(GenPassDesc gp, GenAddr a, const C2<T> *__restrict__ tw, long long ntab, const C2<T> *in,
          C2<T> *out, T scale)
{
    __shared__ C2<T> lds[kGenElems];
    const unsigned bid = blockIdx.x;
    const int grp = (int)(bid % (unsigned)a.ngroups);
    const unsigned rest = bid / (unsigned)a.ngroups;
    const int s = (int)(rest % (unsigned)a.sets);
    const size_t z = rest / (unsigned)a.sets;
    GenCtx<T> c;
    c.gp = &gp;
    c.a = &a;
    c.tw = tw;
    c.ntab = ntab;
    c.in = in;
    c.out = out;
    c.lds = lds;
    c.s = s;
    c.g0 = grp * a.G;
    c.nvalid = min(a.G, a.nl - c.g0);
    c.ibase = z * (size_t)a.pz + (size_t)s * (size_t)a.ps_in + (size_t)c.g0 * (size_t)a.pg_in;
    c.obase = z * (size_t)a.pz + (size_t)s * (size_t)a.ps_out + (size_t)c.g0 * (size_t)a.pg_out;
    c.scale = scale;
    int Ns = 1;
    for (int p = 0; p < gp.np; ++p) {
        const bool first = p == 0, last = p == gp.np - 1;
        switch (gp.rad[p]) {
            case 16: gen_pass<T, 16>(c, Ns, first, last); break;
            case 8: gen_pass<T, 8>(c, Ns, first, last); break;
            case 4: gen_pass<T, 4>(c, Ns, first, last); break;
            case 2: gen_pass<T, 2>(c, Ns, first, last); break;
            case 3: gen_pass<T, 3>(c, Ns, first, last); break;
            case 5: gen_pass<T, 5>(c, Ns, first, last); break;
            case 7: gen_pass<T, 7>(c, Ns, first, last); break;
            case 11: gen_pass<T, 11>(c, Ns, first, last); break;
            default: gen_pass<T, 13>(c, Ns, first, last); break;
        }
        Ns *= gp.rad[p];
    }
}

// Bluestein pre-multiply: w[c*M + m] = x[line l0+c][m] * b[m] for m < N, else 0.
template <typename T>
__global__ void __launch_bounds__(256)
k_blue_pre(const C2<T> *__restrict__ x, size_t base, long long ls, long long es, int l0, int cl,
           int N, int M, const C2<T> *__restrict__ chirp, C2<T> *__restrict__ w)
{
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (size_t)cl * M) return;
    int cc, m;
    if (es == 1) { cc = (int)(idx / M); m = (int)(idx % M); }
    else { m = (int)(idx / cl); cc = (int)(idx % cl); }
    C2<T> v{T(0), T(0)};
    if (m < N) v = cmul(x[base + (size_t)(l0 + cc) * ls + (size_t)m * es], chirp[m]);
    w[(size_t)cc * M + m] = v;
}

// w *= H (the M-point transform of the wrapped conjugate chirp), per line.
template <typename T>
__global__ void __launch_bounds__(256)
k_blue_mul(C2<T> *__restrict__ w, const C2<T> *__restrict__ h, int cl, int M)
{
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (size_t)cl * M) return;
    w[idx] = cmul(w[idx], h[idx % M]);
}

// Bluestein post-multiply: X[k] = b[k] * scale/M * F(w)[(M-k) mod M], k < N.
template <typename T>
__global__ void __launch_bounds__(256)
k_blue_post(const C2<T> *__restrict__ w, int cl, int N, int M, const C2<T> *__restrict__ chirp,
            T scale, C2<T> *__restrict__ y, size_t base, long long ls, long long es, int l0)
{
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (size_t)cl * N) return;
    int cc, k;
    if (es == 1) { cc = (int)(idx / N); k = (int)(idx % N); }
    else { k = (int)(idx / cl); cc = (int)(idx % cl); }
    C2<T> v = cmul(w[(size_t)cc * M + (size_t)((M - k) % M)], chirp[k]);
    v.x *= scale;
    v.y *= scale;
    y[base + (size_t)(l0 + cc) * ls + (size_t)k * es] = v;
}

// ---------------------------------------------------------------------------
// Host side.

// Radices of a direct length (largest first); false if L has a prime factor
// above 13 or needs more passes than the descriptor holds.
bool gen_radices(int L, GenPassDesc &gp)
{
    gp.L = L;
    gp.np = 0;
    if (L < 1 || L > kGenElems) return false;
    int r = L;
    auto push = [&](int x) {
        if (gp.np >= kGenMaxPasses) return false;
        gp.rad[gp.np++] = x;
        return true;
    };
    while (r % 16 == 0)
        if (!push(16)) return false; else r /= 16;
    for (int p2 : {8, 4, 2})
        if (r % p2 == 0) {
            if (!push(p2)) return false;
            r /= p2;
            break;
        }
    for (int p : {3, 5, 7, 11, 13})
        while (r % p == 0) {
            if (!push(p)) return false;
            r /= p;
        }
    return r == 1 && gp.np > 0;
}

bool smooth(long long n)
{
    for (int p : {2, 3, 5, 7, 11, 13})
        while (n % p == 0) n /= p;
    return n == 1;
}

// Split n = n1 * n2 with both direct lengths, n1 closest to sqrt(n); false if none.
bool gen_split(int n, int &n1, int &n2)
{
    int best = 0;
    double bestd = 1e300;
    for (int d = 2; d <= kGenElems && d <= n / 2; ++d) {
        if (n % d) continue;
        const int e = n / d;
        if (e < 2 || e > kGenElems || !smooth(d) || !smooth(e)) continue;
        const double dist = std::fabs(std::log((double)d) - 0.5 * std::log((double)n));
        if (dist < bestd) {
            bestd = dist;
            best = d;
        }
    }
    if (!best) return false;
    n1 = best;
    n2 = n / best;
    return true;
}

// Direct (0) or four-step (1) plan of a smooth length on a given axis kind.
bool plan_smooth(int n, bool rows, int &mode, int &n1, int &n2, GenPassDesc &a, GenPassDesc &b)
{
    const int direct_max = rows ? kGenElems : kGenDirectCols;
    if (n <= direct_max && gen_radices(n, a)) {
        mode = 0;
        n1 = n;
        n2 = 1;
        return true;
    }
    if (!gen_split(n, n1, n2)) {
        // A short smooth column with no split (a prime <= 13 times little):
        // one pass even if narrower than the coalescing target.
        if (n <= kGenElems && gen_radices(n, a)) {
            mode = 0;
            n1 = n;
            n2 = 1;
            return true;
        }
        return false;
    }
    mode = 1;
    return gen_radices(n1, a) && gen_radices(n2, b);
}

// exp(+2 pi i m/n) and the Bluestein tables, in long double.
void host_twiddles(long long n, std::vector<long double> &re_im)
{
    re_im.resize(2 * (size_t)n);
    const long double pi = 3.141592653589793238462643383279502884L;
    for (long long m = 0; m < n; ++m) {
        const long double a = 2.0L * pi * (long double)m / (long double)n;
        re_im[2 * m] = cosl(a);
        re_im[2 * m + 1] = sinl(a);
    }
}

// In-place sign +1 DFT of a power-of-two length (host, long double; radix 2).
void host_fft_pow2(std::vector<long double> &x, long long n)
{
    for (long long i = 1, j = 0; i < n; ++i) {
        long long bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            std::swap(x[2 * i], x[2 * j]);
            std::swap(x[2 * i + 1], x[2 * j + 1]);
        }
    }
    std::vector<long double> tw;
    host_twiddles(n, tw);
    for (long long len = 2; len <= n; len <<= 1) {
        const long long half = len >> 1, step = n / len;
        for (long long s = 0; s < n; s += len)
            for (long long k = 0; k < half; ++k) {
                const long double wr = tw[2 * (k * step)], wi = tw[2 * (k * step) + 1];
                long double *pa = &x[2 * (s + k)], *pb = &x[2 * (s + k + half)];
                const long double br = pb[0] * wr - pb[1] * wi, bi = pb[0] * wi + pb[1] * wr;
                pb[0] = pa[0] - br;
                pb[1] = pa[1] - bi;
                pa[0] += br;
                pa[1] += bi;
            }
    }
}

// Lines of one axis per plane: rows (axis 0) are ny lines of nx contiguous
// elements; columns (axis 1) are nx lines of ny elements at stride nx.
struct AxisGeom {
    int n, nl;
    long long es, ls;
};

AxisGeom axis_geom(int axis, int nx, int ny)
{
    if (axis == 0) return AxisGeom{nx, ny, 1, nx};
    return AxisGeom{ny, nx, nx, 1};
}

size_t csize(int precision) { return precision == RICK_FP64 ? 16 : 8; }

template <typename T>
hipError_t launch_gen(const GenPassDesc &gp, GenAddr a, long long planes, const void *tw,
                      long long ntab, const void *in, void *out, double scale, hipStream_t s)
{
    a.ngroups = (a.nl + a.G - 1) / a.G;
    const unsigned long long blocks = (unsigned long long)a.ngroups * a.sets * planes;
    if (blocks == 0) return hipSuccess;
    if (blocks > 0x7fffffffull) return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL(k_fft_gen<T>, dim3((unsigned)blocks), dim3(kGenThreads), 0, s, gp, a,
                       static_cast<const C2<T> *>(tw), ntab, static_cast<const C2<T> *>(in),
                       static_cast<C2<T> *>(out), (T)scale);
    return hipGetLastError();
}

// One smooth axis transform of `planes` planes (direct: all planes in one
// launch; four-step: per plane through the scratch plane `tmp`).
template <typename T>
hipError_t smooth_axis(const GenAxis &ax, const AxisGeom &ag, bool rows, int planes,
                       long long plane_elems, const void *tw, const C2<T> *in, C2<T> *out,
                       C2<T> *tmp, double scale, hipStream_t s)
{
    const int n = ag.n;
    if (ax.mode == 0) {
        GenAddr a{};
        a.nl = ag.nl;
        a.G = std::max(1, std::min(ag.nl, kGenElems / n));
        a.sets = 1;
        a.pz = plane_elems;
        a.pg_in = a.pg_out = ag.ls;
        a.pe_in = a.pe_out = ag.es;
        a.line_major = rows ? 1 : 0;
        return launch_gen<T>(ax.a, a, planes, tw, n, in, out, scale, s);
    }
    const int N1 = ax.n1, N2 = ax.n2;
    for (int z = 0; z < planes; ++z) {
        const C2<T> *src = in + (size_t)z * plane_elems;
        C2<T> *dst = out + (size_t)z * plane_elems;
        // Pass A: N1-point DFTs over elements n2 + N2*n1 -> tmp at n2 + N2*k1,
        // times W_N^(k1*n2).
        GenAddr A{};
        if (rows) {
            A.nl = N2; A.sets = ag.nl;
            A.pg_in = A.pg_out = 1;
            A.ps_in = A.ps_out = ag.ls;
            A.pe_in = A.pe_out = (long long)N2;
            A.tw_by = 2;
        } else {
            A.nl = ag.nl; A.sets = N2;
            A.pg_in = A.pg_out = ag.ls;
            A.ps_in = A.ps_out = ag.es;
            A.pe_in = A.pe_out = (long long)N2 * ag.es;
            A.tw_by = 1;
        }
        A.G = std::max(1, std::min(A.nl, kGenElems / N1));
        A.line_major = 0;
        hipError_t e = launch_gen<T>(ax.a, A, 1, tw, n, src, tmp, 1.0, s);
        if (e != hipSuccess) return e;
        // Pass B: N2-point DFTs over tmp elements N2*k1 + n2 -> X[k1 + N1*k2].
        GenAddr B{};
        if (rows) {
            B.nl = N1; B.sets = ag.nl;
            B.pg_in = N2; B.pe_in = 1; B.ps_in = ag.ls;
            B.pg_out = 1; B.pe_out = N1; B.ps_out = ag.ls;
            B.line_major = 1;
        } else {
            B.nl = ag.nl; B.sets = N1;
            B.pg_in = B.pg_out = ag.ls;
            B.ps_in = (long long)N2 * ag.es; B.pe_in = ag.es;
            B.ps_out = ag.es; B.pe_out = (long long)N1 * ag.es;
            B.line_major = 0;
        }
        B.G = std::max(1, std::min(B.nl, kGenElems / N2));
        e = launch_gen<T>(ax.b, B, 1, tw, n, tmp, dst, scale, s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

template <typename T>
hipError_t blue_axis(const GenAxis &ax, const AxisGeom &ag, int planes, long long plane_elems,
                     const C2<T> *in, C2<T> *out, void *work, double scale, hipStream_t s)
{
    const int N = ag.n, M = ax.m;
    const int chunk = (int)std::max<size_t>(1, std::min<size_t>((size_t)ag.nl, kBlueChunk / M));
    C2<T> *w1 = static_cast<C2<T> *>(work);
    C2<T> *w2 = w1 + (size_t)chunk * M;
    // The conv FFT's own plan: rows of length M, direct or four-step.
    GenAxis conv{};
    conv.mode = ax.mmode;
    conv.n1 = ax.m1;
    conv.n2 = ax.m2;
    conv.a = ax.ma;
    conv.b = ax.mb;
    for (int z = 0; z < planes; ++z) {
        const size_t pbase = (size_t)z * plane_elems;
        for (int l0 = 0; l0 < ag.nl; l0 += chunk) {
            const int cl = std::min(chunk, ag.nl - l0);
            const size_t nw = (size_t)cl * M;
            hipLaunchKernelGGL(k_blue_pre<T>, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, s,
                               in, pbase, ag.ls, ag.es, l0, cl, N, M,
                               static_cast<const C2<T> *>(ax.chirp), w1);
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
            const AxisGeom cg{M, cl, 1, M};
            e = smooth_axis<T>(conv, cg, true, 1, 0, ax.tw, w1, w1, w2, 1.0, s);
            if (e != hipSuccess) return e;
            hipLaunchKernelGGL(k_blue_mul<T>, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, s,
                               w1, static_cast<const C2<T> *>(ax.hf), cl, M);
            e = hipGetLastError();
            if (e != hipSuccess) return e;
            e = smooth_axis<T>(conv, cg, true, 1, 0, ax.tw, w1, w1, w2, 1.0, s);
            if (e != hipSuccess) return e;
            const size_t ny_ = (size_t)cl * N;
            hipLaunchKernelGGL(k_blue_post<T>, dim3((unsigned)((ny_ + 255) / 256)), dim3(256), 0,
                               s, w1, cl, N, M, static_cast<const C2<T> *>(ax.chirp),
                               (T)(scale / (double)M), out, pbase, ag.ls, ag.es, l0);
            e = hipGetLastError();
            if (e != hipSuccess) return e;
        }
    }
    return hipSuccess;
}

template <typename T>
hipError_t gen_impl(const GenFft &g, int planes, int nx, int ny, const void *in_v, void *out_v,
                    double scale, hipStream_t s, Timing *tm)
{
    const C2<T> *in = static_cast<const C2<T> *>(in_v);
    C2<T> *out = static_cast<C2<T> *>(out_v);
    const long long plane = (long long)nx * ny;
    // Columns first (reading `in`), then rows in place on `out`, scaled.
    for (int axis = 1; axis >= 0; --axis) {
        const GenAxis &ax = g.ax[axis];
        const AxisGeom ag = axis_geom(axis, nx, ny);
        const C2<T> *src = axis == 1 ? in : out;
        const double sc = axis == 0 ? scale : 1.0;
        const int st = axis == 0 ? RICK_STAGE_FFT_ROWS : RICK_STAGE_FFT_COLS;
        tm->begin(st, s);
        hipError_t e;
        if (ax.mode == 2)
            e = blue_axis<T>(ax, ag, planes, plane, src, out, g.work, sc, s);
        else
            e = smooth_axis<T>(ax, ag, axis == 0, planes, plane, ax.tw, src, out,
                               static_cast<C2<T> *>(g.work), sc, s);
        tm->end(st, s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace

bool gen_fft_prepare(int nx, int ny, int precision, GenFft *g)
{
    *g = GenFft{};
    const size_t cs = csize(precision);
    size_t tables = 0, work = 0;
    for (int axis = 0; axis < 2; ++axis) {
        GenAxis &ax = g->ax[axis];
        const AxisGeom ag = axis_geom(axis, nx, ny);
        ax.n = ag.n;
        if (ag.n < 2) return false;
        if (smooth(ag.n) && plan_smooth(ag.n, axis == 0, ax.mode, ax.n1, ax.n2, ax.a, ax.b)) {
            tables += (cs * (size_t)ag.n + 255) / 256 * 256;
            if (ax.mode == 1) work = std::max(work, cs * (size_t)nx * ny);
            continue;
        }
        // Bluestein.
        long long M = 1;
        while (M < 2LL * ag.n - 1) M <<= 1;
        if (M > (1LL << 24)) return false;
        ax.mode = 2;
        ax.m = (int)M;
        if (!plan_smooth(ax.m, true, ax.mmode, ax.m1, ax.m2, ax.ma, ax.mb)) return false;
        tables += 2 * ((cs * (size_t)M + 255) / 256 * 256) + (cs * (size_t)ag.n + 255) / 256 * 256;
        const size_t chunk = std::max<size_t>(1, std::min<size_t>((size_t)ag.nl, kBlueChunk / M));
        work = std::max(work, cs * chunk * (size_t)M * (ax.mmode == 1 ? 2 : 1));
    }
    g->table_bytes = tables;
    g->work_bytes = work;
    return true;
}

hipError_t gen_fft_upload(GenFft *g, int precision, void *tables, void *work)
{
    char *p = static_cast<char *>(tables);
    const size_t cs = csize(precision);
    auto put = [&](const std::vector<long double> &v, size_t n) -> const void * {
        void *dst = p;
        p += (cs * n + 255) / 256 * 256;
        if (precision == RICK_FP64) {
            std::vector<double> d(v.begin(), v.begin() + 2 * n);
            return hipMemcpy(dst, d.data(), 16 * n, hipMemcpyHostToDevice) == hipSuccess ? dst
                                                                                         : nullptr;
        }
        std::vector<float> f(v.begin(), v.begin() + 2 * n);
        return hipMemcpy(dst, f.data(), 8 * n, hipMemcpyHostToDevice) == hipSuccess ? dst
                                                                                    : nullptr;
    };
    std::vector<long double> t;
    for (int axis = 0; axis < 2; ++axis) {
        GenAxis &ax = g->ax[axis];
        if (ax.mode != 2) {
            host_twiddles(ax.n, t);
            if (!(ax.tw = put(t, ax.n))) return hipErrorUnknown;
            continue;
        }
        const long long N = ax.n, M = ax.m;
        host_twiddles(M, t);
        if (!(ax.tw = put(t, M))) return hipErrorUnknown;
        // b[n] = exp(+i pi n^2 / N), with n^2 reduced mod 2N.
        const long double pi = 3.141592653589793238462643383279502884L;
        std::vector<long double> b(2 * N);
        for (long long k = 0; k < N; ++k) {
            const long long q = (long long)((unsigned long long)k * (unsigned long long)k %
                                            (unsigned long long)(2 * N));
            const long double a = pi * (long double)q / (long double)N;
            b[2 * k] = cosl(a);
            b[2 * k + 1] = sinl(a);
        }
        if (!(ax.chirp = put(b, N))) return hipErrorUnknown;
        // H = F_M(h), h[d mod M] = conj(b[|d|]) for |d| < N.
        std::vector<long double> h(2 * M, 0.0L);
        for (long long k = 0; k < N; ++k) {
            h[2 * k] = b[2 * k];
            h[2 * k + 1] = -b[2 * k + 1];
            if (k) {
                h[2 * (M - k)] = b[2 * k];
                h[2 * (M - k) + 1] = -b[2 * k + 1];
            }
        }
        host_fft_pow2(h, M);
        if (!(ax.hf = put(h, M))) return hipErrorUnknown;
    }
    g->work = work;
    return hipSuccess;
}

hipError_t launch_fft_gen(const GenFft &g, int precision, int planes, int nx, int ny,
                          const void *in, void *out, double scale, hipStream_t s, Timing *tm)
{
    Timing off;
    if (tm == nullptr) tm = &off;
    if (precision == RICK_FP64)
        return gen_impl<double>(g, planes, nx, ny, in, out, scale, s, tm);
    return gen_impl<float>(g, planes, nx, ny, in, out, scale, s, tm);
}

}  // namespace rick
