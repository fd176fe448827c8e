// Batched backward C2C FFT of w-planes for gfx950 (the fftw_data stage,
// fourier_transform.c:13-93: per plane a 2-D transform with sign +1, n0 = Ny,
// n1 = Nx, scaled by 1/(Nx*Ny)), operating directly on the interleaved
// plane-major grid layout.
//
// Structure:
//   * a Stockham autosort engine whose pass plan (radix 16 passes, then one
//     pass of radix 2/4/8) is fixed at compile time by the line length L.
//     Each thread holds 16 complex values in registers; a pass reads R values,
//     applies the inter-pass twiddles, does an in-register DFT and writes back
//     through LDS; the first pass reads global memory, the last writes it.
//   * k_fft_rows: lines are grid rows (contiguous u), 4096/L rows per
//     workgroup.
//   * k_fft_cols: lines are the columns of a G-column strip; element n of a
//     line is at row  s*in_set_mul + n*in_stride  and goes to row
//     s*out_set_mul + n*out_stride, optionally multiplied by W_N^(n*s).  With
//     L = Ny this is the whole column transform; long columns use two such
//     passes (N = N1*N2, four-step):
//       A: N1-point DFTs over rows n2 + N2*n1, twiddle W_N^(k1*n2),
//          written to rows n2*N1 + k1 (out of place);
//       B: N2-point DFTs over rows k1 + N1*n2, written back in place, which
//          leaves X[k1 + N1*k2] at row k1 + N1*k2 (natural order).
// Twiddle tables hold exp(+2*pi*i*m/N) for the axis length N; an L-point
// sub-transform uses them with stride N/L.  Per butterfly only W^k, W^2k,
// W^4k, W^8k are loaded; the other powers are products of those.
#include "fft_kernels.h"
#include "rick_internal.h"
#include "rick_math.h"

#include <algorithm>

namespace rick {
namespace {

using namespace fftk;

constexpr int kE = 16;  // complex values per thread

// The fused column/phase pass: complex values per thread for FP64 and FP32
// plane data, and whether FP64 prefetches the next plane into registers.
#ifndef RICK_PHASE_E_F64
#define RICK_PHASE_E_F64 8
#endif
#ifndef RICK_PHASE_E_F32
#define RICK_PHASE_E_F32 8
#endif
#ifndef RICK_PHASE_PREFETCH_F64
#define RICK_PHASE_PREFETCH_F64 0
#endif
// FP64: the w-phase rotation of a pixel advances plane to plane by a fixed
// step (exp(i 2 pi dw (n-1))), one complex multiply per plane instead of a
// sine and cosine.
#ifndef RICK_PHASE_RECUR_F64
#define RICK_PHASE_RECUR_F64 1
#endif
// Complex values per thread of the row and column passes.  Measured on the
// LOFAR-like chain: rows at 8 values took 1.45 / 0.96 ms (FP64 / FP32)
// against 1.35 / 0.90 at 16; FP32 column pass A 1.46 ms at 8 against 1.58 at
// 16 (1024-thread workgroups over 128 columns), FP64 2.69 against 2.64.
#ifndef RICK_ROWS_E_F64
#define RICK_ROWS_E_F64 16
#endif
#ifndef RICK_ROWS_E_F32
#define RICK_ROWS_E_F32 16
#endif
#ifndef RICK_COLS_E_F64
#define RICK_COLS_E_F64 16
#endif
#ifndef RICK_COLS_E_F32
#define RICK_COLS_E_F32 8
#endif
template <typename T>
constexpr int row_values() { return sizeof(T) == 8 ? RICK_ROWS_E_F64 : RICK_ROWS_E_F32; }
template <typename T>
constexpr int col_values() { return sizeof(T) == 8 ? RICK_COLS_E_F64 : RICK_COLS_E_F32; }
constexpr bool kPhasePrefetchF64 = RICK_PHASE_PREFETCH_F64 != 0;
constexpr bool kPhaseRecurF64 = RICK_PHASE_RECUR_F64 != 0;
template <typename T>
constexpr int phase_values(int L)
{
    // Lines shorter than the per-thread count keep 16 (one butterfly per line).
    return L < 16 ? kE : (sizeof(T) == 8 ? RICK_PHASE_E_F64 : RICK_PHASE_E_F32);
}

// Compile-time pass plan of an L-point transform with E values per thread:
// radix-E passes, then one pass of the remaining power of two.
constexpr int nE(int L, int E)
{
    int n = 0;
    while (L >= E && L % E == 0) { L /= E; ++n; }
    return n;
}
constexpr int powE(int n, int E) { return n == 0 ? 1 : E * powE(n - 1, E); }
constexpr int npasses(int L, int E = kE)
{
    return nE(L, E) + (L / powE(nE(L, E), E) > 1 ? 1 : 0);
}
constexpr int radix_of(int L, int p, int E = kE)
{
    return p < nE(L, E) ? E : L / powE(nE(L, E), E);
}
constexpr int ns_of(int L, int p, int E = kE)
{
    return p == 0 ? 1 : ns_of(L, p - 1, E) * radix_of(L, p - 1, E);
}

// LDS row padding of the row transform: RICK_LPAD_MUL extra slots per
// 2^RICK_LPAD_SHIFT elements.
#ifndef RICK_LPAD_SHIFT
#define RICK_LPAD_SHIFT 4
#endif
#ifndef RICK_LPAD_MUL
#define RICK_LPAD_MUL 1
#endif
__host__ __device__ constexpr int lpad(int n) { return n + RICK_LPAD_MUL * (n >> RICK_LPAD_SHIFT); }

// Twiddles W^(r*k*step), r = 1..R-1, from four loads.
template <typename T, int R>
__device__ __forceinline__ void twiddle(C2<T> *v, int k, int step, const C2<T> *__restrict__ tw)
{
    const C2<T> w1 = tw[k * step];
    v[1] = cmul(v[1], w1);
    if constexpr (R >= 4) {
        const C2<T> w2 = tw[2 * k * step];
        const C2<T> w3 = cmul(w1, w2);
        v[2] = cmul(v[2], w2);
        v[3] = cmul(v[3], w3);
        if constexpr (R >= 8) {
            const C2<T> w4 = tw[4 * k * step];
            v[4] = cmul(v[4], w4);
            v[5] = cmul(v[5], cmul(w1, w4));
            v[6] = cmul(v[6], cmul(w2, w4));
            const C2<T> w7 = cmul(w3, w4);
            v[7] = cmul(v[7], w7);
            if constexpr (R == 16) {
                const C2<T> w8 = tw[8 * k * step];
                v[8] = cmul(v[8], w8);
                v[9] = cmul(v[9], cmul(w1, w8));
                v[10] = cmul(v[10], cmul(w2, w8));
                v[11] = cmul(v[11], cmul(w3, w8));
                const C2<T> w12 = cmul(w4, w8);
                v[12] = cmul(v[12], w12);
                v[13] = cmul(v[13], cmul(w1, w12));
                v[14] = cmul(v[14], cmul(w2, w12));
                v[15] = cmul(v[15], cmul(w7, w8));
            }
        }
    }
}

// Addressing of one workgroup's lines.
struct Lines {
    size_t in_base, out_base;   // element offsets of line 0
    size_t in_line, out_line;   // offset between consecutive lines
    size_t in_el, out_el;       // offset between consecutive elements of a line
    int nvalid;                 // lines of this workgroup that exist
    int tw_set;                 // twiddle W_N^(n*tw_set) after the last pass (0: none)
    // Optional tile occupancy of the input (rows only): element n of line g
    // is read only if its tile t = n >> occ_shift is occupied, else it is
    // taken as zero (cells of empty gridder tiles are never written): bit t
    // of occ_mask when occ_mask_on (one line of at most 64 tiles), else
    // occ[g * occ_stride + t] != 0 when occ is set.
    const unsigned char *occ = nullptr;
    int occ_stride = 0, occ_shift = 0;
    bool occ_mask_on = false;
    unsigned long long occ_mask = 0;
    // Optional row liveness of the input (columns): element n is read only if
    // rowlive[n] is non-zero, else it is zero (grid rows without a non-zero
    // cell are skipped by the row transform).
    const unsigned char *rowlive = nullptr;
};

// Whether element n of line g is read from the input (else it is zero).
__device__ __forceinline__ bool elem_live(const Lines &ln, int g, int n)
{
    if (g >= ln.nvalid) return false;
    if (ln.occ_mask_on) {
        if (!((ln.occ_mask >> (n >> ln.occ_shift)) & 1ull)) return false;
    } else if (ln.occ != nullptr && !ln.occ[g * ln.occ_stride + (n >> ln.occ_shift)]) {
        return false;
    }
    return ln.rowlive == nullptr || ln.rowlive[n];
}

// Butterfly b of a pass with L/R butterflies per line: rows map consecutive
// b to consecutive j of one line, columns to consecutive lines.
template <bool ROWS, int LR, int G>
__device__ __forceinline__ void bmap(int b, int &g, int &j)
{
    if constexpr (ROWS) { g = b / LR; j = b % LR; }
    else { g = b % G; j = b / G; }
}

// Where the last pass delivers element n of line g (register slot idx):
// by default to global memory, after the optional four-step twiddle and the
// output scale.
template <typename T>
struct GlobalSink {
    C2<T> *gout;
    const Lines *ln;
    const C2<T> *tw;
    T scale;
    __device__ __forceinline__ void operator()(int g, int n, int idx, C2<T> o) const
    {
        (void)idx;
        if (ln->tw_set) o = cmul(o, tw[(size_t)n * ln->tw_set]);
        o.x *= scale;
        o.y *= scale;
        if (g < ln->nvalid)
            store_stream2(o, &gout[ln->out_base + (size_t)g * ln->out_line + (size_t)n * ln->out_el]);
    }
};

// First-pass input of this thread: element n of line g for butterfly slot
// i*R + r, or zero where the line does not exist or the input is known zero.
template <typename T, int L, int G, bool ROWS, int E = kE>
__device__ __forceinline__ void load_first(C2<T> *vals, int tid, const Lines &ln,
                                           const C2<T> *gin)
{
    constexpr int R = radix_of(L, 0, E), LR = L / R, BPT = E / R;
    constexpr int NT = G * L / E;
#pragma unroll
    for (int i = 0; i < BPT; ++i) {
        int g, j;
        bmap<ROWS, LR, G>(tid + i * NT, g, j);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int n = j + r * LR;
            const bool live = elem_live(ln, g, n);
            const size_t at = ln.in_base + (size_t)g * ln.in_line + (size_t)n * ln.in_el;
            vals[i * R + r] = live ? load_stream2(&gin[at]) : C2<T>{T(0), T(0)};
        }
    }
}

// LDS slot of element n of line g between passes.
template <bool ROWS, int L, int G>
__device__ __forceinline__ int lds_at(int g, int n)
{
    if constexpr (ROWS) return g * (lpad(L) + 1) + lpad(n);
    else return n * G + g;
}

// One pass of the transform and, recursively, the rest.  With PRE the pass's
// inputs are already in vals (preloaded by load_first, or placed by a split
// exchange), so it issues no loads.  With SPLIT the exchange between passes
// moves the real parts through LDS, then the imaginary parts, so the LDS
// buffer holds one T per element (half the bytes) at the cost of two more
// barriers per exchange.
template <typename T, int L, int G, bool ROWS, int P, typename Sink, bool PRE = false,
          bool SPLIT = false, int E = kE>
__device__ __forceinline__ void pass(C2<T> *vals, int tid, const Lines &ln, int ts,
                                     const C2<T> *__restrict__ tw, C2<T> *lds, const C2<T> *gin,
                                     Sink &sink)
{
    constexpr int NP = npasses(L, E);
    if constexpr (P < NP) {
        constexpr int R = radix_of(L, P, E), Ns = ns_of(L, P, E), LR = L / R, BPT = E / R;
        constexpr int NT = G * L / E;
        constexpr bool from_lds = P > 0, to_lds = P + 1 < NP;
        if constexpr (!PRE) {
#pragma unroll
            for (int i = 0; i < BPT; ++i) {
                int g, j;
                bmap<ROWS, LR, G>(tid + i * NT, g, j);
                C2<T> *v = vals + i * R;
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const int n = j + r * LR;
                    if constexpr (from_lds) {
                        v[r] = lds[lds_at<ROWS, L, G>(g, n)];
                    } else {
                        const bool live = elem_live(ln, g, n);
                        v[r] = live ? load_stream2(&gin[ln.in_base + (size_t)g * ln.in_line +
                                                        (size_t)n * ln.in_el])
                                    : C2<T>{T(0), T(0)};
                    }
                }
            }
            if constexpr (from_lds) __syncthreads();
        }
#pragma unroll
        for (int i = 0; i < BPT; ++i) {
            int g, j;
            bmap<ROWS, LR, G>(tid + i * NT, g, j);
            C2<T> *v = vals + i * R;
            if constexpr (Ns > 1) twiddle<T, R>(v, j % Ns, (L / (Ns * R)) * ts, tw);
            dftR<T, R>(v);
            if constexpr (!(to_lds && SPLIT)) {
                const int dbase = (j / Ns) * Ns * R + j % Ns;
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const int n = dbase + r * Ns;
                    if constexpr (to_lds) lds[lds_at<ROWS, L, G>(g, n)] = v[r];
                    else sink(g, n, i * R + r, v[r]);
                }
            }
        }
        if constexpr (to_lds && SPLIT) {
            // Outputs of this pass are in vals; move them to the next pass's
            // layout one component at a time.
            constexpr int R2 = radix_of(L, P + 1, E), LR2 = L / R2, BPT2 = E / R2;
            T *lt = reinterpret_cast<T *>(lds);
#pragma unroll
            for (int comp = 0; comp < 2; ++comp) {
#pragma unroll
                for (int i = 0; i < BPT; ++i) {
                    int g, j;
                    bmap<ROWS, LR, G>(tid + i * NT, g, j);
                    const int dbase = (j / Ns) * Ns * R + j % Ns;
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        const C2<T> o = vals[i * R + r];
                        lt[lds_at<ROWS, L, G>(g, dbase + r * Ns)] = comp == 0 ? o.x : o.y;
                    }
                }
                __syncthreads();
#pragma unroll
                for (int i = 0; i < BPT2; ++i) {
                    int g, j;
                    bmap<ROWS, LR2, G>(tid + i * NT, g, j);
#pragma unroll
                    for (int r = 0; r < R2; ++r) {
                        const T x = lt[lds_at<ROWS, L, G>(g, j + r * LR2)];
                        if (comp == 0) vals[i * R2 + r].x = x;
                        else vals[i * R2 + r].y = x;
                    }
                }
                __syncthreads();
            }
            pass<T, L, G, ROWS, P + 1, Sink, true, true, E>(vals, tid, ln, ts, tw, lds, gin,
                                                            sink);
        } else {
            if constexpr (to_lds) __syncthreads();
            pass<T, L, G, ROWS, P + 1, Sink, false, SPLIT, E>(vals, tid, ln, ts, tw, lds, gin,
                                                              sink);
        }
    }
}

// Lines per workgroup: 4096 elements per workgroup (256 threads), or one
// line of 8192 (512 threads).
constexpr int lines_per_wg(int L) { return L >= 4096 ? 1 : 4096 / L; }
// Columns per workgroup of a plain column pass: FP32 lines of 64 or 128
// elements take 8192 elements (512 threads), so each row segment a
// workgroup reads or writes is 1 KB, as in FP64.
template <typename T>
constexpr int col_lines(int L) { return sizeof(T) == 4 && L >= 64 && L <= 128 ? 8192 / L : lines_per_wg(L); }
constexpr int lds_elems(int L, bool rows)
{
    return rows ? lines_per_wg(L) * (lpad(L) + 1) : lines_per_wg(L) * L;
}
// Occupancy flags per workgroup: G rows x (L/32) tiles of >= 32 cells.
constexpr int oc_bytes(int L) { return lines_per_wg(L) * (L / 32 > 0 ? L / 32 : 1); }

// Occupancy of a sparse grid: counts of the gridder bucket (plane, tile row,
// tile column), zero where the tile was never written, and rows[plane * ny +
// v], non-zero iff grid row v of the plane holds a non-zero cell.  The row
// transform skips the other rows (writes nothing) and later column passes
// read them as zero.
struct Occupancy {
    const uint32_t *counts;
    const unsigned char *rows;
    int tv, tiles_u, tiles_v, ny;
    int tu_shift;   // log2 of the tile width (a power of two)
};

// Row flags of the grid for the column kernels (null: dense input).
struct ColOcc {
    const unsigned char *rows;
    int ny;
};

// Loads the flags of the L input rows of this workgroup (row0 + n * rstride
// of the plane) into LDS and points the input at them.
__device__ __forceinline__ void col_rowlive(const ColOcc &co, int plane, int row0, int rstride,
                                            int L, unsigned char *rl, Lines &ln)
{
    ln.rowlive = nullptr;
    if (co.rows == nullptr) return;
    const unsigned char *pr = co.rows + (size_t)plane * co.ny + row0;
    for (int n = threadIdx.x; n < L; n += blockDim.x) rl[n] = pr[(size_t)n * rstride];
    __syncthreads();
    ln.rowlive = rl;
}

// Rows: workgroup w handles rows w*G .. w*G+G-1 of the planes*ny rows.
// MASK (one row per workgroup, bucket counts of at most 64 tiles per row):
// the row's tile occupancy is a bit mask in scalar registers, so the test of
// each loaded element compiles to a shift and a select.
template <typename T, int L, int E, bool MASK>
__global__ void __launch_bounds__(lines_per_wg(L) * L / E)
k_fft_rows(int rows_total, const C2<T> *__restrict__ tw, const C2<T> *gin, C2<T> *gout, T scale,
           Occupancy oc)
{
    constexpr int G = lines_per_wg(L);
    constexpr int NT = G * L / E;
    constexpr int MAXT = 256;  // tiles per row held in LDS (tile width >= 32, L <= 8192)
    // FP64 exchanges split into real and imaginary halves: one T per element
    // in LDS, so twice the workgroups fit per CU (1.61 -> 1.35 ms on 32
    // planes of 4096^2).  FP32 splits 4096-point rows too (0.658 -> 0.632 ms
    // on the same planes) and keeps whole exchanges elsewhere: splitting
    // measured slower on 8192-point rows (64 planes of 8192^2: 3.9 -> 4.2 ms)
    // and about 2 % slower on the batched jobs' shorter rows.
    constexpr bool kSplit = sizeof(T) == 8 || L == 4096;
    __shared__ C2<T> lds[kSplit ? (lds_elems(L, true) + 1) / 2 : lds_elems(L, true)];
    __shared__ unsigned char occ[oc_bytes(L)];
    Lines ln;
    const size_t row0 = (size_t)blockIdx.x * G;
    ln.in_base = ln.out_base = row0 * L;
    ln.in_line = ln.out_line = L;
    ln.in_el = ln.out_el = 1;
    // One row per workgroup: the grid has exactly rows_total workgroups.
    ln.nvalid = G == 1 ? 1 : (int)min((size_t)G, (size_t)rows_total - row0);
    ln.tw_set = 0;
    if constexpr (MASK) {
        static_assert(G == 1, "the occupancy mask covers one row");
        // The row's flag and its tiles' occupancy are uniform, so no LDS
        // table or barrier is needed.  A row without a non-zero cell is zero
        // and is neither computed nor written.
        if (oc.rows && !oc.rows[row0]) return;
        // Every wave reads the row's (at most 64) tile counts, one per lane;
        // the ballot is the occupancy bit mask.
        const int plane = (int)(row0 / oc.ny), v = (int)(row0 % oc.ny);
        const int t = threadIdx.x & 63;
        const bool occd =
            t < oc.tiles_u &&
            oc.counts[((size_t)plane * oc.tiles_v + v / oc.tv) * oc.tiles_u + t] != 0;
        ln.occ_mask = __ballot(occd);
        ln.occ_mask_on = true;
        ln.occ_shift = oc.tu_shift;
    } else if (oc.rows) {
        // Uniform across the workgroup: none of its rows holds a non-zero
        // cell -> the output rows are zero and are neither computed nor written.
        int live = 0;
        for (int g = threadIdx.x; g < ln.nvalid; g += NT) live |= oc.rows[row0 + g];
        if (!__syncthreads_or(live)) return;
    }
    if (!MASK && oc.counts) {
        for (int i = threadIdx.x; i < G * oc.tiles_u && oc.tiles_u <= MAXT; i += NT) {
            const int g = i / oc.tiles_u, t = i % oc.tiles_u;
            const size_t row = row0 + g;
            unsigned char live = 0;
            if (g < ln.nvalid && (oc.rows == nullptr || oc.rows[row])) {
                const int plane = (int)(row / oc.ny), v = (int)(row % oc.ny);
                live = oc.counts[((size_t)plane * oc.tiles_v + v / oc.tv) * oc.tiles_u + t] != 0;
            }
            occ[i] = live;
        }
        __syncthreads();
        ln.occ = occ;
        ln.occ_stride = oc.tiles_u;
        ln.occ_shift = oc.tu_shift;
    }
    C2<T> vals[E];
    GlobalSink<T> sink{gout, &ln, tw, scale};
    pass<T, L, G, true, 0, GlobalSink<T>, false, kSplit, E>(vals, threadIdx.x, ln, 1, tw, lds,
                                                              gin, sink);
}

// Columns: workgroup (x: strip of G columns, y: row set s, z: plane).
struct ColArgs {
    int nx, ny;
    int in_set_mul, in_stride;     // input row  = s*in_set_mul + n*in_stride
    int out_set_mul, out_stride;   // output row = s*out_set_mul + n*out_stride
    int tw_by_set;                 // twiddle W_N^(n*s) after the DFT (four-step A)
    int ts;                        // table stride for the L-point DFT
};

template <typename T, int L, int G, int E>
__global__ void __launch_bounds__(G * L / E)
k_fft_cols(ColArgs a, const C2<T> *__restrict__ tw, const C2<T> *gin, C2<T> *gout, T scale,
           ColOcc co)
{
    // FP64 only: an FP32 workgroup here (up to 1024 threads) is already at
    // two per CU with whole exchanges, so halving its LDS adds none.
    constexpr bool kSplit = sizeof(T) == 8;
    __shared__ C2<T> lds[kSplit ? G * L / 2 : G * L];
    __shared__ unsigned char rl[L];
    const int c0 = blockIdx.x * G;
    const int s = blockIdx.y;
    const size_t plane = (size_t)blockIdx.z * (size_t)a.nx * (size_t)a.ny;
    Lines ln;
    ln.in_base = plane + (size_t)s * a.in_set_mul * (size_t)a.nx + c0;
    ln.out_base = plane + (size_t)s * a.out_set_mul * (size_t)a.nx + c0;
    ln.in_line = ln.out_line = 1;
    ln.in_el = (size_t)a.in_stride * a.nx;
    ln.out_el = (size_t)a.out_stride * a.nx;
    ln.nvalid = min(G, a.nx - c0);
    ln.tw_set = a.tw_by_set ? s : 0;
    ln.occ = nullptr;
    col_rowlive(co, blockIdx.z, s * a.in_set_mul, a.in_stride, L, rl, ln);
    C2<T> vals[E];
    GlobalSink<T> sink{gout, &ln, tw, scale};
    pass<T, L, G, false, 0, GlobalSink<T>, false, kSplit, E>(vals, threadIdx.x, ln, a.ts, tw,
                                                               lds, gin, sink);
}

// Last column pass fused with the phase correction and plane sum
// (phase_correction.cu:185-250): the workgroup transforms its column strip of
// every w-plane in turn and accumulates the phase-rotated results of each
// output pixel in registers; only the image is written.  Output element n of
// set s is image row  s*vs_mul + n*v_stride.
template <typename T, int E>
struct PhaseSink {
    static constexpr bool kRecur = sizeof(T) == 8 && kPhaseRecurF64;
    double are[E], aim[E];
    // FP64: n of each of this thread's pixels (set once), or, with the
    // recurrence, each pixel's current rotation and its per-plane step.
    double nfv[kRecur ? 1 : E];
    double rc[kRecur ? E : 1], rs[kRecur ? E : 1], sc[kRecur ? E : 1], ss[kRecur ? E : 1];
    const double *nfl;    // FP32: n - 1 of the workgroup's pixels in LDS, [n * G + g]
    int G;
    const double *ltab, *mtab;
    double wterm, dwnorm;
    int u0, s, vs_mul, v_stride, nvalid;
    bool phase;
    __device__ __forceinline__ void operator()(int g, int n, int idx, C2<T> o)
    {
        if (g >= nvalid) return;
        if (!phase) {
            are[idx] += (double)o.x;
            aim[idx] += (double)o.y;
            return;
        }
        if constexpr (kRecur) {
            // (p + iq) * exp(i arg_w); dwnorm * n is applied once at the end.
            const double r = rc[idx], q = rs[idx];
            are[idx] += (double)o.x * r - (double)o.y * q;
            aim[idx] += (double)o.x * q + (double)o.y * r;
            rc[idx] = r * sc[idx] - q * ss[idx];
            rs[idx] = r * ss[idx] + q * sc[idx];
        } else if constexpr (sizeof(T) == 8) {
            phase_add_f64(o.x, o.y, wterm, nfv[idx], dwnorm, are[idx], aim[idx]);
        } else {
            phase_add_f32(o.x, o.y, wterm, nfl[n * G + g], are[idx], aim[idx]);
        }
    }
};

template <typename T, int L, int E>
__global__ void __launch_bounds__(lines_per_wg(L) * L / E)
k_fft_cols_phase(ColArgs a, int planes_in, const C2<T> *__restrict__ tw,
                 const C2<T> *__restrict__ grid, PhaseGeom pg, int vs_mul, int v_stride,
                 double *ore_in, double *oim_in, ColOcc co, const PhaseJob *__restrict__ pj)
{
    // One image (planes 0 .. planes_in-1), or with a batch group's job table
    // the image of job blockIdx.z from its planes of the group's stack.
    int planes = planes_in, plane0 = 0;
    double *__restrict__ ore = ore_in;
    double *__restrict__ oim = oim_in;
    if (pj) {
        const PhaseJob &J = pj[blockIdx.z];
        planes = J.planes;
        plane0 = J.plane0;
        pg.wmin = J.wmin;
        pg.dw = J.dw;
        pg.dwnorm = J.dwnorm;
        pg.phase_on = J.phase_on;
        pg.ltab = J.ltab;
        pg.mtab = J.mtab;
        ore = J.re;
        oim = J.im;
    }
    constexpr int G = lines_per_wg(L);
    constexpr int NT = G * L / E;
    __shared__ C2<T> lds[lds_elems(L, false)];
    // Row flags of two planes: with the FP64 prefetch, plane iw + 1's flags
    // are written while other waves may still be choosing which of plane
    // iw's rows to load, so consecutive planes use different halves.
    __shared__ unsigned char rl[2][L];
    __shared__ double nfl[sizeof(T) == 4 ? G * L : 1];
    const int c0 = blockIdx.x * G;
    const int s = blockIdx.y;
    PhaseSink<T, E> sink;
#pragma unroll
    for (int i = 0; i < E; ++i) sink.are[i] = sink.aim[i] = 0.0;
    sink.ltab = pg.ltab;
    sink.mtab = pg.mtab;
    sink.dwnorm = pg.dwnorm;
    sink.u0 = c0;
    sink.s = s;
    sink.vs_mul = vs_mul;
    sink.v_stride = v_stride;
    sink.nvalid = min(G, a.nx - c0);
    sink.phase = pg.phase_on && planes > 1;
    Lines ln;
    ln.in_line = ln.out_line = 1;
    ln.in_el = (size_t)a.in_stride * a.nx;
    ln.out_el = 0;
    ln.nvalid = sink.nvalid;
    ln.tw_set = 0;
    ln.occ = nullptr;
    ln.rowlive = nullptr;
    ln.out_base = 0;
    // Pixels of this thread in the last pass's element layout.
    constexpr int P = npasses(L, E) - 1, R = radix_of(L, P, E), Ns = ns_of(L, P, E), LR = L / R,
                  BPT = E / R;
    if constexpr (sizeof(T) == 4) {
        // FP32: n - 1 of every pixel once, in LDS (registers are kept for
        // two waves per SIMD).
        for (int e = threadIdx.x; e < G * L; e += NT) {
            const int g = e % G, n = e / G;
            nfl[e] = g < sink.nvalid
                         ? pixel_n(pg.ltab[c0 + g], pg.mtab[s * vs_mul + n * v_stride]) - 1.0
                         : 0.0;
        }
        sink.nfl = nfl;
        sink.G = G;
        __syncthreads();
    } else {
        // FP64: n of every pixel once, in registers.
#pragma unroll
        for (int i = 0; i < BPT; ++i) {
            int g, j;
            bmap<false, LR, G>(threadIdx.x + i * NT, g, j);
            const int dbase = (j / Ns) * Ns * R + j % Ns;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int n = dbase + r * Ns;
                const double nf =
                    g < sink.nvalid ? pixel_n(pg.ltab[c0 + g], pg.mtab[s * vs_mul + n * v_stride])
                                    : 1.0;
                if constexpr (PhaseSink<T, E>::kRecur) {
                    // Plane 0's argument is the reference's (wterm = wmin,
                    // phase_correction.cu:200, 222); each plane adds dw.
                    double a0, a1;
                    {
#pragma clang fp contract(off)
                        a0 = 2.0 * kRefPi * pg.wmin * (nf - 1.0);
                        a1 = 2.0 * kRefPi * pg.dw * (nf - 1.0);
                    }
                    rick_sincos(a0, &sink.rs[i * R + r], &sink.rc[i * R + r]);
                    rick_sincos(a1, &sink.ss[i * R + r], &sink.sc[i * R + r]);
                } else {
                    sink.nfv[i * R + r] = nf;
                }
            }
        }
    }
    auto plane_lines = [&](int iw) {
        const int pw = plane0 + iw;   // plane in the grid's (group's) stack
        if (co.rows) col_rowlive(co, pw, s * a.in_set_mul, a.in_stride, L, rl[iw & 1], ln);
        ln.in_base = (size_t)pw * (size_t)a.nx * (size_t)a.ny +
                     (size_t)s * a.in_set_mul * (size_t)a.nx + c0;
    };
    // FP64 runs one wave per SIMD at this register count: the next plane's
    // inputs are loaded into registers while this plane is transformed.  (The
    // FP32 variant keeps two waves per SIMD instead, which measured faster.)
    constexpr bool kPrefetch = sizeof(T) == 8 && kPhasePrefetchF64;
    C2<T> vals[E];
    if constexpr (kPrefetch) {
        plane_lines(0);
        load_first<T, L, G, false, E>(vals, threadIdx.x, ln, grid);
    }
    for (int iw = 0; iw < planes; ++iw) {
        sink.wterm = pg.wmin + iw * pg.dw;  // phase_correction.cu:200
        if constexpr (kPrefetch) {
            C2<T> nxt[E];
            if (iw + 1 < planes) {
                plane_lines(iw + 1);
                load_first<T, L, G, false, E>(nxt, threadIdx.x, ln, grid);
            }
            pass<T, L, G, false, 0, PhaseSink<T, E>, true, false, E>(vals, threadIdx.x, ln, a.ts,
                                                                    tw, lds, grid, sink);
            __syncthreads();  // LDS is reused by the next plane
#pragma unroll
            for (int e = 0; e < E; ++e) vals[e] = nxt[e];
        } else {
            plane_lines(iw);
            pass<T, L, G, false, 0, PhaseSink<T, E>, false, false, E>(vals, threadIdx.x, ln,
                                                                     a.ts, tw, lds, grid, sink);
            __syncthreads();  // LDS is reused by the next plane
        }
    }
    // Write the accumulated pixels.
#pragma unroll
    for (int i = 0; i < BPT; ++i) {
        int g, j;
        bmap<false, LR, G>(threadIdx.x + i * NT, g, j);
        const int dbase = (j / Ns) * Ns * R + j % Ns;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int n = dbase + r * Ns;
            if (g < sink.nvalid) {
                const size_t pix = (size_t)(s * vs_mul + n * v_stride) * a.nx + (c0 + g);
                double vre = sink.are[i * R + r], vim = sink.aim[i * R + r];
                if constexpr (PhaseSink<T, E>::kRecur || sizeof(T) == 4) {
                    if (sink.phase) {
                        const double nf = pixel_n(pg.ltab[c0 + g], pg.mtab[s * vs_mul + n * v_stride]);
                        vre = vre * pg.dwnorm * nf;
                        vim = vim * pg.dwnorm * nf;
                    }
                }
                store_stream(vre, &ore[pix]);
                store_stream(vim, &oim[pix]);
            }
        }
    }
}

// Launch helpers dispatching on the (power-of-two) line length.
template <typename T>
hipError_t launch_rows(int L, int rows_total, const C2<T> *tw, const C2<T> *in, C2<T> *out,
                       T scale, hipStream_t s, const Occupancy &oc)
{
#define RICK_ROWS(LL)                                                                      \
    case LL: {                                                                             \
        constexpr int G = lines_per_wg(LL);                                                \
        constexpr int E = row_values<T>();                                                 \
        if constexpr (G == 1) {                                                            \
            if (oc.counts && oc.tiles_u <= 64) {                                           \
                hipLaunchKernelGGL((k_fft_rows<T, LL, E, true>), dim3(rows_total),         \
                                   dim3(LL / E), 0, s, rows_total, tw, in, out, scale, oc); \
                return hipGetLastError();                                                  \
            }                                                                              \
        }                                                                                  \
        hipLaunchKernelGGL((k_fft_rows<T, LL, E, false>), dim3((rows_total + G - 1) / G),  \
                           dim3(G * LL / E), 0, s, rows_total, tw, in, out, scale, oc);    \
        return hipGetLastError();                                                          \
    }
    switch (L) {
        RICK_ROWS(16) RICK_ROWS(32) RICK_ROWS(64) RICK_ROWS(128) RICK_ROWS(256)
        RICK_ROWS(512) RICK_ROWS(1024) RICK_ROWS(2048) RICK_ROWS(4096) RICK_ROWS(8192)
        default: return hipErrorInvalidValue;
    }
#undef RICK_ROWS
}

template <typename T>
hipError_t launch_cols(int L, const ColArgs &a, int sets, int planes, const C2<T> *tw,
                       const C2<T> *in, C2<T> *out, hipStream_t s, const ColOcc &co = ColOcc{})
{
#define RICK_COLS(LL)                                                                        \
    case LL: {                                                                               \
        constexpr int G = col_lines<T>(LL);                                                  \
        constexpr int E = col_values<T>();                                                   \
        hipLaunchKernelGGL((k_fft_cols<T, LL, G, E>), dim3((a.nx + G - 1) / G, sets, planes), \
                           dim3(G * LL / E), 0, s, a, tw, in, out, T(1), co);                \
        return hipGetLastError();                                                            \
    }
    switch (L) {
        RICK_COLS(16) RICK_COLS(32) RICK_COLS(64) RICK_COLS(128) RICK_COLS(256)
        RICK_COLS(512)
        default: return hipErrorInvalidValue;
    }
#undef RICK_COLS
}

template <typename T>
hipError_t launch_cols_phase(int L, const ColArgs &a, int sets, int planes, const C2<T> *tw,
                             const C2<T> *grid, const PhaseGeom &pg, int vs_mul, int v_stride,
                             double *re, double *im, hipStream_t s, const ColOcc &co,
                             const PhaseJob *pj, int njobs)
{
#define RICK_CPH(LL)                                                                         \
    case LL: {                                                                               \
        constexpr int G = lines_per_wg(LL);                                                  \
        constexpr int E = phase_values<T>(LL);                                               \
        hipLaunchKernelGGL((k_fft_cols_phase<T, LL, E>),                                     \
                           dim3((a.nx + G - 1) / G, sets, pj ? njobs : 1), dim3(G * LL / E), \
                           0, s, a, planes, tw, grid, pg, vs_mul, v_stride, re, im, co, pj); \
        return hipGetLastError();                                                            \
    }
    switch (L) {
        RICK_CPH(16) RICK_CPH(32) RICK_CPH(64) RICK_CPH(128) RICK_CPH(256) RICK_CPH(512)
        default: return hipErrorInvalidValue;
    }
#undef RICK_CPH
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
    hipError_t e = hipSuccess;

    // ---- columns (v axis, length ny) ----
    if (!fft_needs_tmp(sizeof(T) == 8 ? RICK_FP64 : RICK_FP32, ny)) {
        const ColArgs a{nx, ny, 0, 1, 0, 1, 0, 1};
        tm->begin(RICK_STAGE_FFT_COLS, s);
        e = launch_cols<T>(ny, a, 1, planes, twy, in, out, s);
        tm->end(RICK_STAGE_FFT_COLS, s);
        if (e != hipSuccess) return e;
    } else {
        // Four-step: ny = N1 * N2 with N1 = 64.
        const int N1 = 64, N2 = ny / 64;
        const ColArgs A{nx, ny, 1, N2, N1, 1, 1, N2};   // rows n2 + n1*N2 -> n2*N1 + k1
        const ColArgs B{nx, ny, 1, N1, 1, N1, 0, N1};   // rows k1 + n2*N1 -> k1 + k2*N1
        if (in_v != out_v) {
            tm->begin(RICK_STAGE_FFT_COLS, s);
            e = launch_cols<T>(N1, A, N2, planes, twy, in, out, s);
            tm->end(RICK_STAGE_FFT_COLS, s);
            if (e != hipSuccess) return e;
            tm->begin(RICK_STAGE_FFT_COLS_B, s);
            e = launch_cols<T>(N2, B, N1, planes, twy, out, out, s);
            tm->end(RICK_STAGE_FFT_COLS_B, s);
            if (e != hipSuccess) return e;
        } else {
            // In place: per plane, A: plane -> tmp, B: tmp -> plane.
            if (tmp_v == nullptr) return hipErrorInvalidValue;
            CT *tmp = static_cast<CT *>(tmp_v);
            const size_t pl = (size_t)nx * ny;
            for (int iw = 0; iw < planes; ++iw) {
                tm->begin(RICK_STAGE_FFT_COLS, s);
                e = launch_cols<T>(N1, A, N2, 1, twy, out + iw * pl, tmp, s);
                tm->end(RICK_STAGE_FFT_COLS, s);
                if (e != hipSuccess) return e;
                tm->begin(RICK_STAGE_FFT_COLS_B, s);
                e = launch_cols<T>(N2, B, N1, 1, twy, tmp, out + iw * pl, s);
                tm->end(RICK_STAGE_FFT_COLS_B, s);
                if (e != hipSuccess) return e;
            }
        }
    }
    // ---- rows (u axis, length nx), in place on out, scaled ----
    tm->begin(RICK_STAGE_FFT_ROWS, s);
    e = launch_rows<T>(nx, planes * ny, twx, out, out, (T)scale_d, s, Occupancy{});
    tm->end(RICK_STAGE_FFT_ROWS, s);
    return e;
}

// Image chain: in place, rows first reading only occupied tiles of a sparse
// grid and skipping rows without a non-zero cell, then (for ny > 512)
// four-step pass A in place with rows left in transposed order (skipped rows
// read as zero),
// then the last column pass fused with the phase correction and plane sum,
// writing the image.
template <typename T>
hipError_t image_chain_impl(const FftPlan &p, void *grid_v, double scale_d, hipStream_t s,
                            Timing *tm, const Occupancy &oc, const PhaseGeom &pg, double *re,
                            double *im, const PhaseJob *pj, int njobs)
{
    using CT = C2<T>;
    CT *grid = static_cast<CT *>(grid_v);
    const CT *twx = static_cast<const CT *>(p.tw_x);
    const CT *twy = static_cast<const CT *>(p.tw_y);
    const int nx = p.nx, ny = p.ny, planes = p.planes;
    ColOcc co{};
    if (oc.counts && oc.rows) {
        co.rows = oc.rows;
        co.ny = ny;
    }
    tm->begin(RICK_STAGE_FFT_ROWS, s);
    hipError_t e = launch_rows<T>(nx, planes * ny, twx, grid, grid, (T)scale_d, s, oc);
    tm->end(RICK_STAGE_FFT_ROWS, s);
    if (e != hipSuccess) return e;
    if (!fft_needs_tmp(sizeof(T) == 8 ? RICK_FP64 : RICK_FP32, ny)) {
        const ColArgs a{nx, ny, 0, 1, 0, 1, 0, 1};
        tm->begin(RICK_STAGE_PHASE, s);
        e = launch_cols_phase<T>(ny, a, 1, planes, twy, grid, pg, 0, 1, re, im, s, co, pj,
                                 njobs);
        tm->end(RICK_STAGE_PHASE, s);
        return e;
    }
    // ny = N1 * N2.  FP64 at ny = 8192 splits 128 x 64: the fused pass B
    // repeats its N2-point transform for every plane, so it gets the shorter
    // lines (one LDS exchange at 64 points, two at 128) while pass A, once
    // per plane, takes the 128-point ones (JVLA-like chain, 8192^2 x 64: pass
    // B 16.7 -> 13.9 ms, pass A 17.3 -> 18.0 ms).  FP32 keeps 64 x 128: its
    // 128-point pass A at 8 values per thread (three radix passes) cost more
    // than pass B saved (9.3 -> 11.2 ms against 8.1 -> 7.2 ms).
    const int N1 = (sizeof(T) == 8 && ny > 4096) ? ny / 64 : 64, N2 = ny / N1;
    // A: rows n2 + N2*n1 -> n2 + N2*k1 (in place), twiddle W_N^(k1*n2).
    const ColArgs At{nx, ny, 1, N2, 1, N2, 1, N2};
    // B: rows N2*k1 + n2 -> X[k1 + N1*k2], image row k1 + N1*k2.
    const ColArgs Bt{nx, ny, N2, 1, N2, 1, 0, N1};
    tm->begin(RICK_STAGE_FFT_COLS, s);
    e = launch_cols<T>(N1, At, N2, planes, twy, grid, grid, s, co);
    tm->end(RICK_STAGE_FFT_COLS, s);
    if (e != hipSuccess) return e;
    tm->begin(RICK_STAGE_PHASE, s);
    e = launch_cols_phase<T>(N2, Bt, N1, planes, twy, grid, pg, 1, N1, re, im, s, ColOcc{}, pj,
                             njobs);
    tm->end(RICK_STAGE_PHASE, s);
    return e;
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

hipError_t launch_image_chain(const FftPlan &p, int precision, void *grid, double scale,
                              hipStream_t s, Timing *tm, const GridGeom *occ_geom,
                              const uint32_t *occ_counts, const unsigned char *occ_rows,
                              const PhaseGeom &pg,
                              double *re, double *im, const PhaseJob *jobs, int njobs)
{
    Timing off;
    if (tm == nullptr) tm = &off;
    Occupancy oc{};
    if (occ_geom) {
        if (!occ_counts || p.nx / occ_geom->tu > 256 || occ_geom->tu <= 0 ||
            (occ_geom->tu & (occ_geom->tu - 1)) != 0)
            return hipErrorInvalidValue;
        oc.counts = occ_counts;
        oc.rows = occ_rows;
        oc.tu_shift = __builtin_ctz((unsigned)occ_geom->tu);
        oc.tv = occ_geom->tv;
        oc.tiles_u = occ_geom->tiles_u;
        oc.tiles_v = occ_geom->tiles_v;
        oc.ny = occ_geom->ny;
    }
    if (precision == RICK_FP64)
        return image_chain_impl<double>(p, grid, scale, s, tm, oc, pg, re, im, jobs, njobs);
    return image_chain_impl<float>(p, grid, scale, s, tm, oc, pg, re, im, jobs, njobs);
}

}  // namespace rick
