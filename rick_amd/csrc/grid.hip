// Binned w-stacking gridder for gfx950.
//
// Computes what the reference's wstack() computes (w-stacking.cu:436-541):
//   grid[w][k][j] += sum_rows sum_{f,p} fl32(weight*vis) * K[jKer] * K[kKer]
// over each row's (2KL+1)^2 footprint, with the same index arithmetic
// (IEEE divides, truncation toward zero, +0.5 half-cell offsets, the
// off-centre fabs(u_dist+KL) table look-up), organised for the GPU:
//
//   k_count      rows -> buckets (bucket = w-plane x TUxTV grid tile whose core
//                the row's footprint touches; at most 4 per row).  Each
//                workgroup aggregates the keys of its 16384 rows in an LDS hash
//                table and adds one count per distinct key to the global
//                histogram: device-scope atomics are slow on this chip (its
//                eight L2 slices are not coherent with each other), and 16384
//                consecutive rows touch a few thousand distinct buckets.
//   k_scan_*     multi-block exclusive scans of the histogram -> bucket
//                offsets, the task list (a task is <= rmax records of one
//                bucket, or a zero task), the list of split buckets.
//   k_scatter    forms per row S = sum_{f,p} fl32(weight*vis) (NaN real parts
//                skipped, w-stacking.cu:517) and writes a record {pos_u,
//                pos_v, S} to each of its buckets; slots come from one
//                returning global atomic per distinct key per 2048-row block.
//   k_zero_split clears the tiles of split buckets (overwrite mode).
//   k_tile       one workgroup per task: accumulates its records into an FP64
//                LDS tile (one lane per record, LDS atomics), then writes
//                the tile once with plain, coalesced stores; tasks of split
//                buckets flush with HBM atomics instead.  In overwrite mode
//                each empty bucket gets a zero task that clears its tile; zero
//                tasks sit between the accumulation tasks in bucket order, so
//                the bandwidth-bound clearing shares the launch with the
//                LDS-bound accumulation.  Every grid cell is written exactly
//                once, so no memset of the grid is needed.  For the image
//                chain (sparse overwrite: empty tiles left unwritten) it also
//                flags each grid row it writes a non-zero cell to, so the FFT
//                skips the rest.
//
// The per-cell weight K[jKer]*K[kKer] is bitwise the reference's; the per-cell
// value S*c differs from the reference's sum_{f,p} fl(fl32(w*v)*c) only by
// rounding order (identical when F*P == 1).
#include "rick_internal.h"

#include <atomic>
#include <type_traits>

#pragma clang fp contract(off)

namespace rick {
namespace {

constexpr int kThreads = 256;                 // count / scatter / scan workgroups
constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr int kScanBlock = 4 * kThreads;      // buckets per scan block
// Scatter: 2048-row blocks of 1024 threads, 2 rows in flight per thread (64
// VGPRs: two blocks per CU);
// keys beyond the hash table's reach take a per-record global atomic.
constexpr int kScatterThreads = 1024;
constexpr int kRowsPerThread = 2;
constexpr int kRowsPerBlock = kScatterThreads * kRowsPerThread;
constexpr int kHashBits = 12;
constexpr int kHashSlots = 1 << kHashBits;
constexpr uint32_t kDirect = 0x80000000u;     // item flag: key without a hash slot
// Count: 16384-row blocks, 8192-slot hash with bounded probing, overflow ->
// direct global atomic.
constexpr int kBinThreads = 1024;
constexpr int kBinRowsPerThread = 16;
constexpr int kBinRows = kBinThreads * kBinRowsPerThread;
static_assert(kBinRows % kRowsPerBlock == 0, "a scatter block must not span two batch jobs");
constexpr int kBinHashBits = 13;
constexpr int kBinSlots = 1 << kBinHashBits;
constexpr int kMaxProbe = 32;
constexpr uint32_t kZeroTask = 2u;            // task.w of a zero task
constexpr uint32_t kSplitTask = 3u;           // task.w of a split bucket's task: 3 + split slot
constexpr int kTileThreads = 512;
constexpr int kLdsBytes = 160 * 1024;         // LDS per CU
constexpr uint32_t kPersistSlots = 64;        // task slots per resident workgroup to persist
constexpr int kPadU = 8;  // LDS row padding (elements): rows shift by 8 banks

struct RecF64 {
    double pu, pv, sre, sim;
};
struct RecF32 {
    double pu, pv;
    float sre, sim;
};

template <int PREC>
struct RecOf;
template <>
struct RecOf<RICK_FP64> {
    using type = RecF64;
};
template <>
struct RecOf<RICK_FP32> {
    using type = RecF32;
};

struct Foot {
    int gw, gu, gv, jmin, jmax, kmin, kmax;
};

__device__ __forceinline__ bool in_int_range(double x)
{
    return x > -2147483648.0 && x < 2147483647.0;  // false for NaN
}

// Footprint of a row whose positions (in cells) are pu, pv and w-plane
// coordinate wi; w-stacking.cu:464-476 with the clip made explicit.
// A batch job's rows check their plane against the job's `planes` and land
// at plane0 + gw of the group's plane stack.
__device__ __forceinline__ bool footprint(const GridGeom &g, double pu, double pv, double wi,
                                          Foot &f, int planes, int plane0)
{
    if (!in_int_range(wi) || !in_int_range(pu) || !in_int_range(pv)) return false;
    f.gw = (int)wi;
    f.gu = (int)pu;
    f.gv = (int)pv;
    if (f.gw < 0 || f.gw >= planes) return false;
    f.gw += plane0;
    f.jmin = (f.gu > g.kl - 1) ? f.gu - g.kl : 0;
    f.jmax = (f.gu < g.nx - g.kl) ? f.gu + g.kl : g.nx - 1;
    f.kmin = (f.gv > g.kl - 1) ? f.gv - g.kl : 0;
    f.kmax = (f.gv < g.ny - g.kl) ? f.gv + g.kl : g.ny - 1;
    return f.jmin <= f.jmax && f.kmin <= f.kmax;
}

// u/v range of a record inside the tile kernel (plane comes from the bucket).
__device__ __forceinline__ void uv_range(const GridGeom &g, double pu, double pv, Foot &f)
{
    f.gu = (int)pu;
    f.gv = (int)pv;
    f.jmin = (f.gu > g.kl - 1) ? f.gu - g.kl : 0;
    f.jmax = (f.gu < g.nx - g.kl) ? f.gu + g.kl : g.nx - 1;
    f.kmin = (f.gv > g.kl - 1) ? f.gv - g.kl : 0;
    f.kmax = (f.gv < g.ny - g.kl) ? f.gv + g.kl : g.ny - 1;
}

// A pointer known to address global memory: loads through it are global
// loads, not flat ones (pointers read from a job table are otherwise flat).
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(1))) T *as_global(const T *p)
{
    return (const __attribute__((address_space(1))) T *)p;
}

// x / d, the reference's IEEE division (w-stacking.cu:464-466).  When d is a
// power of two with an exact reciprocal inv (inv != 0), x * inv is that same
// correctly rounded quotient (both are the exact value x * 2^-k, rounded
// once), at the cost of one multiply instead of a division sequence.
__device__ __forceinline__ double div_by(double x, double d, double inv)
{
    return inv != 0.0 ? x * inv : x / d;
}

// The rows a gridding block reads: its job's arrays, row count, w-plane
// spacing and planes, and the block's first row in them.
struct BlockRows {
    const double *uu, *vv, *ww;
    const float *vr, *vi, *wt;
    uint32_t n, base;
    double dw, inv_dw;
    int planes, plane0;
};

// Block whose first virtual row is vrow: a single job (jobs == null), or
// the batch job holding it (jobs sorted by row0, row0 a multiple of the
// block's rows, so a block never spans two jobs).
__device__ __forceinline__ BlockRows block_rows(const GridGeom &g, uint32_t vrow, uint32_t n,
                                                const double *uu, const double *vv,
                                                const double *ww, const float *vr,
                                                const float *vi, const float *wt,
                                                const BatchJob *jobs, int njobs)
{
    if (jobs == nullptr)
        return BlockRows{uu, vv, ww, vr, vi, wt, n, vrow, g.dw, g.inv_dw, g.planes, 0};
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].row0 <= vrow) lo = mid;
        else hi = mid - 1;
    }
    const BatchJob &J = jobs[lo];
    return BlockRows{J.uu, J.vv, J.ww, J.vr, J.vi, J.wt, J.n, vrow - J.row0, J.dw, J.inv_dw,
                     J.planes, J.plane0};
}

// Buckets of a footprint (support <= tile size, so at most 2 tiles per axis;
// cell indices are >= 0, so shifts divide by the tile size).
__device__ __forceinline__ int row_keys(const GridGeom &g, const Foot &f, uint32_t *keys)
{
    const int tu0 = f.jmin >> g.tu_shift, tu1 = f.jmax >> g.tu_shift;
    const int tv0 = f.kmin >> g.tv_shift, tv1 = f.kmax >> g.tv_shift;
    const uint32_t base = (uint32_t)f.gw * (uint32_t)g.tiles_v;
    int n = 0;
    for (int tv = tv0; tv <= tv1; ++tv)
        for (int tu = tu0; tu <= tu1; ++tu)
            keys[n++] = (base + (uint32_t)tv) * (uint32_t)g.tiles_u + (uint32_t)tu;
    return n;
}

// LDS open-addressing hash set of 2^HB bucket keys with at most kMaxProbe
// probes: returns the key's slot, or kEmpty when none was free on its path.
template <int HB>
__device__ __forceinline__ uint32_t hinsert_bounded(uint32_t *hkey, uint32_t key)
{
    uint32_t h = (key * 2654435761u) >> (32 - HB);
    for (int p = 0; p < kMaxProbe; ++p) {
        const uint32_t old = atomicCAS(&hkey[h], kEmpty, key);
        if (old == kEmpty || old == key) return h;
        h = (h + 1) & ((1u << HB) - 1u);
    }
    return kEmpty;
}

// Same, reading the slot before trying to claim it: a key already present
// costs an LDS read instead of a compare-and-swap.
template <int HB>
__device__ __forceinline__ uint32_t hinsert_peek(uint32_t *hkey, uint32_t key)
{
    uint32_t h = (key * 2654435761u) >> (32 - HB);
    for (int p = 0; p < kMaxProbe; ++p) {
        const uint32_t cur = hkey[h];
        if (cur == key) return h;
        if (cur == kEmpty) {
            const uint32_t old = atomicCAS(&hkey[h], kEmpty, key);
            if (old == kEmpty || old == key) return h;
        }
        h = (h + 1) & ((1u << HB) - 1u);
    }
    return kEmpty;
}

__global__ void k_fill_u32(uint32_t *p, uint32_t n, uint32_t v)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        p[i] = v;
}

// 8 waves per SIMD: two 1024-thread blocks per CU (their hash tables take
// 64 KB of LDS each), which needs the SGPR count held to the 8-wave budget.
template <bool BATCH>
__global__ void __launch_bounds__(kBinThreads, 8)
k_count(GridGeom g, uint32_t n_all, const double *__restrict__ uu_all,
        const double *__restrict__ vv_all, const double *__restrict__ ww_all,
        uint32_t *__restrict__ counts, const BatchJob *__restrict__ jobs, int njobs)
{
    __shared__ uint32_t hkey[kBinSlots], hcnt[kBinSlots];
    for (int i = threadIdx.x; i < kBinSlots; i += kBinThreads) {
        hkey[i] = kEmpty;
        hcnt[i] = 0;
    }
    __syncthreads();
    const BlockRows br = block_rows(g, blockIdx.x * (uint32_t)kBinRows, n_all, uu_all, vv_all,
                                    ww_all, nullptr, nullptr, nullptr, BATCH ? jobs : nullptr,
                                    njobs);
    const auto *__restrict__ uu = as_global(br.uu);
    const auto *__restrict__ vv = as_global(br.vv);
    const auto *__restrict__ ww = as_global(br.ww);
    const uint32_t n = br.n, base = br.base;
    // A thread's rows are taken kCountBatch at a time, all their coordinates
    // requested before any is hashed: one memory round trip per batch rather
    // than one per row.
    constexpr int kCountBatch = 4;
    static_assert(kBinRowsPerThread % kCountBatch == 0, "row batches per thread");
    for (int k0 = 0; k0 < kBinRowsPerThread; k0 += kCountBatch) {
        double u[kCountBatch], v[kCountBatch], w[kCountBatch];
#pragma unroll
        for (int j = 0; j < kCountBatch; ++j) {
            const uint32_t i = base + (uint32_t)(k0 + j) * kBinThreads + threadIdx.x;
            u[j] = v[j] = w[j] = 0.0;
            if (i < n) {
                u[j] = uu[i];
                v[j] = vv[i];
                w[j] = ww[i];
            }
        }
#pragma unroll
        for (int j = 0; j < kCountBatch; ++j) {
            const uint32_t i = base + (uint32_t)(k0 + j) * kBinThreads + threadIdx.x;
            Foot f;
            if (i < n && footprint(g, div_by(u[j], g.dx, g.inv_dx), div_by(v[j], g.dx, g.inv_dx),
                                   div_by(w[j], br.dw, br.inv_dw), f, br.planes, br.plane0)) {
                uint32_t keys[4];
                const int nk = row_keys(g, f, keys);
                for (int q = 0; q < nk; ++q) {
                    const uint32_t h = hinsert_peek<kBinHashBits>(hkey, keys[q]);
                    if (h == kEmpty) atomicAdd(&counts[keys[q]], 1u);
                    else atomicAdd(&hcnt[h], 1u);
                }
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kBinSlots; i += kBinThreads) {
        const uint32_t c = hcnt[i];
        if (c) atomicAdd(&counts[hkey[i]], c);
    }
}

// Scan of the histogram in three kernels: per-block sums, a single-block scan
// of those, and the per-block downsweep that emits offsets and tasks.  Five
// quantities are scanned: records, accumulation tasks, split buckets,
// non-empty buckets, zero tasks.  The task array holds the accumulation
// tasks first and the zero tasks after them; the tile kernel interleaves the
// two evenly over its workgroups.
constexpr int kScanQ = 5;
// Tasks of a bucket with c > 0 records.
__device__ __forceinline__ uint32_t ntasks_of(uint32_t c, uint32_t rmax)
{
    return (c + rmax - 1) / rmax;
}

// Tasks of a scan thread's 4 consecutive buckets (nv of them exist): those of
// the non-empty ones, plus, when the grid is overwritten (zero != 0), one zero
// task per g.zrun consecutive empty ones.
__device__ __forceinline__ void group_tasks(const uint32_t *c, int nv, uint32_t rmax, int zero,
                                            int zrun, uint32_t &real, uint32_t &zeros)
{
    real = zeros = 0;
    int run = 0;  // length of the current run of empty buckets
    for (int k = 0; k < nv; ++k) {
        if (c[k]) {
            real += ntasks_of(c[k], rmax);
            run = 0;
        } else {
            if (zero && run % zrun == 0) ++zeros;
            ++run;
        }
    }
}

__global__ void __launch_bounds__(kThreads)
k_scan_reduce(GridGeom g, const uint32_t *__restrict__ counts, uint32_t *__restrict__ blocksum,
              uint32_t nscan, int zero)
{
    __shared__ uint32_t s[kScanQ][kThreads];
    uint32_t r = 0, sp = 0, ne = 0, c[4] = {0, 0, 0, 0};
    const uint32_t b0 = blockIdx.x * kScanBlock + threadIdx.x * 4;
    const int nv = b0 >= g.nb ? 0 : (int)min(4u, g.nb - b0);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (k < nv) {
            c[k] = counts[b0 + k];
            r += c[k];
            sp += c[k] > g.rmax ? 1u : 0u;
            ne += c[k] > 0 ? 1u : 0u;
        }
    }
    uint32_t tr, tz;
    group_tasks(c, nv, g.rmax, zero, g.zrun, tr, tz);
    s[0][threadIdx.x] = r;
    s[1][threadIdx.x] = tr;
    s[2][threadIdx.x] = sp;
    s[3][threadIdx.x] = ne;
    s[4][threadIdx.x] = tz;
    __syncthreads();
    for (int d = kThreads / 2; d > 0; d >>= 1) {
        if ((int)threadIdx.x < d) {
#pragma unroll
            for (int a = 0; a < kScanQ; ++a) s[a][threadIdx.x] += s[a][threadIdx.x + d];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int a = 0; a < kScanQ; ++a) blocksum[a * nscan + blockIdx.x] = s[a][0];
    }
}

constexpr int kTopThreads = 1024;

__global__ void __launch_bounds__(kTopThreads)
k_scan_top(GridGeom g, uint32_t *__restrict__ blocksum, uint32_t nscan,
           uint32_t *__restrict__ offsets, uint32_t *__restrict__ meta)
{
    __shared__ uint32_t s[kScanQ][kTopThreads];
    uint32_t carry[kScanQ] = {0, 0, 0, 0, 0};
    for (uint32_t c0 = 0; c0 < nscan; c0 += kTopThreads) {
        const uint32_t i = c0 + threadIdx.x;
        uint32_t v[kScanQ];
#pragma unroll
        for (int a = 0; a < kScanQ; ++a) {
            v[a] = i < nscan ? blocksum[a * nscan + i] : 0u;
            s[a][threadIdx.x] = v[a];
        }
        __syncthreads();
        for (int d = 1; d < kTopThreads; d <<= 1) {
            uint32_t add[kScanQ] = {0, 0, 0, 0, 0};
            if ((int)threadIdx.x >= d)
#pragma unroll
                for (int a = 0; a < kScanQ; ++a) add[a] = s[a][threadIdx.x - d];
            __syncthreads();
#pragma unroll
            for (int a = 0; a < kScanQ; ++a) s[a][threadIdx.x] += add[a];
            __syncthreads();
        }
#pragma unroll
        for (int a = 0; a < kScanQ; ++a) {
            if (i < nscan) blocksum[a * nscan + i] = carry[a] + s[a][threadIdx.x] - v[a];
            carry[a] += s[a][kTopThreads - 1];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        offsets[g.nb] = carry[0];
        meta[0] = carry[0];             // records
        meta[1] = carry[1] + carry[4];  // tasks (accumulation + zero)
        meta[2] = carry[2];             // split buckets
        meta[3] = carry[3];             // non-empty buckets
        meta[4] = carry[1];             // accumulation tasks (stored first)
        meta[5] = 0;                    // next task for the tile kernel's workgroups
        meta[6] = meta[7] = 0;          // largest |S| of the call (FP32 split buckets)
    }
}

__global__ void __launch_bounds__(kThreads)
k_scan_down(GridGeom g, const uint32_t *__restrict__ counts,
            const uint32_t *__restrict__ blockoff, uint32_t nscan, uint32_t *__restrict__ offsets,
            uint32_t *__restrict__ cursor, uint4 *__restrict__ tasks, uint32_t task_cap,
            uint32_t *__restrict__ split_list, uint32_t split_cap, const uint32_t *__restrict__ meta,
            int zero)
{
    __shared__ uint32_t s[4][kThreads];
    uint32_t c[4] = {0, 0, 0, 0}, v[4] = {0, 0, 0, 0};
    const uint32_t b0 = blockIdx.x * kScanBlock + threadIdx.x * 4;
    const int nv = b0 >= g.nb ? 0 : (int)min(4u, g.nb - b0);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (k < nv) {
            c[k] = counts[b0 + k];
            v[0] += c[k];
            v[2] += c[k] > g.rmax ? 1u : 0u;
        }
    }
    group_tasks(c, nv, g.rmax, zero, g.zrun, v[1], v[3]);
#pragma unroll
    for (int a = 0; a < 4; ++a) s[a][threadIdx.x] = v[a];
    __syncthreads();
    for (int d = 1; d < kThreads; d <<= 1) {
        uint32_t add[4] = {0, 0, 0, 0};
        if ((int)threadIdx.x >= d)
#pragma unroll
            for (int a = 0; a < 4; ++a) add[a] = s[a][threadIdx.x - d];
        __syncthreads();
#pragma unroll
        for (int a = 0; a < 4; ++a) s[a][threadIdx.x] += add[a];
        __syncthreads();
    }
    uint32_t ro = blockoff[blockIdx.x] + s[0][threadIdx.x] - v[0];
    uint32_t to = blockoff[nscan + blockIdx.x] + s[1][threadIdx.x] - v[1];
    uint32_t so = blockoff[2 * nscan + blockIdx.x] + s[2][threadIdx.x] - v[2];
    // Zero tasks follow all accumulation tasks (meta[4] of them).
    uint32_t zo = meta[4] + blockoff[4 * nscan + blockIdx.x] + s[3][threadIdx.x] - v[3];
    int run = 0;
    for (int k = 0; k < nv; ++k) {
        const uint32_t b = b0 + k;
        offsets[b] = ro;
        cursor[b] = ro;
        if (c[k] == 0) {
            // Every zrun-th bucket of a run of empty ones starts a zero task.
            if (zero && run % g.zrun == 0) {
                uint32_t len = 1;
                while (len < (uint32_t)g.zrun && k + (int)len < nv && c[k + len] == 0) ++len;
                if (zo < task_cap) tasks[zo] = make_uint4(b, len, 0u, kZeroTask);
                ++zo;
            }
            ++run;
        } else {
            run = 0;
            const uint32_t nt = ntasks_of(c[k], g.rmax);
            // Tasks of a split bucket carry its slot in split_list (always
            // below split_cap: a split bucket holds more than rmax records).
            const uint32_t tag = nt > 1 ? (so < split_cap ? kSplitTask + so : 1u) : 0u;
            for (uint32_t q = 0; q < nt; ++q) {
                const uint32_t r0 = ro + q * g.rmax, r1 = min(r0 + g.rmax, ro + c[k]);
                if (to + q < task_cap) tasks[to + q] = make_uint4(b, r0, r1, tag);
            }
            if (nt > 1 && so < split_cap) split_list[so++] = b;
            to += nt;
        }
        ro += c[k];
    }
}

// Channel/polarization sums of the 64 rows r0 .. r0+63 of one wave, read
// cooperatively when a row holds F*P = 4L floats (L = 2 or 4): the wave's rows
// are 64*4L contiguous floats, so load j covers rows r0 + j*(64/L) ... with
// lane l holding elements 4(l%L) .. 4(l%L)+3 of row r0 + j*(64/L) + l/L in one
// 16-byte read, and every load instruction reads whole lines.  The running
// sum walks the row's L lanes in element order (shfl_up), so each row's sum is
// formed in the order of the per-row loop in k_scatter (w-stacking.cu:512-527).
// Lane l returns the sums of row r0 + l (zero past n).  widx[c] is the weight
// index (4(l%L) + c) mod P of this lane's element c in every row
// (row_weight_index), computed once per thread.
template <int L>
__device__ __forceinline__ void row_weight_index(uint32_t P, uint32_t (&widx)[4])
{
    const uint32_t q = (threadIdx.x & 63u) % L;
#pragma unroll
    for (int c = 0; c < 4; ++c) widx[c] = P == 1u ? 0u : (4u * q + (uint32_t)c) % P;
}

template <int L>
__device__ __forceinline__ void wave_row_sums(uint32_t r0, uint32_t n, uint32_t P,
                                              const uint32_t (&widx)[4],
                                              const float *__restrict__ vr,
                                              const float *__restrict__ vi,
                                              const float *__restrict__ wt, double &out_re,
                                              double &out_im)
{
    static_assert(L == 2 || L == 4, "four or eight lanes' worth of floats per row");
    constexpr int RPL = 64 / L;  // rows per load
    const int lane = threadIdx.x & 63;
    const int q = lane % L;
    const float4 *r4 = reinterpret_cast<const float4 *>(vr);
    const float4 *i4 = reinterpret_cast<const float4 *>(vi);
    out_re = out_im = 0.0;
#pragma unroll
    for (int j = 0; j < L; ++j) {
        const uint32_t row = r0 + (uint32_t)(j * RPL + lane / L);
        float4 ra = make_float4(0.f, 0.f, 0.f, 0.f), ia = ra;
        float wv[4] = {0.f, 0.f, 0.f, 0.f};
        if (row < n) {
            const size_t e = (size_t)row * L + (size_t)q;
            ra = load_stream4(&r4[e]);
            ia = load_stream4(&i4[e]);
            const float *wrow = wt + (size_t)row * P;
#pragma unroll
            for (int c = 0; c < 4; ++c) wv[c] = wrow[widx[c]];
        }
        const float rr[4] = {ra.x, ra.y, ra.z, ra.w}, ii[4] = {ia.x, ia.y, ia.z, ia.w};
        double run_re = 0.0, run_im = 0.0;
#pragma unroll
        for (int st = 0; st < L; ++st) {
            if (q == st) {
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    if (!isnan(rr[c])) {
                        const float wr = wv[c] * rr[c];
                        const float wi = wv[c] * ii[c];
                        run_re += (double)wr;
                        run_im += (double)wi;
                    }
                }
            }
            if (st + 1 < L) {
                run_re = __shfl_up(run_re, 1, L);
                run_im = __shfl_up(run_im, 1, L);
            }
        }
        // Row j*RPL + m of this load ends in lane m*L + L-1.
        const int src = (lane % RPL) * L + (L - 1);
        const double t_re = __shfl(run_re, src), t_im = __shfl(run_im, src);
        if (lane / RPL == j) {
            out_re = t_re;
            out_im = t_im;
        }
    }
}

// VMODE: 0 = per-row scalar reads, 1 = per-row 16-byte reads (F*P a multiple
// of 4), 2 or 4 = wave-cooperative reads with VMODE lanes per row (F*P = 8 or
// 16); 16-byte modes need 16-byte aligned visibility arrays.
template <int PREC, int VMODE, bool BATCH>
__global__ void __launch_bounds__(kScatterThreads, 8)
k_scatter(GridGeom g, uint32_t n_all, uint32_t F, uint32_t P, const double *__restrict__ uu_all,
          const double *__restrict__ vv_all, const double *__restrict__ ww_all,
          const float *__restrict__ vr_all, const float *__restrict__ vi_all,
          const float *__restrict__ wt_all, uint32_t *__restrict__ cursor,
          typename RecOf<PREC>::type *__restrict__ rec, uint32_t *__restrict__ meta,
          const BatchJob *__restrict__ jobs, int njobs)
{
    const BlockRows br = block_rows(g, blockIdx.x * kRowsPerBlock, n_all, uu_all, vv_all, ww_all,
                                    vr_all, vi_all, wt_all, BATCH ? jobs : nullptr, njobs);
    const double *__restrict__ uu = br.uu;
    const double *__restrict__ vv = br.vv;
    const double *__restrict__ ww = br.ww;
    const float *__restrict__ vr = br.vr;
    const float *__restrict__ vi = br.vi;
    const float *__restrict__ wt = br.wt;
    const uint32_t n = br.n;
    __shared__ uint32_t hkey[kHashSlots], hcnt[kHashSlots], hbase[kHashSlots];
    __shared__ unsigned long long smax[kScatterThreads / 64];
    const uint32_t base = br.base;
    // The read path is chosen on the host, so each instantiation carries one.
    constexpr bool vec4 = VMODE == 1;
    const uint32_t FP = F * P;
    // Every coordinate of the thread's rows (and, with one value per row,
    // its visibility and weight) is requested before any is used, so the
    // loads of both rows overlap instead of each waiting on the arithmetic
    // of the one before.
    double ux[kRowsPerThread], vx[kRowsPerThread], wx[kRowsPerThread];
    float r1[kRowsPerThread], i1[kRowsPerThread], wt1[kRowsPerThread];
#pragma unroll
    for (int k = 0; k < kRowsPerThread; ++k) {
        const uint32_t i = base + k * kScatterThreads + threadIdx.x;
        ux[k] = vx[k] = wx[k] = 0.0;
        r1[k] = i1[k] = wt1[k] = 0.f;
        if (i < n) {
            ux[k] = as_global(uu)[i];
            vx[k] = as_global(vv)[i];
            wx[k] = as_global(ww)[i];
            if (VMODE == 0 && FP == 1u) {
                r1[k] = as_global(vr)[i];
                i1[k] = as_global(vi)[i];
                wt1[k] = as_global(wt)[i];
            }
        }
    }
    double w_re[kRowsPerThread], w_im[kRowsPerThread];
    uint32_t widx[4] = {0u, 0u, 0u, 0u};
    if constexpr (VMODE >= 2) row_weight_index<VMODE>(P, widx);
#pragma unroll
    for (int k = 0; k < kRowsPerThread; ++k) {
        w_re[k] = w_im[k] = 0.0;
        if constexpr (VMODE >= 2)  // whole wave, before any lane leaves
            wave_row_sums<VMODE>(base + k * kScatterThreads + threadIdx.x - (threadIdx.x & 63u),
                                 n, P, widx, vr, vi, wt, w_re[k], w_im[k]);
    }
    for (int i = threadIdx.x; i < kHashSlots; i += kScatterThreads) {
        hkey[i] = kEmpty;
        hcnt[i] = 0;
    }
    __syncthreads();
    double pu[kRowsPerThread], pv[kRowsPerThread], sre[kRowsPerThread], sim[kRowsPerThread];
    uint32_t item[kRowsPerThread][4];
    int nk[kRowsPerThread];
#pragma unroll
    for (int k = 0; k < kRowsPerThread; ++k) {
        const uint32_t i = base + k * kScatterThreads + threadIdx.x;
        nk[k] = 0;
        pu[k] = pv[k] = sre[k] = sim[k] = 0.0;
        Foot f;
        if (i >= n) continue;
        pu[k] = div_by(ux[k], g.dx, g.inv_dx);
        pv[k] = div_by(vx[k], g.dx, g.inv_dx);
        if (!footprint(g, pu[k], pv[k], div_by(wx[k], br.dw, br.inv_dw), f, br.planes,
                       br.plane0))
            continue;
        // Channel/polarization sum, w-stacking.cu:512-527 (weights indexed
        // row*P + p, shared across channels; float product first).
        const size_t vb = (size_t)i * FP;
        const float *wrow = wt + (size_t)i * P;
        double a_re = 0.0, a_im = 0.0;
        if constexpr (VMODE >= 2) {
            a_re = w_re[k];
            a_im = w_im[k];
        } else if (vec4) {
            // Same sum in the same element order, read 16 bytes at a time.
            const float4 *r4 = reinterpret_cast<const float4 *>(vr + vb);
            const float4 *i4 = reinterpret_cast<const float4 *>(vi + vb);
            for (uint32_t e4 = 0; e4 < FP / 4; ++e4) {
                const float4 ra = load_stream4(&r4[e4]), ia = load_stream4(&i4[e4]);
                const float rr[4] = {ra.x, ra.y, ra.z, ra.w}, ii[4] = {ia.x, ia.y, ia.z, ia.w};
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    if (!isnan(rr[c])) {
                        const float wv = wrow[P == 1 ? 0u : (4 * e4 + c) % P];
                        const float wr = wv * rr[c];
                        const float wi = wv * ii[c];
                        a_re += (double)wr;
                        a_im += (double)wi;
                    }
                }
            }
        } else if (FP == 1u) {
            if (!isnan(r1[k])) {
                const float wr = wt1[k] * r1[k];
                const float wi = wt1[k] * i1[k];
                a_re += (double)wr;
                a_im += (double)wi;
            }
        } else {
            for (uint32_t fr = 0; fr < F; ++fr) {
                for (uint32_t p = 0; p < P; ++p) {
                    const size_t q = vb + (size_t)fr * P + p;
                    const float a = vr[q];
                    if (!isnan(a)) {
                        const float wr = wrow[p] * a;
                        const float wi = wrow[p] * vi[q];
                        a_re += (double)wr;
                        a_im += (double)wi;
                    }
                }
            }
        }
        sre[k] = a_re;
        sim[k] = a_im;
        uint32_t keys[4];
        nk[k] = row_keys(g, f, keys);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (q < nk[k]) {
                const uint32_t slot = hinsert_bounded<kHashBits>(hkey, keys[q]);
                if (slot == kEmpty) {
                    item[k][q] = kDirect | keys[q];  // keys < 2^31
                } else {
                    const uint32_t rank = atomicAdd(&hcnt[slot], 1u);
                    item[k][q] = slot | (rank << kHashBits);
                }
            }
        }
    }
    if constexpr (PREC == RICK_FP32) {
        // Largest |S| (as stored, FP32) of the block's rows: bit patterns of
        // non-negative values order like the values, NaN above Inf.
        unsigned long long m = 0ull;
#pragma unroll
        for (int k = 0; k < kRowsPerThread; ++k) {
            if (!nk[k]) continue;
            m = max(m, (unsigned long long)__double_as_longlong(fabs((double)(float)sre[k])));
            m = max(m, (unsigned long long)__double_as_longlong(fabs((double)(float)sim[k])));
        }
        for (int o = 32; o > 0; o >>= 1) m = max(m, (unsigned long long)__shfl_xor(m, o));
        if ((threadIdx.x & 63) == 0) smax[threadIdx.x >> 6] = m;
    }
    __syncthreads();
    if constexpr (PREC == RICK_FP32) {
        if (threadIdx.x == 0) {
            unsigned long long m = 0ull;
            for (int w = 0; w < kScatterThreads / 64; ++w) m = max(m, smax[w]);
            if (m) atomicMax(reinterpret_cast<unsigned long long *>(meta + 6), m);
        }
    }
    // One returning global atomic per distinct key.  A thread's atomics are
    // all issued before any result is stored, so their round trips overlap
    // (in a rolled loop each waits for the one before).
    constexpr int kSlotsPerThread = kHashSlots / kScatterThreads;
    static_assert(kHashSlots % kScatterThreads == 0, "hash slots per thread");
    uint32_t got[kSlotsPerThread];
#pragma unroll
    for (int t = 0; t < kSlotsPerThread; ++t) {
        const int i = threadIdx.x + t * kScatterThreads;
        const uint32_t c = hcnt[i];
        got[t] = c ? atomicAdd(&cursor[hkey[i]], c) : 0u;
    }
#pragma unroll
    for (int t = 0; t < kSlotsPerThread; ++t) hbase[threadIdx.x + t * kScatterThreads] = got[t];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kRowsPerThread; ++k) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (q < nk[k]) {
                const uint32_t it = item[k][q];
                const uint32_t pos = (it & kDirect)
                                         ? atomicAdd(&cursor[it & ~kDirect], 1u)
                                         : hbase[it & (kHashSlots - 1)] + (it >> kHashBits);
                typename RecOf<PREC>::type r;
                r.pu = pu[k];
                r.pv = pv[k];
                if constexpr (PREC == RICK_FP64) {
                    r.sre = sre[k];
                    r.sim = sim[k];
                } else {
                    r.sre = (float)sre[k];
                    r.sim = (float)sim[k];
                }
                rec[pos] = r;
            }
        }
    }
}

// Clears the tiles of buckets that several tasks accumulate into atomically.
template <typename T>
__global__ void __launch_bounds__(kThreads)
k_zero_split(GridGeom g, const uint32_t *__restrict__ split_list, const uint32_t *__restrict__ meta,
             T *__restrict__ grid)
{
    if (blockIdx.x >= meta[2]) return;
    const uint32_t b = split_list[blockIdx.x];
    const uint32_t per_plane = (uint32_t)g.tiles_u * (uint32_t)g.tiles_v;
    const int w = (int)(b / per_plane), rem = (int)(b % per_plane);
    const int tv = rem / g.tiles_u, tu = rem % g.tiles_u;
    const int j0 = tu * g.tu, k0 = tv * g.tv;
    const int nj = min(g.tu, g.nx - j0), nk = min(g.tv, g.ny - k0);
    for (int idx = threadIdx.x; idx < nk * g.tu; idx += kThreads) {
        const int kk = idx / g.tu, jj = idx % g.tu;
        if (jj >= nj) continue;
        const size_t c = ((size_t)w * g.ny + (size_t)(k0 + kk)) * (size_t)g.nx + (size_t)(j0 + jj);
        grid[2 * c] = T(0);
        grid[2 * c + 1] = T(0);
    }
}


// FP32 split buckets (several tasks add into one tile) accumulate in one
// 64-bit fixed-point tile per bucket with integer atomics, so the sum does not
// depend on which records a task holds or the order tasks finish in; the last
// task of the bucket rounds the tile to FP32 once.  Every task must use the
// same scale, from a bound all of them form alike: (records in the bucket) x
// (largest |S| of the whole call, an integer max over the bit patterns of
// non-negative values, taken by k_scatter) x (largest cell weight), mapped
// below 2^50.  Returns -1 when the bound is not finite (NaN/Inf data): the
// bucket's tasks then add FP32 atomically into the (cleared) grid tile.
__device__ __forceinline__ double split_scale_of(const GridGeom &g, unsigned long long maxbits,
                                                 uint32_t count)
{
    const double m = __longlong_as_double((long long)maxbits);
    const double bound = (double)count * m * g.cmax * 1.01;   // margin for FP32 products
    if (!(isfinite(m) && isfinite(bound))) return -1.0;
    if (!(bound > 0.0)) return 1.0;
    int e;
    frexp(bound, &e);                                          // bound < 2^e
    return ldexp(1.0, 50 - e);
}

// 2^52 + 2^51: adding it to x with |x| < 2^51 leaves round(x) in the low
// mantissa bits, so bits(x + M) - bits(M) is round(x) as a two's-complement
// integer (one FMA and one integer subtraction per value).
__device__ __forceinline__ unsigned long long to_fixed(double x, double scale)
{
    constexpr double M = 6755399441055744.0;
    return (unsigned long long)(__double_as_longlong(__fma_rn(x, scale, M)) -
                                __double_as_longlong(M));
}

// Per-task LDS tile accumulation.  Each lane owns one record at a time: it
// forms the S u-weights K[jKer] and, per footprint row, the v-weight K[kKer]
// (the reference's index arithmetic, evaluated 2S times per record instead
// of twice per cell), then adds S*(K[jKer]*K[kKer]) to the S x S cells with
// non-returning LDS atomics, so the 64 lanes of a wave stream independent
// updates.  The FP64 path accumulates in an FP64 LDS tile (ds_add_f64).  The
// FP32 path accumulates 64-bit fixed point (ds_add_u64, about twice the
// throughput of ds_add_f64 here, and order-independent): the task first sums
// |S| over its records, which with the largest cell weight bounds every
// cell's sum by B; values are scaled by a power of two that maps B to 2^50,
// so each contribution is rounded to B * 2^-50 (far below FP32 precision of
// any cell that matters to the image) and the tile is rounded to FP32 once
// when written.  A non-finite bound (NaN/Inf data) falls back to FP64
// atomics.  SMAX >= support bounds the unrolled footprint loops.
// Tiles whose LDS accumulators leave room for at most two workgroups per CU
// run persistent workgroups: the task loop costs about 20 VGPRs, which there
// costs no occupancy, and saves the per-workgroup launch cost of ~10^4 tasks.
// Smaller tiles keep one workgroup per task.
__host__ __device__ constexpr bool tile_persistent(int tu, int tv)
{
    return 3 * (2 * tv * (tu + kPadU) * (int)sizeof(double)) > kLdsBytes;
}

template <int PREC, int TU, int TV, int SMAX, bool HIPREC>
__global__ void __launch_bounds__(kTileThreads)
k_tile(GridGeom g, const uint4 *__restrict__ tasks, uint32_t *__restrict__ meta,
       const typename RecOf<PREC>::type *__restrict__ rec, const double *__restrict__ table,
       void *grid_v, int accumulate, const uint32_t *__restrict__ offsets,
       uint32_t *__restrict__ split_ticket, unsigned long long *__restrict__ split_acc,
       unsigned char *__restrict__ rowlive)
{
    using T = typename std::conditional<PREC == RICK_FP64, double, float>::type;
    using T2 = typename std::conditional<PREC == RICK_FP64, double2, float2>::type;
    constexpr int LDU = TU + kPadU;
    __shared__ double lds[2 * TV * LDU];
    __shared__ T stab[kMaxTable];
    double *sre = lds, *sim = lds + TV * LDU;

    // Workgroup k runs slot k; persistent workgroups (tile_persistent) then
    // take slots gridDim.x + meta[5]++ until all are done.  Slot i is
    // accumulation task f(i) when f(i+1) > f(i), where f(i) = floor(i * Tr / T),
    // else zero task Tr + i - f(i): the two kinds are spread evenly so that
    // clearing overlaps accumulation.
    const uint32_t ntask = meta[1], nreal = meta[4];
    __shared__ uint32_t slot_sh, last_sh;
    uint32_t slot = blockIdx.x;
    for (bool first = true;; first = false) {
        if (!first) {
            if (!tile_persistent(TU, TV)) break;
            // The barrier before the store also orders the previous task's LDS
            // reads before the next task clears the tile.
            __syncthreads();
            if (threadIdx.x == 0) slot_sh = gridDim.x + atomicAdd(&meta[5], 1u);
            __syncthreads();
            // Uniform by construction; readfirstlane keeps the task fields and
            // everything derived from them in scalar registers.
            slot = __builtin_amdgcn_readfirstlane(slot_sh);
        }
        if (slot >= ntask) break;
        const uint32_t fi = (uint32_t)((uint64_t)slot * nreal / ntask);
        const uint32_t fn = (uint32_t)((uint64_t)(slot + 1) * nreal / ntask);
        const uint4 task = tasks[fn > fi ? fi : nreal + slot - fi];
        const uint32_t b = task.x, r0 = task.y, r1 = task.z;
        const bool split = task.w == 1u || task.w >= kSplitTask;
        // FP32 split bucket with a finite bound: the bucket-wide scale, and the
        // tile goes to its integer accumulator (k_split_prep / k_split_finish).
        const uint32_t slot = task.w >= kSplitTask ? task.w - kSplitTask : 0u;
        const double bucket_scale =
            PREC == RICK_FP32 && task.w >= kSplitTask
                ? split_scale_of(g,
                                 *reinterpret_cast<const unsigned long long *>(meta + 6),
                                 offsets[b + 1] - offsets[b])
                : 0.0;
        const bool det = bucket_scale > 0.0;

        const uint32_t per_plane = (uint32_t)g.tiles_u * (uint32_t)g.tiles_v;
        const int w = (int)(b / per_plane), rem = (int)(b % per_plane);
        const int tv = rem / g.tiles_u, tu = rem % g.tiles_u;
        const int j0 = tu * TU, k0 = tv * TV;
        const int nj = min(TU, g.nx - j0), nk = min(TV, g.ny - k0);
        T *grid = static_cast<T *>(grid_v);

        if (task.w == kZeroTask) {
            // A run of task.y empty buckets from b, overwrite mode: clear their tiles.
            T2 z;
            z.x = T(0);
            z.y = T(0);
            T2 *g2 = reinterpret_cast<T2 *>(grid);
            // FP32 cells are 8 bytes: clear two per 16-byte store where the grid
            // allows (16-byte aligned base, even row length and tile width).
            const bool pairs = PREC == RICK_FP32 && (g.nx & 1) == 0 &&
                               (reinterpret_cast<uintptr_t>(grid) & 15u) == 0u;
            for (uint32_t e = 0; e < r0; ++e) {
                const uint32_t be = b + e;
                const int we = (int)(be / per_plane), re = (int)(be % per_plane);
                const int je = (re % g.tiles_u) * TU, ke = (re / g.tiles_u) * TV;
                const int nje = min(TU, g.nx - je), nke = min(TV, g.ny - ke);
                T2 *t0 = g2 + ((size_t)we * g.ny + (size_t)ke) * (size_t)g.nx + (size_t)je;
                if (pairs && (nje & 1) == 0) {
                    float4 *t4 = reinterpret_cast<float4 *>(t0);
                    const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
                    constexpr int HU = TU / 2;
                    for (int idx = threadIdx.x; idx < nke * HU; idx += kTileThreads) {
                        const int kk = idx / HU, jj = idx % HU;
                        if (2 * jj < nje) store_stream4(z4, &t4[(size_t)kk * (size_t)(g.nx / 2) + jj]);
                    }
                    continue;
                }
                for (int idx = threadIdx.x; idx < nke * TU; idx += kTileThreads) {
                    const int kk = idx / TU, jj = idx % TU;
                    if (jj < nje) store_stream2(z, &t0[(size_t)kk * g.nx + jj]);
                }
            }
            continue;
        }

        for (int idx = threadIdx.x; idx < 2 * TV * LDU; idx += kTileThreads) lds[idx] = 0.0;
        for (int idx = threadIdx.x; idx < g.ntable; idx += kTileThreads) stab[idx] = (T)table[idx];
        // FP32: fixed-point scale from the bound of the task's cell sums.
        bool fixed = false;
        double scale = 1.0;
        if (PREC == RICK_FP32 && det) {
            fixed = true;
            scale = bucket_scale;
        } else if constexpr (PREC == RICK_FP32) {
            __shared__ double red[2][kTileThreads / 64];
            double ar = 0.0, ai = 0.0;
            for (uint32_t r = r0 + threadIdx.x; r < r1; r += kTileThreads) {
                ar += fabs((double)rec[r].sre);
                ai += fabs((double)rec[r].sim);
            }
            for (int o = 32; o > 0; o >>= 1) {
                ar += __shfl_xor(ar, o);
                ai += __shfl_xor(ai, o);
            }
            if ((threadIdx.x & 63) == 0) {
                red[0][threadIdx.x >> 6] = ar;
                red[1][threadIdx.x >> 6] = ai;
            }
            __syncthreads();
            double br = 0.0, bi = 0.0;
    #pragma unroll
            for (int q = 0; q < kTileThreads / 64; ++q) {
                br += red[0][q];
                bi += red[1][q];
            }
            const double bound = fmax(br, bi) * g.cmax * 1.01;  // margin for FP32 products
            fixed = isfinite(br) && isfinite(bi) && isfinite(bound);  // (fmax drops NaNs)
            if (fixed && bound > 0.0) {
                int e;
                frexp(bound, &e);                    // bound < 2^e
                scale = ldexp(1.0, 50 - e);          // bound * scale < 2^50
            }
        }
        __syncthreads();

        const int S = g.support, KL = g.kl;
        // Per-cell Gaussian (GAUSS_HI_PRECISION) or table product: a template
        // parameter, so each instantiation carries only its own cell arithmetic.
        constexpr bool hiprec = HIPREC;
        // gauss_kernel_norm constants (w-stacking.cu:254-256).
        const double std22 = 1.0 / (2.0 * 1.0 * 1.0);
        const double gnorm = std22 / kRefPi;
        // Cells of this tile a footprint may touch, intersected with the grid.
        const int jlo = j0, jhi = j0 + nj - 1, klo = k0, khi = k0 + nk - 1;

        // Record r0 + c + 8*lane + wave of each chunk c of 512: consecutive records
        // (often one baseline at consecutive times, i.e. the same cells) go to
        // different waves, so the lanes of one atomic instruction rarely collide.
        const uint32_t roff = (threadIdx.x & 63u) * (kTileThreads / 64) + (threadIdx.x >> 6);
        // The FP32 fixed-point and the FP64-atomic accumulation are separate
        // loops, so the hot one carries no code of the other.
        auto accumulate_all = [&](auto fixed_tag) {
            constexpr bool FIXED = decltype(fixed_tag)::value;
            for (uint32_t c0 = r0; c0 < r1; c0 += kTileThreads) {
                const uint32_t r = c0 + roff;
                if (r >= r1) continue;
                const auto rc = rec[r];
                const double pu = rc.pu, pv = rc.pv;
                const T s_re = rc.sre, s_im = rc.sim;
                Foot f;
                uv_range(g, pu, pv, f);
                const int ja = max(f.jmin, jlo), jb = min(f.jmax, jhi);
                const int ka = max(f.kmin, klo), kb = min(f.kmax, khi);
                const int ju0 = f.gu - KL, kv0 = f.gv - KL;
                // u-weights of the footprint columns (w-stacking.cu:489, 492).
                T wu[SMAX];
                double ud[SMAX];
                uint32_t umask = 0;
        #pragma unroll
                for (int dj = 0; dj < SMAX; ++dj) {
                    const int j = ju0 + dj;
                    wu[dj] = T(0);
                    ud[dj] = 0.0;
                    if (dj < S && j >= ja && j <= jb) {
                        const double u_dist = (double)j + 0.5 - pu;
                        ud[dj] = u_dist;
                        const int jKer = (int)(kOversample * (fabs(u_dist + (double)KL)));
                        if (hiprec) {
                            umask |= 1u << dj;
                        } else if (jKer >= 0 && jKer < g.ntable) {
                            wu[dj] = stab[jKer];
                            umask |= 1u << dj;
                        }
                    }
                }
                if (umask == 0) continue;
        #pragma unroll
                for (int dk = 0; dk < SMAX; ++dk) {
                    const int k = kv0 + dk;
                    if (dk >= S || k < ka || k > kb) continue;
                    // w-stacking.cu:484, 493
                    const double v_dist = (double)k + 0.5 - pv;
                    T wv = T(0);
                    if (!hiprec) {
                        const int kKer = (int)(kOversample * (fabs(v_dist + (double)KL)));
                        if (kKer < 0 || kKer >= g.ntable) continue;
                        wv = stab[kKer];
                    }
                    // LDS index of footprint column 0 on this row; only columns with
                    // their umask bit set (inside the tile) are ever accessed.
                    const int rowbase = (k - k0) * LDU + (ju0 - j0);
        #pragma unroll
                    for (int dj = 0; dj < SMAX; ++dj) {
                        if (!((umask >> dj) & 1u)) continue;
                        // conv_weight = K[jKer] * K[kKer] (w-stacking.cu:499/505), or
                        // gauss_kernel_norm per cell (w-stacking.cu:496).
                        const T c =
                            hiprec ? (T)(gnorm * exp(-((ud[dj] * ud[dj]) + (v_dist * v_dist)) *
                                                     std22))
                                   : wu[dj] * wv;
                        const T are = s_re * c, aim = s_im * c;
                        if constexpr (FIXED) {
                            __hip_atomic_fetch_add(
                                reinterpret_cast<unsigned long long *>(&sre[rowbase + dj]),
                                to_fixed((double)are, scale), __ATOMIC_RELAXED,
                                __HIP_MEMORY_SCOPE_WORKGROUP);
                            __hip_atomic_fetch_add(
                                reinterpret_cast<unsigned long long *>(&sim[rowbase + dj]),
                                to_fixed((double)aim, scale), __ATOMIC_RELAXED,
                                __HIP_MEMORY_SCOPE_WORKGROUP);
                        } else {
                            __hip_atomic_fetch_add(&sre[rowbase + dj], (double)are,
                                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                            __hip_atomic_fetch_add(&sim[rowbase + dj], (double)aim,
                                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                        }
                    }
                }
            }
        };
        if constexpr (PREC == RICK_FP32) {
            if (fixed) accumulate_all(std::true_type{});
            else accumulate_all(std::false_type{});
        } else {
            accumulate_all(std::false_type{});
        }
        __syncthreads();

        if (PREC == RICK_FP32 && det) {
            unsigned long long *acc = split_acc + (size_t)slot * 2 * (TU * TV);
            unsigned long long done = 0ull;
            for (int idx = threadIdx.x; idx < nk * TU; idx += kTileThreads) {
                const int kk = idx / TU, jj = idx % TU;
                if (jj >= nj) continue;
                const int li = kk * LDU + jj;
                const unsigned long long a = reinterpret_cast<const unsigned long long *>(sre)[li];
                const unsigned long long b2 = reinterpret_cast<const unsigned long long *>(sim)[li];
                // Returning atomics: each has been performed (at the device's
                // coherence point) once its old value is back, so all of them
                // are in before the ticket below is taken -- no cache flush.
                if (a) done ^= atomicAdd(&acc[idx], a);
                if (b2) done ^= atomicAdd(&acc[TU * TV + idx], b2);
            }
            // The bucket's last task to get here writes the tile and clears
            // the accumulator and ticket for the next call.
            if (done == 0x5a5a5a5a5a5a5a5aull) last_sh = 2u;   // keeps the old values live
            __syncthreads();
            if (threadIdx.x == 0)
                last_sh = atomicAdd(&split_ticket[slot], 1u) ==
                          ntasks_of(offsets[b + 1] - offsets[b], g.rmax) - 1u;
            __syncthreads();
            if (last_sh) {
                const double inv = 1.0 / bucket_scale;   // exact: a power of two
                T2 *g2 = reinterpret_cast<T2 *>(grid);
                for (int idx = threadIdx.x; idx < nk * TU; idx += kTileThreads) {
                    const int kk = idx / TU, jj = idx % TU;
                    if (jj >= nj) continue;
                    const long long a = (long long)__hip_atomic_exchange(
                        &acc[idx], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    const long long b2 = (long long)__hip_atomic_exchange(
                        &acc[TU * TV + idx], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    T2 v;
                    v.x = (T)((double)a * inv);
                    v.y = (T)((double)b2 * inv);
                    const size_t c =
                        ((size_t)w * g.ny + (size_t)(k0 + kk)) * (size_t)g.nx + (size_t)(j0 + jj);
                    if (accumulate) {
                        const T2 o = g2[c];
                        v.x = o.x + v.x;
                        v.y = o.y + v.y;
                    }
                    store_stream2(v, &g2[c]);
                    if (rowlive && (v.x != T(0) || v.y != T(0)))
                        rowlive[(size_t)w * g.ny + (size_t)(k0 + kk)] = 1;
                }
                if (threadIdx.x == 0)
                    __hip_atomic_store(&split_ticket[slot], 0u, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
            }
            continue;
        }
        for (int idx = threadIdx.x; idx < nk * TU; idx += kTileThreads) {
            const int kk = idx / TU, jj = idx % TU;
            if (jj >= nj) continue;
            const int li = kk * LDU + jj;
            T vre, vim;
            if (PREC == RICK_FP32 && fixed) {
                const double inv = 1.0 / scale;  // exact: a power of two
                vre = (T)((double)reinterpret_cast<const long long *>(sre)[li] * inv);
                vim = (T)((double)reinterpret_cast<const long long *>(sim)[li] * inv);
            } else {
                vre = (T)sre[li];
                vim = (T)sim[li];
            }
            const size_t c =
                ((size_t)w * g.ny + (size_t)(k0 + kk)) * (size_t)g.nx + (size_t)(j0 + jj);
            if (rowlive && (vre != T(0) || vim != T(0)))
                rowlive[(size_t)w * g.ny + (size_t)(k0 + kk)] = 1;
            if (split) {
                if (vre != T(0)) unsafeAtomicAdd(&grid[2 * c], vre);
                if (vim != T(0)) unsafeAtomicAdd(&grid[2 * c + 1], vim);
            } else {
                T2 *g2 = reinterpret_cast<T2 *>(grid) + c;
                T2 v;
                v.x = vre;
                v.y = vim;
                if (accumulate) {
                    const T2 o = *g2;
                    v.x = o.x + v.x;
                    v.y = o.y + v.y;
                }
                store_stream2(v, g2);
            }
        }
    }
}

// Tile grid: one workgroup per task slot, or, for persistent tiles in a
// launch of at least kPersistSlots task slots per resident workgroup, as many
// workgroups as the device holds resident at once (occupancy x CUs, cached
// per device and kernel).  Smaller launches keep one workgroup per slot: they
// often run beside other jobs on other streams, which persistent workgroups
// would hold off the CUs.
template <int PREC, int TU, int TV, int SMAX, bool HIPREC>
uint32_t tile_grid(uint32_t task_cap)
{
    if (!tile_persistent(TU, TV)) return task_cap;
    const void *kernel = reinterpret_cast<const void *>(&k_tile<PREC, TU, TV, SMAX, HIPREC>);
    static std::atomic<uint32_t> resident[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return task_cap;
    uint32_t r = resident[dev].load(std::memory_order_relaxed);
    if (r == 0) {
        int per_cu = 0, cus = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kTileThreads, 0) !=
                hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
                hipSuccess ||
            per_cu <= 0 || cus <= 0)
            return task_cap;
        r = (uint32_t)per_cu * (uint32_t)cus;
        resident[dev].store(r, std::memory_order_relaxed);
    }
    return task_cap / kPersistSlots >= r ? r : task_cap;
}

template <int PREC, int SMAX>
hipError_t launch_tile_s(const GridGeom &g, const GridWork &w, const double *table, void *grid,
                         int accumulate, unsigned char *rowlive, hipStream_t s, Timing *tm)
{
    using Rec = typename RecOf<PREC>::type;
    const Rec *rec = static_cast<const Rec *>(w.records);
    const dim3 blk(kTileThreads);
#define RICK_TILE_LAUNCH(A, B, HI)                                                          \
    hipLaunchKernelGGL((k_tile<PREC, A, B, SMAX, HI>),                                      \
                       dim3(tile_grid<PREC, A, B, SMAX, HI>(w.task_cap)), blk, 0, s,        \
                       g, w.tasks, w.meta, rec, table, grid, accumulate, w.offsets,        \
                       w.split_ticket, w.split_acc, rowlive)
#define RICK_TILE_CASE(A, B)                                                                \
    if (g.tu == A && g.tv == B) {                                                           \
        tm->begin(RICK_STAGE_TILE, s);                                                      \
        if (g.kind == RICK_GAUSS_HI_PRECISION)                                              \
            RICK_TILE_LAUNCH(A, B, true);                                                   \
        else                                                                                \
            RICK_TILE_LAUNCH(A, B, false);                                                  \
        tm->end(RICK_STAGE_TILE, s);                                                        \
        return hipGetLastError();                                                           \
    }
    RICK_TILE_CASE(64, 64)
    RICK_TILE_CASE(64, 32)
    RICK_TILE_CASE(32, 32)
    RICK_TILE_CASE(128, 32)
    RICK_TILE_CASE(32, 64)
#undef RICK_TILE_CASE
#undef RICK_TILE_LAUNCH
    return hipErrorInvalidValue;
}

// Footprint loops are unrolled to g.smax: 7, 9, 11 or 15 (>= the support;
// the plan picks the smallest unless told otherwise).
template <int PREC>
hipError_t launch_tile(const GridGeom &g, const GridWork &w, const double *table, void *grid,
                       int accumulate, unsigned char *rowlive, hipStream_t s, Timing *tm)
{
    switch (g.smax) {
        case 7: return launch_tile_s<PREC, 7>(g, w, table, grid, accumulate, rowlive, s, tm);
        case 9: return launch_tile_s<PREC, 9>(g, w, table, grid, accumulate, rowlive, s, tm);
        case 11: return launch_tile_s<PREC, 11>(g, w, table, grid, accumulate, rowlive, s, tm);
        case 15: return launch_tile_s<PREC, 15>(g, w, table, grid, accumulate, rowlive, s, tm);
        default: return hipErrorInvalidValue;
    }
}

inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Capacities: a footprint (support <= tile size) touches at most 2 tiles per
// axis, so a row lands in at most 4 buckets; a split bucket holds more than
// rmax records, so there are at most 4*rows/(rmax+1) of them; tasks number at
// most nb + 4*rows/rmax.
void capacities(const GridGeom &g, uint32_t max_points, uint32_t *rec_cap, uint32_t *split_cap,
                uint32_t *task_cap, uint32_t *nscan)
{
    const uint64_t recs = 4ull * (uint64_t)max_points;
    *rec_cap = (uint32_t)(recs > 0 ? recs : 1);
    *split_cap = (uint32_t)(recs / ((uint64_t)g.rmax + 1) + 1);
    *task_cap = (uint32_t)((uint64_t)g.nb + recs / g.rmax + 1);
    *nscan = (g.nb + kScanBlock - 1) / kScanBlock;
}

}  // namespace

uint32_t batch_row_align() { return (uint32_t)kBinRows; }

const char *grid_launch_limit(const GridGeom &g, uint32_t max_points)
{
    uint32_t rec_cap, split_cap, task_cap, nscan;
    capacities(g, max_points, &rec_cap, &split_cap, &task_cap, &nscan);
    // A launch holds fewer than 2^32 threads.  Persistent tile launches are
    // capped at the resident workgroups; the others take one workgroup per
    // task slot, and k_zero_split one per split-bucket slot.
    constexpr uint64_t kMaxThreads = 1ull << 32;
    if (!tile_persistent(g.tu, g.tv) && (uint64_t)task_cap * kTileThreads >= kMaxThreads)
        return "too many gridding tasks for one tile launch (fewer tiles or rows per "
               "task needed)";
    if ((uint64_t)split_cap * kThreads >= kMaxThreads)
        return "too many split buckets for one launch (rows_per_task too small)";
    return nullptr;
}

size_t grid_work_bytes(const GridGeom &g, uint32_t max_points, int precision)
{
    uint32_t rec_cap, split_cap, task_cap, nscan;
    capacities(g, max_points, &rec_cap, &split_cap, &task_cap, &nscan);
    const size_t rs = precision == RICK_FP64 ? sizeof(RecF64) : sizeof(RecF32);
    size_t b = 0;
    b += align_up(sizeof(uint32_t) * g.nb, 256);           // counts
    b += align_up((size_t)g.planes * g.ny, 256);           // rowlive (after the counts)
    b += align_up(sizeof(uint32_t) * g.nb, 256);           // cursor
    b += align_up(sizeof(uint32_t) * (g.nb + 1), 256);     // offsets
    b += align_up(sizeof(uint32_t) * kScanQ * nscan, 256); // blocksum
    b += align_up(sizeof(uint4) * task_cap, 256);          // tasks
    b += align_up(sizeof(uint32_t) * split_cap, 256);      // split_list
    b += align_up(sizeof(uint32_t) * 8, 256);              // meta
    b += align_up(rs * rec_cap, 256);                      // records
    if (precision == RICK_FP32) {
        b += align_up(sizeof(uint32_t) * split_cap, 256);                     // split_ticket
        b += align_up(sizeof(unsigned long long) * 2 * (size_t)g.tu * g.tv * split_cap, 256);
    }
    return b;
}

void grid_work_carve(const GridGeom &g, uint32_t max_points, int precision, void *base,
                     GridWork *w)
{
    uint32_t rec_cap, split_cap, task_cap, nscan;
    capacities(g, max_points, &rec_cap, &split_cap, &task_cap, &nscan);
    const size_t rs = precision == RICK_FP64 ? sizeof(RecF64) : sizeof(RecF32);
    char *p = static_cast<char *>(base);
    auto take = [&](size_t bytes) {
        char *q = p;
        p += align_up(bytes, 256);
        return q;
    };
    w->counts = reinterpret_cast<uint32_t *>(take(sizeof(uint32_t) * g.nb));
    w->rowlive = reinterpret_cast<unsigned char *>(take((size_t)g.planes * g.ny));
    w->cursor = reinterpret_cast<uint32_t *>(take(sizeof(uint32_t) * g.nb));
    w->offsets = reinterpret_cast<uint32_t *>(take(sizeof(uint32_t) * (g.nb + 1)));
    w->blocksum = reinterpret_cast<uint32_t *>(take(sizeof(uint32_t) * kScanQ * nscan));
    w->tasks = reinterpret_cast<uint4 *>(take(sizeof(uint4) * task_cap));
    w->split_list = reinterpret_cast<uint32_t *>(take(sizeof(uint32_t) * split_cap));
    w->meta = reinterpret_cast<uint32_t *>(take(sizeof(uint32_t) * 8));
    w->records = take(rs * rec_cap);
    w->split_ticket = nullptr;
    w->split_acc = nullptr;
    if (precision == RICK_FP32) {
        w->split_ticket = reinterpret_cast<uint32_t *>(take(sizeof(uint32_t) * split_cap));
        w->split_acc = reinterpret_cast<unsigned long long *>(
            take(sizeof(unsigned long long) * 2 * (size_t)g.tu * g.tv * split_cap));
    }
    w->rec_cap = rec_cap;
    w->split_cap = split_cap;
    w->task_cap = task_cap;
    w->nscan = nscan;
}

hipError_t launch_grid(const GridGeom &g, const GridWork &w, int precision,
                       const double *table_d, uint32_t n, uint32_t F, uint32_t P,
                       const double *uu, const double *vv, const double *ww, const float *vr,
                       const float *vi, const float *wt, void *grid, int accumulate,
                       hipStream_t s, Timing *tm, bool clear_empty, const BatchJob *jobs,
                       int njobs)
{
    Timing off;
    if (tm == nullptr) tm = &off;
    const uint32_t nblk = (n + kRowsPerBlock - 1) / kRowsPerBlock;
    const uint32_t cblk = (n + kBinRows - 1) / kBinRows;
    // The counts and the row flags that follow them are cleared together.
    const uint32_t fill_words =
        (uint32_t)((reinterpret_cast<const char *>(w.rowlive) - reinterpret_cast<const char *>(w.counts) +
                    align_up((size_t)g.planes * g.ny, 256)) / sizeof(uint32_t));
    const uint32_t fill_blocks = min((fill_words + 255u) / 256u, 4096u);
    // Overwrite of a dense grid: runs of empty buckets get zero tasks.
    const int zero = (!accumulate && clear_empty) ? 1 : 0;
    // Overwrite of a sparse grid: the tile kernel flags the rows it writes
    // non-zero cells to, for the image chain.
    unsigned char *rowlive = (!accumulate && !clear_empty) ? w.rowlive : nullptr;
    tm->begin(RICK_STAGE_COUNT_ROWS, s);
    hipLaunchKernelGGL(k_fill_u32, dim3(fill_blocks), dim3(256), 0, s, w.counts, fill_words, 0u);
    if (n > 0) {
        if (jobs)
            hipLaunchKernelGGL(k_count<true>, dim3(cblk), dim3(kBinThreads), 0, s, g, n, uu, vv,
                               ww, w.counts, jobs, njobs);
        else
            hipLaunchKernelGGL(k_count<false>, dim3(cblk), dim3(kBinThreads), 0, s, g, n, uu, vv,
                               ww, w.counts, jobs, njobs);
    }
    tm->end(RICK_STAGE_COUNT_ROWS, s);
    tm->begin(RICK_STAGE_SCAN, s);
    hipLaunchKernelGGL(k_scan_reduce, dim3(w.nscan), dim3(kThreads), 0, s, g, w.counts,
                       w.blocksum, w.nscan, zero);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kTopThreads), 0, s, g, w.blocksum, w.nscan,
                       w.offsets, w.meta);
    hipLaunchKernelGGL(k_scan_down, dim3(w.nscan), dim3(kThreads), 0, s, g, w.counts,
                       w.blocksum, w.nscan, w.offsets, w.cursor, w.tasks, w.task_cap,
                       w.split_list, w.split_cap, w.meta, zero);
    tm->end(RICK_STAGE_SCAN, s);
    tm->begin(RICK_STAGE_SCATTER, s);
    if (n > 0) {
        const uint32_t FP = F * P;
        const bool vec4 = (FP % 4u) == 0u &&
                          ((reinterpret_cast<uintptr_t>(vr) | reinterpret_cast<uintptr_t>(vi)) &
                           15u) == 0u;   // batch: vr = vi = null, the caller checked the jobs
        const int mode = !vec4 ? 0 : FP == 8u ? 2 : FP == 16u ? 4 : 1;
#define RICK_SCATTER(PR, VM)                                                                  \
    if (jobs)                                                                                 \
        hipLaunchKernelGGL((k_scatter<PR, VM, true>), dim3(nblk), dim3(kScatterThreads), 0, s, \
                           g, n, F, P, uu, vv, ww, vr, vi, wt, w.cursor,                      \
                           static_cast<typename RecOf<PR>::type *>(w.records), w.meta, jobs,  \
                           njobs);                                                            \
    else                                                                                      \
        hipLaunchKernelGGL((k_scatter<PR, VM, false>), dim3(nblk), dim3(kScatterThreads), 0,  \
                           s, g, n, F, P, uu, vv, ww, vr, vi, wt, w.cursor,                   \
                           static_cast<typename RecOf<PR>::type *>(w.records), w.meta, jobs,  \
                           njobs)
#define RICK_SCATTER_MODES(PR)                                                                \
    switch (mode) {                                                                           \
    case 4: RICK_SCATTER(PR, 4); break;                                                       \
    case 2: RICK_SCATTER(PR, 2); break;                                                       \
    case 1: RICK_SCATTER(PR, 1); break;                                                       \
    default: RICK_SCATTER(PR, 0); break;                                                      \
    }
        if (precision == RICK_FP64) {
            RICK_SCATTER_MODES(RICK_FP64)
        } else {
            RICK_SCATTER_MODES(RICK_FP32)
        }
#undef RICK_SCATTER_MODES
#undef RICK_SCATTER
    }
    tm->end(RICK_STAGE_SCATTER, s);
    if (precision == RICK_FP64) {
        if (!accumulate) {
            tm->begin(RICK_STAGE_ZERO_SPLIT, s);
            hipLaunchKernelGGL(k_zero_split<double>, dim3(w.split_cap), dim3(kThreads), 0, s, g,
                               w.split_list, w.meta, static_cast<double *>(grid));
            tm->end(RICK_STAGE_ZERO_SPLIT, s);
        }
        return launch_tile<RICK_FP64>(g, w, table_d, grid, accumulate, rowlive, s, tm);
    }
    // FP32: split buckets accumulate in integer fixed point (order-independent);
    // their grid tiles are cleared for the non-finite fallback.
    if (!accumulate) {
        tm->begin(RICK_STAGE_ZERO_SPLIT, s);
        hipLaunchKernelGGL(k_zero_split<float>, dim3(w.split_cap), dim3(kThreads), 0, s, g,
                           w.split_list, w.meta, static_cast<float *>(grid));
        tm->end(RICK_STAGE_ZERO_SPLIT, s);
    }
    return launch_tile<RICK_FP32>(g, w, table_d, grid, accumulate, rowlive, s, tm);
}

}  // namespace rick
