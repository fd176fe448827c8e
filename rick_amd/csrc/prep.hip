// Device-side data preparation in front of wstack() and behind the FFT:
//
//   rick_normalize_uvw   the NORMALIZE_UVW block of gridding.c:26-82
//   rick_sector_bin      initialize_array(), gridding.c:139-197: rows sorted
//                        into v-slab sectors (one per rank) with halo copies
//   rick_sector_gather   the per-sector input assembly of gridding_cpu.c:
//                        90-138 (rows in reverse list order, v shifted into
//                        the slab), i.e. what one rank passes to wstack()
//   launch_split_complex de-interleaving of a grid for the cube writers
//
// All of this is HBM-bound byte/integer work; the kernels are built for
// coalesced streaming and stable, bit-exact results.
//
// Extrema.  The reference folds min/max with MIN(a,b) = a<b?a:b and
// MAX(a,b) = a>b?a:b starting from 1e20 / 0.0 (gridding.c:37-49).  For inputs
// without NaN, the fold's result is the LAST element that compares equal to
// the extremum (or the start value), which fixes the sign of a zero extremum.
// The parallel reduction carries (value, index) pairs and resolves ties to the
// larger index, so it returns those bits in any reduction order.
#include <algorithm>
#include <cstdio>
#include <mutex>

#include "rick_internal.h"

#pragma clang fp contract(off)

namespace rick {
namespace {

constexpr int kPrepThreads = 256;
constexpr int kExtBlocks = 1024;        // blocks of the extrema pass (max)
constexpr int kSecRows = 4096;          // rows per sector-binning block
constexpr int kSecRowsPerWave = kSecRows / (kPrepThreads / 64);
constexpr int kMaxSlots = 4096;         // nsectors + 1 <= kMaxSlots
constexpr int kAggregateSectors = 4;    // up to this many: wave-aggregated counting
constexpr int kScanChunk = 4096;        // elements per scan block (16 per thread)
constexpr int kScanTopThreads = 1024;
constexpr int kMaxScanChunks = kScanTopThreads * 16;

// ---------------------------------------------------------------------------
// NORMALIZE_UVW.

__device__ inline void take_min(double &v, long long &i, double bv, long long bi)
{
    if (bv < v || (bv == v && bi > i)) {
        v = bv;
        i = bi;
    }
}
__device__ inline void take_max(double &v, long long &i, double bv, long long bi)
{
    if (bv > v || (bv == v && bi > i)) {
        v = bv;
        i = bi;
    }
}

struct Ext6 {
    double v[6];      // min u, v, w, max u, v, w
    long long i[6];
    __device__ void init()
    {
        for (int q = 0; q < 3; ++q) {
            v[q] = 1e20;      // gridding.c:37
            v[q + 3] = 0.0;   // gridding.c:38
            i[q] = i[q + 3] = -1;
        }
    }
    __device__ void merge(const double *bv, const long long *bi)
    {
        for (int q = 0; q < 3; ++q) take_min(v[q], i[q], bv[q], bi[q]);
        for (int q = 3; q < 6; ++q) take_max(v[q], i[q], bv[q], bi[q]);
    }
    // Reduces over the block; the result is valid in thread 0.
    __device__ void block_reduce()
    {
        for (int off = 32; off > 0; off >>= 1) {
            double bv[6];
            long long bi[6];
            for (int q = 0; q < 6; ++q) {
                bv[q] = __shfl_xor(v[q], off);
                bi[q] = __shfl_xor(i[q], off);
            }
            merge(bv, bi);
        }
        __shared__ double sv[kPrepThreads / 64][6];
        __shared__ long long si[kPrepThreads / 64][6];
        const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
        if (lane == 0)
            for (int q = 0; q < 6; ++q) {
                sv[wave][q] = v[q];
                si[wave][q] = i[q];
            }
        __syncthreads();
        if (threadIdx.x == 0)
            for (int w = 1; w < kPrepThreads / 64; ++w) merge(sv[w], si[w]);
    }
};

__global__ void __launch_bounds__(kPrepThreads)
k_uvw_extrema(uint32_t n, const double *__restrict__ uu, const double *__restrict__ vv,
              const double *__restrict__ ww, double *__restrict__ part_v,
              long long *__restrict__ part_i)
{
    Ext6 e;
    e.init();
    const size_t stride = (size_t)gridDim.x * kPrepThreads;
    for (size_t r = (size_t)blockIdx.x * kPrepThreads + threadIdx.x; r < n; r += stride) {
        // This thread visits rows in increasing order, so a tie always
        // goes to the row at hand.
        const double x[3] = {uu[r], vv[r], ww[r]};
        for (int q = 0; q < 3; ++q) {
            if (x[q] <= e.v[q]) {
                e.v[q] = x[q];
                e.i[q] = (long long)r;
            }
            if (x[q] >= e.v[q + 3]) {
                e.v[q + 3] = x[q];
                e.i[q + 3] = (long long)r;
            }
        }
    }
    e.block_reduce();
    if (threadIdx.x == 0)
        for (int q = 0; q < 6; ++q) {
            part_v[q * kExtBlocks + blockIdx.x] = e.v[q];
            part_i[q * kExtBlocks + blockIdx.x] = e.i[q];
        }
}

// fin[0..5] = min u, v, w, max u, v, w; fin[6] = maxg after the margin.
__global__ void __launch_bounds__(kPrepThreads)
k_uvw_finalize(int nblk, const double *__restrict__ part_v, const long long *__restrict__ part_i,
               double *__restrict__ fin, double *__restrict__ extrema)
{
    Ext6 e;
    e.init();
    for (int b = threadIdx.x; b < nblk; b += kPrepThreads) {
        double bv[6];
        long long bi[6];
        for (int q = 0; q < 6; ++q) {
            bv[q] = part_v[q * kExtBlocks + b];
            bi[q] = part_i[q * kExtBlocks + b];
        }
        e.merge(bv, bi);
    }
    e.block_reduce();
    if (threadIdx.x != 0) return;
    // gridding.c:75-79, MAX(a,b) = a>b?a:b.
    const double offset = 0.001;
    const double fu = fabs(e.v[0]), fv = fabs(e.v[1]);
    const double ming = fu > fv ? fu : fv;
    const double gu = fabs(e.v[3]), gv = fabs(e.v[4]);
    double maxg = gu > gv ? gu : gv;
    maxg = maxg > ming ? maxg : ming;
    maxg = maxg + offset * maxg;
    double out[8];
    for (int q = 0; q < 6; ++q) out[q] = e.v[q];
    out[6] = maxg;
    out[7] = 0.0;
    for (int q = 0; q < 8; ++q) {
        fin[q] = out[q];
        if (extrema) extrema[q] = out[q];
    }
}

// gridding.c:80-85.
__global__ void __launch_bounds__(kPrepThreads)
k_uvw_apply(uint32_t n, double *__restrict__ uu, double *__restrict__ vv, double *__restrict__ ww,
            const double *__restrict__ fin)
{
    const double minw = fin[2], maxw = fin[5], maxg = fin[6];
    const size_t stride = (size_t)gridDim.x * kPrepThreads;
    for (size_t r = (size_t)blockIdx.x * kPrepThreads + threadIdx.x; r < n; r += stride) {
        uu[r] = (uu[r] + maxg) / (2.0 * maxg);
        vv[r] = (vv[r] + maxg) / (2.0 * maxg);
        ww[r] = (ww[r] - minw) / (maxw - minw);
    }
}

// ---------------------------------------------------------------------------
// Sector binning.

struct SectorGeom {
    int nsectors;       // slabs (ranks); slots 0..nsectors
    int yaxis;          // slab height in cells
    double dx;
    double supporth;    // ((w_support-1)/2)*dx (init.c:13)
    uint32_t n;         // rows
    uint32_t nb;        // binning blocks
};

// Sectors of one row (gridding.c:156-167): its own slab binphi, the slab
// above when within supporth of the upper border, the slab below when
// within supporth of the lower one.  s[k] = -1 marks no entry.  Rows whose
// binphi falls outside 0..nsectors, and upper neighbours beyond slot
// nsectors, are entries the reference would store out of bounds: they are
// left out and counted in `bad`.
__device__ inline void row_sectors(const SectorGeom &g, double vvh, int s[3], uint32_t &bad)
{
    s[0] = s[1] = s[2] = -1;
    const double b = vvh * (double)g.nsectors;
    if (!(b > -1.0 && b < (double)g.nsectors + 1.0)) {  // also catches NaN
        ++bad;
        return;
    }
    const int binphi = (int)b;
    const double updist = (double)((binphi + 1) * g.yaxis) * g.dx - vvh;
    const double downdist = vvh - (double)(binphi * g.yaxis) * g.dx;
    s[0] = binphi;
    if (updist < g.supporth && updist >= 0.0) {
        if (binphi + 1 <= g.nsectors)
            s[1] = binphi + 1;
        else
            ++bad;
    }
    if (downdist < g.supporth && binphi > 0 && downdist >= 0.0) s[2] = binphi - 1;
}

// Walks the distinct sectors among the wave's entries (s[0..2] per lane,
// -1 = none), wave-uniformly: f(sec, mask, mine) runs once per sector with
// the ballot of the lanes holding an entry for it.  Each round retires at
// least the first pending entry of the lowest pending lane, so the loop ends.
template <typename Fn>
__device__ inline void for_each_sector(const int s[3], Fn &&f)
{
    int pending = (s[0] >= 0 ? 1 : 0) | (s[1] >= 0 ? 2 : 0) | (s[2] >= 0 ? 4 : 0);
    for (;;) {
        const unsigned long long act = __ballot(pending != 0);
        if (act == 0) break;
        const int leader = __ffsll((long long)act) - 1;
        const int first = (pending & 1) ? s[0] : (pending & 2) ? s[1] : s[2];
        const int sec = __builtin_amdgcn_readlane(first, leader);
        int hit = 0;
        for (int j = 0; j < 3; ++j)
            if ((pending >> j & 1) && s[j] == sec) hit = 1 << j;
        const unsigned long long m = __ballot(hit != 0);
        f(sec, m, hit != 0);
        pending &= ~hit;
    }
}

__global__ void __launch_bounds__(kPrepThreads)
k_sector_count(SectorGeom g, const double *__restrict__ vv, uint32_t *__restrict__ counts,
               uint32_t *__restrict__ bad_total)
{
    __shared__ uint32_t hist[kMaxSlots];
    __shared__ uint32_t bad_blk;
    const int slots = g.nsectors + 1;
    for (int s = threadIdx.x; s < slots; s += kPrepThreads) hist[s] = 0;
    if (threadIdx.x == 0) bad_blk = 0;
    __syncthreads();
    uint32_t bad = 0;
    const int lane = threadIdx.x & 63;
    const size_t r0 = (size_t)blockIdx.x * kSecRows;
    for (int k = 0; k < kSecRows / kPrepThreads; ++k) {
        const size_t r = r0 + (size_t)k * kPrepThreads + threadIdx.x;
        int s[3] = {-1, -1, -1};
        if (r < g.n) row_sectors(g, vv[r], s, bad);
        if (g.nsectors <= kAggregateSectors) {
            // Few sectors: most lanes share one; one LDS atomic per sector.
            for_each_sector(s, [&](int sec, unsigned long long m, bool) {
                if (lane == 0) atomicAdd(&hist[sec], (uint32_t)__popcll(m));
            });
        } else {
            for (int j = 0; j < 3; ++j)
                if (s[j] >= 0) atomicAdd(&hist[s[j]], 1u);
        }
    }
    if (bad) atomicAdd(&bad_blk, bad);
    __syncthreads();
    for (int s = threadIdx.x; s < slots; s += kPrepThreads)
        counts[(size_t)s * g.nb + blockIdx.x] = hist[s];
    if (threadIdx.x == 0 && bad_blk) atomicAdd(bad_total, bad_blk);
}

// Exclusive scan of len u32 values: per-chunk sums, one-block scan of the
// sums (writes the grand total to out[len]), per-chunk rescan with its base.
__device__ inline uint32_t wave_incl_scan(uint32_t x)
{
    const int lane = threadIdx.x & 63;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off);
        if (lane >= off) x += y;
    }
    return x;
}

template <int THREADS>
__device__ inline uint32_t block_excl_scan(uint32_t x, uint32_t *total)
{
    __shared__ uint32_t wsum[THREADS / 64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t incl = wave_incl_scan(x);
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    uint32_t base = 0, all = 0;
    for (int w = 0; w < THREADS / 64; ++w) {
        const uint32_t t = wsum[w];
        base += w < wave ? t : 0;
        all += t;
    }
    __syncthreads();
    if (total) *total = all;
    return base + incl - x;
}

__global__ void __launch_bounds__(kPrepThreads)
k_scan_sums(const uint32_t *__restrict__ in, size_t len, uint32_t *__restrict__ sums)
{
    const size_t base = (size_t)blockIdx.x * kScanChunk + (size_t)threadIdx.x * 16;
    uint32_t s = 0;
    for (int k = 0; k < 16; ++k)
        if (base + k < len) s += in[base + k];
    uint32_t tot;
    (void)block_excl_scan<kPrepThreads>(s, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(kScanTopThreads)
k_scan_top(uint32_t *__restrict__ sums, int nchunks, uint32_t *__restrict__ out_total)
{
    uint32_t v[16], s = 0;
    for (int k = 0; k < 16; ++k) {
        const int i = threadIdx.x * 16 + k;
        v[k] = i < nchunks ? sums[i] : 0;
        s += v[k];
    }
    uint32_t tot;
    uint32_t run = block_excl_scan<kScanTopThreads>(s, &tot);
    for (int k = 0; k < 16; ++k) {
        const int i = threadIdx.x * 16 + k;
        if (i < nchunks) sums[i] = run;
        run += v[k];
    }
    if (threadIdx.x == 0) *out_total = tot;
}

__global__ void __launch_bounds__(kPrepThreads)
k_scan_down(const uint32_t *__restrict__ in, size_t len, const uint32_t *__restrict__ sums,
            uint32_t *__restrict__ out)
{
    const size_t base = (size_t)blockIdx.x * kScanChunk + (size_t)threadIdx.x * 16;
    uint32_t v[16], s = 0;
    for (int k = 0; k < 16; ++k) {
        v[k] = base + k < len ? in[base + k] : 0;
        s += v[k];
    }
    uint32_t run = sums[blockIdx.x] + block_excl_scan<kPrepThreads>(s, nullptr);
    for (int k = 0; k < 16; ++k) {
        if (base + k < len) out[base + k] = run;
        run += v[k];
    }
}

// Stable fill: entries of sector s appear in increasing row order, as the
// reference's sequential second pass stores them (gridding.c:178-190).  Each
// wave owns kSecRowsPerWave consecutive rows of the block; its per-sector
// cursors start at the block's offset plus the entries of the waves before
// it, and within a 64-row chunk the lanes that share a sector take
// consecutive slots in lane order.
__global__ void __launch_bounds__(kPrepThreads)
k_sector_fill(SectorGeom g, const double *__restrict__ vv, const uint32_t *__restrict__ offsets,
              uint32_t *__restrict__ rows, uint32_t cap)
{
    extern __shared__ uint32_t cur[];  // [waves][slots]
    const int slots = g.nsectors + 1;
    const int waves = kPrepThreads / 64;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int i = threadIdx.x; i < waves * slots; i += kPrepThreads) cur[i] = 0;
    __syncthreads();
    const size_t r0 = (size_t)blockIdx.x * kSecRows + (size_t)wave * kSecRowsPerWave;
    constexpr int kChunks = kSecRowsPerWave / 64;
    double v[kChunks];   // this lane's rows, read once for both passes
    uint32_t bad = 0;
#pragma unroll
    for (int c = 0; c < kChunks; ++c) {
        const size_t r = r0 + c * 64 + lane;
        v[c] = r < g.n ? vv[r] : 0.0;
    }
    uint32_t *my = cur + wave * slots;
#pragma unroll
    for (int c = 0; c < kChunks; ++c) {
        const size_t r = r0 + c * 64 + lane;
        int s[3] = {-1, -1, -1};
        if (r < g.n) row_sectors(g, v[c], s, bad);
        if (g.nsectors <= kAggregateSectors) {
            for_each_sector(s, [&](int sec, unsigned long long m, bool) {
                if (lane == 0) my[sec] += (uint32_t)__popcll(m);
            });
        } else {
            for (int j = 0; j < 3; ++j)
                if (s[j] >= 0) atomicAdd(&my[s[j]], 1u);
        }
    }
    __syncthreads();
    for (int s = threadIdx.x; s < slots; s += kPrepThreads) {
        uint32_t base = offsets[(size_t)s * g.nb + blockIdx.x];
        for (int w = 0; w < waves; ++w) {
            const uint32_t t = cur[w * slots + s];
            cur[w * slots + s] = base;
            base += t;
        }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < kChunks; ++c) {
        const size_t r = r0 + c * 64 + lane;
        int s[3] = {-1, -1, -1};
        if (r < g.n) row_sectors(g, v[c], s, bad);
        // Lanes sharing a sector take consecutive slots in lane order.
        for_each_sector(s, [&](int sec, unsigned long long m, bool mine) {
            const uint32_t base = my[sec];
            if (mine) {
                const uint32_t rank = __builtin_amdgcn_mbcnt_hi(
                    (uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
                const uint32_t pos = base + rank;
                if (pos < cap) rows[pos] = (uint32_t)r;
            }
            if (lane == 0) my[sec] = base + (uint32_t)__popcll(m);
        });
    }
}

__global__ void k_sector_histo(int slots, uint32_t nb, const uint32_t *__restrict__ offsets,
                               uint32_t *__restrict__ histo, uint32_t *__restrict__ start)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= slots) return;
    const uint32_t a = offsets[(size_t)s * nb], b = offsets[(size_t)(s + 1) * nb];
    start[s] = a;
    histo[s] = b - a;
}

// ---------------------------------------------------------------------------
// Sector gather (gridding_cpu.c:118-138): output row icount takes list entry
// nsec-1-icount.

__global__ void __launch_bounds__(kPrepThreads)
k_gather_rows(const uint32_t *__restrict__ list, uint32_t nsec, uint32_t n, double vshift,
              uint32_t P, const double *__restrict__ uu, const double *__restrict__ vv,
              const double *__restrict__ ww, const float *__restrict__ wt,
              double *__restrict__ uus, double *__restrict__ vvs, double *__restrict__ wws,
              float *__restrict__ wts)
{
    const uint32_t ic = blockIdx.x * kPrepThreads + threadIdx.x;
    if (ic >= nsec) return;
    const uint32_t src = list[nsec - 1 - ic];
    if (src >= n) return;
    uus[ic] = uu[src];
    vvs[ic] = vv[src] - vshift;
    wws[ic] = ww[src];
    for (uint32_t p = 0; p < P; ++p) wts[(size_t)ic * P + p] = wt[(size_t)src * P + p];
}

__global__ void __launch_bounds__(kPrepThreads)
k_gather_vis(const uint32_t *__restrict__ list, uint32_t nsec, uint32_t n, uint32_t FP,
             const float *__restrict__ vr, const float *__restrict__ vi, float *__restrict__ vrs,
             float *__restrict__ vis)
{
    const size_t e = (size_t)blockIdx.x * kPrepThreads + threadIdx.x;
    if (e >= (size_t)nsec * FP) return;
    const uint32_t ic = (uint32_t)(e / FP), k = (uint32_t)(e - (size_t)ic * FP);
    const uint32_t src = list[nsec - 1 - ic];
    if (src >= n) return;
    const size_t from = (size_t)src * FP + k;
    vrs[e] = vr[from];
    vis[e] = vi[from];
}

// Same, four floats per thread when F*P is a multiple of 4 (16-byte aligned
// row segments).
__global__ void __launch_bounds__(kPrepThreads)
k_gather_vis4(const uint32_t *__restrict__ list, uint32_t nsec, uint32_t n, uint32_t FP4,
              const float4 *__restrict__ vr, const float4 *__restrict__ vi,
              float4 *__restrict__ vrs, float4 *__restrict__ vis)
{
    const size_t e = (size_t)blockIdx.x * kPrepThreads + threadIdx.x;
    if (e >= (size_t)nsec * FP4) return;
    const uint32_t ic = (uint32_t)(e / FP4), k = (uint32_t)(e - (size_t)ic * FP4);
    const uint32_t src = list[nsec - 1 - ic];
    if (src >= n) return;
    const size_t from = (size_t)src * FP4 + k;
    vrs[e] = vr[from];
    vis[e] = vi[from];
}

// ---------------------------------------------------------------------------
// Grid de-interleaving for the cube writers.

template <typename T>
__global__ void __launch_bounds__(kPrepThreads)
k_split(const T *__restrict__ g, size_t n, double *__restrict__ re, double *__restrict__ im)
{
    const size_t i = (size_t)blockIdx.x * kPrepThreads + threadIdx.x;
    if (i >= n) return;
    re[i] = (double)g[2 * i];
    im[i] = (double)g[2 * i + 1];
}

#define PREP_HIP(call)                                                                        \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return set_status(e_ == hipErrorOutOfMemory ? RICK_ERR_NOMEM : RICK_ERR_DEVICE,   \
                              "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),          \
                              __FILE__, __LINE__);                                            \
    } while (0)

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

struct SectorLayout {
    uint32_t nb;
    size_t len, nchunks;
    size_t counts, offsets, sums, bad, bytes;  // byte offsets into the workspace
};

bool sector_layout(uint32_t n, int nsectors, SectorLayout &L)
{
    L.nb = n ? (uint32_t)(((uint64_t)n + kSecRows - 1) / kSecRows) : 1;
    L.len = (size_t)(nsectors + 1) * L.nb;
    L.nchunks = (L.len + kScanChunk - 1) / kScanChunk;
    L.counts = 0;
    L.offsets = align256(L.counts + L.len * 4);
    L.sums = align256(L.offsets + (L.len + 1) * 4);
    L.bad = align256(L.sums + L.nchunks * 4);
    L.bytes = align256(L.bad + 4);
    return L.nchunks <= (size_t)kMaxScanChunks;
}

constexpr size_t kNormWorkBytes = (size_t)kExtBlocks * 6 * (sizeof(double) + sizeof(long long)) +
                                  8 * sizeof(double);

// Workspace of rick_sector_bin_noalloc, for callers that cannot allocate: the
// per-(slab, block) counts, their scan and the dropped-entry counter live in
// this array, one copy per device, held until the process ends.  Host code
// never passes the array's name to a kernel (on the host it is a stand-in,
// not a device address): the device address comes from hipGetSymbolAddress,
// once per device, and is kept in g_scratch_ptr.
constexpr size_t kSectorScratchBytes = size_t(4) << 20;
static __device__ __align__(256) unsigned char scratch[kSectorScratchBytes];

constexpr int kMaxDevices = 256;
std::mutex g_scratch_mu;
char *g_scratch_ptr[kMaxDevices];

// Device address of the current device's scratch array.
int sector_scratch(char **out)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess)
        return set_status(RICK_ERR_DEVICE, "hipGetDevice failed: %s", hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices)
        return set_status(RICK_ERR_DEVICE, "device ordinal %d outside 0..%d", dev, kMaxDevices - 1);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr)
            return set_status(RICK_ERR_DEVICE, "hipGetSymbolAddress(scratch) on device %d failed: %s",
                              dev, hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<char *>(p);
    }
    *out = g_scratch_ptr[dev];
    return RICK_OK;
}

// Argument checks shared by the two sector-binning entry points; fills L.
int sector_bin_check(unsigned int num_points, const double *vv, int nsectors, int yaxis, double dx,
                     int w_support, const unsigned int *histo, const unsigned int *start,
                     const unsigned int *rows, SectorLayout &L)
{
    if (nsectors < 1 || nsectors + 1 > kMaxSlots)
        return set_status(RICK_ERR_INVALID, "nsectors %d outside 1..%d", nsectors, kMaxSlots - 1);
    if (yaxis < 1 || !(dx > 0.0) || w_support < 1)
        return set_status(RICK_ERR_INVALID, "yaxis %d, dx %g, w_support %d", yaxis, dx, w_support);
    if (num_points > (1u << 30))
        return set_status(RICK_ERR_INVALID, "%u rows: sector binning handles at most 2^30",
                          num_points);
    if (!histo || !start || (num_points && (!vv || !rows)))
        return set_status(RICK_ERR_INVALID, "null device pointer");
    if (!sector_layout(num_points, nsectors, L))
        return set_status(RICK_ERR_INVALID, "%u rows x %d sectors exceeds the scan capacity",
                          num_points, nsectors);
    return RICK_OK;
}

// Sector binning on workspace w (layout L, at least L.bytes): count, scan,
// fill, per-slot histogram; synchronizes the stream for the totals.
int sector_bin_run(unsigned int num_points, const double *vv, int nsectors, int yaxis, double dx,
                   int w_support, unsigned int *histo, unsigned int *start, unsigned int *rows,
                   unsigned int rows_capacity, char *w, const SectorLayout &L, hipStream_t s,
                   unsigned long *total_entries, unsigned int *dropped)
{
    uint32_t *counts = reinterpret_cast<uint32_t *>(w + L.counts);
    uint32_t *offsets = reinterpret_cast<uint32_t *>(w + L.offsets);
    uint32_t *sums = reinterpret_cast<uint32_t *>(w + L.sums);
    uint32_t *bad = reinterpret_cast<uint32_t *>(w + L.bad);
    SectorGeom g;
    g.nsectors = nsectors;
    g.yaxis = yaxis;
    g.dx = dx;
    g.supporth = (double)((w_support - 1) / 2) * dx;  // init.c:13
    g.n = num_points;
    g.nb = L.nb;
    PREP_HIP(hipMemsetAsync(bad, 0, 4, s));
    hipLaunchKernelGGL(k_sector_count, dim3(L.nb), dim3(kPrepThreads), 0, s, g, vv, counts, bad);
    PREP_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_scan_sums, dim3((unsigned)L.nchunks), dim3(kPrepThreads), 0, s, counts,
                       L.len, sums);
    PREP_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kScanTopThreads), 0, s, sums, (int)L.nchunks,
                       offsets + L.len);
    PREP_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_scan_down, dim3((unsigned)L.nchunks), dim3(kPrepThreads), 0, s, counts,
                       L.len, sums, offsets);
    PREP_HIP(hipGetLastError());
    const int slots = nsectors + 1;
    const size_t lds = (size_t)(kPrepThreads / 64) * slots * sizeof(uint32_t);
    hipLaunchKernelGGL(k_sector_fill, dim3(L.nb), dim3(kPrepThreads), lds, s, g, vv, offsets,
                       rows, rows_capacity);
    PREP_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_sector_histo, dim3((slots + 255) / 256), dim3(256), 0, s, slots, L.nb,
                       offsets, histo, start);
    PREP_HIP(hipGetLastError());
    uint32_t host[2];
    PREP_HIP(hipMemcpyAsync(&host[0], offsets + L.len, 4, hipMemcpyDeviceToHost, s));
    PREP_HIP(hipMemcpyAsync(&host[1], bad, 4, hipMemcpyDeviceToHost, s));
    PREP_HIP(hipStreamSynchronize(s));
    if (total_entries) *total_entries = host[0];
    if (dropped) *dropped = host[1];
    if (host[0] > rows_capacity)
        return set_status(RICK_ERR_INVALID, "rows capacity %u < %u sector entries", rows_capacity,
                          host[0]);
    return RICK_OK;
}

}  // namespace

hipError_t launch_split_complex(const void *grid, int precision, size_t n, double *re, double *im,
                                hipStream_t s)
{
    if (n == 0) return hipSuccess;
    const dim3 grd((unsigned)((n + kPrepThreads - 1) / kPrepThreads));
    if (precision == RICK_FP32)
        hipLaunchKernelGGL(k_split<float>, grd, dim3(kPrepThreads), 0, s,
                           static_cast<const float *>(grid), n, re, im);
    else
        hipLaunchKernelGGL(k_split<double>, grd, dim3(kPrepThreads), 0, s,
                           static_cast<const double *>(grid), n, re, im);
    return hipGetLastError();
}

}  // namespace rick

using namespace rick;

extern "C" {

size_t rick_normalize_uvw_workspace_bytes(void) { return kNormWorkBytes; }

int rick_normalize_uvw(unsigned int num_points, double *uu, double *vv, double *ww,
                       double *extrema, void *workspace, void *stream)
{
    clear_status();
    if (!workspace || (num_points && (!uu || !vv || !ww)))
        return set_status(RICK_ERR_INVALID, "null device pointer");
    hipStream_t s = static_cast<hipStream_t>(stream);
    char *w = static_cast<char *>(workspace);
    double *part_v = reinterpret_cast<double *>(w);
    long long *part_i = reinterpret_cast<long long *>(w + (size_t)kExtBlocks * 6 * sizeof(double));
    double *fin = reinterpret_cast<double *>(w + (size_t)kExtBlocks * 6 * 16);
    const uint32_t n = num_points;
    const int nblk = (int)std::min<uint64_t>(kExtBlocks,
                                             std::max<uint64_t>(1, (n + 2047ull) / 2048ull));
    hipLaunchKernelGGL(k_uvw_extrema, dim3(nblk), dim3(kPrepThreads), 0, s, n, uu, vv, ww,
                       part_v, part_i);
    PREP_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_uvw_finalize, dim3(1), dim3(kPrepThreads), 0, s, nblk, part_v, part_i,
                       fin, extrema);
    PREP_HIP(hipGetLastError());
    if (n) {
        const unsigned ab = (unsigned)std::min<uint64_t>(8192, (n + kPrepThreads - 1) / kPrepThreads);
        hipLaunchKernelGGL(k_uvw_apply, dim3(ab), dim3(kPrepThreads), 0, s, n, uu, vv, ww, fin);
        PREP_HIP(hipGetLastError());
    }
    return RICK_OK;
}

size_t rick_sector_workspace_bytes(unsigned int num_points, int nsectors)
{
    if (nsectors < 1 || nsectors + 1 > kMaxSlots) return 0;
    SectorLayout L;
    return sector_layout(num_points, nsectors, L) ? L.bytes : 0;
}

int rick_sector_bin(unsigned int num_points, const double *vv, int nsectors, int yaxis, double dx,
                    int w_support, unsigned int *histo, unsigned int *start, unsigned int *rows,
                    unsigned int rows_capacity, void *workspace, void *stream,
                    unsigned long *total_entries, unsigned int *dropped)
{
    clear_status();
    SectorLayout L;
    int rc = sector_bin_check(num_points, vv, nsectors, yaxis, dx, w_support, histo, start, rows,
                              L);
    if (rc != RICK_OK) return rc;
    if (!workspace) return set_status(RICK_ERR_INVALID, "null device pointer");
    return sector_bin_run(num_points, vv, nsectors, yaxis, dx, w_support, histo, start, rows,
                          rows_capacity, static_cast<char *>(workspace), L,
                          static_cast<hipStream_t>(stream), total_entries, dropped);
}

size_t rick_sector_bin_noalloc_max_workspace(void) { return kSectorScratchBytes; }

int rick_sector_bin_noalloc(unsigned int num_points, const double *vv, int nsectors, int yaxis,
                            double dx, int w_support, unsigned int *histo, unsigned int *start,
                            unsigned int *rows, unsigned int rows_capacity, void *stream,
                            unsigned long *total_entries, unsigned int *dropped)
{
    clear_status();
    SectorLayout L;
    int rc = sector_bin_check(num_points, vv, nsectors, yaxis, dx, w_support, histo, start, rows,
                              L);
    if (rc != RICK_OK) return rc;
    if (L.bytes > kSectorScratchBytes)
        return set_status(RICK_ERR_INVALID,
                          "%u rows x %d sectors need %zu bytes of workspace; the allocation-free "
                          "entry holds %zu (use rick_sector_bin with a caller workspace)",
                          num_points, nsectors, L.bytes, kSectorScratchBytes);
    char *w = nullptr;
    rc = sector_scratch(&w);
    if (rc != RICK_OK) return rc;
    return sector_bin_run(num_points, vv, nsectors, yaxis, dx, w_support, histo, start, rows,
                          rows_capacity, w, L, static_cast<hipStream_t>(stream), total_entries,
                          dropped);
}

int rick_sector_gather(const unsigned int *list, unsigned int nsec, int isector, int yaxis,
                       double dx, unsigned int num_points, unsigned int freq_per_chan,
                       unsigned int polarizations, const double *uu, const double *vv,
                       const double *ww, const float *weight, const float *vis_real,
                       const float *vis_img, double *uus, double *vvs, double *wws,
                       float *weightss, float *visreals, float *visimgs, void *stream)
{
    clear_status();
    if (isector < 0 || yaxis < 1 || freq_per_chan == 0 || polarizations == 0)
        return set_status(RICK_ERR_INVALID, "isector %d, yaxis %d, F %u, P %u", isector, yaxis,
                          freq_per_chan, polarizations);
    if (nsec == 0) return RICK_OK;
    if (!list || !uu || !vv || !ww || !weight || !vis_real || !vis_img || !uus || !vvs || !wws ||
        !weightss || !visreals || !visimgs)
        return set_status(RICK_ERR_INVALID, "null device pointer");
    hipStream_t s = static_cast<hipStream_t>(stream);
    // gridding_cpu.c:24 shift = dx*yaxis; vvs = vv - isector*shift (line 121).
    const double shift = (double)(dx * yaxis);
    const double vshift = (double)(unsigned)isector * shift;
    hipLaunchKernelGGL(k_gather_rows, dim3((nsec + kPrepThreads - 1) / kPrepThreads),
                       dim3(kPrepThreads), 0, s, list, nsec, num_points, vshift, polarizations,
                       uu, vv, ww, weight, uus, vvs, wws, weightss);
    PREP_HIP(hipGetLastError());
    const uint32_t FP = freq_per_chan * polarizations;
    const bool vec = FP % 4 == 0 &&
                     ((uintptr_t)vis_real | (uintptr_t)vis_img | (uintptr_t)visreals |
                      (uintptr_t)visimgs) % 16 == 0;
    const size_t ne = (size_t)nsec * (vec ? FP / 4 : FP);
    const dim3 grd((unsigned)((ne + kPrepThreads - 1) / kPrepThreads));
    if (vec)
        hipLaunchKernelGGL(k_gather_vis4, grd, dim3(kPrepThreads), 0, s, list, nsec, num_points,
                           FP / 4, reinterpret_cast<const float4 *>(vis_real),
                           reinterpret_cast<const float4 *>(vis_img),
                           reinterpret_cast<float4 *>(visreals),
                           reinterpret_cast<float4 *>(visimgs));
    else
        hipLaunchKernelGGL(k_gather_vis, grd, dim3(kPrepThreads), 0, s, list, nsec, num_points,
                           FP, vis_real, vis_img, visreals, visimgs);
    PREP_HIP(hipGetLastError());
    return RICK_OK;
}

}  // extern "C"
