// Binned w-stacking gridder for gfx950.
//
// Computes what the reference's wstack() computes (w-stacking.cu:436-541):
//   grid[w][k][j] += sum_rows sum_{f,p} fl32(weight*vis) * K[jKer] * K[kKer]
// over each row's (2KL+1)^2 footprint, with the same index arithmetic
// (IEEE divides, truncation toward zero, +0.5 half-cell offsets, the
// off-centre fabs(u_dist+KL) table look-up), but organised for the GPU:
//
//   k_count    one thread per row: w-plane and footprint -> up to 4 buckets
//              (bucket = w-plane x TUxTV grid tile whose core the footprint
//              touches); wave-aggregated counter increments.
//   k_scan     one workgroup: exclusive scans -> bucket offsets, task offsets
//              (a task is <= rmax rows of one bucket), list of split buckets.
//   k_scatter  one thread per row: the channel/polarization sum
//              S = sum_{f,p} fl32(weight*vis) (skip NaN real parts,
//              w-stacking.cu:517) is formed once per row and a compact
//              record {pos_u, pos_v, S} is written to each of its buckets.
//   k_zero_split  clears the tiles of split buckets (overwrite mode).
//   k_tile     one workgroup per task: accumulates its rows into an LDS tile
//              (one lane per footprint cell, LDS atomics), then writes the
//              tile once with plain, coalesced stores; tasks of split buckets
//              flush with HBM atomics instead.  Every grid cell is written by
//              exactly one tile, so no prior memset of the grid is needed.
//
// The per-cell weight K[jKer]*K[kKer] is bitwise the reference's; the
// per-cell value S*c differs from the reference's sum_{f,p} fl(fl32(w*v)*c)
// only by rounding order (identical when F*P == 1).
#include "rick_internal.h"

#pragma clang fp contract(off)

namespace rick {
namespace {

constexpr int kThreads = 256;
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
__device__ __forceinline__ bool footprint(const GridGeom &g, double pu, double pv, double wi,
                                          Foot &f)
{
    if (!in_int_range(wi) || !in_int_range(pu) || !in_int_range(pv)) return false;
    f.gw = (int)wi;
    f.gu = (int)pu;
    f.gv = (int)pv;
    if (f.gw < 0 || f.gw >= g.planes) return false;
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

__device__ __forceinline__ uint32_t bucket_id(const GridGeom &g, int w, int tv, int tu)
{
    return ((uint32_t)w * (uint32_t)g.tiles_v + (uint32_t)tv) * (uint32_t)g.tiles_u + (uint32_t)tu;
}

// Wave-aggregated atomicAdd(&ctr[key], 1) for every active lane; returns the
// counter value this lane obtained.  Must be called by the whole wave.
__device__ __forceinline__ uint32_t agg_inc(uint32_t *ctr, uint32_t key, bool active)
{
    const int lane = (int)__lane_id();
    unsigned long long pending = __ballot(active);
    uint32_t res = 0;
    while (pending) {
        const int leader = __ffsll(pending) - 1;
        const uint32_t lkey = (uint32_t)__shfl((int)key, leader);
        const unsigned long long peers = __ballot(active && key == lkey) & pending;
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(&ctr[lkey], (uint32_t)__popcll(peers));
        base = (uint32_t)__shfl((int)base, leader);
        if ((peers >> lane) & 1ull)
            res = base + (uint32_t)__popcll(peers & ((1ull << lane) - 1ull));
        pending &= ~peers;
    }
    return res;
}

__global__ void k_fill_u32(uint32_t *p, uint32_t n, uint32_t v)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void __launch_bounds__(kThreads)
k_count(GridGeom g, uint32_t n, const double *__restrict__ uu, const double *__restrict__ vv,
        const double *__restrict__ ww, uint32_t *__restrict__ counts)
{
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    Foot f{};
    bool act = i < n;
    if (act) act = footprint(g, uu[i] / g.dx, vv[i] / g.dx, ww[i] / g.dw, f);
    const int tu0 = act ? f.jmin / g.tu : 0, tu1 = act ? f.jmax / g.tu : 0;
    const int tv0 = act ? f.kmin / g.tv : 0, tv1 = act ? f.kmax / g.tv : 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int du = q & 1, dv = q >> 1;
        const bool a = act && tu0 + du <= tu1 && tv0 + dv <= tv1;
        const uint32_t key = a ? bucket_id(g, f.gw, tv0 + dv, tu0 + du) : 0u;
        agg_inc(counts, key, a);
    }
}

// Single-workgroup scan over nb buckets.
constexpr int kScanThreads = 1024;

__global__ void __launch_bounds__(kScanThreads)
k_scan(GridGeom g, const uint32_t *__restrict__ counts, uint32_t *__restrict__ offsets,
       uint32_t *__restrict__ cursor, uint32_t *__restrict__ task_off,
       uint32_t *__restrict__ split_list, uint32_t split_cap, uint32_t *__restrict__ meta)
{
    __shared__ uint32_t s_rows[kScanThreads], s_tasks[kScanThreads], s_split[kScanThreads];
    const uint32_t t = threadIdx.x;
    const uint32_t nb = g.nb;
    const uint32_t per = (nb + kScanThreads - 1) / kScanThreads;
    const uint32_t b0 = t * per;
    const uint32_t b1 = min(nb, b0 + per);
    uint32_t rows = 0, tasks = 0, split = 0;
    for (uint32_t b = b0; b < b1; ++b) {
        const uint32_t c = counts[b];
        const uint32_t nt = c == 0 ? 1u : (c + g.rmax - 1) / g.rmax;
        rows += c;
        tasks += nt;
        split += nt > 1 ? 1u : 0u;
    }
    s_rows[t] = rows;
    s_tasks[t] = tasks;
    s_split[t] = split;
    __syncthreads();
    // Hillis-Steele inclusive scan.
    for (uint32_t d = 1; d < kScanThreads; d <<= 1) {
        uint32_t a = 0, b = 0, c = 0;
        if (t >= d) {
            a = s_rows[t - d];
            b = s_tasks[t - d];
            c = s_split[t - d];
        }
        __syncthreads();
        s_rows[t] += a;
        s_tasks[t] += b;
        s_split[t] += c;
        __syncthreads();
    }
    uint32_t ro = s_rows[t] - rows, to = s_tasks[t] - tasks, so = s_split[t] - split;
    for (uint32_t b = b0; b < b1; ++b) {
        const uint32_t c = counts[b];
        const uint32_t nt = c == 0 ? 1u : (c + g.rmax - 1) / g.rmax;
        offsets[b] = ro;
        cursor[b] = ro;
        task_off[b] = to;
        if (nt > 1 && so < split_cap) split_list[so] = b;
        ro += c;
        to += nt;
        so += nt > 1 ? 1u : 0u;
    }
    if (t == kScanThreads - 1) {
        offsets[nb] = s_rows[t];
        task_off[nb] = s_tasks[t];
        meta[0] = s_rows[t];
        meta[1] = s_tasks[t];
        meta[2] = s_split[t];
    }
}

template <int PREC>
__global__ void __launch_bounds__(kThreads)
k_scatter(GridGeom g, uint32_t n, uint32_t F, uint32_t P, const double *__restrict__ uu,
          const double *__restrict__ vv, const double *__restrict__ ww,
          const float *__restrict__ vr, const float *__restrict__ vi,
          const float *__restrict__ wt, uint32_t *__restrict__ cursor,
          typename RecOf<PREC>::type *__restrict__ rec)
{
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    Foot f{};
    double pu = 0.0, pv = 0.0;
    bool act = i < n;
    if (act) {
        pu = uu[i] / g.dx;
        pv = vv[i] / g.dx;
        act = footprint(g, pu, pv, ww[i] / g.dw, f);
    }
    // Channel/polarization sum, w-stacking.cu:512-527 (weights indexed
    // row*P + p, shared across channels; float product first).
    double sre = 0.0, sim = 0.0;
    if (act) {
        const size_t vb = (size_t)i * F * P;
        const float *wrow = wt + (size_t)i * P;
        for (uint32_t fr = 0; fr < F; ++fr) {
            for (uint32_t p = 0; p < P; ++p) {
                const size_t k = vb + (size_t)fr * P + p;
                const float a = vr[k];
                if (!isnan(a)) {
                    const float wr = wrow[p] * a;
                    const float wi = wrow[p] * vi[k];
                    sre += (double)wr;
                    sim += (double)wi;
                }
            }
        }
    }
    const int tu0 = act ? f.jmin / g.tu : 0, tu1 = act ? f.jmax / g.tu : 0;
    const int tv0 = act ? f.kmin / g.tv : 0, tv1 = act ? f.kmax / g.tv : 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int du = q & 1, dv = q >> 1;
        const bool a = act && tu0 + du <= tu1 && tv0 + dv <= tv1;
        const uint32_t key = a ? bucket_id(g, f.gw, tv0 + dv, tu0 + du) : 0u;
        const uint32_t slot = agg_inc(cursor, key, a);
        if (a) {
            typename RecOf<PREC>::type r;
            r.pu = pu;
            r.pv = pv;
            if constexpr (PREC == RICK_FP64) {
                r.sre = sre;
                r.sim = sim;
            } else {
                r.sre = (float)sre;
                r.sim = (float)sim;
            }
            rec[slot] = r;
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

__device__ __forceinline__ double readlane_f64(double x, int l)
{
    const uint64_t b = __double_as_longlong(x);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), l);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
__device__ __forceinline__ float readlane_f32(float x, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}

// Per-task LDS tile accumulation.  PREC selects FP64 LDS atomics or FP32 LDS
// atomics with an exactly captured rounding error (TwoSum on the value the
// returning atomic saw) accumulated in a second FP32 array.
template <int PREC, int TU, int TV>
__global__ void __launch_bounds__(kThreads)
k_tile(GridGeom g, const uint32_t *__restrict__ offsets, const uint32_t *__restrict__ task_off,
       const uint32_t *__restrict__ meta, const typename RecOf<PREC>::type *__restrict__ rec,
       const double *__restrict__ table, void *grid_v, int accumulate)
{
    using T = typename std::conditional<PREC == RICK_FP64, double, float>::type;
    constexpr int LDU = TU + kPadU;
    constexpr int NARR = PREC == RICK_FP64 ? 2 : 4;
    __shared__ T lds[NARR * TV * LDU];
    __shared__ T stab[kMaxTable];
    T *sre = lds, *sim = lds + TV * LDU;
    T *cre = lds + 2 * TV * LDU, *cim = lds + 3 * TV * LDU;  // FP32 only

    const uint32_t task = blockIdx.x;
    if (task >= meta[1]) return;
    // Bucket of this task: largest b with task_off[b] <= task (every bucket
    // has at least one task, so task_off is strictly increasing).
    uint32_t lo = 0, hi = g.nb;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (task_off[mid] <= task) lo = mid; else hi = mid;
    }
    const uint32_t b = lo;
    const uint32_t chunk = task - task_off[b];
    const bool split = task_off[b + 1] - task_off[b] > 1;
    const uint32_t r0 = offsets[b] + chunk * g.rmax;
    const uint32_t r1 = min(r0 + g.rmax, offsets[b + 1]);

    const uint32_t per_plane = (uint32_t)g.tiles_u * (uint32_t)g.tiles_v;
    const int w = (int)(b / per_plane), rem = (int)(b % per_plane);
    const int tv = rem / g.tiles_u, tu = rem % g.tiles_u;
    const int j0 = tu * TU, k0 = tv * TV;
    const int nj = min(TU, g.nx - j0), nk = min(TV, g.ny - k0);

    for (int idx = threadIdx.x; idx < NARR * TV * LDU; idx += kThreads) lds[idx] = T(0);
    for (int idx = threadIdx.x; idx < g.ntable; idx += kThreads) stab[idx] = (T)table[idx];
    __syncthreads();

    const int lane = (int)(threadIdx.x & 63);
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int S = g.support, KL = g.kl, ncell = S * S;
    // gauss_kernel_norm constants (w-stacking.cu:254-256).
    const double std22 = 1.0 / (2.0 * 1.0 * 1.0);
    const double gnorm = std22 / kRefPi;

    for (uint32_t base = r0 + (uint32_t)wave * 64u; base < r1; base += 4u * 64u) {
        // Each lane fetches one record; rows are then broadcast one by one.
        const uint32_t my = base + (uint32_t)lane;
        double mpu = 0.0, mpv = 0.0;
        T msr = T(0), msi = T(0);
        if (my < r1) {
            const auto r = rec[my];
            mpu = r.pu;
            mpv = r.pv;
            msr = r.sre;
            msi = r.sim;
        }
        const int nrows = (int)min(64u, r1 - base);
        for (int q = 0; q < nrows; ++q) {
            const double pu = readlane_f64(mpu, q);
            const double pv = readlane_f64(mpv, q);
            T s_re, s_im;
            if constexpr (PREC == RICK_FP64) {
                s_re = readlane_f64(msr, q);
                s_im = readlane_f64(msi, q);
            } else {
                s_re = readlane_f32(msr, q);
                s_im = readlane_f32(msi, q);
            }
            Foot f;
            uv_range(g, pu, pv, f);
            for (int cell = lane; cell < ncell; cell += 64) {
                const int dk = cell / S, dj = cell - dk * S;
                const int j = f.gu - KL + dj, k = f.gv - KL + dk;
                if (j < f.jmin || j > f.jmax || k < f.kmin || k > f.kmax) continue;
                if (j < j0 || j >= j0 + nj || k < k0 || k >= k0 + nk) continue;
                // w-stacking.cu:484, 489, 492-493
                const double v_dist = (double)k + 0.5 - pv;
                const double u_dist = (double)j + 0.5 - pu;
                T c;
                if (g.kind == RICK_GAUSS_HI_PRECISION) {
                    c = (T)(gnorm * exp(-((u_dist * u_dist) + (v_dist * v_dist)) * std22));
                } else {
                    const int jKer = (int)(kOversample * (fabs(u_dist + (double)KL)));
                    const int kKer = (int)(kOversample * (fabs(v_dist + (double)KL)));
                    if (jKer < 0 || jKer >= g.ntable || kKer < 0 || kKer >= g.ntable) continue;
                    c = stab[jKer] * stab[kKer];
                }
                const int li = (k - k0) * LDU + (j - j0);
                const T are = s_re * c, aim = s_im * c;
                if constexpr (PREC == RICK_FP64) {
                    __hip_atomic_fetch_add(&sre[li], are, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    __hip_atomic_fetch_add(&sim[li], aim, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                } else {
                    // TwoSum of (old, x): the atomic stored fl(old + x); its
                    // rounding error is recovered exactly and accumulated.
                    const float o1 = __hip_atomic_fetch_add(&sre[li], are, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    const float o2 = __hip_atomic_fetch_add(&sim[li], aim, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    const float s1 = o1 + are, b1 = s1 - o1;
                    const float e1 = (o1 - (s1 - b1)) + (are - b1);
                    const float s2 = o2 + aim, b2 = s2 - o2;
                    const float e2 = (o2 - (s2 - b2)) + (aim - b2);
                    if (e1 != 0.0f) __hip_atomic_fetch_add(&cre[li], e1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    if (e2 != 0.0f) __hip_atomic_fetch_add(&cim[li], e2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                }
            }
        }
    }
    __syncthreads();

    T *grid = static_cast<T *>(grid_v);
    for (int idx = threadIdx.x; idx < nk * TU; idx += kThreads) {
        const int kk = idx / TU, jj = idx % TU;
        if (jj >= nj) continue;
        const int li = kk * LDU + jj;
        T vre = sre[li], vim = sim[li];
        if constexpr (PREC == RICK_FP32) {
            vre = vre + cre[li];
            vim = vim + cim[li];
        }
        const size_t c = ((size_t)w * g.ny + (size_t)(k0 + kk)) * (size_t)g.nx + (size_t)(j0 + jj);
        if (split) {
            if (vre != T(0)) unsafeAtomicAdd(&grid[2 * c], vre);
            if (vim != T(0)) unsafeAtomicAdd(&grid[2 * c + 1], vim);
        } else {
            using T2 = typename std::conditional<PREC == RICK_FP64, double2, float2>::type;
            T2 *g2 = reinterpret_cast<T2 *>(grid) + c;
            T2 v;
            v.x = vre;
            v.y = vim;
            if (accumulate) {
                const T2 o = *g2;
                v.x = o.x + v.x;
                v.y = o.y + v.y;
            }
            *g2 = v;
        }
    }
}

template <int PREC>
hipError_t launch_tile(const GridGeom &g, const GridWork &w, const double *table, void *grid,
                       int accumulate, hipStream_t s, Timing *tm)
{
    using Rec = typename RecOf<PREC>::type;
    const Rec *rec = static_cast<const Rec *>(w.records);
    const dim3 grd(w.task_cap), blk(kThreads);
#define RICK_TILE_CASE(A, B)                                                                 \
    if (g.tu == A && g.tv == B) {                                                            \
        tm->begin(RICK_STAGE_TILE, s);                                                       \
        hipLaunchKernelGGL((k_tile<PREC, A, B>), grd, blk, 0, s, g, w.offsets, w.task_off,  \
                           w.meta, rec, table, grid, accumulate);                            \
        tm->end(RICK_STAGE_TILE, s);                                                         \
        return hipGetLastError();                                                            \
    }
    RICK_TILE_CASE(64, 64)
    RICK_TILE_CASE(64, 32)
    RICK_TILE_CASE(32, 32)
    RICK_TILE_CASE(128, 32)
    RICK_TILE_CASE(32, 64)
#undef RICK_TILE_CASE
    return hipErrorInvalidValue;
}

inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

}  // namespace

// Capacities: a footprint (support <= tile size) touches at most 2 tiles per
// axis, so a row lands in at most 4 buckets; a split bucket holds more than
// rmax rows, so there are at most 4*rows/(rmax+1) of them; tasks number at
// most nb + 4*rows/rmax.
static void capacities(const GridGeom &g, uint32_t max_points, uint32_t *rec_cap,
                       uint32_t *split_cap, uint32_t *task_cap)
{
    const uint64_t recs = 4ull * (uint64_t)max_points;
    *rec_cap = (uint32_t)(recs > 0 ? recs : 1);
    *split_cap = (uint32_t)(recs / ((uint64_t)g.rmax + 1) + 1);
    *task_cap = (uint32_t)((uint64_t)g.nb + recs / g.rmax + 1);
}

size_t grid_work_bytes(const GridGeom &g, uint32_t max_points, int precision)
{
    uint32_t rec_cap, split_cap, task_cap;
    capacities(g, max_points, &rec_cap, &split_cap, &task_cap);
    const size_t rs = precision == RICK_FP64 ? sizeof(RecF64) : sizeof(RecF32);
    size_t b = 0;
    b += align_up(sizeof(uint32_t) * g.nb, 256);           // counts
    b += align_up(sizeof(uint32_t) * g.nb, 256);           // cursor
    b += align_up(sizeof(uint32_t) * (g.nb + 1), 256);     // offsets
    b += align_up(sizeof(uint32_t) * (g.nb + 1), 256);     // task_off
    b += align_up(sizeof(uint32_t) * split_cap, 256);      // split_list
    b += align_up(sizeof(uint32_t) * 4, 256);              // meta
    b += align_up(rs * rec_cap, 256);                      // records
    return b;
}

void grid_work_carve(const GridGeom &g, uint32_t max_points, int precision, void *base,
                     GridWork *w)
{
    uint32_t rec_cap, split_cap, task_cap;
    capacities(g, max_points, &rec_cap, &split_cap, &task_cap);
    const size_t rs = precision == RICK_FP64 ? sizeof(RecF64) : sizeof(RecF32);
    char *p = static_cast<char *>(base);
    auto take = [&](size_t bytes) {
        char *q = p;
        p += align_up(bytes, 256);
        return q;
    };
    w->counts = reinterpret_cast<uint32_t *>(take(sizeof(uint32_t) * g.nb));
    w->cursor = reinterpret_cast<uint32_t *>(take(sizeof(uint32_t) * g.nb));
    w->offsets = reinterpret_cast<uint32_t *>(take(sizeof(uint32_t) * (g.nb + 1)));
    w->task_off = reinterpret_cast<uint32_t *>(take(sizeof(uint32_t) * (g.nb + 1)));
    w->split_list = reinterpret_cast<uint32_t *>(take(sizeof(uint32_t) * split_cap));
    w->meta = reinterpret_cast<uint32_t *>(take(sizeof(uint32_t) * 4));
    w->records = take(rs * rec_cap);
    w->rec_cap = rec_cap;
    w->split_cap = split_cap;
    w->task_cap = task_cap;
}

hipError_t launch_grid(const GridGeom &g, const GridWork &w, int precision,
                       const double *table_d, uint32_t n, uint32_t F, uint32_t P,
                       const double *uu, const double *vv, const double *ww, const float *vr,
                       const float *vi, const float *wt, void *grid, int accumulate,
                       hipStream_t s, Timing *tm)
{
    Timing off;
    if (tm == nullptr) tm = &off;
    const uint32_t nblk = (n + kThreads - 1) / kThreads;
    const uint32_t fill_blocks = min((g.nb + 255u) / 256u, 4096u);
    tm->begin(RICK_STAGE_COUNT_ROWS, s);
    hipLaunchKernelGGL(k_fill_u32, dim3(fill_blocks), dim3(256), 0, s, w.counts, g.nb, 0u);
    if (n > 0)
        hipLaunchKernelGGL(k_count, dim3(nblk), dim3(kThreads), 0, s, g, n, uu, vv, ww, w.counts);
    tm->end(RICK_STAGE_COUNT_ROWS, s);
    tm->begin(RICK_STAGE_SCAN, s);
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(kScanThreads), 0, s, g, w.counts, w.offsets,
                       w.cursor, w.task_off, w.split_list, w.split_cap, w.meta);
    tm->end(RICK_STAGE_SCAN, s);
    if (precision == RICK_FP64) {
        tm->begin(RICK_STAGE_SCATTER, s);
        if (n > 0)
            hipLaunchKernelGGL(k_scatter<RICK_FP64>, dim3(nblk), dim3(kThreads), 0, s, g, n, F, P,
                               uu, vv, ww, vr, vi, wt, w.cursor,
                               static_cast<RecF64 *>(w.records));
        tm->end(RICK_STAGE_SCATTER, s);
        tm->begin(RICK_STAGE_ZERO, s);
        if (!accumulate)
            hipLaunchKernelGGL(k_zero_split<double>, dim3(w.split_cap), dim3(kThreads), 0, s, g,
                               w.split_list, w.meta, static_cast<double *>(grid));
        tm->end(RICK_STAGE_ZERO, s);
        return launch_tile<RICK_FP64>(g, w, table_d, grid, accumulate, s, tm);
    }
    tm->begin(RICK_STAGE_SCATTER, s);
    if (n > 0)
        hipLaunchKernelGGL(k_scatter<RICK_FP32>, dim3(nblk), dim3(kThreads), 0, s, g, n, F, P, uu,
                           vv, ww, vr, vi, wt, w.cursor, static_cast<RecF32 *>(w.records));
    tm->end(RICK_STAGE_SCATTER, s);
    tm->begin(RICK_STAGE_ZERO, s);
    if (!accumulate)
        hipLaunchKernelGGL(k_zero_split<float>, dim3(w.split_cap), dim3(kThreads), 0, s, g,
                           w.split_list, w.meta, static_cast<float *>(grid));
    tm->end(RICK_STAGE_ZERO, s);
    return launch_tile<RICK_FP32>(g, w, table_d, grid, accumulate, s, tm);
}

}  // namespace rick
