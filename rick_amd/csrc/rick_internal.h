// Internal declarations shared by the gfx950 kernels and the C ABI layer.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <utility>
#include <vector>

#include "../../include/rick.h"

namespace rick {

// Status reporting shared by the C ABI entry points (rick_last_status /
// rick_last_error): set_status records code and message and returns code.
int set_status(int code, const char *fmt, ...);
void clear_status();

// De-interleaves n complex grid elements (FP64 or FP32) into two double
// arrays on stream s.
hipError_t launch_split_complex(const void *grid, int precision, size_t n, double *re, double *im,
                                hipStream_t s);

// Streaming (non-temporal) stores for the large write streams: the grid
// and the transformed planes are many times the last-level cache and are not
// read again before being evicted, so keeping them out of the caches leaves
// those for data that is re-read (the next call's coordinates, twiddles,
// records).  The builtin takes native vector types, hence the conversions.
template <typename S>
__device__ __forceinline__ void store_stream(S v, S *p)
{
    __builtin_nontemporal_store(v, p);
}
template <typename V2>
__device__ __forceinline__ void store_stream2(V2 v, V2 *p)
{
    using S = decltype(v.x);
    typedef S nv __attribute__((ext_vector_type(2)));
    const nv x = {v.x, v.y};
    __builtin_nontemporal_store(x, reinterpret_cast<nv *>(p));
}
__device__ __forceinline__ void store_stream4(float4 v, float4 *p)
{
    typedef float nv __attribute__((ext_vector_type(4)));
    const nv x = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(x, reinterpret_cast<nv *>(p));
}
// Streaming loads for data read once (the planes a transform pass reads, the
// visibilities the scatter sums).
template <typename V2>
__device__ __forceinline__ V2 load_stream2(const V2 *p)
{
    using S = decltype(p->x);
    typedef S nv __attribute__((ext_vector_type(2)));
    const nv x = __builtin_nontemporal_load(reinterpret_cast<const nv *>(p));
    V2 r;
    r.x = x.x;
    r.y = x.y;
    return r;
}
__device__ __forceinline__ float4 load_stream4(const float4 *p)
{
    typedef float nv __attribute__((ext_vector_type(4)));
    const nv x = __builtin_nontemporal_load(reinterpret_cast<const nv *>(p));
    return make_float4(x.x, x.y, x.z, x.w);
}

// Table oversampling of the reference (w-stacking.cu:253).
constexpr int kOversample = 5;
// Truncated pi literal of the reference (w-stacking.h:6).
constexpr double kRefPi = 3.14159265359;
// Maximum convolution support handled by the gridder.
constexpr int kMaxSupport = 15;
constexpr int kMaxTable = kOversample * kMaxSupport;

// Optional per-stage kernel timing with HIP events recorded on the launch
// stream (stage ids are the RICK_STAGE_* values of rick.h).
struct Timing {
    bool on = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev[RICK_STAGE_COUNT];
    size_t used[RICK_STAGE_COUNT] = {};
    void begin(int st, hipStream_t s)
    {
        if (!on) return;
        if (used[st] == ev[st].size()) {
            hipEvent_t a = nullptr, b = nullptr;
            (void)hipEventCreate(&a);
            (void)hipEventCreate(&b);
            ev[st].emplace_back(a, b);
        }
        (void)hipEventRecord(ev[st][used[st]].first, s);
    }
    void end(int st, hipStream_t s)
    {
        if (!on) return;
        (void)hipEventRecord(ev[st][used[st]].second, s);
        used[st]++;
    }
};

// Geometry of one gridding problem.
struct GridGeom {
    int planes, nx, ny;
    int support, kl;        // kl = (support-1)/2 (w-stacking.cu:252)
    int kind;               // RICK_GAUSS ...
    double dx, dw;
    int tu, tv;             // tile size (cells), powers of two
    int tu_shift, tv_shift; // log2(tu), log2(tv)
    int tiles_u, tiles_v;   // tiles per axis
    uint32_t nb;            // buckets = planes*tiles_u*tiles_v
    uint32_t rmax;          // rows per task
    int ntable;
    double cmax;            // largest per-cell convolution weight (a bound)
    int zrun;               // empty buckets cleared per zero task (1..4)
    int smax;               // unrolled footprint length of the tile kernel (>= support)
    // 1/dx and 1/dw when dx / dw are powers of two (then x * inv is exactly
    // the IEEE quotient x / d), else 0 (divide).
    double inv_dx, inv_dw;
};

// 1/d if d is a power of two whose reciprocal is exact, else 0.
inline double exact_inverse(double d)
{
    int e = 0;
    if (!(d > 0.0) || __builtin_frexp(d, &e) != 0.5) return 0.0;
    const double inv = 1.0 / d;
    return (__builtin_isfinite(inv) && 1.0 / inv == d) ? inv : 0.0;
}

// Workspace of the binned gridder, carved out of one device allocation.
struct GridWork {
    uint32_t *counts;       // [nb]  records per bucket
    uint32_t *cursor;       // [nb]  scatter cursor (starts at the bucket offset)
    uint32_t *offsets;      // [nb+1] exclusive scan of counts
    uint32_t *blocksum;     // [5*nscan] per scan block: records, accumulation tasks, split,
                            //   non-empty, zero tasks
    uint4 *tasks;           // [task_cap] (bucket, first record, end record, split)
    uint32_t *split_list;   // [split_cap] buckets handled by more than one task
    uint32_t *meta;         // [8]: records, tasks (all), split buckets, non-empty buckets,
                            //   accumulation tasks (stored before the zero tasks), the
                            //   tile kernel's next task slot (persistent workgroups)
    void *records;          // [rec_cap] sorted row records
    // [planes*ny] per grid row of each plane: non-zero iff the tile kernel
    // wrote a non-zero cell to it (sparse overwrite only; cleared with the
    // counts, whose allocation it directly follows).
    unsigned char *rowlive;
    // FP32 only: per split slot, the count of its tasks done and its 64-bit
    // fixed-point accumulator tile (re then im, tu*tv each); both are zero
    // between calls (cleared at plan creation, then by the bucket's last
    // task).  Null for FP64.
    uint32_t *split_ticket;         // [split_cap]
    unsigned long long *split_acc;  // [split_cap * 2 * tu * tv]
    uint32_t rec_cap, split_cap, task_cap, nscan;
};

// One job of a batch group (jobs sharing grid size, kernel, support, tile
// shape and channels/polarizations): its rows, w-plane spacing and the
// group's w-planes it owns.  Rows are numbered in a virtual row space in
// which job j starts at row0 (a multiple of the count kernel's block rows),
// so every gridding block belongs to one job.
struct BatchJob {
    const double *uu, *vv, *ww;
    const float *vr, *vi, *wt;
    uint32_t n;             // rows
    uint32_t row0;          // first virtual row
    int plane0, planes;     // the job's planes in the group's plane stack
    double dw, inv_dw;      // inv_dw: exact_inverse(dw)
};
// Rows per block of the count kernel: batch jobs start at multiples of it.
uint32_t batch_row_align();

size_t grid_work_bytes(const GridGeom &g, uint32_t max_points, int precision);
// Null if every gridding launch of a plan with this geometry and capacity
// fits HIP's launch limits, else why not.
const char *grid_launch_limit(const GridGeom &g, uint32_t max_points);
void grid_work_carve(const GridGeom &g, uint32_t max_points, int precision, void *base,
                     GridWork *w);

// Launches the binned gridder.  Returns a hipError_t.
// With jobs (device array of njobs BatchJob), num_points is the size of the
// virtual row space and the input pointers are the jobs' own; the scatter's
// read mode then assumes 16-byte aligned visibility arrays when F*P is a
// multiple of 4 (checked by the caller).
hipError_t launch_grid(const GridGeom &g, const GridWork &w, int precision,
                       const double *table_d, uint32_t num_points, uint32_t F, uint32_t P,
                       const double *uu, const double *vv, const double *ww,
                       const float *vr, const float *vi, const float *wt, void *grid,
                       int accumulate, hipStream_t s, Timing *tm, bool clear_empty = true,
                       const BatchJob *jobs = nullptr, int njobs = 0);

// FFT: per-plane 2-D backward transform of planes x ny x nx complex, scaled.
struct FftPlan {
    int planes, nx, ny;
    const void *tw_x, *tw_y;   // device twiddle tables exp(+2 pi i m/N), N entries
};
// Long columns (ny above 512) use two passes; an in-place transform stages
// each plane through tmp (one plane).  Output in natural row order.
bool fft_needs_tmp(int precision, int ny);
hipError_t launch_fft(const FftPlan &p, int precision, const void *in, void *out, void *tmp,
                      double scale, hipStream_t s, Timing *tm);

// FFT of any axis lengths >= 2 (mixed radix 2..16, 3, 5, 7, 11, 13 in
// LDS; four-step for long or strided axes; Bluestein for other primes).
constexpr int kGenMaxPasses = 16;
struct GenPassDesc {
    int L, np;                      // line length and its passes
    int rad[kGenMaxPasses];         // radix of each pass
};
struct GenAxis {
    int n;                          // axis length
    int mode;                       // 0 direct, 1 four-step, 2 Bluestein
    int n1, n2;                     // four-step split n = n1*n2
    GenPassDesc a, b;               // direct: a; four-step: a (n1), b (n2)
    int m, mmode, m1, m2;           // Bluestein: convolution length and its plan
    GenPassDesc ma, mb;
    const void *tw;                 // device exp(+2 pi i k/n) [n] (Bluestein: [m])
    const void *chirp, *hf;         // Bluestein: exp(+i pi k^2/n) [n], F_m(chirp*) [m]
};
struct GenFft {
    GenAxis ax[2];                  // 0: rows (length nx), 1: columns (length ny)
    size_t table_bytes, work_bytes; // device bytes of the tables and of the scratch
    void *work;
};
// Plans both axes; false if an axis is shorter than 2 or too long.
bool gen_fft_prepare(int nx, int ny, int precision, GenFft *g);
// Writes the tables (synchronously) at `tables` (table_bytes) and records the
// scratch `work` (work_bytes).
hipError_t gen_fft_upload(GenFft *g, int precision, void *tables, void *work);
// Per-plane backward 2-D transform, in == out allowed, scaled.
hipError_t launch_fft_gen(const GenFft &g, int precision, int planes, int nx, int ny,
                          const void *in, void *out, double scale, hipStream_t s, Timing *tm);

// Phase correction + plane sum.
struct PhaseGeom {
    int planes, nx, ny;
    double wmin, dw, dwnorm;
    int phase_on;
    const double *ltab, *mtab;  // device: sin(xcoord*resolution) per u / per v index
};
hipError_t launch_phase(const PhaseGeom &p, int precision, const void *gridss, double *re,
                        double *im, int accumulate, hipStream_t s, Timing *tm);

// Image chain after gridding, in place on grid (clobbered) and without
// scratch: row FFT reading only the tiles whose gridder bucket count
// (occ_counts, geometry occ_geom; null: dense grid) is non-zero and skipping
// grid rows whose flag in occ_rows (planes*ny, written by the tile kernel) is
// zero, column FFT whose last pass is fused with the phase correction and plane sum
// into the images (overwritten).  Scales by `scale` (1/(Nx*Ny)).
// Phase geometry and output of one job of a batch group (its planes in the
// group's plane stack and its images).
struct PhaseJob {
    int plane0, planes;
    double wmin, dw, dwnorm;
    int phase_on;
    const double *ltab, *mtab;
    double *re, *im;
};
hipError_t launch_image_chain(const FftPlan &p, int precision, void *grid, double scale,
                              hipStream_t s, Timing *tm, const GridGeom *occ_geom,
                              const uint32_t *occ_counts, const unsigned char *occ_rows,
                              const PhaseGeom &pg, double *re, double *im,
                              const PhaseJob *jobs = nullptr,
                              int njobs = 0);

}  // namespace rick
