// C ABI of librick_hip.so: device-resident plans and the reference's drop-in
// entry points (wstack, phase_correction) plus rick_fft_planes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "rick_internal.h"
#include "rick_tables.h"

using namespace rick;

namespace {

thread_local int g_status = RICK_OK;
thread_local char g_msg[512] = "";

std::atomic<int> g_kernel{RICK_KAISERBESSEL};  // reference Makefile default (Makefile:66-70)
std::atomic<int> g_phase_on{1};                // PHASE_ON (Makefile:56)

int set_error(int code, const char *fmt, ...)
{
    g_status = code;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_msg, sizeof g_msg, fmt, ap);
    va_end(ap);
    return code;
}
void clear_error()
{
    g_status = RICK_OK;
    g_msg[0] = '\0';
}

}  // namespace

namespace rick {
int set_status(int code, const char *fmt, ...)
{
    g_status = code;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_msg, sizeof g_msg, fmt, ap);
    va_end(ap);
    return code;
}
void clear_status() { clear_error(); }
}  // namespace rick

namespace {

#define RICK_HIP(call)                                                                        \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return set_error(e_ == hipErrorOutOfMemory ? RICK_ERR_NOMEM : RICK_ERR_DEVICE,    \
                             "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, \
                             __LINE__);                                                       \
    } while (0)

bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

struct TileShape {
    int tu, tv;
};
const TileShape kTiles[] = {{64, 64}, {64, 32}, {32, 32}, {128, 32}, {32, 64}};

}  // namespace

struct rick_plan_s {
    rick_plan_desc d{};
    GridGeom geom{};
    GridWork work{};
    FftPlan fft{};
    PhaseGeom phase{};
    void *dmem = nullptr;
    size_t dbytes = 0;
    double *table_d = nullptr;
    void *tmp_plane = nullptr;
    bool fft_ok = false;    // power-of-two axes in [16, 8192]: the Stockham engine of fft.hip
    bool gen_ok = false;    // any other axis lengths: the mixed-radix / Bluestein engine
    GenFft gen{};
    bool grid_ok = false;
    Timing timing;
};

extern "C" {

int rick_last_status(void) { return g_status; }
const char *rick_last_error(void) { return g_msg; }
const char *rick_version(void) { return "rick-mi355x 0.1 (gfx950)"; }

int rick_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int rick_set_kernel(int kind)
{
    clear_error();
    if (kind < RICK_GAUSS || kind > RICK_GAUSS_HI_PRECISION)
        return set_error(RICK_ERR_INVALID, "unknown kernel kind %d", kind);
    g_kernel = kind;
    return RICK_OK;
}
int rick_get_kernel(void) { return g_kernel; }
int rick_set_phase(int on)
{
    clear_error();
    g_phase_on = on ? 1 : 0;
    return RICK_OK;
}
int rick_get_phase(void) { return g_phase_on; }

int rick_plan_destroy(rick_plan p)
{
    if (!p) return RICK_OK;
    int dev = 0;
    hipGetDevice(&dev);
    hipSetDevice(p->d.device);
    if (p->dmem) hipFree(p->dmem);
    for (auto &v : p->timing.ev)
        for (auto &e : v) {
            hipEventDestroy(e.first);
            hipEventDestroy(e.second);
        }
    hipSetDevice(dev);
    delete p;
    return RICK_OK;
}

int rick_plan_create(rick_plan *out, const rick_plan_desc *desc)
{
    clear_error();
    if (!out || !desc) return set_error(RICK_ERR_INVALID, "null plan or descriptor");
    *out = nullptr;
    const rick_plan_desc &d = *desc;
    if (d.num_w_planes < 1 || d.grid_size_x < 1 || d.grid_size_y < 1)
        return set_error(RICK_ERR_INVALID, "bad grid %d x %d x %d planes", d.grid_size_x,
                         d.grid_size_y, d.num_w_planes);
    if (d.kernel < RICK_GAUSS || d.kernel > RICK_GAUSS_HI_PRECISION)
        return set_error(RICK_ERR_INVALID, "unknown kernel kind %d", d.kernel);
    if (d.w_support < 1 || d.w_support > kMaxSupport || (d.w_support % 2) == 0)
        return set_error(RICK_ERR_INVALID, "w_support must be odd and in [1, %d], got %d",
                         kMaxSupport, d.w_support);
    if (d.precision != RICK_FP64 && d.precision != RICK_FP32)
        return set_error(RICK_ERR_INVALID, "unknown precision %d", d.precision);
    if (!(d.dx > 0.0) || !(d.dw > 0.0))
        return set_error(RICK_ERR_INVALID, "dx and dw must be positive");
    if (d.xaxistot < 0 || d.yaxistot < 0)
        return set_error(RICK_ERR_INVALID, "negative image totals %d x %d", d.xaxistot,
                         d.yaxistot);

    auto p = std::make_unique<rick_plan_s>();
    p->d = d;
    int tu = d.tile_u ? d.tile_u : 64, tv = d.tile_v ? d.tile_v : 64;
    bool tile_ok = false;
    for (const auto &t : kTiles) tile_ok |= (t.tu == tu && t.tv == tv);
    if (!tile_ok) return set_error(RICK_ERR_INVALID, "unsupported tile %d x %d", tu, tv);
    if (tu < d.w_support || tv < d.w_support)
        return set_error(RICK_ERR_INVALID, "tile %d x %d smaller than support %d", tu, tv,
                         d.w_support);
    GridGeom &g = p->geom;
    g.planes = d.num_w_planes;
    g.nx = d.grid_size_x;
    g.ny = d.grid_size_y;
    g.support = d.w_support;
    g.kl = (d.w_support - 1) / 2;
    g.kind = d.kernel;
    g.dx = d.dx;
    g.dw = d.dw;
    g.inv_dx = exact_inverse(d.dx);
    g.inv_dw = exact_inverse(d.dw);
    g.tu = tu;
    g.tv = tv;
    g.tu_shift = __builtin_ctz((unsigned)tu);
    g.tv_shift = __builtin_ctz((unsigned)tv);
    g.tiles_u = (g.nx + tu - 1) / tu;
    g.tiles_v = (g.ny + tv - 1) / tv;
    const uint64_t nb = (uint64_t)g.planes * g.tiles_u * g.tiles_v;
    if (nb >= (1ull << 31)) return set_error(RICK_ERR_INVALID, "too many grid tiles");
    g.nb = (uint32_t)nb;
    g.rmax = d.rows_per_task > 0 ? (uint32_t)d.rows_per_task : 4096u;
    g.ntable = kTableOversample * d.w_support;
    // ~64-128 KB of grid per zero task measured best (LOFAR-like grid).
    g.zrun = d.precision == RICK_FP64 ? 2 : 4;
    // Unrolled footprint loops of the tile kernel: the instantiated lengths.
    {
        const int unrolls[] = {7, 9, 11, 15};
        g.smax = 0;
        for (int u : unrolls)
            if (u >= d.w_support && (d.footprint_unroll == 0 ? g.smax == 0 : u == d.footprint_unroll))
                g.smax = u;
        if (g.smax == 0)
            return set_error(RICK_ERR_INVALID,
                             "footprint_unroll %d: must be 0 or one of 7, 9, 11, 15 and >= "
                             "w_support %d", d.footprint_unroll, d.w_support);
    }
    if ((uint64_t)d.max_points * 4ull >= (1ull << 32))
        return set_error(RICK_ERR_INVALID, "max_points %u too large (limit 2^30)", d.max_points);

    p->grid_ok = d.max_points > 0;
    if (p->grid_ok) {
        if (const char *why = grid_launch_limit(g, d.max_points))
            return set_error(RICK_ERR_INVALID, "%d x %d x %d grid, %d x %d tiles, %u points: %s",
                             g.nx, g.ny, g.planes, tu, tv, d.max_points, why);
    }
    p->fft_ok = is_pow2(g.nx) && is_pow2(g.ny) && g.nx >= 16 && g.ny >= 16 && g.nx <= 8192 &&
                g.ny <= 8192;
    if (!p->fft_ok) p->gen_ok = gen_fft_prepare(g.nx, g.ny, d.precision, &p->gen);

    // Device memory: [table | twiddles x | twiddles y | l table | m table |
    //                 tmp plane | gridding workspace].
    const size_t csz = d.precision == RICK_FP64 ? 16 : 8;
    auto al = [](size_t x) { return (x + 255) / 256 * 256; };
    size_t off = 0;
    const size_t o_table = off; off += al(sizeof(double) * kMaxTable);
    const size_t o_twx = off; off += al(csz * (size_t)g.nx);
    const size_t o_twy = off; off += al(csz * (size_t)g.ny);
    const size_t o_l = off; off += al(sizeof(double) * (size_t)g.nx);
    const size_t o_m = off; off += al(sizeof(double) * (size_t)g.ny);
    const size_t o_tmp = off;
    if (p->fft_ok && fft_needs_tmp(d.precision, g.ny)) off += al(csz * (size_t)g.nx * g.ny);
    const size_t o_gen_tab = off;
    if (p->gen_ok) off += al(p->gen.table_bytes);
    const size_t o_gen_work = off;
    if (p->gen_ok) off += al(p->gen.work_bytes);
    const size_t o_work = off;
    if (p->grid_ok) off += grid_work_bytes(g, d.max_points, d.precision);

    int prev = 0;
    hipGetDevice(&prev);
    RICK_HIP(hipSetDevice(d.device));
    hipError_t e = hipMalloc(&p->dmem, off);
    if (e != hipSuccess) {
        hipSetDevice(prev);
        return set_error(RICK_ERR_NOMEM, "hipMalloc(%zu) failed: %s", off, hipGetErrorString(e));
    }
    p->dbytes = off;
    char *base = static_cast<char *>(p->dmem);
    p->table_d = reinterpret_cast<double *>(base + o_table);
    if (p->fft_ok && fft_needs_tmp(d.precision, g.ny)) p->tmp_plane = base + o_tmp;
    if (p->grid_ok) grid_work_carve(g, d.max_points, d.precision, base + o_work, &p->work);
    // The FP32 split-bucket accumulators and tickets start at zero (the
    // gridder leaves them zero after every call).
    if (p->grid_ok) e = hipMemset(base + o_work, 0, grid_work_bytes(g, d.max_points, d.precision));
    if (e != hipSuccess) {
        hipFree(p->dmem);
        hipSetDevice(prev);
        return set_error(RICK_ERR_DEVICE, "workspace clear failed: %s", hipGetErrorString(e));
    }

    // Host tables -> device (synchronous; plan creation is not on a hot path).
    std::vector<double> h(kMaxTable, 0.0);
    if (d.kernel != RICK_GAUSS_HI_PRECISION) kernel_table(d.kernel, d.w_support, h.data());
    // Largest per-cell weight: K[j]*K[k] <= max|K|^2, or the hi-precision
    // Gaussian's norm (w-stacking.cu:254-256, 496).
    double kmax = 0.0;
    for (int i = 0; i < g.ntable; ++i) kmax = std::max(kmax, std::fabs(h[i]));
    g.cmax = d.kernel == RICK_GAUSS_HI_PRECISION ? (1.0 / (2.0 * 1.0 * 1.0)) / kRefPi : kmax * kmax;
    e = hipMemcpy(p->table_d, h.data(), sizeof(double) * kMaxTable, hipMemcpyHostToDevice);
    std::vector<double> tw;
    auto put_tw = [&](int n, size_t o) {
        tw.assign(2 * (size_t)n, 0.0);
        twiddles(n, tw.data());
        if (d.precision == RICK_FP64)
            return hipMemcpy(base + o, tw.data(), 16 * (size_t)n, hipMemcpyHostToDevice);
        std::vector<float> f(tw.begin(), tw.end());
        return hipMemcpy(base + o, f.data(), 8 * (size_t)n, hipMemcpyHostToDevice);
    };
    if (e == hipSuccess) e = put_tw(g.nx, o_twx);
    if (e == hipSuccess) e = put_tw(g.ny, o_twy);
    std::vector<double> lt(g.nx), mt(g.ny);
    pixel_sines(g.nx, d.xaxistot > 0 ? d.xaxistot : g.nx, d.resolution, lt.data());
    pixel_sines(g.ny, d.yaxistot > 0 ? d.yaxistot : g.ny, d.resolution, mt.data());
    if (e == hipSuccess)
        e = hipMemcpy(base + o_l, lt.data(), sizeof(double) * g.nx, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(base + o_m, mt.data(), sizeof(double) * g.ny, hipMemcpyHostToDevice);
    if (e == hipSuccess && p->gen_ok)
        e = gen_fft_upload(&p->gen, d.precision, base + o_gen_tab, base + o_gen_work);
    hipSetDevice(prev);
    if (e != hipSuccess) {
        hipFree(p->dmem);
        return set_error(RICK_ERR_DEVICE, "table upload failed: %s", hipGetErrorString(e));
    }
    p->fft = FftPlan{g.planes, g.nx, g.ny, base + o_twx, base + o_twy};
    PhaseGeom &ph = p->phase;
    ph.planes = g.planes;
    ph.nx = g.nx;
    ph.ny = g.ny;
    ph.wmin = d.wmin;
    // phase_correction.cu:100-102
    ph.dw = (d.wmax - d.wmin) / (double)d.num_w_planes;
    ph.dwnorm = ph.dw / (d.wmax - d.wmin);
    ph.phase_on = d.phase_on;
    ph.ltab = reinterpret_cast<const double *>(base + o_l);
    ph.mtab = reinterpret_cast<const double *>(base + o_m);
    *out = p.release();
    return RICK_OK;
}

size_t rick_plan_device_bytes(rick_plan p) { return p ? p->dbytes : 0; }

size_t rick_plan_grid_bytes(rick_plan p)
{
    if (!p) return 0;
    const size_t csz = p->d.precision == RICK_FP64 ? 16 : 8;
    return csz * (size_t)p->geom.planes * (size_t)p->geom.nx * (size_t)p->geom.ny;
}

int rick_plan_grid(rick_plan p, unsigned int n, unsigned int F, unsigned int P, const double *uu,
                   const double *vv, const double *ww, const float *vr, const float *vi,
                   const float *wt, void *grid, int accumulate, void *stream)
{
    clear_error();
    if (!p) return set_error(RICK_ERR_INVALID, "null plan");
    if (!p->grid_ok) return set_error(RICK_ERR_INVALID, "plan has no gridding capacity");
    if (n > p->d.max_points)
        return set_error(RICK_ERR_INVALID, "num_points %u exceeds plan capacity %u", n,
                         p->d.max_points);
    if (n > 0 && (F == 0 || P == 0))
        return set_error(RICK_ERR_INVALID, "freq_per_chan and polarizations must be >= 1");
    if (!grid) return set_error(RICK_ERR_INVALID, "null grid");
    if (n > 0 && (!uu || !vv || !ww || !vr || !vi || !wt))
        return set_error(RICK_ERR_INVALID, "null input array");
    RICK_HIP(hipSetDevice(p->d.device));
    RICK_HIP(launch_grid(p->geom, p->work, p->d.precision, p->table_d, n, F, P, uu, vv, ww, vr,
                         vi, wt, grid, accumulate, static_cast<hipStream_t>(stream),
                         &p->timing));
    return RICK_OK;
}

int rick_plan_fft(rick_plan p, const void *in, void *out, void *stream)
{
    clear_error();
    if (!p) return set_error(RICK_ERR_INVALID, "null plan");
    if (!p->fft_ok && !p->gen_ok)
        return set_error(RICK_ERR_UNSUPPORTED, "no FFT plan for %d x %d (axes must be >= 2)",
                         p->geom.nx, p->geom.ny);
    if (!in || !out) return set_error(RICK_ERR_INVALID, "null buffer");
    RICK_HIP(hipSetDevice(p->d.device));
    // fourier_transform.c:24
    const double norm = 1.0 / (double)((double)p->geom.nx * (double)p->geom.ny);
    const hipStream_t s = static_cast<hipStream_t>(stream);
    if (p->fft_ok)
        RICK_HIP(launch_fft(p->fft, p->d.precision, in, out, p->tmp_plane, norm, s, &p->timing));
    else
        RICK_HIP(launch_fft_gen(p->gen, p->d.precision, p->geom.planes, p->geom.nx, p->geom.ny,
                                in, out, norm, s, &p->timing));
    return RICK_OK;
}

int rick_plan_phase(rick_plan p, const void *gridss, double *re, double *im, int accumulate,
                    void *stream)
{
    clear_error();
    if (!p) return set_error(RICK_ERR_INVALID, "null plan");
    if (!gridss || !re || !im) return set_error(RICK_ERR_INVALID, "null buffer");
    if (p->d.phase_on && p->geom.planes > 1 && !(p->d.wmax != p->d.wmin))
        return set_error(RICK_ERR_INVALID, "wmax == wmin: the w-plane spacing is undefined");
    RICK_HIP(hipSetDevice(p->d.device));
    RICK_HIP(launch_phase(p->phase, p->d.precision, gridss, re, im, accumulate,
                          static_cast<hipStream_t>(stream), &p->timing));
    return RICK_OK;
}

int rick_plan_profile(rick_plan p, int enable)
{
    clear_error();
    if (!p) return set_error(RICK_ERR_INVALID, "null plan");
    p->timing.on = enable != 0;
    for (auto &u : p->timing.used) u = 0;
    return RICK_OK;
}

int rick_plan_stage_times(rick_plan p, double *ms, int *launches, int n)
{
    clear_error();
    if (!p) return set_error(RICK_ERR_INVALID, "null plan");
    RICK_HIP(hipSetDevice(p->d.device));
    for (int st = 0; st < RICK_STAGE_COUNT; ++st) {
        double sum = 0.0;
        for (size_t k = 0; k < p->timing.used[st]; ++k) {
            const auto &e = p->timing.ev[st][k];
            RICK_HIP(hipEventSynchronize(e.second));
            float t = 0.0f;
            RICK_HIP(hipEventElapsedTime(&t, e.first, e.second));
            sum += t;
        }
        if (st < n) {
            if (ms) ms[st] = sum;
            if (launches) launches[st] = (int)p->timing.used[st];
        }
        p->timing.used[st] = 0;
    }
    return RICK_OK;
}

int rick_plan_grid_stats(rick_plan p, unsigned int *stats, int n)
{
    clear_error();
    if (!p || !stats) return set_error(RICK_ERR_INVALID, "null argument");
    if (!p->grid_ok) return set_error(RICK_ERR_INVALID, "plan has no gridding capacity");
    RICK_HIP(hipSetDevice(p->d.device));
    RICK_HIP(hipDeviceSynchronize());
    uint32_t meta[4] = {0, 0, 0, 0};
    RICK_HIP(hipMemcpy(meta, p->work.meta, sizeof meta, hipMemcpyDeviceToHost));
    for (int i = 0; i < n && i < 4; i++) stats[i] = meta[i];
    return RICK_OK;
}

int rick_plan_image(rick_plan p, unsigned int n, unsigned int F, unsigned int P, const double *uu,
                    const double *vv, const double *ww, const float *vr, const float *vi,
                    const float *wt, void *grid, double *re, double *im, void *stream)
{
    clear_error();
    if (!p) return set_error(RICK_ERR_INVALID, "null plan");
    if (!p->grid_ok) return set_error(RICK_ERR_INVALID, "plan has no gridding capacity");
    if (!p->fft_ok && !p->gen_ok)
        return set_error(RICK_ERR_UNSUPPORTED, "no FFT plan for %d x %d (axes must be >= 2)",
                         p->geom.nx, p->geom.ny);
    if (n > p->d.max_points)
        return set_error(RICK_ERR_INVALID, "num_points %u exceeds plan capacity %u", n,
                         p->d.max_points);
    if (n > 0 && (F == 0 || P == 0))
        return set_error(RICK_ERR_INVALID, "freq_per_chan and polarizations must be >= 1");
    if (!grid || !re || !im) return set_error(RICK_ERR_INVALID, "null buffer");
    if (n > 0 && (!uu || !vv || !ww || !vr || !vi || !wt))
        return set_error(RICK_ERR_INVALID, "null input array");
    if (p->d.phase_on && p->geom.planes > 1 && !(p->d.wmax != p->d.wmin))
        return set_error(RICK_ERR_INVALID, "wmax == wmin: the w-plane spacing is undefined");
    RICK_HIP(hipSetDevice(p->d.device));
    const hipStream_t s = static_cast<hipStream_t>(stream);
    const double norm = 1.0 / (double)((double)p->geom.nx * (double)p->geom.ny);
    if (!p->fft_ok) {
        // Other axis lengths: dense grid, the mixed-radix / Bluestein FFT in
        // place, then the phase correction and plane sum.
        RICK_HIP(launch_grid(p->geom, p->work, p->d.precision, p->table_d, n, F, P, uu, vv, ww,
                             vr, vi, wt, grid, 0, s, &p->timing, true));
        RICK_HIP(launch_fft_gen(p->gen, p->d.precision, p->geom.planes, p->geom.nx, p->geom.ny,
                                grid, grid, norm, s, &p->timing));
        RICK_HIP(launch_phase(p->phase, p->d.precision, grid, re, im, 0, s, &p->timing));
        return RICK_OK;
    }
    // Sparse grid: tiles of empty buckets are left unwritten; the row FFT
    // reads only occupied tiles (bucket counts) and treats the rest as zero.
    // (A row of more than 256 tiles exceeds the row FFT's occupancy table;
    // such grids are cleared densely instead.)
    const bool sparse = p->geom.nx / p->geom.tu <= 256;
    RICK_HIP(launch_grid(p->geom, p->work, p->d.precision, p->table_d, n, F, P, uu, vv, ww, vr,
                         vi, wt, grid, 0, s, &p->timing, !sparse));
    RICK_HIP(launch_image_chain(p->fft, p->d.precision, grid, norm, s, &p->timing,
                                sparse ? &p->geom : nullptr, sparse ? p->work.counts : nullptr,
                                sparse ? p->work.rowlive : nullptr, p->phase, re, im));
    return RICK_OK;
}


// ---------------------------------------------------------------------------
// Batched imaging: one plan over the group's stacked w-planes, plus the job
// tables the gridding and phase kernels read.

}  // extern "C"

struct rick_batch_s {
    rick_plan plan = nullptr;
    void *dmem = nullptr;
    size_t dbytes = 0;
    BatchJob *jobs_d = nullptr;
    PhaseJob *pjobs_d = nullptr;
    void *grid = nullptr;
    int njobs = 0;
    uint32_t vrows = 0;
    unsigned int F = 1, P = 1;
};

extern "C" {

int rick_batch_create(rick_batch *out, const rick_batch_desc *desc, const rick_batch_job *jobs,
                      int njobs)
{
    clear_error();
    if (!out || !desc || !jobs) return set_error(RICK_ERR_INVALID, "null argument");
    *out = nullptr;
    const rick_batch_desc &d = *desc;
    const int N = d.grid_size;
    if (!(is_pow2(N) && N >= 16 && N <= 8192))
        return set_error(RICK_ERR_UNSUPPORTED,
                         "batched imaging takes power-of-two grids in [16, 8192], got %d", N);
    if (njobs < 1) return set_error(RICK_ERR_INVALID, "no jobs");
    if (d.freq_per_chan == 0 || d.polarizations == 0)
        return set_error(RICK_ERR_INVALID, "freq_per_chan and polarizations must be >= 1");
    const unsigned int FP = d.freq_per_chan * d.polarizations;
    const uint64_t align = batch_row_align();
    std::vector<BatchJob> bj(njobs);
    std::vector<PhaseJob> pj(njobs);
    uint64_t vrows = 0;
    int planes = 0;
    for (int j = 0; j < njobs; ++j) {
        const rick_batch_job &J = jobs[j];
        if (J.num_w_planes < 1) return set_error(RICK_ERR_INVALID, "job %d: no w-planes", j);
        if (J.num_points > 0 && (!J.uu || !J.vv || !J.ww || !J.vis_real || !J.vis_img ||
                                 !J.weight))
            return set_error(RICK_ERR_INVALID, "job %d: null input array", j);
        if (!J.image_real || !J.image_imag)
            return set_error(RICK_ERR_INVALID, "job %d: null image", j);
        if (FP % 4 == 0 &&
            ((reinterpret_cast<uintptr_t>(J.vis_real) | reinterpret_cast<uintptr_t>(J.vis_img)) &
             15u))
            return set_error(RICK_ERR_INVALID, "job %d: visibility arrays not 16-byte aligned", j);
        if (d.phase_on && J.num_w_planes > 1 && !(J.wmax != J.wmin))
            return set_error(RICK_ERR_INVALID, "job %d: wmax == wmin", j);
        if (!(J.dw >= 0.0)) return set_error(RICK_ERR_INVALID, "job %d: negative dw", j);
        const double jdw = J.dw > 0.0 ? J.dw : 1.0 / J.num_w_planes;   // init.c:12
        bj[j] = BatchJob{J.uu, J.vv, J.ww, J.vis_real, J.vis_img, J.weight, J.num_points,
                         (uint32_t)vrows, planes, J.num_w_planes, jdw, exact_inverse(jdw)};
        // phase_correction.cu:100-102
        const double dwp = (J.wmax - J.wmin) / (double)J.num_w_planes;
        pj[j] = PhaseJob{planes, J.num_w_planes, J.wmin, dwp, dwp / (J.wmax - J.wmin),
                         d.phase_on, nullptr, nullptr, J.image_real, J.image_imag};
        vrows += ((uint64_t)J.num_points + align - 1) / align * align;
        planes += J.num_w_planes;
        if (vrows > (1ull << 30))
            return set_error(RICK_ERR_INVALID, "batch rows (padded) exceed 2^30");
    }
    rick_plan_desc pd{};
    pd.num_w_planes = planes;
    pd.grid_size_x = pd.grid_size_y = N;
    pd.kernel = d.kernel;
    pd.w_support = d.w_support;
    pd.dx = 1.0 / N;               // init.c:11
    pd.dw = 1.0;                   // each job carries its own
    pd.precision = d.precision;
    pd.max_points = (unsigned int)std::max<uint64_t>(vrows, 1);
    pd.resolution = 1.0;
    pd.wmin = 0.0;
    pd.wmax = 1.0;
    pd.phase_on = d.phase_on;
    pd.tile_u = d.tile_u;
    pd.tile_v = d.tile_v;
    pd.rows_per_task = d.rows_per_task;
    pd.device = d.device;
    pd.footprint_unroll = d.footprint_unroll;
    auto b = std::make_unique<rick_batch_s>();
    int rc = rick_plan_create(&b->plan, &pd);
    if (rc != RICK_OK) return rc;
    const size_t csz = d.precision == RICK_FP64 ? 16 : 8;
    auto al = [](size_t x) { return (x + 255) / 256 * 256; };
    size_t off = 0;
    const size_t o_jobs = off; off += al(sizeof(BatchJob) * njobs);
    const size_t o_pj = off; off += al(sizeof(PhaseJob) * njobs);
    const size_t o_lm = off; off += al(sizeof(double) * 2 * (size_t)N * njobs);
    const size_t o_grid = off; off += al(csz * (size_t)planes * N * N);
    int prev = 0;
    hipGetDevice(&prev);
    hipError_t e = hipSetDevice(d.device);
    if (e == hipSuccess) e = hipMalloc(&b->dmem, off);
    if (e != hipSuccess) {
        hipSetDevice(prev);
        rick_plan_destroy(b->plan);
        return set_error(RICK_ERR_NOMEM, "batch: hipMalloc(%zu) failed: %s", off,
                         hipGetErrorString(e));
    }
    char *base = static_cast<char *>(b->dmem);
    b->dbytes = off;
    b->jobs_d = reinterpret_cast<BatchJob *>(base + o_jobs);
    b->pjobs_d = reinterpret_cast<PhaseJob *>(base + o_pj);
    b->grid = base + o_grid;
    double *lm = reinterpret_cast<double *>(base + o_lm);
    std::vector<double> tab(2 * (size_t)N);
    for (int j = 0; j < njobs && e == hipSuccess; ++j) {
        pixel_sines(N, N, jobs[j].resolution, tab.data());          // phase_correction.cu:204-209
        pixel_sines(N, N, jobs[j].resolution, tab.data() + N);
        pj[j].ltab = lm + 2 * (size_t)N * j;
        pj[j].mtab = lm + 2 * (size_t)N * j + N;
        e = hipMemcpy(lm + 2 * (size_t)N * j, tab.data(), sizeof(double) * 2 * N,
                      hipMemcpyHostToDevice);
    }
    if (e == hipSuccess)
        e = hipMemcpy(b->jobs_d, bj.data(), sizeof(BatchJob) * njobs, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(b->pjobs_d, pj.data(), sizeof(PhaseJob) * njobs, hipMemcpyHostToDevice);
    hipSetDevice(prev);
    if (e != hipSuccess) {
        hipFree(b->dmem);
        rick_plan_destroy(b->plan);
        return set_error(RICK_ERR_DEVICE, "batch: table upload failed: %s", hipGetErrorString(e));
    }
    b->njobs = njobs;
    b->vrows = (uint32_t)vrows;
    b->F = d.freq_per_chan;
    b->P = d.polarizations;
    *out = b.release();
    return RICK_OK;
}

int rick_batch_run(rick_batch b, void *stream)
{
    clear_error();
    if (!b) return set_error(RICK_ERR_INVALID, "null batch");
    rick_plan p = b->plan;
    RICK_HIP(hipSetDevice(p->d.device));
    const hipStream_t s = static_cast<hipStream_t>(stream);
    const bool sparse = p->geom.nx / p->geom.tu <= 256;
    RICK_HIP(launch_grid(p->geom, p->work, p->d.precision, p->table_d, b->vrows, b->F, b->P,
                         nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, b->grid, 0, s,
                         &p->timing, !sparse, b->jobs_d, b->njobs));
    const double norm = 1.0 / (double)((double)p->geom.nx * (double)p->geom.ny);
    RICK_HIP(launch_image_chain(p->fft, p->d.precision, b->grid, norm, s, &p->timing,
                                sparse ? &p->geom : nullptr, sparse ? p->work.counts : nullptr,
                                sparse ? p->work.rowlive : nullptr, p->phase, nullptr, nullptr,
                                b->pjobs_d, b->njobs));
    return RICK_OK;
}

int rick_batch_destroy(rick_batch b)
{
    if (!b) return RICK_OK;
    int dev = 0;
    hipGetDevice(&dev);
    hipSetDevice(b->plan->d.device);
    if (b->dmem) hipFree(b->dmem);
    hipSetDevice(dev);
    rick_plan_destroy(b->plan);
    delete b;
    return RICK_OK;
}

size_t rick_batch_device_bytes(rick_batch b)
{
    return b ? b->dbytes + rick_plan_device_bytes(b->plan) : 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Drop-in layer: per-device context with cached plans and buffers.

namespace {

struct DevBuf {
    void *p = nullptr;
    size_t n = 0;
    int ensure(size_t bytes)
    {
        if (bytes <= n) return RICK_OK;
        if (p) hipFree(p);
        p = nullptr;
        n = 0;
        if (hipMalloc(&p, bytes) != hipSuccess)
            return set_error(RICK_ERR_NOMEM, "hipMalloc(%zu) failed", bytes);
        n = bytes;
        return RICK_OK;
    }
};

struct Ctx {
    std::mutex mu;
    int device = 0;
    hipStream_t stream = nullptr;
    rick_plan grid_plan = nullptr, phase_plan = nullptr, fft_plan = nullptr;
    DevBuf in, grid, out, img;
    void *pinned[2] = {nullptr, nullptr};
    size_t pinned_bytes = 0;
};

constexpr size_t kStage = 64ull << 20;  // bytes per D2H staging chunk

Ctx *context_for(int rank)
{
    static std::mutex mu;
    static std::vector<std::unique_ptr<Ctx>> ctxs;
    const int ndev = rick_device_count();
    if (ndev <= 0) {
        set_error(RICK_ERR_DEVICE, "no HIP device found");
        return nullptr;
    }
    const int dev = ((rank % ndev) + ndev) % ndev;
    std::lock_guard<std::mutex> lk(mu);
    if ((int)ctxs.size() < ndev) ctxs.resize(ndev);
    if (!ctxs[dev]) {
        auto c = std::make_unique<Ctx>();
        c->device = dev;
        if (hipSetDevice(dev) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess) {
            set_error(RICK_ERR_DEVICE, "cannot set up device %d", dev);
            return nullptr;
        }
        ctxs[dev] = std::move(c);
    }
    hipSetDevice(dev);
    return ctxs[dev].get();
}

bool same_plan(const rick_plan p, const rick_plan_desc &d)
{
    if (!p) return false;
    const rick_plan_desc &a = p->d;
    return a.num_w_planes == d.num_w_planes && a.grid_size_x == d.grid_size_x &&
           a.grid_size_y == d.grid_size_y && a.kernel == d.kernel && a.w_support == d.w_support &&
           a.dx == d.dx && a.dw == d.dw && a.precision == d.precision &&
           a.max_points >= d.max_points && a.resolution == d.resolution && a.wmin == d.wmin &&
           a.wmax == d.wmax && a.phase_on == d.phase_on && a.device == d.device &&
           a.tile_u == d.tile_u && a.tile_v == d.tile_v && a.rows_per_task == d.rows_per_task &&
           a.xaxistot == d.xaxistot && a.yaxistot == d.yaxistot;
}

int ensure_plan(rick_plan *slot, const rick_plan_desc &d)
{
    if (same_plan(*slot, d)) return RICK_OK;
    rick_plan_destroy(*slot);
    *slot = nullptr;
    return rick_plan_create(slot, &d);
}

// dst[i] += src[i] with a few host threads.
void host_add(double *dst, const double *src, size_t n, int nthreads)
{
    nthreads = std::max(1, std::min(nthreads, 64));
    if (n < (1u << 16) || nthreads == 1) {
        for (size_t i = 0; i < n; i++) dst[i] += src[i];
        return;
    }
    std::vector<std::thread> th;
    const size_t per = (n + nthreads - 1) / nthreads;
    for (int t = 0; t < nthreads; t++) {
        const size_t a = t * per, b = std::min(n, a + per);
        if (a >= b) break;
        th.emplace_back([=] {
            for (size_t i = a; i < b; i++) dst[i] += src[i];
        });
    }
    for (auto &x : th) x.join();
}

// host[0..n) += device[0..n) through double-buffered pinned staging.
int accumulate_to_host(Ctx &c, double *host, const double *dev, size_t n, int nthreads)
{
    if (!c.pinned[0]) {
        for (int k = 0; k < 2; k++)
            if (hipHostMalloc(&c.pinned[k], kStage, hipHostMallocDefault) != hipSuccess)
                return set_error(RICK_ERR_NOMEM, "hipHostMalloc(%zu) failed", kStage);
        c.pinned_bytes = kStage;
    }
    const size_t per = kStage / sizeof(double);
    hipEvent_t ev[2];
    for (auto &e : ev) RICK_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    size_t nchunks = (n + per - 1) / per;
    auto issue = [&](size_t k) -> hipError_t {
        const size_t a = k * per, b = std::min(n, a + per);
        hipError_t e = hipMemcpyAsync(c.pinned[k & 1], dev + a, (b - a) * sizeof(double),
                                      hipMemcpyDeviceToHost, c.stream);
        if (e == hipSuccess) e = hipEventRecord(ev[k & 1], c.stream);
        return e;
    };
    int st = RICK_OK;
    if (nchunks > 0 && issue(0) != hipSuccess) st = set_error(RICK_ERR_DEVICE, "D2H failed");
    for (size_t k = 0; st == RICK_OK && k < nchunks; k++) {
        if (k + 1 < nchunks && issue(k + 1) != hipSuccess) {
            st = set_error(RICK_ERR_DEVICE, "D2H failed");
            break;
        }
        if (hipEventSynchronize(ev[k & 1]) != hipSuccess) {
            st = set_error(RICK_ERR_DEVICE, "D2H sync failed");
            break;
        }
        const size_t a = k * per, b = std::min(n, a + per);
        host_add(host + a, static_cast<const double *>(c.pinned[k & 1]), b - a, nthreads);
    }
    hipStreamSynchronize(c.stream);
    for (auto &e : ev) hipEventDestroy(e);
    return st;
}

void report(const char *fn)
{
    if (g_status != RICK_OK) fprintf(stderr, "[rick] %s: %s\n", fn, g_msg);
}

}  // namespace

extern "C" {

void wstack(int num_w_planes, unsigned int num_points, unsigned int freq_per_chan,
            unsigned int polarizations, double *uu, double *vv, double *ww, float *vis_real,
            float *vis_img, float *weight, double dx, double dw, int w_support, int grid_size_x,
            int grid_size_y, double *grid, int num_threads, int rank)
{
    clear_error();
    auto body = [&]() -> int {
        if (num_w_planes < 1 || grid_size_x < 1 || grid_size_y < 1 || !grid)
            return set_error(RICK_ERR_INVALID, "bad grid arguments");
        if (num_points == 0) return RICK_OK;
        if (freq_per_chan == 0 || polarizations == 0)
            return set_error(RICK_ERR_INVALID, "freq_per_chan and polarizations must be >= 1");
        Ctx *c = context_for(rank);
        if (!c) return g_status;
        std::lock_guard<std::mutex> lk(c->mu);
        rick_plan_desc d{};
        d.num_w_planes = num_w_planes;
        d.grid_size_x = grid_size_x;
        d.grid_size_y = grid_size_y;
        d.kernel = g_kernel;
        d.w_support = w_support;
        d.dx = dx;
        d.dw = dw;
        d.precision = RICK_FP64;
        d.max_points = num_points;
        d.device = c->device;
        if (c->grid_plan && same_plan(c->grid_plan, d)) d.max_points = c->grid_plan->d.max_points;
        int st = ensure_plan(&c->grid_plan, d);
        if (st != RICK_OK) return st;
        const size_t n = num_points, nv = n * freq_per_chan * polarizations,
                     nw = n * polarizations;
        auto al = [](size_t x) { return (x + 255) / 256 * 256; };
        const size_t o_u = 0, o_v = al(8 * n), o_w = o_v + al(8 * n), o_r = o_w + al(8 * n),
                     o_i = o_r + al(4 * nv), o_t = o_i + al(4 * nv), tot = o_t + al(4 * nw);
        if ((st = c->in.ensure(tot)) != RICK_OK) return st;
        const size_t gbytes = rick_plan_grid_bytes(c->grid_plan);
        if ((st = c->grid.ensure(gbytes)) != RICK_OK) return st;
        char *b = static_cast<char *>(c->in.p);
        RICK_HIP(hipMemcpyAsync(b + o_u, uu, 8 * n, hipMemcpyHostToDevice, c->stream));
        RICK_HIP(hipMemcpyAsync(b + o_v, vv, 8 * n, hipMemcpyHostToDevice, c->stream));
        RICK_HIP(hipMemcpyAsync(b + o_w, ww, 8 * n, hipMemcpyHostToDevice, c->stream));
        RICK_HIP(hipMemcpyAsync(b + o_r, vis_real, 4 * nv, hipMemcpyHostToDevice, c->stream));
        RICK_HIP(hipMemcpyAsync(b + o_i, vis_img, 4 * nv, hipMemcpyHostToDevice, c->stream));
        RICK_HIP(hipMemcpyAsync(b + o_t, weight, 4 * nw, hipMemcpyHostToDevice, c->stream));
        st = rick_plan_grid(c->grid_plan, num_points, freq_per_chan, polarizations,
                            reinterpret_cast<double *>(b + o_u),
                            reinterpret_cast<double *>(b + o_v),
                            reinterpret_cast<double *>(b + o_w),
                            reinterpret_cast<float *>(b + o_r), reinterpret_cast<float *>(b + o_i),
                            reinterpret_cast<float *>(b + o_t), c->grid.p, 0, c->stream);
        if (st != RICK_OK) return st;
        return accumulate_to_host(*c, grid, static_cast<double *>(c->grid.p), gbytes / 8,
                                  num_threads);
    };
    body();
    report("wstack");
}

void phase_correction(double *gridss, double *image_real, double *image_imag, int xaxis, int yaxis,
                      int num_w_planes, int xaxistot, int yaxistot, double resolution, double wmin,
                      double wmax, int num_threads, int rank)
{
    clear_error();
    auto body = [&]() -> int {
        if (!gridss || !image_real || !image_imag || xaxis < 1 || yaxis < 1 ||
            num_w_planes < 1 || xaxistot < 1 || yaxistot < 1)
            return set_error(RICK_ERR_INVALID, "bad phase_correction arguments");
        Ctx *c = context_for(rank);
        if (!c) return g_status;
        std::lock_guard<std::mutex> lk(c->mu);
        rick_plan_desc d{};
        d.num_w_planes = num_w_planes;
        d.grid_size_x = xaxis;
        d.grid_size_y = yaxis;
        d.kernel = RICK_GAUSS;
        d.w_support = 1;
        d.dx = 1.0;
        d.dw = 1.0;
        d.precision = RICK_FP64;
        d.max_points = 0;
        d.resolution = resolution;
        d.wmin = wmin;
        d.wmax = wmax;
        d.phase_on = g_phase_on;
        d.device = c->device;
        d.xaxistot = xaxistot;
        d.yaxistot = yaxistot;
        int st = ensure_plan(&c->phase_plan, d);
        if (st != RICK_OK) return st;
        const size_t gbytes = rick_plan_grid_bytes(c->phase_plan);
        const size_t npix = (size_t)xaxis * yaxis;
        if ((st = c->grid.ensure(gbytes)) != RICK_OK) return st;
        if ((st = c->img.ensure(2 * 8 * npix)) != RICK_OK) return st;
        RICK_HIP(hipMemcpyAsync(c->grid.p, gridss, gbytes, hipMemcpyHostToDevice, c->stream));
        double *re = static_cast<double *>(c->img.p), *im = re + npix;
        st = rick_plan_phase(c->phase_plan, c->grid.p, re, im, 0, c->stream);
        if (st != RICK_OK) return st;
        if ((st = accumulate_to_host(*c, image_real, re, npix, num_threads)) != RICK_OK)
            return st;
        return accumulate_to_host(*c, image_imag, im, npix, num_threads);
    };
    body();
    report("phase_correction");
}

void rick_fft_planes(int num_w_planes, int grid_size_x, int grid_size_y, int xaxis, int yaxis,
                     const double *grid, double *gridss, int rank)
{
    clear_error();
    auto body = [&]() -> int {
        if (!grid || !gridss || num_w_planes < 1)
            return set_error(RICK_ERR_INVALID, "bad rick_fft_planes arguments");
        if (xaxis != grid_size_x || yaxis != grid_size_y)
            return set_error(RICK_ERR_UNSUPPORTED, "multi-rank slabs are not supported");
        Ctx *c = context_for(rank);
        if (!c) return g_status;
        std::lock_guard<std::mutex> lk(c->mu);
        rick_plan_desc d{};
        d.num_w_planes = num_w_planes;
        d.grid_size_x = grid_size_x;
        d.grid_size_y = grid_size_y;
        d.kernel = RICK_GAUSS;
        d.w_support = 1;
        d.dx = 1.0;
        d.dw = 1.0;
        d.precision = RICK_FP64;
        d.device = c->device;
        int st = ensure_plan(&c->fft_plan, d);
        if (st != RICK_OK) return st;
        const size_t gbytes = rick_plan_grid_bytes(c->fft_plan);
        if ((st = c->grid.ensure(gbytes)) != RICK_OK) return st;
        if ((st = c->out.ensure(gbytes)) != RICK_OK) return st;
        RICK_HIP(hipMemcpyAsync(c->grid.p, grid, gbytes, hipMemcpyHostToDevice, c->stream));
        st = rick_plan_fft(c->fft_plan, c->grid.p, c->out.p, c->stream);
        if (st != RICK_OK) return st;
        RICK_HIP(hipMemcpyAsync(gridss, c->out.p, gbytes, hipMemcpyDeviceToHost, c->stream));
        RICK_HIP(hipStreamSynchronize(c->stream));
        return RICK_OK;
    };
    body();
    report("rick_fft_planes");
}

}  // extern "C"
