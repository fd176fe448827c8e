/*
 * librick_hip.so -- MI355X (gfx950) implementation of the RICK imaging chain:
 * w-stacking gridding -> per-w-plane 2-D C2C FFT -> w-phase correction and
 * plane sum.
 *
 * Two layers, both plain C (no HIP or torch types in any signature):
 *
 *  1. Drop-in entry points with the reference's own signatures and semantics
 *     (host pointers, results accumulated into caller-zeroed buffers):
 *     wstack(), phase_correction(), plus rick_fft_planes() as the comm-free
 *     stand-in for the FFT stage.  A C program linked against the reference's
 *     w-stacking.o / phase_correction.o links against this library instead.
 *
 *  2. A device-resident plan API (device pointers, caller's HIP stream passed
 *     as void*) that never allocates after rick_plan_create(), so its calls can
 *     be captured in a HIP graph and run back to back without host syncs.
 *
 * Grid layout everywhere: w-plane major, then v (rows, Ny), then u (Nx,
 * contiguous), complex interleaved re/im: element (u, v, w) of a plane set is
 * at 2*(u + v*Nx + w*Nx*Ny) (w-stacking.cu:491).  FP64 grids are double[2*n],
 * FP32 grids float[2*n].  Images are two separate double[Nx*Ny] arrays
 * (fourier_transform.c:209-210).
 */
#ifndef RICK_H
#define RICK_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------- */
/* Status codes.                                                             */
enum {
    RICK_OK = 0,
    RICK_ERR_INVALID = 1,     /* bad argument (sizes, kernel, wmin == wmax ...) */
    RICK_ERR_UNSUPPORTED = 2, /* valid for the reference, not implemented here */
    RICK_ERR_DEVICE = 3,      /* HIP runtime error / no device                 */
    RICK_ERR_NOMEM = 4        /* device or host allocation failed               */
};

/* Convolution kernels, the reference's compile-time GAUSS / KAISERBESSEL /
 * GAUSS_HI_PRECISION (Makefile:66-70). */
enum { RICK_GAUSS = 0, RICK_KAISERBESSEL = 1, RICK_GAUSS_HI_PRECISION = 2 };

/* Arithmetic of the grid, FFT and phase stages. */
enum { RICK_FP64 = 0, RICK_FP32 = 1 };

/* ------------------------------------------------------------------------- */
/* 1. Drop-in entry points.                                                   */

/* Replaces wstack(), w-stacking.h:39-57 / w-stacking.cu:211-543 (C build).
 * grid is a host array of 2*num_w_planes*grid_size_x*grid_size_y doubles and
 * is accumulated into.  dx is used for both axes (w-stacking.cu:464-465).
 * The convolution kernel is the one set by rick_set_kernel() (default
 * KAISERBESSEL, the reference Makefile's default).  num_threads sizes the
 * host-side accumulate; rank selects device rank % device_count
 * (w-stacking.hip.cpp:269).  Rows whose w-plane index is outside
 * [0, num_w_planes) are dropped.  Errors: see rick_last_error(). */
void wstack(int num_w_planes, unsigned int num_points, unsigned int freq_per_chan,
            unsigned int polarizations, double *uu, double *vv, double *ww,
            float *vis_real, float *vis_img, float *weight, double dx, double dw,
            int w_support, int grid_size_x, int grid_size_y, double *grid,
            int num_threads, int rank);

/* Replaces phase_correction(), w-stacking.h:70-83 / phase_correction.cu:97-328.
 * gridss: host, 2*num_w_planes*xaxis*yaxis doubles; image_real/image_imag:
 * host, xaxis*yaxis doubles each, accumulated into.  PHASE_ON behaviour is
 * set by rick_set_phase() (default on, Makefile:56).  A rank's slab
 * (xaxis x yaxis of xaxistot x yaxistot, fourier_transform.c:215) takes its
 * pixel coordinates as the reference does: local iu, iv against the totals
 * (phase_correction.cu:204-209).  wmax == wmin is rejected (the reference
 * divides 0/0, phase_correction.cu:102). */
void phase_correction(double *gridss, double *image_real, double *image_imag,
                      int xaxis, int yaxis, int num_w_planes, int xaxistot,
                      int yaxistot, double resolution, double wmin, double wmax,
                      int num_threads, int rank);

/* Comm-free replacement of the FFT stage: fftw_data() (fourier_transform.c:
 * 13-93, FFTW_BACKWARD, n0 = Ny, n1 = Nx, scaled by 1/(Nx*Ny)) and cuda_fft()
 * (cuda_fft.cu:62-199) for one rank.  Reads grid, writes gridss (host arrays,
 * same layout; may alias).  Any axis length >= 2 (power-of-two axes in
 * [16, 8192] take the Stockham engine; other lengths mixed radix 2..16, 3,
 * 5, 7, 11, 13, four-step for long or strided axes, Bluestein for larger
 * prime factors).  Single rank: xaxis == grid_size_x and yaxis ==
 * grid_size_y. */
void rick_fft_planes(int num_w_planes, int grid_size_x, int grid_size_y,
                     int xaxis, int yaxis, const double *grid, double *gridss,
                     int rank);

/* Drop-in configuration (the reference fixes these at compile time). */
int rick_set_kernel(int kind);      /* RICK_GAUSS / _KAISERBESSEL / _GAUSS_HI_PRECISION */
int rick_get_kernel(void);
int rick_set_phase(int phase_on);   /* 0 = plain plane sum, as without PHASE_ON */
int rick_get_phase(void);

/* Status of the last call made on this host thread. */
int rick_last_status(void);
const char *rick_last_error(void);
int rick_device_count(void);
const char *rick_version(void);

/* ------------------------------------------------------------------------- */
/* 2. Device-resident plan API (no counterpart in the reference, which has   */
/*    only host-pointer entry points; this is its allocation-free, stream-   */
/*    ordered core, used by the drop-in layer above).                         */

typedef struct rick_plan_s *rick_plan;

typedef struct {
    int num_w_planes;
    int grid_size_x;        /* Nx (u, contiguous) */
    int grid_size_y;        /* Ny (v)             */
    int kernel;             /* RICK_GAUSS ...     */
    int w_support;          /* odd, 1..15         */
    double dx, dw;          /* as passed to wstack: usually 1/Nx, 1/planes */
    int precision;          /* RICK_FP64 / RICK_FP32 */
    unsigned int max_points;    /* capacity: rows per rick_plan_grid call */
    /* Phase stage (phase_correction.cu:97-102). */
    double resolution, wmin, wmax;
    int phase_on;
    /* Gridder tiling; 0 selects the built-in defaults. */
    int tile_u, tile_v;
    int rows_per_task;
    int device;             /* HIP device ordinal the plan lives on */
    /* Phase stage of one rank's slab: the image is grid_size_x x grid_size_y
     * pixels of an xaxistot x yaxistot image, pixel coordinates taken as the
     * reference does, local index against the totals (phase_correction.cu:
     * 204-209).  0 = the plan's own grid size. */
    int xaxistot, yaxistot;
    /* Gridder: length of the unrolled footprint loops, one of 7, 9, 11, 15
     * and >= w_support (a tuned choice, rick_amd.autotune); 0 = the smallest
     * that fits. */
    int footprint_unroll;
} rick_plan_desc;

int rick_plan_create(rick_plan *plan, const rick_plan_desc *desc);
int rick_plan_destroy(rick_plan plan);
/* Device bytes held by the plan (tables + gridding workspace). */
size_t rick_plan_device_bytes(rick_plan plan);
/* Bytes of one grid (2*planes*Nx*Ny elements of the plan's precision). */
size_t rick_plan_grid_bytes(rick_plan plan);

/* Gridding (wstack) into a device grid.  accumulate = 0 overwrites every cell
 * of grid (no prior clear needed); 1 adds to its contents.  num_points <=
 * max_points.  All arrays are device pointers; stream is a hipStream_t. */
int rick_plan_grid(rick_plan plan, unsigned int num_points, unsigned int freq_per_chan,
                   unsigned int polarizations, const double *uu, const double *vv,
                   const double *ww, const float *vis_real, const float *vis_img,
                   const float *weight, void *grid, int accumulate, void *stream);

/* Per-plane backward 2-D FFT scaled by 1/(Nx*Ny); in == out allowed. */
int rick_plan_fft(rick_plan plan, const void *in, void *out, void *stream);

/* Phase correction and plane sum into device images (double[Nx*Ny] each). */
int rick_plan_phase(rick_plan plan, const void *gridss, double *image_real,
                    double *image_imag, int accumulate, void *stream);

/* Per-stage kernel timing.  While enabled, every kernel the plan launches is
 * bracketed by HIP events on its stream; rick_plan_stage_times() waits for the
 * recorded work and returns, per stage, the summed kernel time (ms) and the
 * number of launches since the previous query, then resets. */
enum {
    RICK_STAGE_COUNT_ROWS = 0, /* bucket histogram (incl. clearing it)        */
    RICK_STAGE_SCAN = 1,       /* bucket / task offsets                        */
    RICK_STAGE_SCATTER = 2,    /* per-row channel sums -> sorted records       */
    RICK_STAGE_ZERO = 3,       /* reserved: empty tiles are cleared by zero    */
                               /* tasks inside stage 4                         */
    RICK_STAGE_TILE = 4,       /* LDS-tile accumulation, grid write, clearing  */
    RICK_STAGE_FFT_COLS = 5,   /* column FFT (single pass, or four-step A)     */
    RICK_STAGE_FFT_COLS_B = 6, /* column FFT four-step B                       */
    RICK_STAGE_FFT_ROWS = 7,   /* row FFT (+ 1/(Nx*Ny))                         */
    RICK_STAGE_PHASE = 8,      /* phase correction + plane sum                 */
    RICK_STAGE_ZERO_SPLIT = 9, /* clear tiles of split buckets                 */
    RICK_STAGE_COUNT = 10
};
int rick_plan_profile(rick_plan plan, int enable);
int rick_plan_stage_times(rick_plan plan, double *ms, int *launches, int n);

/* Binning statistics of the last rick_plan_grid call (waits for it):
 * stats[0] = row records written (rows x tiles each row touches),
 * stats[1] = gridding tasks (workgroups), stats[2] = buckets split over
 * several tasks, stats[3] = non-empty buckets.  n <= 4 entries are written. */
int rick_plan_grid_stats(rick_plan plan, unsigned int *stats, int n);

/* Whole chain: grid (overwrite) -> FFT in place -> phase into the images
 * (overwrite).  grid is caller scratch of rick_plan_grid_bytes(). */
int rick_plan_image(rick_plan plan, unsigned int num_points, unsigned int freq_per_chan,
                    unsigned int polarizations, const double *uu, const double *vv,
                    const double *ww, const float *vis_real, const float *vis_img,
                    const float *weight, void *grid, double *image_real,
                    double *image_imag, void *stream);

/* Batched imaging: many independent images of one grid size, imaged by one
 * launch per stage (no counterpart in the reference, which images one grid
 * per run).  Jobs share the grid size, kernel, support, precision and
 * channels/polarizations; they differ in rows, w-planes and phase geometry.
 * The group's w-planes are stacked into one grid; gridding bins every job's
 * rows in one pass (each block of rows belongs to one job), the tile kernel
 * runs one task list, the FFT passes run over all planes, and the last
 * column pass writes each job's image from its own planes.  All pointers of
 * the jobs are device pointers, fixed at creation: rick_batch_run() reads
 * their current contents and allocates nothing, so it can be captured in a
 * HIP graph and replayed. */
typedef struct {
    unsigned int num_points;                 /* rows */
    const double *uu, *vv, *ww;              /* normalised as for wstack()   */
    const float *vis_real, *vis_img, *weight;
    int num_w_planes;
    double dw;                               /* 0: 1/num_w_planes (init.c:12) */
    double resolution, wmin, wmax;           /* phase_correction.cu:97-102   */
    double *image_real, *image_imag;         /* Nx*Ny each, overwritten      */
} rick_batch_job;

typedef struct {
    int grid_size;                           /* Nx = Ny: power of two, 16..8192 */
    int kernel, w_support, precision, phase_on;
    unsigned int freq_per_chan, polarizations;
    int tile_u, tile_v, rows_per_task, footprint_unroll;   /* 0: defaults */
    int device;
} rick_batch_desc;

typedef struct rick_batch_s *rick_batch;
int rick_batch_create(rick_batch *batch, const rick_batch_desc *desc, const rick_batch_job *jobs,
                      int njobs);
int rick_batch_run(rick_batch batch, void *stream);
int rick_batch_destroy(rick_batch batch);
size_t rick_batch_device_bytes(rick_batch batch);

/* ------------------------------------------------------------------------- */
/* 3. Data on either side of the chain: the binMS input format, device-side  */
/*    NORMALIZE_UVW and sector binning in front of wstack(), and the image   */
/*    and cube writers behind phase_correction().                            */

/* The twelve values of meta.txt that the reference parses (struct meta,
 * allvars.h; read by readMetaData, init.c:300-333). */
typedef struct {
    unsigned int nmeasures;      /* rows                                   */
    unsigned long nvis;          /* rows * freq_per_chan * polarisations   */
    unsigned int freq_per_chan;
    unsigned int polarisations;
    unsigned int ntimes;
    double dt, thours;
    unsigned int baselines;
    double uvmin, uvmax, wmin, wmax;
} rick_binms_meta;

/* Replaces readMetaData() (init.c:300-333): parses <dir>/meta.txt, the first
 * twelve whitespace-separated values in the reference's order and formats
 * (%u %lu %u %u %u %lf %lf %u %lf %lf %lf %lf); later lines are ignored.
 * Fails (instead of exit(1)) when the file is missing or short. */
int rick_binms_read_meta(const char *dir, rick_binms_meta *meta);

/* Replaces the row split of metaData_calculation() (init.c:335-372): the last
 * `drop` rows are left out (the reference drops 1000), the rest is split into
 * `size` equal parts, the remainder going to the last rank. */
int rick_binms_partition(const rick_binms_meta *meta, int rank, int size, unsigned int drop,
                         unsigned long *startrow, unsigned int *nrows);

/* Replaces readData() (init.c:424-490): reads rows [startrow, startrow+nrows)
 * of ucoord.bin, vcoord.bin, wcoord.bin (f64), weights.bin (f32, P per row),
 * visibilities_real.bin and visibilities_img.bin (f32, F*P per row) under
 * dir.  device < 0: the six arrays are host memory; device >= 0: they are
 * device memory on that device and each file is streamed through two pinned
 * staging buffers.  Short files are an error. */
int rick_binms_read(const char *dir, unsigned int freq_per_chan, unsigned int polarizations,
                    unsigned long startrow, unsigned int nrows, double *uu, double *vv, double *ww,
                    float *weight, float *vis_real, float *vis_img, int device);

/* Writes a binMS directory (host arrays) in the layout of the reference's
 * converter, scripts/create_binMS.py:210-245: the six raw arrays and a
 * meta.txt of 15 lines (the twelve values above, then the weight array's
 * shape: rows, polarisations, 1.0).  meta->nvis must be rows*F*P. */
int rick_binms_write(const char *dir, const rick_binms_meta *meta, const double *uu,
                     const double *vv, const double *ww, const float *weight,
                     const float *vis_real, const float *vis_img);

/* Replaces the raw image output of WRITE_IMAGE (fourier_transform.c:280-289):
 * each file is created ("wb") and the xaxis*yaxis doubles are written at byte
 * rank*xaxis*yaxis*8.  Host arrays. */
int rick_write_image(const char *real_path, const char *imag_path, const double *image_real,
                     const double *image_imag, int xaxis, int yaxis, int rank);

/* Replaces the WRITE_DATA cube writers write_gridded_data() (gridding.c:
 * 210-282) and write_fftw_data() (fourier_transform.c:134-201) for one rank:
 * the real and imaginary parts of a grid (layout above; FP64 or FP32, written
 * as doubles) go to two files, plane after plane.  device < 0: grid is host
 * memory; device >= 0: device memory on that device, de-interleaved there. */
int rick_write_cube(const char *real_path, const char *imag_path, const void *grid, int precision,
                    int num_w_planes, int grid_size_x, int grid_size_y, int device);

/* NORMALIZE_UVW (gridding.c:26-82) on device arrays, in place, stream-
 * ordered: u,v -> (x + maxg)/(2*maxg) with maxg the largest |u|,|v| extremum
 * plus 0.1 %, w -> (w - wmin)/(wmax - wmin), extrema folded from 1e20 (min)
 * and 0.0 (max) as the reference does.  extrema (device double[8], may be
 * NULL) receives min u, v, w, max u, v, w, maxg, 0.  workspace: device
 * memory of rick_normalize_uvw_workspace_bytes().  Bitwise the reference's
 * results for NaN-free input; all-equal w gives NaN as in the reference. */
size_t rick_normalize_uvw_workspace_bytes(void);
int rick_normalize_uvw(unsigned int num_points, double *uu, double *vv, double *ww,
                       double *extrema, void *workspace, void *stream);

/* Sector binning of initialize_array() (gridding.c:139-197) on the device:
 * each row goes to slab binphi = (int)(v*nsectors) and to the slab above /
 * below when within ((w_support-1)/2)*dx of its border.  Outputs (device):
 * histo[nsectors+1] and start[nsectors+1] (the reference's histo_send and
 * the offsets of each sector's list in rows), rows[] the concatenated lists,
 * each in increasing row order as the reference fills them.  Synchronizes
 * the stream; *total_entries receives the number of entries and *dropped the
 * entries the reference would store out of bounds (binphi outside
 * 0..nsectors), which are left out.  Fails if rows_capacity is too small
 * (3*num_points always suffices).  nsectors <= 4095, num_points <= 2^30.
 * workspace: rick_sector_workspace_bytes(num_points, nsectors) bytes. */
size_t rick_sector_workspace_bytes(unsigned int num_points, int nsectors);
int rick_sector_bin(unsigned int num_points, const double *vv, int nsectors, int yaxis, double dx,
                    int w_support, unsigned int *histo, unsigned int *start, unsigned int *rows,
                    unsigned int rows_capacity, void *workspace, void *stream,
                    unsigned long *total_entries, unsigned int *dropped);

/* rick_sector_bin() without a workspace argument and without any allocation.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (tight serving loops, harnesses that forbid hipMalloc).  The
 * workspace is a fixed array in the library's device image, one per device,
 * of rick_sector_bin_noalloc_max_workspace() bytes (4 MiB); calls whose
 * rick_sector_workspace_bytes(num_points, nsectors) exceeds it fail with
 * RICK_ERR_INVALID.  Every call on one device shares that array: calls that
 * overlap on one device (other host threads, other streams) race and must be
 * serialized by the caller.  Otherwise identical to rick_sector_bin(),
 * including the stream synchronization. */
size_t rick_sector_bin_noalloc_max_workspace(void);
int rick_sector_bin_noalloc(unsigned int num_points, const double *vv, int nsectors, int yaxis,
                            double dx, int w_support, unsigned int *histo, unsigned int *start,
                            unsigned int *rows, unsigned int rows_capacity, void *stream,
                            unsigned long *total_entries, unsigned int *dropped);

/* Per-sector input assembly of gridding_cpu.c:90-138 on the device: output
 * row i takes row list[nsec-1-i] (the reference walks the list backwards),
 * v shifted by isector*dx*yaxis; weights (P per row) and visibilities (F*P
 * per row) follow their row.  The result is what rank isector passes to
 * wstack().  List entries >= num_points are skipped.  Stream-ordered. */
int rick_sector_gather(const unsigned int *list, unsigned int nsec, int isector, int yaxis,
                       double dx, unsigned int num_points, unsigned int freq_per_chan,
                       unsigned int polarizations, const double *uu, const double *vv,
                       const double *ww, const float *weight, const float *vis_real,
                       const float *vis_img, double *uus, double *vvs, double *wws,
                       float *weightss, float *visreals, float *visimgs, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* RICK_H */
