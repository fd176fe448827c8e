/*
 * CPU oracle for the RICK imaging chain: wstack -> per-w-plane backward C2C FFT
 * -> phase_correction.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness checker for the
 * HIP implementation in rick_amd/ and the timed CPU baseline in bench.py.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it;
 * the product (librick_hip.so, rick_amd/) never links or calls it.
 *
 * Every function restates the reference algorithm operation by operation (same
 * expression order, no FP contraction) so that, run single-threaded, it is
 * bit-identical to the reference's own CPU branch compiled from
 * w-stacking.cu / phase_correction.cu with -ffp-contract=off.  Deliberate
 * differences (all documented at the function):
 *   - 64-bit grid/image indices (the reference's 32-bit ones overflow at
 *     2*planes*Nx*Ny >= 2^32, w-stacking.cu:491, phase_correction.cu:198);
 *   - rows whose w-plane index falls outside [0, planes) are dropped (the
 *     reference writes them past the grid into scratch slack, init.c:391-409);
 *   - footprints are clipped to the grid and table look-ups outside the table
 *     contribute zero (the reference reads out of bounds for pos < 0).
 */
#ifndef RICK_ORACLE_H
#define RICK_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Convolution kernel choice; mirrors the reference's compile-time macros
 * GAUSS / KAISERBESSEL / GAUSS_HI_PRECISION (Makefile:66-70). */
enum { ORACLE_GAUSS = 0, ORACLE_KAISERBESSEL = 1, ORACLE_GAUSS_HI_PRECISION = 2 };

/* bessel0, w-stacking.cu:54-64 (series for I0 stopped at ds <= sum*precision). */
double oracle_bessel0(double x, double precision);

/* 1-D oversampled kernel table of 5*support doubles.
 * GAUSS: makeGaussKernel (w-stacking.cu:35-51) with std22 = 0.5.
 * KAISERBESSEL: makeKaiserBesselKernel (w-stacking.cu:65-89) with alpha 8.6,
 * oversampling 1, no sinc.  Returns the table length, or -1 for a bad kind. */
int oracle_kernel_table(int kind, int support, double *table);

/* wstack, CPU branch (w-stacking.cu:211-267 set-up, 436-541 loop).
 * grid is accumulated into (+=), layout 2*(u + v*Nx + w*Nx*Ny) re,im. */
void oracle_wstack(int kind, int num_w_planes, unsigned int num_points,
                   unsigned int freq_per_chan, unsigned int polarizations,
                   const double *uu, const double *vv, const double *ww,
                   const float *vis_real, const float *vis_img,
                   const float *weight, double dx, double dw, int w_support,
                   int grid_size_x, int grid_size_y, double *grid,
                   int num_threads);

/* fftw_data (fourier_transform.c:13-93) for a single rank: per w-plane, a 2-D
 * backward C2C transform (n0 = Ny rows, n1 = Nx contiguous; sign +1) scaled by
 * 1/(Nx*Ny), read from grid and written to gridss (same interleaved layout).
 * FFTW is not available; the transform is an iterative radix-2 FFT for
 * power-of-two axes and a direct DFT otherwise. */
void oracle_fft_planes(int num_w_planes, int grid_size_x, int grid_size_y,
                       const double *grid, double *gridss, int num_threads);

/* phase_correction, CPU branch (phase_correction.cu:97-102, 185-250).
 * image_real/image_imag are accumulated into (+=).  phase_on selects the
 * PHASE_ON branch (default in the reference Makefile:56). */
void oracle_phase_correction(const double *gridss, double *image_real,
                             double *image_imag, int xaxis, int yaxis,
                             int num_w_planes, int xaxistot, int yaxistot,
                             double resolution, double wmin, double wmax,
                             int phase_on, int num_threads);

/* NORMALIZE_UVW block of gridding() (gridding.c:21-83), in place. */
void oracle_normalize_uvw(size_t n, double *uu, double *vv, double *ww);

/* initialize_array (gridding.c:139-197): the two sequential passes over the
 * rows that count (histo[nsectors+1]) and fill the per-sector row lists,
 * concatenated in rows[] with sector s starting at start[s].  Entries the
 * reference would store out of bounds (binphi outside 0..nsectors, an upper
 * neighbour beyond slot nsectors) are left out and counted; returns that
 * count.  rows must hold 3*n entries. */
unsigned int oracle_sector_bin(size_t n, const double *vv, int nsectors, int yaxis, double dx,
                               int w_support, unsigned int *histo, unsigned int *start,
                               unsigned int *rows);

/* The per-sector assembly loop of gridding_cpu.c:90-138: the list walked
 * backwards, v shifted by isector*dx*yaxis, weights and visibilities copied
 * with their rows. */
void oracle_sector_gather(const unsigned int *list, unsigned int nsec, int isector, int yaxis,
                          double dx, unsigned int freq_per_chan, unsigned int polarizations,
                          const double *uu, const double *vv, const double *ww,
                          const float *weight, const float *vis_real, const float *vis_img,
                          double *uus, double *vvs, double *wws, float *weightss,
                          float *visreals, float *visimgs);

#ifdef __cplusplus
}
#endif
#endif
