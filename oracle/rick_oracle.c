/*
 * CPU oracle for the RICK imaging chain.  TEST INFRASTRUCTURE ONLY: the checker
 * for the HIP path and the timed CPU baseline; the product never calls it.
 *
 * Compiled with -ffp-contract=off so each expression rounds exactly as the
 * reference's C source reads.  See rick_oracle.h for the function contracts
 * and the deliberate, documented departures from the reference.
 */
#include "rick_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Truncated pi literal used by the reference (w-stacking.h:6, allvars.h:47). */
#define REF_PI 3.14159265359
/* Table oversampling factor (w-stacking.cu:253). */
#define OVERSAMPLE 5

double oracle_bessel0(double x, double precision)
{
    /* w-stacking.cu:54-64 */
    double d = 0.0, ds = 1.0, sum = 1.0;
    do {
        d += 2.0;
        ds *= x * x / (d * d);
        sum += ds;
    } while (ds > sum * precision);
    return sum;
}

int oracle_kernel_table(int kind, int support, double *kernel)
{
    const int n = OVERSAMPLE * support, mid = n / 2;
    if (kind == ORACLE_GAUSS) {
        /* makeGaussKernel, w-stacking.cu:35-51, called with std22 = 1/(2*1*1)
         * (w-stacking.cu:254-255, 260). */
        const double std22 = 1.0 / (2.0 * 1.0 * 1.0);
        const double norm = std22 / REF_PI;
        for (int i = 0; i != mid + 1; i++) {
            double term = (double)i / (double)OVERSAMPLE;
            kernel[mid + i] = sqrt(norm) * exp(-(term * term) * std22);
        }
        for (int i = 0; i != mid; i++) kernel[i] = kernel[n - 1 - i];
        return n;
    }
    if (kind == ORACLE_KAISERBESSEL) {
        /* makeKaiserBesselKernel, w-stacking.cu:65-89, with alpha = 8.6,
         * overSamplingFactor = 1, withSinc = 0 (w-stacking.cu:263-266): the
         * sinc factor is the constant filterRatio = 1/overSamplingFactor. */
        const double overSamplingFactor = 1.0, alpha = 8.6;
        const double filterRatio = 1.0 / overSamplingFactor;
        const double normFactor = overSamplingFactor / oracle_bessel0(alpha, 1e-8);
        for (int i = 0; i != mid + 1; i++) {
            double term = (double)i / mid;
            kernel[mid + i] = filterRatio *
                              oracle_bessel0(alpha * sqrt(1.0 - (term * term)), 1e-8) *
                              normFactor;
        }
        for (int i = 0; i != mid; i++) kernel[i] = kernel[n - 1 - i];
        return n;
    }
    return -1;
}

void oracle_wstack(int kind, int num_w_planes, unsigned int num_points,
                   unsigned int freq_per_chan, unsigned int polarizations,
                   const double *uu, const double *vv, const double *ww,
                   const float *vis_real, const float *vis_img,
                   const float *weight, double dx, double dw, int w_support,
                   int grid_size_x, int grid_size_y, double *grid,
                   int num_threads)
{
    /* Set-up, w-stacking.cu:252-267. */
    const int KernelLen = (w_support - 1) / 2;
    const int increaseprecision = OVERSAMPLE;
    const double std22 = 1.0 / (2.0 * 1.0 * 1.0);
    const double norm = std22 / REF_PI;
    const int ntable = increaseprecision * w_support;
    double *convkernel = (double *)malloc((size_t)ntable * sizeof(double));
    if (kind != ORACLE_GAUSS_HI_PRECISION) oracle_kernel_table(kind, w_support, convkernel);

#ifdef _OPENMP
    if (num_threads > 0) omp_set_num_threads(num_threads);
#else
    (void)num_threads;
#endif

    /* Row loop, w-stacking.cu:447-536. */
#pragma omp parallel for schedule(static)
    for (long long il = 0; il < (long long)num_points; il++) {
        const size_t i = (size_t)il;
        const size_t visindex = i * freq_per_chan * polarizations;

        /* w-stacking.cu:464-470: IEEE divides, then truncation toward zero. */
        double pos_u = uu[i] / dx;
        double pos_v = vv[i] / dx;
        double ww_i = ww[i] / dw;
        /* Conversions of out-of-range or NaN values are undefined in C; such
         * rows are dropped (the reference never produces them from
         * normalised inputs). */
        if (!(ww_i > -2147483648.0 && ww_i < 2147483647.0)) continue;
        if (!(pos_u > -2147483648.0 && pos_u < 2147483647.0)) continue;
        if (!(pos_v > -2147483648.0 && pos_v < 2147483647.0)) continue;
        int grid_w = (int)ww_i;
        int grid_u = (int)pos_u;
        int grid_v = (int)pos_v;
        /* w = 1.0 after NORMALIZE_UVW gives grid_w == planes; the reference
         * writes that row into slack past the grid which the FFT never reads,
         * i.e. it is dropped from the image. */
        if (grid_w < 0 || grid_w >= num_w_planes) continue;

        /* w-stacking.cu:473-476, evaluated in signed arithmetic and clipped. */
        int jmin = (grid_u > KernelLen - 1) ? grid_u - KernelLen : 0;
        int jmax = (grid_u < grid_size_x - KernelLen) ? grid_u + KernelLen : grid_size_x - 1;
        int kmin = (grid_v > KernelLen - 1) ? grid_v - KernelLen : 0;
        int kmax = (grid_v < grid_size_y - KernelLen) ? grid_v + KernelLen : grid_size_y - 1;
        if (jmax > grid_size_x - 1) jmax = grid_size_x - 1;
        if (kmax > grid_size_y - 1) kmax = grid_size_y - 1;

        for (int k = kmin; k <= kmax; k++) {
            /* w-stacking.cu:484 (half-cell offset on v, as on u). */
            double v_dist = (double)k + 0.5 - pos_v;
            for (int j = jmin; j <= jmax; j++) {
                double u_dist = (double)j + 0.5 - pos_u;
                /* w-stacking.cu:491, widened to 64 bits. */
                size_t iKer = 2 * ((size_t)j + (size_t)k * (size_t)grid_size_x +
                                   (size_t)grid_w * (size_t)grid_size_x * (size_t)grid_size_y);
                /* w-stacking.cu:492-493: off-centre table look-up via fabs. */
                int jKer = (int)(increaseprecision * (fabs(u_dist + (double)KernelLen)));
                int kKer = (int)(increaseprecision * (fabs(v_dist + (double)KernelLen)));
                double conv_weight;
                if (kind == ORACLE_GAUSS_HI_PRECISION) {
                    /* gauss_kernel_norm, w-stacking.cu:28-33. */
                    conv_weight = norm * exp(-((u_dist * u_dist) + (v_dist * v_dist)) * std22);
                } else {
                    /* w-stacking.cu:499/505; entries outside the table weigh 0. */
                    if (jKer < 0 || jKer >= ntable || kKer < 0 || kKer >= ntable) continue;
                    conv_weight = convkernel[jKer] * convkernel[kKer];
                }
                /* w-stacking.cu:508-527: float product first, then times the
                 * double weight; rows whose real part is NaN are skipped. */
                double add_term_real = 0.0;
                double add_term_img = 0.0;
                size_t ifine = visindex;
                for (unsigned int ifreq = 0; ifreq < freq_per_chan; ifreq++) {
                    size_t iweight = visindex / freq_per_chan;
                    for (unsigned int ipol = 0; ipol < polarizations; ipol++) {
                        if (!isnan(vis_real[ifine])) {
                            float wr = weight[iweight] * vis_real[ifine];
                            float wi = weight[iweight] * vis_img[ifine];
                            add_term_real += wr * conv_weight;
                            add_term_img += wi * conv_weight;
                        }
                        ifine++;
                        iweight++;
                    }
                }
                /* w-stacking.cu:529-532 */
#pragma omp atomic
                grid[iKer] += add_term_real;
#pragma omp atomic
                grid[iKer + 1] += add_term_img;
            }
        }
    }
    free(convkernel);
}

/* ---------------------------------------------------------------------- */
/* CPU FFT used in place of FFTW (absent from this environment).          */

static int is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

/* Twiddles exp(+2*pi*i*k/n), k < n, from long-double cos/sin (correctly
 * rounded to double in practice). */
static double *make_twiddles(int n)
{
    double *tw = (double *)malloc(2 * (size_t)n * sizeof(double));
    for (int k = 0; k < n; k++) {
        long double a = 2.0L * 3.141592653589793238462643383279502884L * (long double)k / (long double)n;
        tw[2 * k] = (double)cosl(a);
        tw[2 * k + 1] = (double)sinl(a);
    }
    return tw;
}

/* In-place backward (sign +1) transform of one contiguous complex line. */
static void line_fft(double *x, int n, const double *tw, double *tmp)
{
    if (is_pow2(n)) {
        /* Iterative radix-2 decimation in time with bit-reversal permutation. */
        for (int i = 1, j = 0; i < n; i++) {
            int bit = n >> 1;
            for (; j & bit; bit >>= 1) j ^= bit;
            j ^= bit;
            if (i < j) {
                double tr = x[2 * i], ti = x[2 * i + 1];
                x[2 * i] = x[2 * j]; x[2 * i + 1] = x[2 * j + 1];
                x[2 * j] = tr; x[2 * j + 1] = ti;
            }
        }
        for (int len = 2; len <= n; len <<= 1) {
            const int half = len >> 1, step = n / len;
            for (int s = 0; s < n; s += len) {
                for (int k = 0; k < half; k++) {
                    const double wr = tw[2 * (k * step)], wi = tw[2 * (k * step) + 1];
                    double *a = x + 2 * (s + k), *b = x + 2 * (s + k + half);
                    double br = b[0] * wr - b[1] * wi;
                    double bi = b[0] * wi + b[1] * wr;
                    b[0] = a[0] - br; b[1] = a[1] - bi;
                    a[0] = a[0] + br; a[1] = a[1] + bi;
                }
            }
        }
    } else {
        /* Direct DFT with exact twiddle indices. */
        for (int k = 0; k < n; k++) {
            double sr = 0.0, si = 0.0;
            for (int j = 0; j < n; j++) {
                const size_t m = ((size_t)j * (size_t)k) % (size_t)n;
                const double wr = tw[2 * m], wi = tw[2 * m + 1];
                sr += x[2 * j] * wr - x[2 * j + 1] * wi;
                si += x[2 * j] * wi + x[2 * j + 1] * wr;
            }
            tmp[2 * k] = sr; tmp[2 * k + 1] = si;
        }
        memcpy(x, tmp, 2 * (size_t)n * sizeof(double));
    }
}

void oracle_fft_planes(int num_w_planes, int grid_size_x, int grid_size_y,
                       const double *grid, double *gridss, int num_threads)
{
    const int nx = grid_size_x, ny = grid_size_y;
    const size_t plane = (size_t)nx * (size_t)ny;
    /* fourier_transform.c:24 */
    const double norm = 1.0 / (double)(grid_size_x * (double)grid_size_y);
    double *twx = make_twiddles(nx), *twy = make_twiddles(ny);
#ifdef _OPENMP
    if (num_threads > 0) omp_set_num_threads(num_threads);
#else
    (void)num_threads;
#endif
    for (int iw = 0; iw < num_w_planes; iw++) {
        const double *src = grid + 2 * plane * (size_t)iw;
        double *dst = gridss + 2 * plane * (size_t)iw;
        memcpy(dst, src, 2 * plane * sizeof(double));
        /* Transform along x (contiguous, n1 = Nx). */
#pragma omp parallel
        {
            double *tmp = (double *)malloc(2 * (size_t)(nx > ny ? nx : ny) * sizeof(double));
            double *col = (double *)malloc(2 * (size_t)ny * sizeof(double));
#pragma omp for schedule(static)
            for (int iv = 0; iv < ny; iv++) line_fft(dst + 2 * (size_t)iv * nx, nx, twx, tmp);
            /* Transform along y (stride Nx, n0 = Ny). */
#pragma omp for schedule(static)
            for (int iu = 0; iu < nx; iu++) {
                for (int iv = 0; iv < ny; iv++) {
                    col[2 * iv] = dst[2 * ((size_t)iv * nx + iu)];
                    col[2 * iv + 1] = dst[2 * ((size_t)iv * nx + iu) + 1];
                }
                line_fft(col, ny, twy, tmp);
                for (int iv = 0; iv < ny; iv++) {
                    /* fourier_transform.c:73-74 */
                    dst[2 * ((size_t)iv * nx + iu)] = norm * col[2 * iv];
                    dst[2 * ((size_t)iv * nx + iu) + 1] = norm * col[2 * iv + 1];
                }
            }
            free(col);
            free(tmp);
        }
    }
    free(twx);
    free(twy);
}

void oracle_phase_correction(const double *gridss, double *image_real,
                             double *image_imag, int xaxis, int yaxis,
                             int num_w_planes, int xaxistot, int yaxistot,
                             double resolution, double wmin, double wmax,
                             int phase_on, int num_threads)
{
    /* phase_correction.cu:100-102 */
    const double dw = (wmax - wmin) / (double)num_w_planes;
    const double dwnorm = dw / (wmax - wmin);
#ifdef _OPENMP
    if (num_threads > 0) omp_set_num_threads(num_threads);
#else
    (void)num_threads;
#endif
    /* phase_correction.cu:191-249 */
#pragma omp parallel for collapse(2) schedule(static)
    for (int iw = 0; iw < num_w_planes; iw++) {
        for (int iv = 0; iv < yaxis; iv++) {
            for (int iu = 0; iu < xaxis; iu++) {
                /* Indices widened to 64 bits (the reference's unsigned int
                 * overflows for 2*planes*x*y >= 2^32). */
                size_t index = 2 * ((size_t)iu + (size_t)iv * (size_t)xaxis +
                                    (size_t)xaxis * (size_t)yaxis * (size_t)iw);
                size_t img_index = (size_t)iu + (size_t)iv * (size_t)xaxis;
                double wterm = wmin + iw * dw;
                if (phase_on && num_w_planes > 1) {
                    /* fftshift-style pixel index, then sin(index*resolution). */
                    double xcoord = (double)(iu - xaxistot / 2);
                    if (xcoord < 0.0) xcoord = (double)(iu + xaxistot / 2);
                    xcoord = sin(xcoord * resolution);
                    double ycoord = (double)(iv - yaxistot / 2);
                    if (ycoord < 0.0) ycoord = (double)(iv + yaxistot / 2);
                    ycoord = sin(ycoord * resolution);
                    double radius2 = (xcoord * xcoord + ycoord * ycoord);
                    /* The xcoord <= 1 branch (213-220) is overwritten by
                     * 222-223 and does not affect the result. */
                    double preal = cos(2.0 * REF_PI * wterm * (sqrt(1 - radius2) - 1.0));
                    double pimag = sin(2.0 * REF_PI * wterm * (sqrt(1 - radius2) - 1.0));
                    double p = gridss[index], q = gridss[index + 1];
                    double r = preal, s = pimag;
#pragma omp atomic
                    image_real[img_index] += (p * r - q * s) * dwnorm * sqrt(1 - radius2);
#pragma omp atomic
                    image_imag[img_index] += (p * s + q * r) * dwnorm * sqrt(1 - radius2);
                } else {
#pragma omp atomic
                    image_real[img_index] += gridss[index];
#pragma omp atomic
                    image_imag[img_index] += gridss[index + 1];
                }
            }
        }
    }
}

void oracle_normalize_uvw(size_t n, double *uu, double *vv, double *ww)
{
    /* gridding.c:26-82 (single rank: the MPI_Allreduce is the identity). */
    double minu = 1e20, minv = 1e20, minw = 1e20;
    double maxu = 0, maxv = 0, maxw = 0;
    for (size_t i = 0; i < n; i++) {
        minu = minu < uu[i] ? minu : uu[i];
        minv = minv < vv[i] ? minv : vv[i];
        minw = minw < ww[i] ? minw : ww[i];
        maxu = maxu > uu[i] ? maxu : uu[i];
        maxv = maxv > vv[i] ? maxv : vv[i];
        maxw = maxw > ww[i] ? maxw : ww[i];
    }
    const double offset = 0.001;
    double ming = fabs(minu) > fabs(minv) ? fabs(minu) : fabs(minv);
    double maxg = fabs(maxu) > fabs(maxv) ? fabs(maxu) : fabs(maxv);
    maxg = maxg > ming ? maxg : ming;
    maxg = maxg + offset * maxg;
    for (size_t i = 0; i < n; i++) {
        uu[i] = (uu[i] + maxg) / (2.0 * maxg);
        vv[i] = (vv[i] + maxg) / (2.0 * maxg);
        ww[i] = (ww[i] - minw) / (maxw - minw);
    }
}

/* Slots of one row, in the order the reference tests them (gridding.c:
 * 156-167); -1 marks no entry.  Returns the out-of-bounds entries skipped. */
static unsigned int sector_slots(double vvh, int nsectors, int yaxis, double dx, double supporth,
                                 int s[3])
{
    s[0] = s[1] = s[2] = -1;
    const double b = vvh * nsectors;
    if (!(b > -1.0 && b < (double)nsectors + 1.0)) return 1;
    const int binphi = (int)b;
    const double updist = (double)((binphi + 1) * yaxis) * dx - vvh;
    const double downdist = vvh - (double)(binphi * yaxis) * dx;
    unsigned int bad = 0;
    s[0] = binphi;
    if (updist < supporth && updist >= 0.0) {
        if (binphi + 1 <= nsectors) s[1] = binphi + 1;
        else bad++;
    }
    if (downdist < supporth && binphi > 0 && downdist >= 0.0) s[2] = binphi - 1;
    return bad;
}

unsigned int oracle_sector_bin(size_t n, const double *vv, int nsectors, int yaxis, double dx,
                               int w_support, unsigned int *histo, unsigned int *start,
                               unsigned int *rows)
{
    const double supporth = (double)((w_support - 1) / 2) * dx;  /* init.c:13 */
    unsigned int bad = 0;
    int s[3];
    if (nsectors < 1) return 0;
    for (int k = 0; k <= nsectors; k++) histo[k] = 0;
    /* First pass: histo_send (gridding.c:150-168). */
    for (size_t i = 0; i < n; i++) {
        bad += sector_slots(vv[i], nsectors, yaxis, dx, supporth, s);
        for (int j = 0; j < 3; j++)
            if (s[j] >= 0) histo[s[j]]++;
    }
    unsigned int acc = 0;
    for (int k = 0; k <= nsectors; k++) {
        start[k] = acc;
        acc += histo[k];
    }
    /* Second pass: sectorarray[s][counter[s]++] = row (gridding.c:178-190). */
    unsigned int *counter = calloc((size_t)(unsigned)nsectors + 1u, sizeof *counter);
    for (size_t i = 0; i < n; i++) {
        sector_slots(vv[i], nsectors, yaxis, dx, supporth, s);
        for (int j = 0; j < 3; j++)
            if (s[j] >= 0) rows[start[s[j]] + counter[s[j]]++] = (unsigned int)i;
    }
    free(counter);
    return bad;
}

void oracle_sector_gather(const unsigned int *list, unsigned int nsec, int isector, int yaxis,
                          double dx, unsigned int freq_per_chan, unsigned int polarizations,
                          const double *uu, const double *vv, const double *ww,
                          const float *weight, const float *vis_real, const float *vis_img,
                          double *uus, double *vvs, double *wws, float *weightss,
                          float *visreals, float *visimgs)
{
    const double shift = (double)(dx * yaxis);     /* gridding_cpu.c:24 */
    const unsigned int isec = (unsigned int)isector;
    size_t icount = 0, ip = 0, inu = 0;
    for (long iphi = (long)nsec - 1; iphi >= 0; iphi--) {
        const size_t il = list[iphi];
        uus[icount] = uu[il];
        vvs[icount] = vv[il] - isec * shift;
        wws[icount] = ww[il];
        for (unsigned int p = 0; p < polarizations; p++)
            weightss[ip++] = weight[il * polarizations + p];
        for (unsigned int f = 0; f < polarizations * freq_per_chan; f++) {
            visreals[inu] = vis_real[il * polarizations * freq_per_chan + f];
            visimgs[inu] = vis_img[il * polarizations * freq_per_chan + f];
            inu++;
        }
        icount++;
    }
}
