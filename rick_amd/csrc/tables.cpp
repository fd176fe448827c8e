// Host-side tables of the imaging chain, computed with the C library so they
// are bitwise what the reference computes on the CPU.
#include <cmath>
#include <cstdlib>
#include <vector>

#include "rick_tables.h"

namespace rick {

// I0 series of the reference (w-stacking.cu:54-64).
double bessel0(double x, double precision)
{
    double d = 0.0, ds = 1.0, sum = 1.0;
    do {
        d += 2.0;
        ds *= x * x / (d * d);
        sum += ds;
    } while (ds > sum * precision);
    return sum;
}

int kernel_table(int kind, int support, double *kernel)
{
    const int n = kTableOversample * support, mid = n / 2;
    if (kind == 0) {
        // makeGaussKernel (w-stacking.cu:35-51) with std22 = 1/(2*1*1).
        const double std22 = 1.0 / (2.0 * 1.0 * 1.0);
        const double norm = std22 / kTablePi;
        for (int i = 0; i != mid + 1; i++) {
            const double term = (double)i / (double)kTableOversample;
            kernel[mid + i] = std::sqrt(norm) * std::exp(-(term * term) * std22);
        }
    } else if (kind == 1) {
        // makeKaiserBesselKernel (w-stacking.cu:65-89): alpha 8.6,
        // oversampling factor 1, no sinc -> constant filter ratio 1.
        const double overSamplingFactor = 1.0, alpha = 8.6;
        const double filterRatio = 1.0 / overSamplingFactor;
        const double normFactor = overSamplingFactor / bessel0(alpha, 1e-8);
        for (int i = 0; i != mid + 1; i++) {
            const double term = (double)i / mid;
            kernel[mid + i] =
                filterRatio * bessel0(alpha * std::sqrt(1.0 - (term * term)), 1e-8) * normFactor;
        }
    } else {
        return -1;
    }
    for (int i = 0; i != mid; i++) kernel[i] = kernel[n - 1 - i];
    return n;
}

void pixel_sines(int n, int ntot, double resolution, double *out)
{
    // phase_correction.cu:204-209: fftshift-style index, then sin().
    for (int i = 0; i < n; i++) {
        double c = (double)(i - ntot / 2);
        if (c < 0.0) c = (double)(i + ntot / 2);
        out[i] = std::sin(c * resolution);
    }
}

void twiddles(int n, double *re_im)
{
    for (int m = 0; m < n; m++) {
        const long double a =
            2.0L * 3.141592653589793238462643383279502884L * (long double)m / (long double)n;
        re_im[2 * m] = (double)cosl(a);
        re_im[2 * m + 1] = (double)sinl(a);
    }
}

}  // namespace rick
