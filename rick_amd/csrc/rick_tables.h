// Host-side table builders (plain C++, no HIP).
#pragma once

namespace rick {

constexpr int kTableOversample = 5;
constexpr double kTablePi = 3.14159265359;  // w-stacking.h:6

double bessel0(double x, double precision);
// 5*support doubles; kind 0 = Gauss, 1 = Kaiser-Bessel.  Returns n or -1.
int kernel_table(int kind, int support, double *kernel);
// sin(index*resolution) for the fftshift-style pixel index of
// phase_correction.cu:204-209, for n pixels of an axis of ntot.
void pixel_sines(int n, int ntot, double resolution, double *out);
// exp(+2*pi*i*m/n), m < n, interleaved re/im.
void twiddles(int n, double *re_im);

}  // namespace rick
