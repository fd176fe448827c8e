// W-phase correction and plane sum for gfx950 (phase_correction(),
// phase_correction.cu:97-102 and 185-250).
//
// One thread per image pixel (iu, iv) walks all w-planes of its column of the
// plane stack and accumulates in registers, so the image is written once with
// no atomics (the reference issues two atomics per plane and pixel).  Per
// pixel, everything that does not depend on the plane is hoisted:
//   l = sin(xcoord*resolution), m = sin(ycoord*resolution) come from tables
//   computed on the host with the C library's sin (bitwise the reference's
//   values), radius2 = l*l + m*m and n = sqrt(1 - radius2) are formed once;
// per plane the phase 2*PI*wterm*(n - 1) is formed with the reference's
// operation order and no contraction (1 - radius2 cancels heavily, so an FMA
// here would move the phase by up to ~1e-11 rad at |w| ~ 1e4).
//
// FP32 variant: the plane stack is float; the phase argument and its
// reduction to [-pi, pi] stay FP64, the rotation and accumulation of the
// plane terms use FP32 sin/cos and an FP64 accumulator.
#include "rick_internal.h"
#include "rick_math.h"

#pragma clang fp contract(off)

namespace rick {
namespace {

constexpr int kBlk = 256;

template <int PREC>
__global__ void __launch_bounds__(kBlk)
k_phase(PhaseGeom g, const void *__restrict__ gridss_v, double *__restrict__ ore,
        double *__restrict__ oim, int accumulate)
{
    const int iu = blockIdx.x * kBlk + threadIdx.x;
    const int iv = blockIdx.y;
    if (iu >= g.nx) return;
    const size_t pix = (size_t)iv * g.nx + iu;
    const size_t plane = (size_t)g.nx * g.ny;
    double acc_re = 0.0, acc_im = 0.0;
    const bool phase = g.phase_on && g.planes > 1;
    const double n = phase ? pixel_n(g.ltab[iu], g.mtab[iv]) : 1.0;
    for (int iw = 0; iw < g.planes; ++iw) {
        const size_t idx = pix + (size_t)iw * plane;
        if constexpr (PREC == RICK_FP64) {
            const double2 v = reinterpret_cast<const double2 *>(gridss_v)[idx];
            if (phase) phase_add_f64(v.x, v.y, g.wmin + iw * g.dw, n, g.dwnorm, acc_re, acc_im);
            else {
                acc_re += v.x;
                acc_im += v.y;
            }
        } else {
            const float2 v = reinterpret_cast<const float2 *>(gridss_v)[idx];
            if (phase) phase_add_f32(v.x, v.y, g.wmin + iw * g.dw, n - 1.0, acc_re, acc_im);
            else {
                acc_re += (double)v.x;
                acc_im += (double)v.y;
            }
        }
    }
    if (PREC == RICK_FP32 && phase) {
        // phase_add_f32 leaves the pixel's factor dwnorm * n to the sum.
        acc_re = acc_re * g.dwnorm * n;
        acc_im = acc_im * g.dwnorm * n;
    }
    if (accumulate) {
        acc_re = ore[pix] + acc_re;
        acc_im = oim[pix] + acc_im;
    }
    ore[pix] = acc_re;
    oim[pix] = acc_im;
}

}  // namespace

hipError_t launch_phase(const PhaseGeom &g, int precision, const void *gridss, double *re,
                        double *im, int accumulate, hipStream_t s, Timing *tm)
{
    const dim3 grd((g.nx + kBlk - 1) / kBlk, g.ny), blk(kBlk);
    if (tm) tm->begin(RICK_STAGE_PHASE, s);
    if (precision == RICK_FP64)
        hipLaunchKernelGGL(k_phase<RICK_FP64>, grd, blk, 0, s, g, gridss, re, im, accumulate);
    else
        hipLaunchKernelGGL(k_phase<RICK_FP32>, grd, blk, 0, s, g, gridss, re, im, accumulate);
    if (tm) tm->end(RICK_STAGE_PHASE, s);
    return hipGetLastError();
}

}  // namespace rick
