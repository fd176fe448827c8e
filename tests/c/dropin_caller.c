/* A C caller of the imaging chain through the reference's entry points,
 * linked with -lrick_hip as a program built against the reference's
 * w-stacking.o / phase_correction.o would be: wstack() -> rick_fft_planes()
 * (the FFT stage) -> phase_correction(), host arrays throughout, with the
 * caller zeroing the grid and images as init.c:392 and
 * fourier_transform.c:209-210 do.
 *
 *   dropin_caller --version     prints rick_version() (no device needed)
 *   dropin_caller OUTDIR        runs the chain on seeded inputs and writes
 *                               OUTDIR/inputs.bin (uu, vv, ww f64; vis_real,
 *                               vis_img, weight f32), grid.bin, image.bin
 *                               (real then imaginary f64). */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "rick.h"

enum { NX = 128, NY = 96, PLANES = 4, F = 2, P = 1, SUPPORT = 7, ROWS = 20000 };

static uint64_t state = 0x2545F4914F6CDD1Dull;
static double uniform(void)
{
    state = state * 6364136223846793005ull + 1442695040888963407ull;
    return (double)(state >> 11) * (1.0 / 9007199254740992.0);
}

static int failed(const char *what)
{
    if (rick_last_status() == RICK_OK) return 0;
    fprintf(stderr, "%s: %s\n", what, rick_last_error());
    return 1;
}

static int put(FILE *f, const void *p, size_t bytes)
{
    return fwrite(p, 1, bytes, f) == bytes ? 0 : 1;
}

int main(int argc, char **argv)
{
    if (argc > 1 && strcmp(argv[1], "--version") == 0) {
        printf("%s\n", rick_version());
        return 0;
    }
    if (argc < 2) {
        fprintf(stderr, "usage: %s OUTDIR | --version\n", argv[0]);
        return 2;
    }
    const size_t n = ROWS, nv = n * F * P, cells = (size_t)2 * PLANES * NX * NY;
    double *uu = malloc(n * 8), *vv = malloc(n * 8), *ww = malloc(n * 8);
    float *vr = malloc(nv * 4), *vi = malloc(nv * 4), *wt = malloc(n * P * 4);
    double *grid = calloc(cells, 8), *gridss = calloc(cells, 8);
    double *re = calloc((size_t)NX * NY, 8), *im = calloc((size_t)NX * NY, 8);
    if (!uu || !vv || !ww || !vr || !vi || !wt || !grid || !gridss || !re || !im) return 3;
    for (size_t i = 0; i < n; i++) {
        uu[i] = 0.02 + 0.96 * uniform();
        vv[i] = 0.02 + 0.70 * uniform();   /* dx = 1/NX for both axes: v < NY/NX */
        ww[i] = 0.999 * uniform();
        wt[i] = (float)(0.5 + uniform());
    }
    for (size_t i = 0; i < nv; i++) {
        vr[i] = (float)(uniform() - 0.5);
        vi[i] = (float)(uniform() - 0.5);
    }
    const double resolution = 1.0 / (3.0 * NX), wmin = -150.0, wmax = 250.0;
    rick_set_kernel(RICK_KAISERBESSEL);
    rick_set_phase(1);
    wstack(PLANES, ROWS, F, P, uu, vv, ww, vr, vi, wt, 1.0 / NX, 1.0 / PLANES, SUPPORT, NX, NY,
           grid, 4, 0);
    if (failed("wstack")) return 1;
    rick_fft_planes(PLANES, NX, NY, NX, NY, grid, gridss, 0);
    if (failed("rick_fft_planes")) return 1;
    phase_correction(gridss, re, im, NX, NY, PLANES, NX, NY, resolution, wmin, wmax, 4, 0);
    if (failed("phase_correction")) return 1;

    char path[4096];
    int bad = 0;
    snprintf(path, sizeof path, "%s/inputs.bin", argv[1]);
    FILE *f = fopen(path, "wb");
    if (!f) return 4;
    bad |= put(f, uu, n * 8) | put(f, vv, n * 8) | put(f, ww, n * 8);
    bad |= put(f, vr, nv * 4) | put(f, vi, nv * 4) | put(f, wt, n * P * 4);
    fclose(f);
    snprintf(path, sizeof path, "%s/grid.bin", argv[1]);
    if (!(f = fopen(path, "wb"))) return 4;
    bad |= put(f, grid, cells * 8);
    fclose(f);
    snprintf(path, sizeof path, "%s/image.bin", argv[1]);
    if (!(f = fopen(path, "wb"))) return 4;
    bad |= put(f, re, (size_t)NX * NY * 8) | put(f, im, (size_t)NX * NY * 8);
    fclose(f);
    printf("dropin_caller: %dx%dx%d, %d rows, %s\n", NX, NY, PLANES, ROWS,
           bad ? "write failed" : "ok");
    return bad ? 5 : 0;
}
