/* Compiles the reference's own w-stacking.h and this library's rick.h in one
 * translation unit: C accepts a repeated function declaration only when the
 * two prototypes are compatible, so this builds only while wstack() and
 * phase_correction() keep the reference's C signatures (w-stacking.h:39-57,
 * 70-83).  RICK_PLANT_MISMATCH adds a deliberately wrong prototype, which
 * must fail to compile (the check has teeth). */
#include "w-stacking.h"
#include "rick.h"

#ifdef RICK_PLANT_MISMATCH
void phase_correction(double *, double *, double *, int, int, int, int, int, double, double,
                      float, int, int);
#endif

int main(void)
{
    void (*w)(int, unsigned int, unsigned int, unsigned int, double *, double *, double *,
              float *, float *, float *, double, double, int, int, int, double *, int, int) =
        wstack;
    void (*p)(double *, double *, double *, int, int, int, int, int, double, double, double,
              int, int) = phase_correction;
    return (w != 0 && p != 0) ? 0 : 1;
}
