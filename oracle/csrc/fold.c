/* Oracle of the device coordinate fold (TEST INFRASTRUCTURE ONLY).
 *
 * Restates fold_rescale of include/cufinufft/spreadinterp.hpp:61-89:
 *   r = fma(x, 1/(2 pi), 1/2)            rounded to nearest
 *   r = r - floor(r)                     rounded toward -inf
 *   return r * n                         rounded toward -inf
 * Here the directed roundings are done by switching the FPU rounding mode
 * with fesetround, an implementation independent of the library's
 * exact-error-term emulation. Built with -frounding-math.
 */
#include <fenv.h>
#include <math.h>
#include <stdint.h>

#pragma STDC FENV_ACCESS ON

void oracle_fold_rescale_f(const float *x, int64_t m, int n, float *out) {
  const float x2pi = 0.159154943091895345554011992339482617f;
  const int old = fegetround();
  for (int64_t i = 0; i < m; ++i) {
    fesetround(FE_TONEAREST);
    volatile float r = fmaf(x[i], x2pi, 0.5f);
    volatile float fl = floorf(r);
    fesetround(FE_DOWNWARD);
    volatile float d = r - fl;
    volatile float p = d * (float)n;
    out[i] = p;
  }
  fesetround(old);
}

void oracle_fold_rescale_d(const double *x, int64_t m, int n, double *out) {
  const double x2pi = 0.159154943091895345554011992339482617;
  const int old = fegetround();
  for (int64_t i = 0; i < m; ++i) {
    fesetround(FE_TONEAREST);
    volatile double r = fma(x[i], x2pi, 0.5);
    volatile double fl = floor(r);
    fesetround(FE_DOWNWARD);
    volatile double d = r - fl;
    volatile double p = d * (double)n;
    out[i] = p;
  }
  fesetround(old);
}
