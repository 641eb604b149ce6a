/* Extensions of the MI355X NUFFT library beyond the cuFINUFFT interface.
 *
 * Each entry point says which reference interface it mirrors, or that it has
 * none. Same conventions as cufinufft.h: device pointers, integer return
 * codes from finufft_errors.h.
 */
#ifndef FINUFFT_AMD_H
#define FINUFFT_AMD_H

#include <cufinufft.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Adjoint of the plan's transform on the plan's points (types 1 and 2): a
 * type-1 plan applies the type-2 transform with the opposite sign (c <- fk),
 * a type-2 plan the type-1 transform with the opposite sign (fk <- c).
 * GPU counterpart of the reference CPU finufft_execute_adjoint
 * (include/finufft/finufft_eitherprec.h:65). */
FINUFFT_EXPORT int cufinufft_execute_adjoint(cufinufft_plan d_plan, hipDoubleComplex *d_c,
                                             hipDoubleComplex *d_fk);
FINUFFT_EXPORT int cufinufftf_execute_adjoint(cufinufftf_plan d_plan, hipFloatComplex *d_c,
                                              hipFloatComplex *d_fk);

/* Plan introspection; no reference counterpart (the reference exposes these
 * values only through opts.debug printing, src/cuda/makeplan.cu:116-118).
 * iout[16]: type, dim, ns, nf1, nf2, nf3, gpu_method, binsize x/y/z,
 *           batchsize, M, number of bins, sub-column side of the column-window
 *           spreader (0 when unused), whole-bin parity-colored spreading (0/1),
 *           number of dense-bin subproblems.
 * dout[4]:  beta, upsampfac, 0, 0. */
FINUFFT_EXPORT int finufft_amd_plan_info(const void *plan, int is_double, int64_t *iout,
                                         double *dout);

/* Copies the point sort of the last setpts into device arrays: perm[M] (the
 * sorted order of the points) and, when binstart is not NULL,
 * binstart[nbins + 1]. No reference counterpart (the reference keeps its
 * idxnupts private); used to test the sort against the oracle. */
FINUFFT_EXPORT int finufft_amd_copy_sort(const void *plan, int is_double, int *perm,
                                         int *binstart);

/* Per-stage GPU timing of execute (HIP events on the plan's stream); no
 * reference counterpart (the reference prints CPU-side timings under
 * opts.debug). set: on != 0 enables recording. get: waits for the recorded
 * events, writes the milliseconds spent since the last get in ms[8] as
 * {zero, spread, fft, deconvolve/amplify, interp, type-3 phases, 0, 0}, and
 * resets. */
FINUFFT_EXPORT int finufft_amd_set_stage_timing(void *plan, int is_double, int on);
FINUFFT_EXPORT int finufft_amd_get_stage_times(void *plan, int is_double, double *ms);

/* ---- plan-time math, exposed for tests and tools (host only, no GPU) ---- */

/* Smallest 2,3,5-smooth multiple of f that is >= n; reference:
 * finufft::common::next235 (src/common/utils.cpp:118-121). */
FINUFFT_EXPORT int64_t finufft_amd_next235(int64_t n, int64_t f);

/* n-node Gauss-Legendre rule on [-1,1], nodes ascending; reference:
 * finufft::common::gaussquad (src/common/utils.cpp:18-58). */
FINUFFT_EXPORT void finufft_amd_gaussquad(int n, double *x, double *w);

/* Kernel width and ES beta chosen for tol at upsampfac (is_double selects the
 * precision the tolerance is held in); reference: setup_spreadinterp
 * (src/cuda/makeplan.cu:43-119). Returns an error code (7 or 8 as makeplan). */
FINUFFT_EXPORT int finufft_amd_kernel_shape(double tol, int is_double, double upsampfac,
                                            int kerevalmeth, int *ns, double *beta);

/* Horner coefficients (nc rows of ns values, row k = z^k) of the kernel for
 * sigma = 2 or 1.25; reference tables include/cufinufft/contrib/
 * ker_horner_allw_loop.inc and ker_lowupsampfac_horner_allw_loop.inc, made by
 * devel/gen_all_horner_C_code.m. Returns the number of rows written. */
FINUFFT_EXPORT int finufft_amd_horner_coeffs(int ns, double upsampfac, double *out, int nc_max);

/* The device fold of a coordinate into [0, n) with the GPU rounding rules,
 * evaluated on the host with the same code; reference: fold_rescale
 * (include/cufinufft/spreadinterp.hpp:61-89). */
FINUFFT_EXPORT void finufft_amd_fold_rescale_f(const float *x, int64_t m, int n, float *out);
FINUFFT_EXPORT void finufft_amd_fold_rescale_d(const double *x, int64_t m, int n, double *out);

/* Fine grid length for m modes (types 1/2); reference: set_nf_type12
 * (src/cuda/fft.cu:166-179). */
FINUFFT_EXPORT int64_t finufft_amd_fine_grid_len(int64_t m, double upsampfac, int ns);

#ifdef __cplusplus
}
#endif

#endif
