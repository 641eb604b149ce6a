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

/* The type-2 direction of a type 1 or 2 plan (the transform of a type-2
 * plan, the adjoint of a type-1 plan) on the modes fk, with the derivatives
 * of its output with respect to every point coordinate through the
 * derivative of the spreading kernel: c[ntransf][M] (may be NULL) and
 * dc[dim][ntransf][M], dimension 0 = x. No reference counterpart (the
 * reference has no gradient code); it is the coordinate part of the backward
 * pass of both transform types. */
FINUFFT_EXPORT int cufinufft_interp_grad(cufinufft_plan d_plan, hipDoubleComplex *d_c,
                                         hipDoubleComplex *d_dc, hipDoubleComplex *d_fk);
FINUFFT_EXPORT int cufinufftf_interp_grad(cufinufftf_plan d_plan, hipFloatComplex *d_c,
                                          hipFloatComplex *d_dc, hipFloatComplex *d_fk);

/* Fused execution of many independent problems of one type. No reference
 * counterpart: the reference runs one plan per problem. plans[nplans] are
 * type-2 plans, or type-1 plans (not the 3-D fp32 column-window spreader's),
 * (cufinufft_plan or cufinufftf_plan as is_double says) with points set,
 * sharing type, dimension, ntransf, isign, tolerance and upsampfac, made with
 * gpu_method 0 or 2. Problems of equal fine-grid shape are run in chunks
 * whose grids fit arena_bytes (0: 192 MiB, a size meant to stay in the GPU's
 * last-level cache from stage to stage); the batch allocates one chunk's
 * grids. The plans must outlive the batch and keep their points. execute
 * takes host arrays of nplans device pointers: c[p] holds the ntransf x M_p
 * values of problem p (type 2: output; type 1: input), fk[p] its ntransf x
 * N_p modes (type 2: input; type 1: output). The whole schedule is recorded
 * into a HIP graph on the first execute and replayed afterwards, on the
 * stream of plans[0]. */
typedef void *finufft_amd_batch;
FINUFFT_EXPORT int finufft_amd_batch_makeplan(int is_double, int nplans, void *const *plans,
                                              size_t arena_bytes, finufft_amd_batch *out);
FINUFFT_EXPORT int finufft_amd_batch_execute(finufft_amd_batch batch, void *const *c,
                                             void *const *fk);
FINUFFT_EXPORT int finufft_amd_batch_destroy(finufft_amd_batch batch);
/* iout[7]: number of problems, vectors per round, distinct fine-grid shapes,
 * 1 when the schedule runs as a recorded graph, when recording failed the
 * stage (1 begin, 2 a call inside the capture, 3 end, 4 instantiate) and its
 * HIP error code, and the number of chunks. use_graph = 0 runs direct
 * launches instead (for comparison). */
FINUFFT_EXPORT int finufft_amd_batch_info(finufft_amd_batch batch, int64_t *iout);
FINUFFT_EXPORT int finufft_amd_batch_set_graph(finufft_amd_batch batch, int use_graph);

/* The learned cost model's network (no reference counterpart; the
 * reference picks methods from fixed tables, src/cuda/heuristics.cu:270-280).
 * A multilayer perceptron with FINUFFT_AMD_MLP_IN inputs, two ReLU layers of
 * FINUFFT_AMD_MLP_HIDDEN units and FINUFFT_AMD_MLP_OUT linear outputs; the
 * parameters are one packed fp32 array W1[H][IN] b1[H] W2[H][H] b2[H]
 * W3[OUT][H] b3[OUT]. All arrays are device arrays, samples row-major;
 * stream is a hipStream_t (NULL: the null stream).
 *   forward: y[n][OUT] from x[n][IN];
 *   grad:    adds to grad[NPARAM] and *loss the gradient and value of
 *            mean_b 0.5 sum_j mask[b][j] (y[b][j] - target[b][j])^2 (grad and
 *            loss must be zeroed by the caller);
 *   adam:    one Adam step (beta1 0.9, beta2 0.999, eps 1e-8) at step >= 1. */
#define FINUFFT_AMD_MLP_IN 11
#define FINUFFT_AMD_MLP_HIDDEN 32
#define FINUFFT_AMD_MLP_OUT 8
#define FINUFFT_AMD_MLP_NPARAM                                                       \
  (FINUFFT_AMD_MLP_HIDDEN * FINUFFT_AMD_MLP_IN + FINUFFT_AMD_MLP_HIDDEN +            \
   FINUFFT_AMD_MLP_HIDDEN * FINUFFT_AMD_MLP_HIDDEN + FINUFFT_AMD_MLP_HIDDEN +        \
   FINUFFT_AMD_MLP_OUT * FINUFFT_AMD_MLP_HIDDEN + FINUFFT_AMD_MLP_OUT)
FINUFFT_EXPORT int finufft_amd_mlp_forward(const float *params, int n, const float *x, float *y,
                                           void *stream);
FINUFFT_EXPORT int finufft_amd_mlp_grad(const float *params, int n, const float *x,
                                        const float *target, const float *mask, float *grad,
                                        float *loss, void *stream);
FINUFFT_EXPORT int finufft_amd_mlp_adam(float *params, const float *grad, float *m1, float *m2,
                                        int step, float lr, void *stream);

/* Point-density statistics for the cost model; no reference counterpart
 * (the reference's method tables do not look at the points,
 * src/cuda/heuristics.cu:270-281). x, y, z: device coordinate arrays of M
 * points (fp32, or fp64 with is_double), as for setpts. Each coordinate is
 * folded into [0, g) and the points counted per cell of the g^dim grid
 * (g^dim <= 2^20); stats[3] receives max / mean, 99th percentile / mean and
 * the occupied fraction of the cell counts. Synchronizes `stream` (a
 * hipStream_t, NULL: the null stream). */
FINUFFT_EXPORT int finufft_amd_density_stats(int dim, int64_t M, const void *x, const void *y,
                                             const void *z, int is_double, int g,
                                             double *stats, void *stream);

/* The statistics of finufft_amd_density_stats without allocating: no
 * counterpart in the reference; for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops). The counts
 * live in a library-owned device array of 2^18 cells (1 MiB per GPU, held
 * until the process ends), so g^dim <= 2^18 (else 21, invalid argument); the
 * three statistics are written to the DEVICE array d_stats[3] on `stream`,
 * with no host synchronization. Calls that overlap on one GPU share the
 * array and race: serialize them (one stream). */
FINUFFT_EXPORT int finufft_amd_density_stats_noalloc(int dim, int64_t M, const void *x,
                                                     const void *y, const void *z, int is_double,
                                                     int g, double *d_stats, void *stream);

/* Plan introspection; no reference counterpart (the reference exposes these
 * values only through opts.debug printing, src/cuda/makeplan.cu:116-118).
 * iout[16]: type, dim, ns, nf1, nf2, nf3, gpu_method, binsize x/y/z,
 *           batchsize, M, number of bins, sub-column side of the column-window
 *           spreader (0 when unused), whole-bin parity-colored spreading (0/1),
 *           number of dense-bin subproblems.
 * dout[4]:  beta, upsampfac, points staged per batch by output-driven
 *           spreading (gpu_method 3; else 0), 1 when the grid <-> modes
 *           transforms run as pruned line passes (else rocFFT, 0). */
/* Grid <-> modes transforms of the plans made after the call (no reference
 * counterpart; the reference always runs a full cuFFT transform plus a
 * separate deconvolution / amplification pass, src/cuda/execute.cu:105-204).
 * path 0 (the default): on 3-D grids whose sides are powers of two in
 * [16, 4096], three pruned line-transform passes with the deconvolution or
 * amplification folded in; rocFFT otherwise. path 1: rocFFT and the separate
 * pass for every plan. Process-wide; other values -> 21 (invalid argument). */
FINUFFT_EXPORT int finufft_amd_set_fft_path(int path);

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

/* LDS / VALU throughput probe on the current device; no reference
 * counterpart (a measurement helper: the spreading kernels' choice of LDS
 * operations rests on these rates). mode: 0 ds_add_f32, 1 ds_add_f64,
 * 2 ds_add_u32, 3 float read-modify-write (ds_read + v_add + ds_write),
 * 4 ds_read_b32, 5 ds_read_b32 at one address per wave, 6 ds_read_b128,
 * 7 ds_read_b128 at one address per wave, 8 v_fma_f32. Runs blocks
 * workgroups of 256 lanes, reps iterations each; writes the elapsed
 * milliseconds (the single-iteration launch subtracted) and the number of
 * wave-level operations timed. */
FINUFFT_EXPORT int finufft_amd_lds_probe(int mode, int blocks, int reps, double *ms,
                                         double *wave_ops);

/* ---- plan-time math, exposed for tests and tools (host only, no GPU) ---- */

/* Smallest 2,3,5-smooth multiple of f that is >= n; reference:
 * finufft::common::next235 (src/common/utils.cpp:118-121). */
FINUFFT_EXPORT int64_t finufft_amd_next235(int64_t n, int64_t f);

/* n-node Gauss-Legendre rule on [-1,1], nodes ascending; reference:
 * finufft::common::gaussquad (src/common/utils.cpp:18-58). */
FINUFFT_EXPORT void finufft_amd_gaussquad(int n, double *x, double *w);

/* Kernel width, shape parameter and family (0 ES, 1 PSWF) that a plan of
 * this dimension and type chooses for tol at upsampfac (is_double selects the
 * precision the tolerance is held in). kerevalmeth 1 gives the PSWF kernel
 * with the width law of the reference CPU library (theoretical_kernel_ns and
 * set_kernel_shape_given_ns, src/common/kernel.cpp:60-150); 0 and 2 give the
 * ES kernel with the reference GPU library's rule (setup_spreadinterp,
 * src/cuda/makeplan.cu:43-119). Returns an error code (7, 8 or 24 as
 * makeplan). */
FINUFFT_EXPORT int finufft_amd_kernel_shape(double tol, int is_double, double upsampfac,
                                            int kerevalmeth, int dim, int type, int *ns,
                                            double *beta, int *formula);

/* Horner coefficients (nc rows of ns values, row k = z^k) of the ES kernel for
 * sigma = 2 or 1.25; reference tables include/cufinufft/contrib/
 * ker_horner_allw_loop.inc and ker_lowupsampfac_horner_allw_loop.inc, made by
 * devel/gen_all_horner_C_code.m. Returns the number of rows written. */
FINUFFT_EXPORT int finufft_amd_horner_coeffs(int ns, double upsampfac, double *out, int nc_max);

/* Horner coefficients (same layout) of the PSWF kernel of shape parameter
 * beta; reference: precompute_horner_coeffs (include/finufft/makeplan.hpp:
 * 204-315). Returns the number of rows written. */
FINUFFT_EXPORT int finufft_amd_horner_coeffs_pswf(int ns, double beta, double *out, int nc_max);

/* psi_0^c(z) normalised to 1 at z = 0 (0 for |z| > 1) at n points; reference:
 * finufft::common::PSWF0 (include/finufft_common/pswf.h:25-61,
 * src/common/pswf.cpp). */
FINUFFT_EXPORT void finufft_amd_pswf(double c, const double *z, int64_t n, double *out);

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
