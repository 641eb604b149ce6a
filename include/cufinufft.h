/* Public C interface of the MI355X NUFFT library: a drop-in replacement for the
 * cuFINUFFT guru and simple interfaces.
 *
 * Every function keeps the name, argument order and meaning of the reference
 * declaration it replaces (include/cufinufft.h:16-186 of flatironinstitute/
 * finufft); only cuFloatComplex / cuDoubleComplex become hipFloatComplex /
 * hipDoubleComplex, which have the same 8 / 16 byte layout. Pointers to points,
 * strengths and modes are device pointers. Return values are the codes of
 * finufft_errors.h (0 = success).
 *
 * C callers compile with hipcc, or with a host compiler and
 * -D__HIP_PLATFORM_AMD__ -I/opt/rocm/include.
 */
#ifndef CUFINUFFT_H
#define CUFINUFFT_H

#include <hip/hip_complex.h>
#include <stdint.h>

#include <cufinufft_opts.h>
#include <finufft_errors.h>

#define FINUFFT_EXPORT __attribute__((visibility("default")))

typedef struct cufinufft_plan_s *cufinufft_plan;
typedef struct cufinufft_fplan_s *cufinufftf_plan;

#ifdef __cplusplus
extern "C" {
#endif

/* reference: include/cufinufft.h:18 (defaults at src/cuda/c_interface.cpp:153-172) */
FINUFFT_EXPORT void cufinufft_default_opts(cufinufft_opts *opts);

/* reference: include/cufinufft.h:20-25 */
FINUFFT_EXPORT int cufinufft_makeplan(int type, int dim, const int64_t *n_modes, int iflag,
                                      int ntr, double eps, cufinufft_plan *d_plan_ptr,
                                      const cufinufft_opts *opts);
FINUFFT_EXPORT int cufinufftf_makeplan(int type, int dim, const int64_t *n_modes, int iflag,
                                       int ntr, float eps, cufinufftf_plan *d_plan_ptr,
                                       const cufinufft_opts *opts);

/* reference: include/cufinufft.h:27-32 */
FINUFFT_EXPORT int cufinufft_setpts(cufinufft_plan d_plan, int64_t M, const double *d_x,
                                    const double *d_y, const double *d_z, int N,
                                    const double *d_s, const double *d_t, const double *d_u);
FINUFFT_EXPORT int cufinufftf_setpts(cufinufftf_plan d_plan, int64_t M, const float *d_x,
                                     const float *d_y, const float *d_z, int N,
                                     const float *d_s, const float *d_t, const float *d_u);

/* reference: include/cufinufft.h:34-37 */
FINUFFT_EXPORT int cufinufft_execute(cufinufft_plan d_plan, hipDoubleComplex *d_c,
                                     hipDoubleComplex *d_fk);
FINUFFT_EXPORT int cufinufftf_execute(cufinufftf_plan d_plan, hipFloatComplex *d_c,
                                      hipFloatComplex *d_fk);

/* reference: include/cufinufft.h:39-40 */
FINUFFT_EXPORT int cufinufft_destroy(cufinufft_plan d_plan);
FINUFFT_EXPORT int cufinufftf_destroy(cufinufftf_plan d_plan);

/* Simple (one-shot) interfaces: 3 dims x 3 types x {single, many} x {double,
 * float}; reference: include/cufinufft.h:45-186. Each makes a plan, sets the
 * points, executes once and destroys the plan. */
/* dimension 1: reference include/cufinufft.h:45-89 */
FINUFFT_EXPORT int cufinufft1d1many(int n_transf, int64_t nj, const double *xj, const hipDoubleComplex *cj, int iflag, double eps, int64_t ms, hipDoubleComplex *fk, const cufinufft_opts *opts);
FINUFFT_EXPORT int cufinufftf1d1many(int n_transf, int64_t nj, const float *xj, const hipFloatComplex *cj, int iflag, float eps, int64_t ms, hipFloatComplex *fk, const cufinufft_opts *opts);
FINUFFT_EXPORT int cufinufft1d1(int64_t nj, const double *xj, const hipDoubleComplex *cj, int iflag, double eps, int64_t ms, hipDoubleComplex *fk, const cufinufft_opts *opts);
FINUFFT_EXPORT int cufinufftf1d1(int64_t nj, const float *xj, const hipFloatComplex *cj, int iflag, float eps, int64_t ms, hipFloatComplex *fk, const cufinufft_opts *opts);
FINUFFT_EXPORT int cufinufft1d2many(int n_transf, int64_t nj, const double *xj, hipDoubleComplex *cj, int iflag, double eps, int64_t ms, const hipDoubleComplex *fk, const cufinufft_opts *opts);
FINUFFT_EXPORT int cufinufftf1d2many(int n_transf, int64_t nj, const float *xj, hipFloatComplex *cj, int iflag, float eps, int64_t ms, const hipFloatComplex *fk, const cufinufft_opts *opts);
FINUFFT_EXPORT int cufinufft1d2(int64_t nj, const double *xj, hipDoubleComplex *cj, int iflag, double eps, int64_t ms, const hipDoubleComplex *fk, const cufinufft_opts *opts);
FINUFFT_EXPORT int cufinufftf1d2(int64_t nj, const float *xj, hipFloatComplex *cj, int iflag, float eps, int64_t ms, const hipFloatComplex *fk, const cufinufft_opts *opts);
FINUFFT_EXPORT int cufinufft1d3many(int n_transf, int64_t nj, const double *xj, const hipDoubleComplex *cj, int iflag, double eps, int64_t nk, const double *s, hipDoubleComplex *fk, const cufinufft_opts *opts);
FINUFFT_EXPORT int cufinufftf1d3many(int n_transf, int64_t nj, const float *xj, const hipFloatComplex *cj, int iflag, float eps, int64_t nk, const float *s, hipFloatComplex *fk, const cufinufft_opts *opts);
FINUFFT_EXPORT int cufinufft1d3(int64_t nj, const double *xj, const hipDoubleComplex *cj, int iflag, double eps, int64_t nk, const double *s, hipDoubleComplex *fk, const cufinufft_opts *opts);
FINUFFT_EXPORT int cufinufftf1d3(int64_t nj, const float *xj, const hipFloatComplex *cj, int iflag, float eps, int64_t nk, const float *s, hipFloatComplex *fk, const cufinufft_opts *opts);
/* dimension 2: reference include/cufinufft.h:93-137 */
FINUFFT_EXPORT int cufinufft2d1many(int n_transf, int64_t nj, const double *xj, const double *yj, const hipDoubleComplex *cj, int iflag, double eps, int64_t ms, int64_t mt, hipDoubleComplex *fk, const cufinufft_opts *opts);
FINUFFT_EXPORT int cufinufftf2d1many(int n_transf, int64_t nj, const float *xj, const float *yj, const hipFloatComplex *cj, int iflag, float eps, int64_t ms, int64_t mt, hipFloatComplex *fk, const cufinufft_opts *opts);
FINUFFT_EXPORT int cufinufft2d1(int64_t nj, const double *xj, const double *yj, const hipDoubleComplex *cj, int iflag, double eps, int64_t ms, int64_t mt, hipDoubleComplex *fk, const cufinufft_opts *opts);
FINUFFT_EXPORT int cufinufftf2d1(int64_t nj, const float *xj, const float *yj, const hipFloatComplex *cj, int iflag, float eps, int64_t ms, int64_t mt, hipFloatComplex *fk, const cufinufft_opts *opts);
FINUFFT_EXPORT int cufinufft2d2many(int n_transf, int64_t nj, const double *xj, const double *yj, hipDoubleComplex *cj, int iflag, double eps, int64_t ms, int64_t mt, const hipDoubleComplex *fk, const cufinufft_opts *opts);
FINUFFT_EXPORT int cufinufftf2d2many(int n_transf, int64_t nj, const float *xj, const float *yj, hipFloatComplex *cj, int iflag, float eps, int64_t ms, int64_t mt, const hipFloatComplex *fk, const cufinufft_opts *opts);
FINUFFT_EXPORT int cufinufft2d2(int64_t nj, const double *xj, const double *yj, hipDoubleComplex *cj, int iflag, double eps, int64_t ms, int64_t mt, const hipDoubleComplex *fk, const cufinufft_opts *opts);
FINUFFT_EXPORT int cufinufftf2d2(int64_t nj, const float *xj, const float *yj, hipFloatComplex *cj, int iflag, float eps, int64_t ms, int64_t mt, const hipFloatComplex *fk, const cufinufft_opts *opts);
FINUFFT_EXPORT int cufinufft2d3many(int n_transf, int64_t nj, const double *xj, const double *yj, const hipDoubleComplex *cj, int iflag, double eps, int64_t nk, const double *s, const double *t, hipDoubleComplex *fk, const cufinufft_opts *opts);
FINUFFT_EXPORT int cufinufftf2d3many(int n_transf, int64_t nj, const float *xj, const float *yj, const hipFloatComplex *cj, int iflag, float eps, int64_t nk, const float *s, const float *t, hipFloatComplex *fk, const cufinufft_opts *opts);
FINUFFT_EXPORT int cufinufft2d3(int64_t nj, const double *xj, const double *yj, const hipDoubleComplex *cj, int iflag, double eps, int64_t nk, const double *s, const double *t, hipDoubleComplex *fk, const cufinufft_opts *opts);
FINUFFT_EXPORT int cufinufftf2d3(int64_t nj, const float *xj, const float *yj, const hipFloatComplex *cj, int iflag, float eps, int64_t nk, const float *s, const float *t, hipFloatComplex *fk, const cufinufft_opts *opts);
/* dimension 3: reference include/cufinufft.h:141-185 */
FINUFFT_EXPORT int cufinufft3d1many(int n_transf, int64_t nj, const double *xj, const double *yj, const double *zj, const hipDoubleComplex *cj, int iflag, double eps, int64_t ms, int64_t mt, int64_t mu, hipDoubleComplex *fk, const cufinufft_opts *opts);
FINUFFT_EXPORT int cufinufftf3d1many(int n_transf, int64_t nj, const float *xj, const float *yj, const float *zj, const hipFloatComplex *cj, int iflag, float eps, int64_t ms, int64_t mt, int64_t mu, hipFloatComplex *fk, const cufinufft_opts *opts);
FINUFFT_EXPORT int cufinufft3d1(int64_t nj, const double *xj, const double *yj, const double *zj, const hipDoubleComplex *cj, int iflag, double eps, int64_t ms, int64_t mt, int64_t mu, hipDoubleComplex *fk, const cufinufft_opts *opts);
FINUFFT_EXPORT int cufinufftf3d1(int64_t nj, const float *xj, const float *yj, const float *zj, const hipFloatComplex *cj, int iflag, float eps, int64_t ms, int64_t mt, int64_t mu, hipFloatComplex *fk, const cufinufft_opts *opts);
FINUFFT_EXPORT int cufinufft3d2many(int n_transf, int64_t nj, const double *xj, const double *yj, const double *zj, hipDoubleComplex *cj, int iflag, double eps, int64_t ms, int64_t mt, int64_t mu, const hipDoubleComplex *fk, const cufinufft_opts *opts);
FINUFFT_EXPORT int cufinufftf3d2many(int n_transf, int64_t nj, const float *xj, const float *yj, const float *zj, hipFloatComplex *cj, int iflag, float eps, int64_t ms, int64_t mt, int64_t mu, const hipFloatComplex *fk, const cufinufft_opts *opts);
FINUFFT_EXPORT int cufinufft3d2(int64_t nj, const double *xj, const double *yj, const double *zj, hipDoubleComplex *cj, int iflag, double eps, int64_t ms, int64_t mt, int64_t mu, const hipDoubleComplex *fk, const cufinufft_opts *opts);
FINUFFT_EXPORT int cufinufftf3d2(int64_t nj, const float *xj, const float *yj, const float *zj, hipFloatComplex *cj, int iflag, float eps, int64_t ms, int64_t mt, int64_t mu, const hipFloatComplex *fk, const cufinufft_opts *opts);
FINUFFT_EXPORT int cufinufft3d3many(int n_transf, int64_t nj, const double *xj, const double *yj, const double *zj, const hipDoubleComplex *cj, int iflag, double eps, int64_t nk, const double *s, const double *t, const double *u, hipDoubleComplex *fk, const cufinufft_opts *opts);
FINUFFT_EXPORT int cufinufftf3d3many(int n_transf, int64_t nj, const float *xj, const float *yj, const float *zj, const hipFloatComplex *cj, int iflag, float eps, int64_t nk, const float *s, const float *t, const float *u, hipFloatComplex *fk, const cufinufft_opts *opts);
FINUFFT_EXPORT int cufinufft3d3(int64_t nj, const double *xj, const double *yj, const double *zj, const hipDoubleComplex *cj, int iflag, double eps, int64_t nk, const double *s, const double *t, const double *u, hipDoubleComplex *fk, const cufinufft_opts *opts);
FINUFFT_EXPORT int cufinufftf3d3(int64_t nj, const float *xj, const float *yj, const float *zj, const hipFloatComplex *cj, int iflag, float eps, int64_t nk, const float *s, const float *t, const float *u, hipFloatComplex *fk, const cufinufft_opts *opts);

#ifdef __cplusplus
}
#endif

#endif
