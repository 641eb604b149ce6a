/* A C caller of the public interface (include/cufinufft.h), compiled as C
 * with the host compiler: the guru sequence makeplan / setpts / execute /
 * destroy and the one-shot 3-D type-1 call in single precision, and a 1-D
 * type-2 one-shot call in double precision, each checked against a direct
 * sum at a few outputs; then the error cases of the reference's
 * test/cuda/cufinufft_error_handling.cu. Exits 0 and prints PASS when every
 * check holds. The reference's own C caller of the same interface is
 * test/cuda/public_api_test.c. */
#include <cufinufft.h>
#include <hip/hip_runtime_api.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#define PI_D 3.14159265358979323846

#define CHECK(cond, what)                          \
  do {                                             \
    if (!(cond)) {                                 \
      fprintf(stderr, "FAIL: %s\n", what);         \
      return 1;                                    \
    }                                              \
  } while (0)

static double urand(unsigned *s) {
  *s = *s * 1664525u + 1013904223u;
  return (double)(*s >> 8) / 16777216.0;
}

static int type1_3d_float(void) {
  const int64_t M = 5000, N[3] = {12, 10, 8};
  const int64_t NT = N[0] * N[1] * N[2];
  unsigned seed = 1;
  float *x = malloc(sizeof(float) * M), *y = malloc(sizeof(float) * M),
        *z = malloc(sizeof(float) * M);
  hipFloatComplex *c = malloc(sizeof(hipFloatComplex) * M);
  hipFloatComplex *fk = malloc(sizeof(hipFloatComplex) * NT);
  for (int64_t j = 0; j < M; ++j) {
    x[j] = (float)(PI_D * (2 * urand(&seed) - 1));
    y[j] = (float)(PI_D * (2 * urand(&seed) - 1));
    z[j] = (float)(PI_D * (2 * urand(&seed) - 1));
    c[j] = make_hipFloatComplex((float)(2 * urand(&seed) - 1), (float)(2 * urand(&seed) - 1));
  }
  float *dx, *dy, *dz;
  hipFloatComplex *dc, *dfk;
  CHECK(hipMalloc((void **)&dx, sizeof(float) * M) == hipSuccess, "hipMalloc");
  CHECK(hipMalloc((void **)&dy, sizeof(float) * M) == hipSuccess, "hipMalloc");
  CHECK(hipMalloc((void **)&dz, sizeof(float) * M) == hipSuccess, "hipMalloc");
  CHECK(hipMalloc((void **)&dc, sizeof(hipFloatComplex) * M) == hipSuccess, "hipMalloc");
  CHECK(hipMalloc((void **)&dfk, sizeof(hipFloatComplex) * NT) == hipSuccess, "hipMalloc");
  hipMemcpy(dx, x, sizeof(float) * M, hipMemcpyHostToDevice);
  hipMemcpy(dy, y, sizeof(float) * M, hipMemcpyHostToDevice);
  hipMemcpy(dz, z, sizeof(float) * M, hipMemcpyHostToDevice);
  hipMemcpy(dc, c, sizeof(hipFloatComplex) * M, hipMemcpyHostToDevice);

  for (int pass = 0; pass < 2; ++pass) {
    if (pass == 0) { /* guru interface */
      cufinufftf_plan plan;
      cufinufft_opts opts;
      cufinufft_default_opts(&opts);
      CHECK(cufinufftf_makeplan(1, 3, N, 1, 1, 1e-5f, &plan, &opts) == 0, "makeplan");
      CHECK(cufinufftf_setpts(plan, M, dx, dy, dz, 0, NULL, NULL, NULL) == 0, "setpts");
      CHECK(cufinufftf_execute(plan, dc, dfk) == 0, "execute");
      CHECK(cufinufftf_destroy(plan) == 0, "destroy");
    } else { /* one-shot */
      CHECK(cufinufftf3d1(M, dx, dy, dz, dc, 1, 1e-5f, N[0], N[1], N[2], dfk, NULL) == 0,
            "cufinufftf3d1");
    }
    hipMemcpy(fk, dfk, sizeof(hipFloatComplex) * NT, hipMemcpyDeviceToHost);
    /* modes k = -N/2 .. , x fastest; f[k] = sum_j c_j exp(+i k.x_j) */
    double num = 0, den = 0;
    for (int t = 0; t < 24; ++t) {
      const int64_t m = (int64_t)(urand(&seed) * NT) % NT;
      const int64_t k1 = m % N[0] - N[0] / 2, k2 = (m / N[0]) % N[1] - N[1] / 2,
                    k3 = m / (N[0] * N[1]) - N[2] / 2;
      double re = 0, im = 0;
      for (int64_t j = 0; j < M; ++j) {
        const double ph = k1 * (double)x[j] + k2 * (double)y[j] + k3 * (double)z[j];
        const double cr = hipCrealf(c[j]), ci = hipCimagf(c[j]);
        re += cr * cos(ph) - ci * sin(ph);
        im += cr * sin(ph) + ci * cos(ph);
      }
      const double dr = hipCrealf(fk[m]) - re, di = hipCimagf(fk[m]) - im;
      num += dr * dr + di * di;
      den += re * re + im * im;
    }
    const double err = sqrt(num / den);
    printf("3-D type 1 fp32 (%s): relerr %.2e\n", pass ? "one-shot" : "guru", err);
    CHECK(err < 25 * 1e-5, "3-D type 1 accuracy");
  }
  hipFree(dx);
  hipFree(dy);
  hipFree(dz);
  hipFree(dc);
  hipFree(dfk);
  free(x);
  free(y);
  free(z);
  free(c);
  free(fk);
  return 0;
}

static int type2_1d_double(void) {
  const int64_t M = 3000, N = 64;
  unsigned seed = 7;
  double *x = malloc(sizeof(double) * M);
  hipDoubleComplex *fk = malloc(sizeof(hipDoubleComplex) * N);
  hipDoubleComplex *c = malloc(sizeof(hipDoubleComplex) * M);
  for (int64_t j = 0; j < M; ++j) x[j] = PI_D * (2 * urand(&seed) - 1);
  for (int64_t k = 0; k < N; ++k)
    fk[k] = make_hipDoubleComplex(2 * urand(&seed) - 1, 2 * urand(&seed) - 1);
  double *dx;
  hipDoubleComplex *dc, *dfk;
  CHECK(hipMalloc((void **)&dx, sizeof(double) * M) == hipSuccess, "hipMalloc");
  CHECK(hipMalloc((void **)&dc, sizeof(hipDoubleComplex) * M) == hipSuccess, "hipMalloc");
  CHECK(hipMalloc((void **)&dfk, sizeof(hipDoubleComplex) * N) == hipSuccess, "hipMalloc");
  hipMemcpy(dx, x, sizeof(double) * M, hipMemcpyHostToDevice);
  hipMemcpy(dfk, fk, sizeof(hipDoubleComplex) * N, hipMemcpyHostToDevice);
  CHECK(cufinufft1d2(M, dx, dc, -1, 1e-10, N, dfk, NULL) == 0, "cufinufft1d2");
  hipMemcpy(c, dc, sizeof(hipDoubleComplex) * M, hipMemcpyDeviceToHost);
  double num = 0, den = 0;
  for (int64_t j = 0; j < M; j += 97) {
    double re = 0, im = 0;
    for (int64_t k = 0; k < N; ++k) {
      const double ph = -(double)(k - N / 2) * x[j];
      re += hipCreal(fk[k]) * cos(ph) - hipCimag(fk[k]) * sin(ph);
      im += hipCreal(fk[k]) * sin(ph) + hipCimag(fk[k]) * cos(ph);
    }
    const double dr = hipCreal(c[j]) - re, di = hipCimag(c[j]) - im;
    num += dr * dr + di * di;
    den += re * re + im * im;
  }
  const double err = sqrt(num / den);
  printf("1-D type 2 fp64 (one-shot): relerr %.2e\n", err);
  CHECK(err < 25 * 1e-10, "1-D type 2 accuracy");
  hipFree(dx);
  hipFree(dc);
  hipFree(dfk);
  free(x);
  free(fk);
  free(c);
  return 0;
}

static int errors(void) {
  const int64_t N[3] = {16, 16, 16};
  cufinufftf_plan plan = NULL;
  cufinufft_opts opts;
  CHECK(cufinufftf_makeplan(1, 0, N, 1, 1, 1e-5f, &plan, NULL) == FINUFFT_ERR_DIM_NOTVALID, "dim");
  CHECK(cufinufftf_makeplan(4, 1, N, 1, 1, 1e-5f, &plan, NULL) == FINUFFT_ERR_TYPE_NOTVALID,
        "type");
  CHECK(cufinufftf_makeplan(1, 1, N, 1, 0, 1e-5f, &plan, NULL) == FINUFFT_ERR_NTRANS_NOTVALID,
        "ntrans");
  const int64_t Nbad[3] = {0, 1, 1};
  CHECK(cufinufftf_makeplan(1, 1, Nbad, 1, 1, 1e-5f, &plan, NULL) == FINUFFT_ERR_NDATA_NOTVALID,
        "nmodes");
  cufinufft_default_opts(&opts);
  opts.upsampfac = 0.9;
  opts.gpu_kerevalmeth = 0;
  CHECK(cufinufftf_makeplan(1, 2, N, 1, 1, 1e-5f, &plan, &opts) ==
            FINUFFT_ERR_UPSAMPFAC_TOO_SMALL,
        "upsampfac");
  opts.gpu_kerevalmeth = 1;
  CHECK(cufinufftf_makeplan(1, 2, N, 1, 1, 1e-5f, &plan, &opts) == FINUFFT_ERR_HORNER_WRONG_BETA,
        "horner beta");
  cufinufft_default_opts(&opts);
  opts.gpu_method = 2;
  opts.gpu_binsizex = opts.gpu_binsizey = opts.gpu_binsizez = 1024;
  CHECK(cufinufftf_makeplan(1, 3, N, 1, 1, 1e-5f, &plan, &opts) ==
            FINUFFT_ERR_INSUFFICIENT_SHMEM,
        "shmem");
  CHECK(cufinufft_destroy(NULL) == FINUFFT_ERR_PLAN_NOTVALID, "destroy null");
  CHECK(cufinufftf_makeplan(3, 1, N, 1, 1, 1e-5f, &plan, NULL) == 0, "makeplan type 3");
  CHECK(cufinufftf_setpts(plan, 1, NULL, NULL, NULL, 1, NULL, NULL, NULL) ==
            FINUFFT_ERR_INVALID_ARGUMENT,
        "type 3 null targets");
  cufinufftf_destroy(plan);
  CHECK(cufinufftf_makeplan(3, 1, N, 1, 1, 1e-5f, &plan, NULL) == 0, "makeplan type 3");
  CHECK(cufinufftf_setpts(plan, 1, NULL, NULL, NULL, -1, NULL, NULL, NULL) ==
            FINUFFT_ERR_NUM_NU_PTS_INVALID,
        "type 3 negative N");
  cufinufftf_destroy(plan);
  printf("error codes as the reference\n");
  return 0;
}

int main(void) {
  if (type1_3d_float() || type2_1d_double() || errors()) return 1;
  printf("PASS\n");
  return 0;
}
