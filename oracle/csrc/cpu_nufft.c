/* CPU restatement of the NUFFT spread / interpolate / deconvolve stages
 * (TEST INFRASTRUCTURE and CPU BASELINE ONLY; never linked into the library).
 *
 * Follows the reference algorithm: points folded to [0, nf) (the plain fold of
 * include/finufft/simd.hpp:318-326, evaluated in double), first node
 * s = ceil(x - ns/2), kernel values from the ES kernel by Horner polynomials
 * (coefficients passed in, row k = z^k, as include/cufinufft/spreadinterp.hpp:
 * 113-134) and tensor-product spreading with periodic wrap
 * (include/finufft/spreadinterp.hpp); deconvolution divides by the kernel's
 * Fourier series (src/cuda/execute.cu:24-63, modeord 0).
 *
 * Threads: interpolation is parallel over points. Spreading splits the
 * slowest grid axis into slabs of SLAB planes; points are bucketed by slab and
 * even slabs are spread in parallel, then odd slabs, so that no two threads
 * ever write the same plane (SLAB > 2 ceil(ns/2)).
 * Grids and strengths are interleaved complex float (re, im).
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define MAXNS 16
#define SLAB 20

static inline double fold01(double x, int n) {
  const double t = x * 0.15915494309189533577 + 0.5;
  double r = (t - floor(t)) * n;
  if (r >= n) r -= n;
  if (!(r >= 0.0 && r < n)) r = 0.0;
  return r;
}

static inline void kernel_vals(double r, int ns, int nc, const double *coef, int *s0,
                               float *ker) {
  const int s = (int)ceil(r - 0.5 * ns);
  const double z = 2.0 * ((double)s - r) + (ns - 1);
  for (int i = 0; i < ns; ++i) {
    double v = coef[(nc - 1) * ns + i];
    for (int k = nc - 2; k >= 0; --k) v = v * z + coef[k * ns + i];
    ker[i] = (float)v;
  }
  *s0 = s;
}

static void spread_one(int dim, const int *nf, const double *r, const float *cv, int ns, int nc,
                       const double *coef, float *fw) {
  float ker[3][MAXNS];
  int st[3] = {0, 0, 0};
  for (int d = 0; d < dim; ++d) kernel_vals(r[d], ns, nc, coef, &st[d], ker[d]);
  int ix[MAXNS];
  for (int a = 0; a < ns; ++a) {
    int i = st[0] + a;
    ix[a] = i < 0 ? i + nf[0] : (i >= nf[0] ? i - nf[0] : i);
  }
  const int nz = dim > 2 ? ns : 1, ny = dim > 1 ? ns : 1;
  for (int e = 0; e < nz; ++e) {
    int iz = dim > 2 ? st[2] + e : 0;
    if (dim > 2) iz = iz < 0 ? iz + nf[2] : (iz >= nf[2] ? iz - nf[2] : iz);
    const float kz = dim > 2 ? ker[2][e] : 1.0f;
    for (int b = 0; b < ny; ++b) {
      int iy = dim > 1 ? st[1] + b : 0;
      if (dim > 1) iy = iy < 0 ? iy + nf[1] : (iy >= nf[1] ? iy - nf[1] : iy);
      const float w = kz * (dim > 1 ? ker[1][b] : 1.0f);
      const float wr = cv[0] * w, wi = cv[1] * w;
      float *row = fw + 2 * ((int64_t)iz * nf[1] + iy) * nf[0];
      for (int a = 0; a < ns; ++a) {
        row[2 * ix[a]] += wr * ker[0][a];
        row[2 * ix[a] + 1] += wi * ker[0][a];
      }
    }
  }
}

/* fw (zeroed by the caller, nf[0]*nf[1]*nf[2] complex) += spread of c. */
void cpu_spread_f(int dim, int64_t M, const float *x, const float *y, const float *z,
                  const float *c, const int *nf, int ns, int nc, const double *coef, float *fw,
                  int nthreads) {
  const float *xyz[3] = {x, y, z};
  const int slow = dim - 1;
  const int nslab = (nf[slow] + SLAB - 1) / SLAB;
  int *slab = (int *)malloc(sizeof(int) * M);
  int64_t *start = (int64_t *)calloc(nslab + 1, sizeof(int64_t));
  int64_t *order = (int64_t *)malloc(sizeof(int64_t) * M);
  for (int64_t j = 0; j < M; ++j) {
    const double r = fold01((double)xyz[slow][j], nf[slow]);
    int s = (int)ceil(r - 0.5 * ns);
    if (s < 0) s += nf[slow];
    slab[j] = s / SLAB;
    start[slab[j] + 1]++;
  }
  for (int k = 0; k < nslab; ++k) start[k + 1] += start[k];
  {
    int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (nslab + 1));
    memcpy(fill, start, sizeof(int64_t) * (nslab + 1));
    for (int64_t j = 0; j < M; ++j) order[fill[slab[j]]++] = j;
    free(fill);
  }
  /* Phases 0 and 1: even / odd slabs whose planes do not wrap past nf (two
   * slabs of one parity are >= SLAB planes apart and touch at most
   * SLAB + ns - 1 planes each). Phase 2: the slabs that wrap, one at a time. */
  for (int phase = 0; phase < 3; ++phase) {
#pragma omp parallel for schedule(dynamic, 1) num_threads(phase == 2 ? 1 : nthreads)
    for (int k = 0; k < nslab; ++k) {
      const int hi = (k + 1) * SLAB < nf[slow] ? (k + 1) * SLAB : nf[slow];
      const int wraps = hi + ns - 1 > nf[slow];
      if (phase == 2 ? !wraps : (wraps || (k % 2) != phase)) continue;
      for (int64_t q = start[k]; q < start[k + 1]; ++q) {
        const int64_t j = order[q];
        double r[3];
        for (int d = 0; d < dim; ++d) r[d] = fold01((double)xyz[d][j], nf[d]);
        spread_one(dim, nf, r, c + 2 * j, ns, nc, coef, fw);
      }
    }
  }
  free(slab);
  free(start);
  free(order);
}

/* c[j] = interpolation of fw at point j. */
void cpu_interp_f(int dim, int64_t M, const float *x, const float *y, const float *z,
                  float *c, const int *nf, int ns, int nc, const double *coef, const float *fw,
                  int nthreads) {
  const float *xyz[3] = {x, y, z};
#pragma omp parallel for schedule(static) num_threads(nthreads)
  for (int64_t j = 0; j < M; ++j) {
    float ker[3][MAXNS];
    int st[3] = {0, 0, 0};
    for (int d = 0; d < dim; ++d)
      kernel_vals(fold01((double)xyz[d][j], nf[d]), ns, nc, coef, &st[d], ker[d]);
    int ix[MAXNS];
    for (int a = 0; a < ns; ++a) {
      int i = st[0] + a;
      ix[a] = i < 0 ? i + nf[0] : (i >= nf[0] ? i - nf[0] : i);
    }
    float ar = 0, ai = 0;
    const int nz = dim > 2 ? ns : 1, ny = dim > 1 ? ns : 1;
    for (int e = 0; e < nz; ++e) {
      int iz = dim > 2 ? st[2] + e : 0;
      if (dim > 2) iz = iz < 0 ? iz + nf[2] : (iz >= nf[2] ? iz - nf[2] : iz);
      float pr = 0, pi = 0;
      for (int b = 0; b < ny; ++b) {
        int iy = dim > 1 ? st[1] + b : 0;
        if (dim > 1) iy = iy < 0 ? iy + nf[1] : (iy >= nf[1] ? iy - nf[1] : iy);
        const float *row = fw + 2 * ((int64_t)iz * nf[1] + iy) * nf[0];
        float rr = 0, ri = 0;
        for (int a = 0; a < ns; ++a) {
          rr += row[2 * ix[a]] * ker[0][a];
          ri += row[2 * ix[a] + 1] * ker[0][a];
        }
        const float ky = dim > 1 ? ker[1][b] : 1.0f;
        pr += rr * ky;
        pi += ri * ky;
      }
      const float kz = dim > 2 ? ker[2][e] : 1.0f;
      ar += pr * kz;
      ai += pi * kz;
    }
    c[2 * j] = ar;
    c[2 * j + 1] = ai;
  }
}

/* Type 1: fk[k] = fw[wrap(k)] / prod ker_d[|k_d|]; type 2 (to_grid != 0):
 * fw[wrap(k)] = fk[k] / prod ker_d[|k_d|] (fw zeroed by the caller).
 * ker_d has nf[d]/2 + 1 entries; modes in CMCL order, x fastest. */
void cpu_deconv_f(int dim, const int *ms, const int *nf, const double *k0, const double *k1,
                  const double *k2, float *fw, float *fk, int to_grid, int nthreads) {
  const double *ker[3] = {k0, k1, k2};
  int m[3] = {1, 1, 1}, n[3] = {1, 1, 1};
  for (int d = 0; d < dim; ++d) {
    m[d] = ms[d];
    n[d] = nf[d];
  }
#pragma omp parallel for collapse(2) schedule(static) num_threads(nthreads)
  for (int c2 = 0; c2 < m[2]; ++c2)
    for (int c1 = 0; c1 < m[1]; ++c1) {
      const int f2 = dim > 2 ? c2 - m[2] / 2 : 0, f1 = dim > 1 ? c1 - m[1] / 2 : 0;
      const int w2 = f2 < 0 ? f2 + n[2] : f2, w1 = f1 < 0 ? f1 + n[1] : f1;
      const double kk = (dim > 2 ? ker[2][abs(f2)] : 1.0) * (dim > 1 ? ker[1][abs(f1)] : 1.0);
      for (int c0 = 0; c0 < m[0]; ++c0) {
        const int f0 = c0 - m[0] / 2;
        const int w0 = f0 < 0 ? f0 + n[0] : f0;
        const double kv = kk * ker[0][abs(f0)];
        const int64_t fi = ((int64_t)c2 * m[1] + c1) * m[0] + c0;
        const int64_t wi = ((int64_t)w2 * n[1] + w1) * n[0] + w0;
        if (to_grid) {
          fw[2 * wi] = (float)(fk[2 * fi] / kv);
          fw[2 * wi + 1] = (float)(fk[2 * fi + 1] / kv);
        } else {
          fk[2 * fi] = (float)(fw[2 * wi] / kv);
          fk[2 * fi + 1] = (float)(fw[2 * wi + 1] / kv);
        }
      }
    }
}
