// Element-wise and small kernels of the plan (see plan_kernels.hpp).
#include "plan_kernels.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

#include "common.hpp"
#include "host_math.hpp"

namespace fa {
namespace {

constexpr int EW_T = 256;

inline dim3 ew_grid(int64_t n, int blk = 1) {
  const int64_t b = (n + EW_T - 1) / EW_T;
  return dim3(unsigned(std::max<int64_t>(1, std::min<int64_t>(b, 1 << 20))), unsigned(blk));
}

template <typename T>
__global__ void fseries_kernel(int dim, int nf0, int nf1, int nf2, int q, const double *f,
                               const double *z, T *o0, T *o1, T *o2) {
  const int d = blockIdx.y;
  const int nf = d == 0 ? nf0 : (d == 1 ? nf1 : nf2);
  T *out = d == 0 ? o0 : (d == 1 ? o1 : o2);
  const double *fd = f + d * MAX_NQUAD;
  const double *zd = z + d * MAX_NQUAD;
  const double w = 2.0 * PI / nf;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i <= nf / 2; i += gridDim.x * blockDim.x) {
    double s = 0;
    for (int n = 0; n < q; ++n) s += fd[n] * 2.0 * cos(double(i) * w * zd[n]);
    out[i] = T((i % 2) ? -s : s);
  }
}

struct ModeMap {
  int dim, modeord;
  int ms[3], nf[3];
  double rms[3], rnf[3];  // reciprocals (index arithmetic by multiplication)
};

// Frequency of mode index k along one axis of length m.
__device__ __forceinline__ int freq_of(int k, int m, int modeord) {
  return modeord == 0 ? k - m / 2 : (k < m - m / 2 ? k : k - m);
}

// q = n / d and n - q d for 0 <= n, d < 2^31, with rd = 1 / d in double:
// the product is within 2^-21 of n / d, so one correction step is exact.
__device__ __forceinline__ int divmod31(int n, int d, double rd, int &r) {
  int q = int(double(n) * rd);
  r = n - q * d;
  if (r < 0) {
    --q;
    r += d;
  } else if (r >= d) {
    ++q;
    r -= d;
  }
  return q;
}

template <typename T>
using vec2_t = T __attribute__((ext_vector_type(2)));

// Fine-grid cell g (< 2^31) -> its mode index kidx and the product kv of the
// kernel's Fourier coefficients, or false outside the modes. The outer axes
// are tested first, so that a row outside the modes costs one test.
template <typename T, int DIM>
__device__ __forceinline__ bool cell_to_mode(int g, const int *nf, const double *rnf, const int *ms,
                                             int modeord, const gptr<const T> *ker, int &kidx,
                                             T &kv) {
  int c[DIM];
  int rem = g;
#pragma unroll
  for (int d = 0; d < DIM - 1; ++d) rem = divmod31(rem, nf[d], rnf[d], c[d]);
  c[DIM - 1] = rem;
  kidx = 0;
  kv = T(1);
#pragma unroll
  for (int d = DIM - 1; d >= 0; --d) {
    const int m = ms[d], gd = c[d];
    int fr;
    if (gd < m - m / 2)
      fr = gd;
    else if (gd >= nf[d] - m / 2)
      fr = gd - nf[d];
    else
      return false;
    const int k = modeord == 0 ? fr + m / 2 : (fr >= 0 ? fr : m + fr);
    kv *= ker[d][fr >= 0 ? fr : -fr];
    kidx = kidx * m + k;
  }
  return true;
}

// Mode i (< 2^31) -> its fine-grid cell and the kernel product.
template <typename T, int DIM>
__device__ __forceinline__ int mode_to_cell(int i, const ModeMap &mm, const gptr<const T> *ker,
                                            T &kv) {
  int k[DIM];
  int rem = i;
#pragma unroll
  for (int d = 0; d < DIM - 1; ++d) rem = divmod31(rem, mm.ms[d], mm.rms[d], k[d]);
  k[DIM - 1] = rem;
  int widx = 0;
  kv = T(1);
#pragma unroll
  for (int d = DIM - 1; d >= 0; --d) {
    const int fr = freq_of(k[d], mm.ms[d], mm.modeord);
    kv *= ker[d][fr >= 0 ? fr : -fr];
    widx = widx * mm.nf[d] + (fr >= 0 ? fr : mm.nf[d] + fr);
  }
  return widx;
}

template <typename T, int DIM>
__global__ void __launch_bounds__(EW_T)
    deconv_to_modes_kernel(ModeMap mm, const T *k0, const T *k1, const T *k2,
                           const cplx_t<T> *fw, cplx_t<T> *fk, int64_t nmodes, int64_t nftot) {
  const int t = blockIdx.y;
  const gptr<const T> ker[3] = {to_global(k0), to_global(k1), to_global(k2)};
  const gptr<const vec2_t<T>> src =
      to_global(reinterpret_cast<const vec2_t<T> *>(fw)) + int64_t(t) * nftot;
  const gptr<vec2_t<T>> dst = to_global(reinterpret_cast<vec2_t<T> *>(fk)) + int64_t(t) * nmodes;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < nmodes;
       i += int64_t(gridDim.x) * blockDim.x) {
    T kv;
    const int widx = mode_to_cell<T, DIM>(int(i), mm, ker, kv);
    const vec2_t<T> v = src[widx];
    vec2_t<T> o;
    o.x = v.x / kv;
    o.y = v.y / kv;
    dst[i] = o;
  }
}

template <typename T, int DIM>
__global__ void __launch_bounds__(EW_T)
    amplify_to_grid_kernel(ModeMap mm, const T *k0, const T *k1, const T *k2, cplx_t<T> *fw,
                           const cplx_t<T> *fk, int64_t nmodes, int64_t nftot) {
  const int t = blockIdx.y;
  const gptr<const T> ker[3] = {to_global(k0), to_global(k1), to_global(k2)};
  const gptr<const vec2_t<T>> src =
      to_global(reinterpret_cast<const vec2_t<T> *>(fk)) + int64_t(t) * nmodes;
  const gptr<vec2_t<T>> dst = to_global(reinterpret_cast<vec2_t<T> *>(fw)) + int64_t(t) * nftot;
  for (int64_t g = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; g < nftot;
       g += int64_t(gridDim.x) * blockDim.x) {
    int kidx;
    T kv;
    vec2_t<T> out = {T(0), T(0)};
    if (cell_to_mode<T, DIM>(int(g), mm.nf, mm.rnf, mm.ms, mm.modeord, ker, kidx, kv)) {
      const vec2_t<T> v = src[kidx];
      out.x = v.x / kv;
      out.y = v.y / kv;
    }
    dst[g] = out;
  }
}

template <typename T, int DIM>
__global__ void __launch_bounds__(EW_T)
    amplify_batch_kernel(const AmpItem<T> *items, const cplx_t<T> *const *fk_base,
                         cplx_t<T> *const *fw_base, int nprob, int64_t fk_vec_offset,
                         int64_t total_blocks) {
  const int64_t b = blockIdx.x;
  if (b >= total_blocks) return;
  // The block's problem: the last item whose first block is <= b.
  int lo = 0, hi = nprob - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (items[mid].blk_begin <= b)
      lo = mid;
    else
      hi = mid - 1;
  }
  lo = __builtin_amdgcn_readfirstlane(lo);  // uniform: scalar loads of the descriptor
  const AmpItem<T> it = items[lo];
  const gptr<const T> ker[3] = {to_global(it.ker[0]), to_global(it.ker[1]),
                                to_global(it.ker[2])};
  const int64_t local = b - it.blk_begin;
  const int t = int(local / it.chunks);
  // Grid indices fit 31 bits (the plan caps a vector's grid there).
  const int g0 = int(local % it.chunks) * AMP_CHUNK;
  const int g1 = g0 + AMP_CHUNK < it.nftot ? g0 + AMP_CHUNK : int(it.nftot);
  const gptr<const vec2_t<T>> fk =
      to_global(reinterpret_cast<const vec2_t<T> *>(fk_base[lo])) +
      (fk_vec_offset + t) * it.nmodes;
  const gptr<vec2_t<T>> fw = to_global(reinterpret_cast<vec2_t<T> *>(fw_base[lo])) +
                             int64_t(t) * it.nftot;
  for (int o = int(threadIdx.x); o < g1 - g0; o += EW_T) {
    const int g = g0 + o;
    int kidx;
    T kv;
    vec2_t<T> out = {T(0), T(0)};
    if (cell_to_mode<T, DIM>(g, it.nf, it.rnf, it.ms, it.modeord, ker, kidx, kv)) {
      const vec2_t<T> v = fk[kidx];
      out.x = v.x / kv;
      out.y = v.y / kv;
    }
    fw[g] = out;
  }
}

// The reverse of amplify_batch_kernel for type-1 batches: the mode cells of
// every grid of the round, divided by the kernel's coefficients, to fk (the
// blocks and their cells as there; cells outside the modes are skipped).
template <typename T, int DIM>
__global__ void __launch_bounds__(EW_T)
    deconv_batch_kernel(const AmpItem<T> *items, const cplx_t<T> *const *fw_base,
                        cplx_t<T> *const *fk_base, int nprob, int64_t fk_vec_offset,
                        int64_t total_blocks) {
  const int64_t b = blockIdx.x;
  if (b >= total_blocks) return;
  int lo = 0, hi = nprob - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (items[mid].blk_begin <= b)
      lo = mid;
    else
      hi = mid - 1;
  }
  lo = __builtin_amdgcn_readfirstlane(lo);
  const AmpItem<T> it = items[lo];
  const gptr<const T> ker[3] = {to_global(it.ker[0]), to_global(it.ker[1]),
                                to_global(it.ker[2])};
  const int64_t local = b - it.blk_begin;
  const int t = int(local / it.chunks);
  const int g0 = int(local % it.chunks) * AMP_CHUNK;
  const int g1 = g0 + AMP_CHUNK < it.nftot ? g0 + AMP_CHUNK : int(it.nftot);
  const gptr<const vec2_t<T>> fw = to_global(reinterpret_cast<const vec2_t<T> *>(fw_base[lo])) +
                                   int64_t(t) * it.nftot;
  const gptr<vec2_t<T>> fk = to_global(reinterpret_cast<vec2_t<T> *>(fk_base[lo])) +
                             (fk_vec_offset + t) * it.nmodes;
  for (int o = int(threadIdx.x); o < g1 - g0; o += EW_T) {
    const int g = g0 + o;
    int kidx;
    T kv;
    if (cell_to_mode<T, DIM>(g, it.nf, it.rnf, it.ms, it.modeord, ker, kidx, kv)) {
      const vec2_t<T> v = fw[g];
      vec2_t<T> out;
      out.x = v.x / kv;
      out.y = v.y / kv;
      fk[kidx] = out;
    }
  }
}

template <typename T>
__global__ void zero_kernel(cplx_t<T> *p, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    p[i] = cplx_t<T>{0, 0};
}

template <typename T>
__global__ void minmax_kernel(const T *a, int n, double *lo, double *hi) {
  __shared__ double slo[EW_T], shi[EW_T];
  double l = INFINITY, h = -INFINITY;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const double v = double(a[i]);
    l = fmin(l, v);
    h = fmax(h, v);
  }
  slo[threadIdx.x] = l;
  shi[threadIdx.x] = h;
  __syncthreads();
  for (int s = EW_T / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      slo[threadIdx.x] = fmin(slo[threadIdx.x], slo[threadIdx.x + s]);
      shi[threadIdx.x] = fmax(shi[threadIdx.x], shi[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    lo[blockIdx.x] = slo[0];
    hi[blockIdx.x] = shi[0];
  }
}

template <typename T>
__global__ void affine_kernel(const T *in, T *out, int n, T negcenter, T scale) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    out[i] = (in[i] + negcenter) * scale;
}

template <typename T>
__global__ void t3_prephase_kernel(int dim, const T *x, const T *y, const T *z, double D0,
                                   double D1, double D2, int sign, int M, cplx_t<T> *pre) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < M;
       i += int64_t(gridDim.x) * blockDim.x) {
    double ph = D0 * double(x[i]);
    if (dim > 1) ph += D1 * double(y[i]);
    if (dim > 2) ph += D2 * double(z[i]);
    double sn, cs;
    sincos(ph, &sn, &cs);
    pre[i] = cplx_t<T>{T(cs), T(sign * sn)};
  }
}

template <typename T>
__global__ void t3_deconv_kernel(int dim, const T *s0, const T *s1, const T *s2, const T *p0,
                                 const T *p1, const T *p2, double C0, double C1, double C2,
                                 double D0, double D1, double D2, int sign, int use_phase,
                                 int q, const double *f, const double *z, int N,
                                 cplx_t<T> *dec) {
  const T *sp[3] = {p0, p1, p2};
  for (int64_t k = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; k < N;
       k += int64_t(gridDim.x) * blockDim.x) {
    double ph = 1.0;
    for (int d = 0; d < dim; ++d) {
      const double sk = double(sp[d][k]);
      double acc = 0;
      for (int n = 0; n < q; ++n) acc += f[n] * 2.0 * cos(sk * z[n]);
      ph *= acc;
    }
    const double inv = 1.0 / ph;
    double re = inv, im = 0;
    if (use_phase) {
      double phase = C0 * (double(s0[k]) - D0);
      if (dim > 1) phase += C1 * (double(s1[k]) - D1);
      if (dim > 2) phase += C2 * (double(s2[k]) - D2);
      double sn, cs;
      sincos(phase, &sn, &cs);
      re = cs * inv;
      im = sign * sn * inv;
    }
    dec[k] = cplx_t<T>{T(re), T(im)};
  }
}

template <typename T>
__global__ void cmul_kernel(const cplx_t<T> *a, const cplx_t<T> *b, cplx_t<T> *out, int n) {
  const int64_t t = blockIdx.y;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const cplx_t<T> u = a[t * n + i], v = b[i];
    out[t * n + i] = cplx_t<T>{u.x * v.x - u.y * v.y, u.x * v.y + u.y * v.x};
  }
}

}  // namespace

template <typename T>
void fseries(int dim, const int nf[3], int q, const double *f, const double *z, T *const out[3],
             hipStream_t stream) {
  int nmax = 0;
  for (int d = 0; d < dim; ++d) nmax = std::max(nmax, nf[d] / 2 + 1);
  dim3 grid(std::max(1, std::min((nmax + EW_T - 1) / EW_T, 4096)), dim);
  fseries_kernel<T><<<grid, EW_T, 0, stream>>>(dim, nf[0], nf[1], nf[2], q, f, z, out[0],
                                               dim > 1 ? out[1] : out[0],
                                               dim > 2 ? out[2] : out[0]);
  launch_check("fseries_kernel");
}

template <typename T>
void deconvolve(int dim, int modeord, bool to_modes, const int ms[3], const int nf[3],
                const T *const ker[3], cplx_t<T> *fw, cplx_t<T> *fk, int blk,
                hipStream_t stream) {
  ModeMap mm{dim, modeord, {ms[0], ms[1], ms[2]}, {nf[0], nf[1], nf[2]}, {}, {}};
  for (int d = 0; d < 3; ++d) {
    mm.rms[d] = 1.0 / double(ms[d]);
    mm.rnf[d] = 1.0 / double(nf[d]);
  }
  const int64_t nmodes = int64_t(ms[0]) * ms[1] * ms[2];
  const int64_t nftot = int64_t(nf[0]) * nf[1] * nf[2];
  const T *k1 = dim > 1 ? ker[1] : ker[0], *k2 = dim > 2 ? ker[2] : ker[0];
  if (to_modes) {
    auto k = dim == 1 ? deconv_to_modes_kernel<T, 1>
                      : (dim == 2 ? deconv_to_modes_kernel<T, 2> : deconv_to_modes_kernel<T, 3>);
    k<<<ew_grid(nmodes, blk), EW_T, 0, stream>>>(mm, ker[0], k1, k2, fw, fk, nmodes, nftot);
    launch_check("deconv_to_modes_kernel");
  } else {
    auto k = dim == 1 ? amplify_to_grid_kernel<T, 1>
                      : (dim == 2 ? amplify_to_grid_kernel<T, 2> : amplify_to_grid_kernel<T, 3>);
    k<<<ew_grid(nftot, blk), EW_T, 0, stream>>>(mm, ker[0], k1, k2, fw, fk, nmodes, nftot);
    launch_check("amplify_to_grid_kernel");
  }
}

template <typename T>
void amplify_batch(const AmpItem<T> *items, const cplx_t<T> *const *fk_base,
                   cplx_t<T> *const *fw_base, int nprob, int dim, int64_t fk_vec_offset,
                   int64_t total_blocks, hipStream_t stream) {
  if (total_blocks <= 0 || nprob <= 0) return;
  if (total_blocks > INT32_MAX) fail(FINUFFT_ERR_MAXNALLOC, "amplify_batch: grid too large");
  auto k = dim == 1 ? amplify_batch_kernel<T, 1>
                    : (dim == 2 ? amplify_batch_kernel<T, 2> : amplify_batch_kernel<T, 3>);
  k<<<unsigned(total_blocks), EW_T, 0, stream>>>(items, fk_base, fw_base, nprob, fk_vec_offset,
                                                 total_blocks);
  launch_check("amplify_batch_kernel");
}

template <typename T>
void deconv_batch(const AmpItem<T> *items, const cplx_t<T> *const *fw_base,
                  cplx_t<T> *const *fk_base, int nprob, int dim, int64_t fk_vec_offset,
                  int64_t total_blocks, hipStream_t stream) {
  if (total_blocks <= 0 || nprob <= 0) return;
  if (total_blocks > INT32_MAX) fail(FINUFFT_ERR_MAXNALLOC, "deconv_batch: grid too large");
  auto k = dim == 1 ? deconv_batch_kernel<T, 1>
                    : (dim == 2 ? deconv_batch_kernel<T, 2> : deconv_batch_kernel<T, 3>);
  k<<<unsigned(total_blocks), EW_T, 0, stream>>>(items, fw_base, fk_base, nprob, fk_vec_offset,
                                                 total_blocks);
  launch_check("deconv_batch_kernel");
}

template <typename T>
void zero(cplx_t<T> *p, int64_t n, hipStream_t stream) {
  if (n <= 0) return;
  zero_kernel<T><<<ew_grid(n), EW_T, 0, stream>>>(p, n);
  launch_check("zero_kernel");
}

template <typename T>
void minmax(const T *a, int n, double *lo, double *hi, hipStream_t stream) {
  const int nb = std::max(1, std::min((n + EW_T - 1) / EW_T, 1024));
  dbuf<double> tmp(4 * size_t(nb));
  minmax_kernel<T><<<nb, EW_T, 0, stream>>>(a, n, tmp.get(), tmp.get() + nb);
  launch_check("minmax_kernel");
  std::vector<double> h(2 * size_t(nb));
  hip_check(hipMemcpyAsync(h.data(), tmp.get(), 2 * nb * sizeof(double), hipMemcpyDeviceToHost,
                           stream),
            "hipMemcpyAsync(minmax)");
  hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize(minmax)");
  double l = INFINITY, u = -INFINITY;
  for (int i = 0; i < nb; ++i) {
    l = std::min(l, h[i]);
    u = std::max(u, h[nb + i]);
  }
  *lo = l;
  *hi = u;
}

template <typename T>
void affine(const T *in, T *out, int n, double center, double scale, hipStream_t stream) {
  if (n <= 0) return;
  affine_kernel<T><<<ew_grid(n), EW_T, 0, stream>>>(in, out, n, T(-center), T(scale));
  launch_check("affine_kernel");
}

template <typename T>
void t3_prephase(int dim, const T *const xyz[3], const double D[3], int sign, int M,
                 cplx_t<T> *pre, hipStream_t stream) {
  if (M <= 0) return;
  t3_prephase_kernel<T><<<ew_grid(M), EW_T, 0, stream>>>(dim, xyz[0], xyz[1], xyz[2], D[0],
                                                         D[1], D[2], sign, M, pre);
  launch_check("t3_prephase_kernel");
}

template <typename T>
void t3_deconv(int dim, const T *const stu[3], const T *const stup[3], const double C[3],
               const double D[3], int sign, int q, const double *f, const double *z, int N,
               cplx_t<T> *dec, hipStream_t stream) {
  if (N <= 0) return;
  const bool finite = std::isfinite(C[0]) && std::isfinite(C[1]) && std::isfinite(C[2]);
  const bool nonzero = C[0] != 0 || C[1] != 0 || C[2] != 0;
  t3_deconv_kernel<T><<<ew_grid(N), EW_T, 0, stream>>>(
      dim, stu[0], stu[1], stu[2], stup[0], stup[1], stup[2], C[0], C[1], C[2], D[0], D[1], D[2],
      sign, (finite && nonzero) ? 1 : 0, q, f, z, N, dec);
  launch_check("t3_deconv_kernel");
}

template <typename T>
void cmul(const cplx_t<T> *a, const cplx_t<T> *b, cplx_t<T> *out, int n, int blk,
          hipStream_t stream) {
  if (n <= 0 || blk <= 0) return;
  cmul_kernel<T><<<ew_grid(n, blk), EW_T, 0, stream>>>(a, b, out, n);
  launch_check("cmul_kernel");
}

#define FA_INST(T)                                                                             \
  template void fseries<T>(int, const int[3], int, const double *, const double *,            \
                           T *const[3], hipStream_t);                                          \
  template void deconvolve<T>(int, int, bool, const int[3], const int[3], const T *const[3],  \
                              cplx_t<T> *, cplx_t<T> *, int, hipStream_t);                     \
  template void zero<T>(cplx_t<T> *, int64_t, hipStream_t);                                    \
  template void amplify_batch<T>(const AmpItem<T> *, const cplx_t<T> *const *,                 \
                                 cplx_t<T> *const *, int, int, int64_t, int64_t, hipStream_t); \
  template void deconv_batch<T>(const AmpItem<T> *, const cplx_t<T> *const *,           \
                               cplx_t<T> *const *, int, int, int64_t, int64_t, hipStream_t); \
  template void minmax<T>(const T *, int, double *, double *, hipStream_t);                    \
  template void affine<T>(const T *, T *, int, double, double, hipStream_t);                   \
  template void t3_prephase<T>(int, const T *const[3], const double[3], int, int, cplx_t<T> *, \
                               hipStream_t);                                                   \
  template void t3_deconv<T>(int, const T *const[3], const T *const[3], const double[3],       \
                             const double[3], int, int, const double *, const double *, int,   \
                             cplx_t<T> *, hipStream_t);                                        \
  template void cmul<T>(const cplx_t<T> *, const cplx_t<T> *, cplx_t<T> *, int, int,           \
                        hipStream_t);
FA_INST(float)
FA_INST(double)
#undef FA_INST

}  // namespace fa
