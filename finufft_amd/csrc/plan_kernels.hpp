// Element-wise and small kernels of the plan: kernel Fourier series,
// deconvolution / amplification between the fine grid and the modes, the
// type-3 rescale / phase factors, min/max reductions and grid zeroing.
#pragma once

#include <hip/hip_runtime.h>

#include "spreadinterp.hpp"

namespace fa {

// Up to three 1-D series of the kernel's Fourier transform at integer
// frequencies i = 0..nf_d/2 (type 1/2):
//   out_d[i] = (-1)^i sum_n f_n 2 cos(2 pi i z_n / nf_d)
// (src/cuda/fft.cu:29-49; the sign flip places the grid origin at x = -pi).
// f and z hold q nodes per dimension at stride MAX_NQUAD; evaluated in double.
template <typename T>
void fseries(int dim, const int nf[3], int q, const double *f, const double *z, T *const out[3],
             hipStream_t stream);

// fk <-> fw for blk vectors. Type 1 (to_modes): fk[k] = fw[w(k)] / prod_d
// ker_d[|k_d|]. Type 2: fw[w(k)] = fk[k] / prod_d ker_d[|k_d|] and every other
// fw cell is set to 0 (so no separate clearing pass is needed).
// modeord 0: k from -N/2 upward; 1: FFT order (src/cuda/execute.cu:24-63).
template <typename T>
void deconvolve(int dim, int modeord, bool to_modes, const int ms[3], const int nf[3],
                const T *const ker[3], cplx_t<T> *fw, cplx_t<T> *fk, int blk,
                hipStream_t stream);

// One problem of a fused amplification launch (many independent type-2
// problems): modes fk (vector 0 of the round, nmodes per vector) -> fine grid
// fw (nftot per vector), for nvec vectors. blk_begin is the first block of
// this problem in the launch; every problem gets nvec * chunks blocks of
// AMP_CHUNK cells.
template <typename T>
struct AmpItem {
  int dim, modeord;
  int ms[3], nf[3];
  double rnf[3];  // 1 / nf (index arithmetic by reciprocal multiplication)
  const T *ker[3];
  int64_t nmodes, nftot;
  int64_t blk_begin;
  int chunks;  // blocks per vector
};
constexpr int AMP_CHUNK = 4096;

// Amplifies all problems of a round in one launch. fk_base[p] / fw_base[p]
// are the problem's vector-0 pointers of the round; total_blocks is the sum
// of nvec * chunks. Writes every fine-grid cell (zeros outside the modes).
template <typename T>
void amplify_batch(const AmpItem<T> *items, const cplx_t<T> *const *fk_base,
                   cplx_t<T> *const *fw_base, int nprob, int dim, int64_t fk_vec_offset,
                   int64_t total_blocks, hipStream_t stream);

// Type-1 counterpart of amplify_batch (same items and blocks): fk_base[p]
// (vector fk_vec_offset of the round) receives the mode cells of the grids
// fw_base[p] divided by the kernel's Fourier coefficients.
template <typename T>
void deconv_batch(const AmpItem<T> *items, const cplx_t<T> *const *fw_base,
                  cplx_t<T> *const *fk_base, int nprob, int dim, int64_t fk_vec_offset,
                  int64_t total_blocks, hipStream_t stream);

// Sets n complex values to zero (graph-capture friendly, no memset node).
template <typename T>
void zero(cplx_t<T> *p, int64_t n, hipStream_t stream);

// lo/hi of a device array (n >= 1); synchronizes the stream.
template <typename T>
void minmax(const T *a, int n, double *lo, double *hi, hipStream_t stream);

// ----- type 3 (src/cuda/setpts.cu:172-293) -----
// out[i] = (in[i] - center) * inv_scale
template <typename T>
void affine(const T *in, T *out, int n, double center, double scale, hipStream_t stream);
// pre[j] = exp(i sign (D . x_j)); phases accumulated in double.
template <typename T>
void t3_prephase(int dim, const T *const xyz[3], const double D[3], int sign, int M,
                 cplx_t<T> *pre, hipStream_t stream);
// dec[k] = exp(i sign C.(s_k - D)) / prod_d phihat_d(s'_k), with
// phihat(s) = sum_n f_n 2 cos(s z_n) (fft.cu:57-76); C = 0 skips the phase.
template <typename T>
void t3_deconv(int dim, const T *const stu[3], const T *const stup[3], const double C[3],
               const double D[3], int sign, int q, const double *f, const double *z, int N,
               cplx_t<T> *dec, hipStream_t stream);
// a[t*n + i] *= b[i] for t < blk, or out[t*n + i] = a[t*n + i] * b[i].
template <typename T>
void cmul(const cplx_t<T> *a, const cplx_t<T> *b, cplx_t<T> *out, int n, int blk,
          hipStream_t stream);

}  // namespace fa
