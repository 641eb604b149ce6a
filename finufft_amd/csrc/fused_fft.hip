// Pruned, fused amplification + FFT of 2-D type-2 grids (see fused_fft.hpp).
#include "fused_fft.hpp"

#include <cmath>

#include "common.hpp"
#include "host_math.hpp"

namespace fa {

bool radix_plan(int n, RadixPlan &out) {
  out = RadixPlan{};
  if (n < 1) return false;
  std::vector<int> r;
  while (n % 4 == 0) {
    r.push_back(4);
    n /= 4;
  }
  while (n % 2 == 0) {
    r.push_back(2);
    n /= 2;
  }
  while (n % 3 == 0) {
    r.push_back(3);
    n /= 3;
  }
  while (n % 5 == 0) {
    r.push_back(5);
    n /= 5;
  }
  if (n != 1 || r.size() > 16) return false;
  for (size_t i = 0; i < r.size(); ++i) out.code |= uint64_t(r[i]) << (4 * i);
  out.npass = int(r.size());
  return true;
}

template <typename T>
bool fused_fft_shape(const int nf[2], int &cw, int &lcw, int &rw, int &lrw, size_t &lds_col,
                     size_t &lds_row) {
  const int Ec = fft_col_elems<T>(), Er = fft_row_elems<T>();
  RadixPlan rp;
  for (int d = 0; d < 2; ++d)
    if (nf[d] > FFT_MAX_N || !radix_plan(nf[d], rp)) return false;
  cw = 1;
  lcw = 0;
  while (cw < 16 && 2 * cw * nf[1] <= Ec) {
    cw *= 2;
    ++lcw;
  }
  rw = 1;
  lrw = 0;
  while (rw < 16 && 2 * rw * nf[0] <= Er) {
    rw *= 2;
    ++lrw;
  }
  const size_t cell = 2 * sizeof(T);
  lds_col = (2 * size_t(cw) * nf[1] + FFT_TW_LO + FFT_TW_HI) * cell;
  lds_row = (2 * size_t(rw) * nf[0] + FFT_TW_LO + FFT_TW_HI) * cell;
  return true;
}

template <typename T>
void fft_twiddles(int n, std::vector<T> &out) {
  out.resize(2 * size_t(n));
  for (int j = 0; j < n; ++j) {
    const double a = 2.0 * PI * double(j) / double(n);
    out[2 * j] = T(std::cos(a));
    out[2 * j + 1] = T(std::sin(a));
  }
}

namespace {

template <typename T>
using v2 = T __attribute__((ext_vector_type(2)));

template <typename T>
__device__ __forceinline__ v2<T> cmul(v2<T> a, v2<T> b) {
  v2<T> r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}

// q = a / d, r = a - q d for 0 <= a < 2^24, rd = 1 / d in float: the
// product is within 2^-6 of a / d, so one correction step is exact.
__device__ __forceinline__ int divmod24(int a, int d, float rd, int &r) {
  int q = int(float(a) * rd);
  r = a - q * d;
  if (r < 0) {
    --q;
    r += d;
  } else if (r >= d) {
    ++q;
    r -= d;
  }
  return q;
}

// In-register DFT of length R with exponent sign s: u[f] <- sum_q u[q] w^(f q),
// w = exp(s 2 pi i / R).
template <typename T, int R>
__device__ __forceinline__ void small_dft(v2<T> *u, int s) {
  if constexpr (R == 2) {
    const v2<T> a = u[0], b = u[1];
    u[0] = a + b;
    u[1] = a - b;
  } else if constexpr (R == 4) {
    const v2<T> t0 = u[0] + u[2], t1 = u[0] - u[2], t2 = u[1] + u[3];
    const v2<T> d = u[1] - u[3];
    v2<T> t3;  // s i d
    t3.x = -T(s) * d.y;
    t3.y = T(s) * d.x;
    u[0] = t0 + t2;
    u[2] = t0 - t2;
    u[1] = t1 + t3;
    u[3] = t1 - t3;
  } else if constexpr (R == 3) {
    const T c = T(-0.5), sn = T(s) * T(0.86602540378443864676);
    const v2<T> t = u[1] + u[2], d = u[1] - u[2];
    const v2<T> m = u[0] + c * t;
    v2<T> id;  // sn i d
    id.x = -sn * d.y;
    id.y = sn * d.x;
    u[0] = u[0] + t;
    u[1] = m + id;
    u[2] = m - id;
  } else {  // R == 5
    const T c1 = T(0.30901699437494742410), c2 = T(-0.80901699437494742410);
    const T s1 = T(s) * T(0.95105651629515357212), s2 = T(s) * T(0.58778525229247312917);
    const v2<T> a1 = u[1] + u[4], b1 = u[1] - u[4], a2 = u[2] + u[3], b2 = u[2] - u[3];
    const v2<T> m1 = u[0] + c1 * a1 + c2 * a2, m2 = u[0] + c2 * a1 + c1 * a2;
    const v2<T> n1 = s1 * b1 + s2 * b2, n2 = s2 * b1 - s1 * b2;
    v2<T> in1, in2;  // i n1, i n2
    in1.x = -n1.y;
    in1.y = n1.x;
    in2.x = -n2.y;
    in2.y = n2.x;
    u[0] = u[0] + a1 + a2;
    u[1] = m1 + in1;
    u[4] = m1 - in1;
    u[2] = m2 + in2;
    u[3] = m2 - in2;
  }
}

// One Stockham pass of radix R over cnt transforms of length n held in LDS
// (element (pos, c) at pos * cnt + c when cfast, else c * n + pos): for
// butterfly i of a transform, with p the product of the previous radices and
// k = i mod p, the inputs src[i + r n / R] are twiddled by w_n^(r k n / (p R)),
// transformed, and stored to dst[(i - k) R + k + r p]. lp = log2 p when p is
// a power of two, else -1.
template <typename T, int R>
__device__ void stockham_pass(const v2<T> *src, v2<T> *dst, int n, int cnt, int lcnt, bool cfast,
                              int p, int lp, const v2<T> *tw, int sign) {
  const int nR = n / R;
  const float rnR = 1.0f / float(nR), rp = 1.0f / float(p);
  const int tstep = n / (p * R);
  const int total = cnt * nR;
  const int ostep = cfast ? p * cnt : p;
#pragma unroll 1
  for (int b = threadIdx.x; b < total; b += FFT_THREADS) {
    int c, i;
    if (cfast) {
      c = b & (cnt - 1);
      i = b >> lcnt;
    } else {
      c = divmod24(b, nR, rnR, i);
    }
    int k;
    if (lp >= 0)
      k = i & (p - 1);
    else
      divmod24(i, p, rp, k);
    v2<T> u[R];
    const int ib = cfast ? i * cnt + c : c * n + i;
    const int istep = cfast ? nR * cnt : nR;
#pragma unroll
    for (int r = 0; r < R; ++r) u[r] = src[ib + r * istep];
#pragma unroll
    for (int r = 1; r < R; ++r) {
      const int t = r * k * tstep;  // < n
      v2<T> w = cmul(tw[t & (FFT_TW_LO - 1)], tw[FFT_TW_LO + (t >> 6)]);
      w.y *= T(sign);
      u[r] = cmul(u[r], w);
    }
    small_dft<T, R>(u, sign);
    const int j = (i - k) * R + k;
    const int ob = cfast ? j * cnt + c : c * n + j;
#pragma unroll
    for (int r = 0; r < R; ++r) dst[ob + r * ostep] = u[r];
  }
}

// cnt transforms of length n, starting in a, with the twiddle tables tw in
// LDS; returns the buffer holding the result (a or b). Ends with a barrier.
template <typename T>
__device__ v2<T> *lds_fft(v2<T> *a, v2<T> *b, int n, int cnt, int lcnt, bool cfast, RadixPlan rp,
                          const v2<T> *tw, int sign) {
  int p = 1, lp = 0;
  for (int s = 0; s < rp.npass; ++s) {
    const int R = int((rp.code >> (4 * s)) & 15);
    switch (R) {
      case 4: stockham_pass<T, 4>(a, b, n, cnt, lcnt, cfast, p, lp, tw, sign); break;
      case 2: stockham_pass<T, 2>(a, b, n, cnt, lcnt, cfast, p, lp, tw, sign); break;
      case 3: stockham_pass<T, 3>(a, b, n, cnt, lcnt, cfast, p, lp, tw, sign); break;
      default: stockham_pass<T, 5>(a, b, n, cnt, lcnt, cfast, p, lp, tw, sign); break;
    }
    p *= R;
    lp = (lp >= 0 && (R == 2 || R == 4)) ? lp + (R == 4 ? 2 : 1) : -1;
    v2<T> *t = a;
    a = b;
    b = t;
    __syncthreads();
  }
  return a;
}

// Blocks are dealt to the 8 XCDs of the GPU in turn (block b on XCD b mod 8),
// each XCD with its own L2. Renumbering so that every XCD takes a contiguous
// range of logical blocks keeps neighbouring blocks, whose writes share cache
// lines, in one L2.
__device__ __forceinline__ int64_t xcd_block(int64_t b, int64_t nblocks) {
  constexpr int NX = 8;
  const int64_t per = nblocks / NX, rem = nblocks % NX;
  const int64_t x = b % NX, q = b / NX;
  return x * per + (x < rem ? x : rem) + q;
}

// Stages the two twiddle tables of length n from the full table twg (no
// barrier: the caller's first barrier covers it).
template <typename T>
__device__ __forceinline__ void stage_twiddles(v2<T> *tw, gptr<const v2<T>> twg, int n) {
  for (int j = threadIdx.x; j < FFT_TW_LO + FFT_TW_HI; j += FFT_THREADS) {
    const int src = j < FFT_TW_LO ? j : FFT_TW_LO * (j - FFT_TW_LO);
    if (src < n) tw[j] = twg[src];
  }
}

// The item of block b (the last whose first block is <= b), wave-uniform.
template <typename T, bool COLS>
__device__ __forceinline__ int find_item(const FusedFftItem<T> *items, int nprob, int64_t b) {
  int lo = 0, hi = nprob - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if ((COLS ? items[mid].col_begin : items[mid].row_begin) <= b)
      lo = mid;
    else
      hi = mid - 1;
  }
  return __builtin_amdgcn_readfirstlane(lo);
}

// Mode index along an axis of m modes for frequency f (modeord 0: from -m/2;
// 1: FFT order).
__device__ __forceinline__ int mode_index(int f, int m, int modeord) {
  return modeord == 0 ? f + m / 2 : (f >= 0 ? f : m + f);
}

template <typename T>
__global__ void __launch_bounds__(FFT_THREADS)
    fused_cols_kernel(const FusedFftItem<T> *items, const cplx_t<T> *const *fk_base,
                      cplx_t<T> *const *fw_base, int nprob, int64_t fk_vec_offset, int sign,
                      int64_t total_blocks) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  if (int64_t(blockIdx.x) >= total_blocks) return;
  const int64_t blk = xcd_block(blockIdx.x, total_blocks);
  const int lo = find_item<T, true>(items, nprob, blk);
  const FusedFftItem<T> it = items[lo];
  const int64_t local = blk - it.col_begin;
  const int t = int(local / it.ncol);
  const int g = int(local % it.ncol);
  const int m0 = it.ms[0], m1 = it.ms[1], nf0 = it.nf[0], nf1 = it.nf[1];
  const int cw = it.cw, lcw = it.lcw;
  const int h0 = m0 - m0 / 2, h1 = m1 - m1 / 2;
  const int slot0 = g * cw;
  const int ncols = min(cw, m0 - slot0);
  const gptr<const v2<T>> fk = to_global(reinterpret_cast<const v2<T> *>(fk_base[lo])) +
                               (fk_vec_offset + t) * it.nmodes;
  const gptr<v2<T>> fw = to_global(reinterpret_cast<v2<T> *>(fw_base[lo])) + int64_t(t) * it.nftot;
  const gptr<const T> k0 = to_global(it.rker[0]), k1 = to_global(it.rker[1]);
  v2<T> *A = reinterpret_cast<v2<T> *>(lds_raw);
  v2<T> *B = A + cw * nf1;
  v2<T> *TW = B + cw * nf1;
  stage_twiddles<T>(TW, to_global(reinterpret_cast<const v2<T> *>(it.tw[1])), nf1);
  // element (gy, c) at gy * cw + c: the amplified modes of column slot0 + c,
  // zero outside the mode rows
  for (int e = threadIdx.x; e < nf1 * cw; e += FFT_THREADS) {
    const int gy = e >> lcw, c = e & (cw - 1);
    v2<T> val = {T(0), T(0)};
    const int fy = gy < h1 ? gy : (gy >= nf1 - m1 / 2 ? gy - nf1 : nf1);
    if (c < ncols && fy != nf1) {
      const int s = slot0 + c;
      const int fx = s < h0 ? s : s - m0;
      const int kx = mode_index(fx, m0, it.modeord), ky = mode_index(fy, m1, it.modeord);
      val = fk[int64_t(ky) * m0 + kx] * (k1[fy >= 0 ? fy : -fy] * k0[fx >= 0 ? fx : -fx]);
    }
    A[e] = val;
  }
  __syncthreads();
  const v2<T> *res = lds_fft<T>(A, B, nf1, cw, lcw, true, it.rp[1], TW, sign);
  for (int e = threadIdx.x; e < nf1 * cw; e += FFT_THREADS) {
    const int gy = e >> lcw, c = e & (cw - 1);
    if (c >= ncols) continue;
    const int s = slot0 + c;
    const int gx = s < h0 ? s : nf0 - m0 + s;
    fw[int64_t(gy) * nf0 + gx] = res[e];
  }
}

template <typename T>
__global__ void __launch_bounds__(FFT_THREADS)
    fused_rows_kernel(const FusedFftItem<T> *items, cplx_t<T> *const *fw_base, int nprob,
                      int sign, int64_t total_blocks) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  if (int64_t(blockIdx.x) >= total_blocks) return;
  const int64_t blk = xcd_block(blockIdx.x, total_blocks);
  const int lo = find_item<T, false>(items, nprob, blk);
  const FusedFftItem<T> it = items[lo];
  const int64_t local = blk - it.row_begin;
  const int t = int(local / it.nrow);
  const int g = int(local % it.nrow);
  const int m0 = it.ms[0], nf0 = it.nf[0], nf1 = it.nf[1];
  const int rw = it.rw;
  const int h0 = m0 - m0 / 2;
  const int row0 = g * rw;
  const int nr = min(rw, nf1 - row0);
  const gptr<v2<T>> fw = to_global(reinterpret_cast<v2<T> *>(fw_base[lo])) + int64_t(t) * it.nftot +
                         int64_t(row0) * nf0;
  v2<T> *A = reinterpret_cast<v2<T> *>(lds_raw);
  v2<T> *B = A + rw * nf0;
  v2<T> *TW = B + rw * nf0;
  stage_twiddles<T>(TW, to_global(reinterpret_cast<const v2<T> *>(it.tw[0])), nf0);
  const float rn0 = 1.0f / float(nf0);
  const int lrw = it.lrw;
  // element (x, r) at x * rw + r (rows interleaved, so that the butterflies
  // index with shifts): the mode columns of row row0 + r, read row by row
  for (int e = threadIdx.x; e < rw * nf0; e += FFT_THREADS) {
    int x;
    const int r = divmod24(e, nf0, rn0, x);
    v2<T> val = {T(0), T(0)};
    if (r < nr && (x < h0 || x >= nf0 - m0 / 2)) val = fw[e];
    A[(x << lrw) + r] = val;
  }
  __syncthreads();
  const v2<T> *res = lds_fft<T>(A, B, nf0, rw, lrw, true, it.rp[0], TW, sign);
  for (int e = threadIdx.x; e < nr * nf0; e += FFT_THREADS) {
    int x;
    const int r = divmod24(e, nf0, rn0, x);
    fw[e] = res[(x << lrw) + r];
  }
}

}  // namespace

template <typename T>
void fused_fft_prepare() {
  // the largest launch: 2 buffers of FFT_MAX_N elements and the tables
  const int lds = int((2 * size_t(FFT_MAX_N) + FFT_TW_LO + FFT_TW_HI) * 2 * sizeof(T));
  hip_check(hipFuncSetAttribute(reinterpret_cast<const void *>(&fused_cols_kernel<T>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, lds),
            "hipFuncSetAttribute(fused_cols_kernel)");
  hip_check(hipFuncSetAttribute(reinterpret_cast<const void *>(&fused_rows_kernel<T>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, lds),
            "hipFuncSetAttribute(fused_rows_kernel)");
}

template <typename T>
void fused_fft_cols(const FusedFftItem<T> *items, const cplx_t<T> *const *fk_base,
                    cplx_t<T> *const *fw_base, int nprob, int64_t fk_vec_offset, int sign,
                    int64_t total_blocks, size_t lds, hipStream_t stream) {
  if (total_blocks <= 0 || nprob <= 0) return;
  if (total_blocks > INT32_MAX) fail(FINUFFT_ERR_MAXNALLOC, "fused_fft_cols: grid too large");
  fused_cols_kernel<T><<<unsigned(total_blocks), FFT_THREADS, lds, stream>>>(
      items, fk_base, fw_base, nprob, fk_vec_offset, sign, total_blocks);
  launch_check("fused_cols_kernel");
}

template <typename T>
void fused_fft_rows(const FusedFftItem<T> *items, cplx_t<T> *const *fw_base, int nprob, int sign,
                    int64_t total_blocks, size_t lds, hipStream_t stream) {
  if (total_blocks <= 0 || nprob <= 0) return;
  if (total_blocks > INT32_MAX) fail(FINUFFT_ERR_MAXNALLOC, "fused_fft_rows: grid too large");
  fused_rows_kernel<T><<<unsigned(total_blocks), FFT_THREADS, lds, stream>>>(
      items, fw_base, nprob, sign, total_blocks);
  launch_check("fused_rows_kernel");
}

#define FA_INST(T)                                                                            \
  template bool fused_fft_shape<T>(const int[2], int &, int &, int &, int &, size_t &,       \
                                   size_t &);                                                 \
  template void fused_fft_prepare<T>();                                                       \
  template void fft_twiddles<T>(int, std::vector<T> &);                                      \
  template void fused_fft_cols<T>(const FusedFftItem<T> *, const cplx_t<T> *const *,         \
                                  cplx_t<T> *const *, int, int64_t, int, int64_t, size_t,    \
                                  hipStream_t);                                               \
  template void fused_fft_rows<T>(const FusedFftItem<T> *, cplx_t<T> *const *, int, int,     \
                                  int64_t, size_t, hipStream_t);
FA_INST(float)
FA_INST(double)
#undef FA_INST

}  // namespace fa
