// Pruned 3-D line transforms of a plan's fine grid (see fft3.hpp).
#include "fft3.hpp"

#include <cmath>
#include <vector>

#include "common.hpp"
#include "host_math.hpp"

namespace fa {

namespace {

template <typename T>
using c2 = T __attribute__((ext_vector_type(2)));

template <typename T>
__device__ __forceinline__ c2<T> f3_cmul(c2<T> a, c2<T> b) {
  c2<T> r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}

// In-register DFT of length R (2, 4 or 8) with exponent sign s:
// u[f] <- sum_q u[q] exp(s 2 pi i f q / R).
template <typename T, int R>
__device__ __forceinline__ void f3_dft(c2<T> *u, T s) {
  if constexpr (R == 2) {
    const c2<T> a = u[0], b = u[1];
    u[0] = a + b;
    u[1] = a - b;
  } else if constexpr (R == 4) {
    const c2<T> t0 = u[0] + u[2], t1 = u[0] - u[2], t2 = u[1] + u[3], d = u[1] - u[3];
    c2<T> t3;  // s i d
    t3.x = -s * d.y;
    t3.y = s * d.x;
    u[0] = t0 + t2;
    u[2] = t0 - t2;
    u[1] = t1 + t3;
    u[3] = t1 - t3;
  } else {
    static_assert(R == 8, "radix 2, 4 or 8");
    // even / odd halves: X[f] = E[f] + w^f O[f], X[f + 4] = E[f] - w^f O[f],
    // w = exp(s i pi / 4)
    c2<T> e[4] = {u[0], u[2], u[4], u[6]}, o[4] = {u[1], u[3], u[5], u[7]};
    f3_dft<T, 4>(e, s);
    f3_dft<T, 4>(o, s);
    const T h = T(0.70710678118654752440);
    c2<T> t1, t2, t3;
    t1.x = h * (o[1].x - s * o[1].y);  // (1 + s i) / sqrt 2
    t1.y = h * (o[1].y + s * o[1].x);
    t2.x = -s * o[2].y;  // s i
    t2.y = s * o[2].x;
    t3.x = -h * (o[3].x + s * o[3].y);  // (-1 + s i) / sqrt 2
    t3.y = h * (s * o[3].x - o[3].y);
    u[0] = e[0] + o[0];
    u[4] = e[0] - o[0];
    u[1] = e[1] + t1;
    u[5] = e[1] - t1;
    u[2] = e[2] + t2;
    u[6] = e[2] - t2;
    u[3] = e[3] + t3;
    u[7] = e[3] - t3;
  }
}

// Radix of the Stockham step that follows the radices multiplying to P
// (N a power of two): 8 while 8 divides what is left, then 4 or 2.
template <int N, int P>
constexpr int f3_radix() {
  return (N / P) >= 8 ? 8 : N / P;
}

// Cell <-> compact mode index along an axis of n cells and m modes
// (h = m - m / 2 non-negative frequencies first): -1 for a cell outside the
// modes.
__device__ __forceinline__ int f3_compact(int g, int n, int m) {
  const int h = m - m / 2;
  return g < h ? g : (g >= n - m / 2 ? g - n + m : -1);
}

// One Stockham step of radix R after radices with product P, over the G
// lines of a workgroup: butterfly i of a line takes cells i + r N/R (r < R)
// through in(pos, line), twiddles them by w_N^(r k N/(P R)) (k = i mod P),
// transforms them and hands them to out(pos, line, value) at cells
// (i - k) R + k + r P. ROWS: butterflies of one line are on consecutive
// threads (else lines are). SYNC: a barrier between the reads and the writes
// (the step reads and writes the same LDS block).
template <typename T, int N, int G, int R, int P, bool ROWS, bool SYNC, class In, class Out>
__device__ __forceinline__ void f3_step(const c2<T> *tw, T s, In &&in, Out &&out) {
  constexpr int NR = N / R;
  constexpr int NB = G * NR;
  constexpr int PER = (NB + F3_THREADS - 1) / F3_THREADS;
  constexpr int TSTEP = N / (P * R);
  c2<T> u[PER][R];
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int b = int(threadIdx.x) + j * F3_THREADS;
    if (NB % F3_THREADS == 0 || b < NB) {
      const int i = ROWS ? b % NR : b / G;
      const int line = ROWS ? b / NR : b % G;
#pragma unroll
      for (int r = 0; r < R; ++r) u[j][r] = in(i + r * NR, line);
      if constexpr (P > 1) {
        const int k = i % P;
#pragma unroll
        for (int r = 1; r < R; ++r) {
          c2<T> w = tw[r * k * TSTEP];
          w.y *= s;
          u[j][r] = f3_cmul(u[j][r], w);
        }
      }
      f3_dft<T, R>(u[j], s);
    }
  }
  if constexpr (SYNC) __syncthreads();
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int b = int(threadIdx.x) + j * F3_THREADS;
    if (NB % F3_THREADS == 0 || b < NB) {
      const int i = ROWS ? b % NR : b / G;
      const int line = ROWS ? b / NR : b % G;
      const int k = i % P;
      const int o = (i - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) out(o + r * P, line, u[j][r]);
    }
  }
}

// The middle steps (through LDS) from the step after radix product P up to
// the last, which is left to the caller.
template <typename T, int N, int G, int P, bool ROWS, class Lay>
__device__ __forceinline__ void f3_middle(c2<T> *lds, const c2<T> *tw, T s, Lay &&lay) {
  constexpr int R = f3_radix<N, P>();
  if constexpr (P * R < N) {
    auto in = [&](int pos, int line) { return lds[lay(pos, line)]; };
    auto out = [&](int pos, int line, c2<T> v) { lds[lay(pos, line)] = v; };
    f3_step<T, N, G, R, P, ROWS, true>(tw, s, in, out);
    __syncthreads();
    f3_middle<T, N, G, P * R, ROWS>(lds, tw, s, lay);
  }
}

// The radix product before the last step.
template <int N, int P = 1>
constexpr int f3_last_p() {
  constexpr int R = f3_radix<N, P>();
  if constexpr (P * R == N)
    return P;
  else
    return f3_last_p<N, P * R>();
}

// Transforms G lines: first(pos, line) gives the input cells, last(pos,
// line, v) takes the output cells; the steps between run through lds laid
// out by lay(pos, line). tw (N values in LDS) must be staged before the
// first barrier of this function.
template <typename T, int N, int G, bool ROWS, class Lay, class First, class Last>
__device__ __forceinline__ void f3_transform(c2<T> *lds, const c2<T> *tw, T s, Lay &&lay,
                                             First &&first, Last &&last) {
  constexpr int R0 = f3_radix<N, 1>();
  static_assert(R0 < N, "at least two steps");
  auto to_lds = [&](int pos, int line, c2<T> v) { lds[lay(pos, line)] = v; };
  f3_step<T, N, G, R0, 1, ROWS, false>(tw, s, first, to_lds);
  __syncthreads();
  f3_middle<T, N, G, R0, ROWS>(lds, tw, s, lay);
  constexpr int PL = f3_last_p<N>();
  constexpr int RL = N / PL;
  auto from_lds = [&](int pos, int line) { return lds[lay(pos, line)]; };
  f3_step<T, N, G, RL, PL, ROWS, false>(tw, s, from_lds, last);
}

template <typename T, int N>
struct F3Shape {
  static constexpr int G = f3_lines<T>(N);
  // Column layout: cell pos of line l at (pos + pos / 8) G + l. The first
  // step writes cells 8 i + r of the 64 / G butterflies i of a wave; the
  // gap of G cells per 8 puts consecutive i in different LDS banks when
  // fewer than 256 bytes of lines share a position (G below 32 in fp32).
  static constexpr int PADW = G < 256 / int(2 * sizeof(T)) ? G : 0;
  static constexpr int COLS = N * G + (N / 8) * PADW;
  static constexpr int NP = N + N / 8;                  // row layout pitch
  static constexpr int ROWS = G * NP;
};

template <typename T>
__device__ __forceinline__ void f3_stage_tw(c2<T> *tw, const T *g, int n) {
  const c2<T> *src = reinterpret_cast<const c2<T> *>(g);
  for (int j = threadIdx.x; j < n; j += F3_THREADS) tw[j] = src[j];
}

// Strided-axis pass: workgroup (ablk, b) of vector blockIdx.y transforms
// lines a0 .. a0 + G - 1 of outer index b.
template <typename T, int N>
__global__ void __launch_bounds__(F3_THREADS) fft3_cols_kernel(F3Cols<T> p) {
  using S = F3Shape<T, N>;
  constexpr int G = S::G;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  c2<T> *lds = reinterpret_cast<c2<T> *>(lds_raw);
  c2<T> *tw = lds + S::COLS;
  f3_stage_tw<T>(tw, p.tw, N);
  const int b = int(blockIdx.x) / p.nblk_a;
  const int64_t a0 = int64_t(int(blockIdx.x) - b * p.nblk_a) * G;
  const int t = blockIdx.y;
  const c2<T> *in = reinterpret_cast<const c2<T> *>(p.in) + t * p.in_v + b * p.in_b + a0;
  c2<T> *out = reinterpret_cast<c2<T> *>(p.out) + t * p.out_v + b * p.out_b + a0;
  const int64_t A = p.A;
  const int nl = int(min(int64_t(G), A - a0));  // lines of this workgroup
  const int m = p.m;
  const bool icomp = p.in_compact != 0, ocomp = p.out_compact != 0;
  auto lay = [](int pos, int line) { return pos * G + line + (pos >> 3) * S::PADW; };
  auto first = [&](int pos, int line) {
    c2<T> v = {T(0), T(0)};
    const int c = icomp ? f3_compact(pos, N, m) : pos;
    if (line < nl && c >= 0) v = in[int64_t(c) * A + line];
    return v;
  };
  auto last = [&](int pos, int line, c2<T> v) {
    const int c = ocomp ? f3_compact(pos, N, m) : pos;
    if (line < nl && c >= 0) out[int64_t(c) * A + line] = v;
  };
  f3_transform<T, N, G, false>(lds, tw, p.sign, lay, first, last);
}

// x pass over rows of the compact (cy, cz) modes.
template <typename T, int N>
__global__ void __launch_bounds__(F3_THREADS) fft3_rows_kernel(F3Rows<T> p) {
  using S = F3Shape<T, N>;
  constexpr int G = S::G;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  c2<T> *lds = reinterpret_cast<c2<T> *>(lds_raw);
  c2<T> *tw = lds + S::ROWS;
  int64_t *rowoff = reinterpret_cast<int64_t *>(tw + N);  // fk offset of the row
  T *rker = reinterpret_cast<T *>(rowoff + G);            // 1 / (k1[|fy|] k2[|fz|])
  f3_stage_tw<T>(tw, p.tw, N);
  const int r0 = int(blockIdx.x) * G;
  const int t = blockIdx.y;
  const int m0 = p.ms[0], m1 = p.ms[1], m2 = p.ms[2];
  const int nl = min(G, p.nrows - r0);
  if (int(threadIdx.x) < nl) {
    const int r = r0 + int(threadIdx.x);
    const int cy = r % m1, cz = r / m1;
    const int h1 = m1 - m1 / 2, h2 = m2 - m2 / 2;
    const int fy = cy < h1 ? cy : cy - m1, fz = cz < h2 ? cz : cz - m2;
    const int ky = p.modeord == 0 ? fy + m1 / 2 : cy;
    const int kz = p.modeord == 0 ? fz + m2 / 2 : cz;
    rowoff[threadIdx.x] = (int64_t(kz) * m1 + ky) * m0;
    rker[threadIdx.x] = p.ker[1][fy < 0 ? -fy : fy] * p.ker[2][fz < 0 ? -fz : fz];
  }
  __syncthreads();
  c2<T> *s2 = reinterpret_cast<c2<T> *>(p.s2) + t * p.s2_v + int64_t(r0) * N;
  c2<T> *fk = reinterpret_cast<c2<T> *>(p.fk) + t * p.fk_v;
  const T *k0 = p.ker[0];
  const int h0 = m0 - m0 / 2;
  const int modeord = p.modeord;
  // mode index along x of cell g (or -1), and the divisor k0[|fx|] k1 k2
  auto xmode = [&](int g, int line, T &q) {
    const int c = f3_compact(g, N, m0);
    if (c < 0) return -1;
    const int fx = c < h0 ? c : c - m0;
    q = T(1) / (k0[fx < 0 ? -fx : fx] * rker[line]);
    return modeord == 0 ? fx + m0 / 2 : c;
  };
  auto lay = [](int pos, int line) { return line * S::NP + pos + (pos >> 3); };
  if (p.to_modes) {
    auto first = [&](int pos, int line) {
      c2<T> v = {T(0), T(0)};
      if (line < nl) v = s2[int64_t(line) * N + pos];
      return v;
    };
    auto last = [&](int pos, int line, c2<T> v) {
      T q;
      const int kx = line < nl ? xmode(pos, line, q) : -1;
      if (kx >= 0) fk[rowoff[line] + kx] = v * q;
    };
    f3_transform<T, N, G, true>(lds, tw, p.sign, lay, first, last);
  } else {
    auto first = [&](int pos, int line) {
      c2<T> v = {T(0), T(0)};
      T q;
      const int kx = line < nl ? xmode(pos, line, q) : -1;
      if (kx >= 0) v = fk[rowoff[line] + kx] * q;
      return v;
    };
    auto last = [&](int pos, int line, c2<T> v) {
      if (line < nl) s2[int64_t(line) * N + pos] = v;
    };
    f3_transform<T, N, G, true>(lds, tw, p.sign, lay, first, last);
  }
}

}  // namespace

template <typename T>
size_t f3_lds_cols(int n) {
  const int g = f3_lines<T>(n);
  const int padw = g < int(256 / (2 * sizeof(T))) ? g : 0;
  return (size_t(n) * g + size_t(n / 8) * padw + n) * 2 * sizeof(T);
}

template <typename T>
size_t f3_lds_rows(int n) {
  const int g = f3_lines<T>(n);
  return (size_t(g) * (n + n / 8) + n) * 2 * sizeof(T) + size_t(g) * (sizeof(int64_t) + sizeof(T));
}

bool fft3_supported(int dim, const int nf[3]) {
  if (dim != 3) return false;
  for (int d = 0; d < 3; ++d)
    if (nf[d] < F3_MIN_N || nf[d] > F3_MAX_N || (nf[d] & (nf[d] - 1)) != 0) return false;
  return true;
}

template <typename T>
F3Passes<T> f3_passes(int to_modes, const int ms[3], const int nf[3], int modeord,
                      cplx_t<T> *fw, cplx_t<T> *fk, cplx_t<T> *s1, const T *const ker[3],
                      const T *const tw[3], int sign) {
  F3Passes<T> q{};
  const int64_t nftot = int64_t(nf[0]) * nf[1] * nf[2];
  const int64_t s1v = int64_t(nf[0]) * nf[1] * ms[2];
  const int64_t nmodes = int64_t(ms[0]) * ms[1] * ms[2];
  const T sg = T(sign > 0 ? 1 : -1);
  // z lines: fw (full z) <-> S1 (compact z); lines are the nf0 nf1 cells of a plane
  F3Cols<T> &z = q.z;
  z.A = int64_t(nf[0]) * nf[1];
  z.nb = 1;
  z.m = ms[2];
  z.n = nf[2];
  z.tw = tw[2];
  z.sign = sg;
  // y lines: S1 (full y) <-> S2 in fw (compact y); lines are the nf0 cells
  // of a row, outer index the compact z plane
  F3Cols<T> &y = q.y;
  y.A = nf[0];
  y.nb = ms[2];
  y.m = ms[1];
  y.n = nf[1];
  y.tw = tw[1];
  y.sign = sg;
  // x rows: S2 in fw <-> fk
  F3Rows<T> &x = q.x;
  x.s2 = fw;
  x.fk = fk;
  x.s2_v = nftot;
  x.fk_v = nmodes;
  for (int d = 0; d < 3; ++d) {
    x.ms[d] = ms[d];
    x.ker[d] = ker[d];
  }
  x.modeord = modeord;
  x.to_modes = to_modes;
  x.nrows = ms[1] * ms[2];
  x.n = nf[0];
  x.tw = tw[0];
  x.sign = sg;
  if (to_modes) {
    z.in = fw, z.out = s1, z.in_v = nftot, z.out_v = s1v;
    z.in_compact = 0, z.out_compact = 1;
    y.in = s1, y.out = fw, y.in_v = s1v, y.out_v = nftot;
    y.in_b = int64_t(nf[0]) * nf[1], y.out_b = int64_t(nf[0]) * ms[1];
    y.in_compact = 0, y.out_compact = 1;
  } else {
    y.in = fw, y.out = s1, y.in_v = nftot, y.out_v = s1v;
    y.in_b = int64_t(nf[0]) * ms[1], y.out_b = int64_t(nf[0]) * nf[1];
    y.in_compact = 1, y.out_compact = 0;
    z.in = s1, z.out = fw, z.in_v = s1v, z.out_v = nftot;
    z.in_compact = 1, z.out_compact = 0;
  }
  return q;
}

// ---- HIP launches

namespace {

template <typename T, int N>
void f3_cols_n(const F3Cols<T> &p, unsigned gx, unsigned gy, size_t lds, hipStream_t s) {
  const void *k = reinterpret_cast<const void *>(&fft3_cols_kernel<T, N>);
  hip_check(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)),
            "hipFuncSetAttribute(fft3_cols_kernel)");
  fft3_cols_kernel<T, N><<<dim3(gx, gy), F3_THREADS, lds, s>>>(p);
  launch_check("fft3_cols_kernel");
}

template <typename T, int N>
void f3_rows_n(const F3Rows<T> &p, unsigned gx, unsigned gy, size_t lds, hipStream_t s) {
  const void *k = reinterpret_cast<const void *>(&fft3_rows_kernel<T, N>);
  hip_check(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)),
            "hipFuncSetAttribute(fft3_rows_kernel)");
  fft3_rows_kernel<T, N><<<dim3(gx, gy), F3_THREADS, lds, s>>>(p);
  launch_check("fft3_rows_kernel");
}

// Calls LAUNCH<n> for the power of two n in [F3_MIN_N, F3_MAX_N].
#define F3_DISPATCH(n, LAUNCH, ...)                                          \
  switch (n) {                                                               \
    case 16: LAUNCH<T, 16>(__VA_ARGS__); break;                              \
    case 32: LAUNCH<T, 32>(__VA_ARGS__); break;                              \
    case 64: LAUNCH<T, 64>(__VA_ARGS__); break;                              \
    case 128: LAUNCH<T, 128>(__VA_ARGS__); break;                            \
    case 256: LAUNCH<T, 256>(__VA_ARGS__); break;                            \
    case 512: LAUNCH<T, 512>(__VA_ARGS__); break;                            \
    case 1024: LAUNCH<T, 1024>(__VA_ARGS__); break;                          \
    case 2048: LAUNCH<T, 2048>(__VA_ARGS__); break;                          \
    case 4096: LAUNCH<T, 4096>(__VA_ARGS__); break;                          \
    default: fail(FINUFFT_ERR_INVALID_ARGUMENT, "fft3: unsupported line length"); \
  }

// One strided-axis pass over blk vectors.
template <typename T>
void f3_cols(const F3Cols<T> &p, int blk, hipStream_t s) {
  const int n = p.n;
  const int64_t nb = p.nb;
  const int g = f3_lines<T>(n);
  const int64_t nblk_a = (p.A + g - 1) / g;
  if (nblk_a * nb > INT32_MAX || blk > 65535 || blk < 1 || nb < 1)
    fail(FINUFFT_ERR_MAXNALLOC, "fft3: grid too large");
  F3Cols<T> q = p;
  q.nblk_a = int(nblk_a);
  const size_t lds = f3_lds_cols<T>(n);
  F3_DISPATCH(n, f3_cols_n, q, unsigned(nblk_a * nb), unsigned(blk), lds, s)
}

// The x pass over blk vectors.
template <typename T>
void f3_rows(const F3Rows<T> &p, int blk, hipStream_t s) {
  const int n = p.n;
  const int g = f3_lines<T>(n);
  const int64_t nblk = (int64_t(p.nrows) + g - 1) / g;
  if (nblk > INT32_MAX || nblk < 1 || blk > 65535 || blk < 1)
    fail(FINUFFT_ERR_MAXNALLOC, "fft3: grid too large");
  const size_t lds = f3_lds_rows<T>(n);
  F3_DISPATCH(n, f3_rows_n, p, unsigned(nblk), unsigned(blk), lds, s)
}

#undef F3_DISPATCH

}  // namespace

template <typename T>
void Fft3<T>::setup(const int ms[3], const int nf[3], int batch, int modeord) {
  if (!fft3_supported(3, nf)) fail(FINUFFT_ERR_INVALID_ARGUMENT, "fft3: unsupported grid");
  for (int d = 0; d < 3; ++d) {
    ms_[d] = ms[d];
    nf_[d] = nf[d];
    if (ms[d] < 1 || ms[d] > nf[d]) fail(FINUFFT_ERR_INVALID_ARGUMENT, "fft3: modes exceed grid");
    std::vector<T> h(2 * size_t(nf[d]));
    for (int j = 0; j < nf[d]; ++j) {
      const double a = 2.0 * PI * double(j) / double(nf[d]);
      h[2 * j] = T(std::cos(a));
      h[2 * j + 1] = T(std::sin(a));
    }
    tw_[d].resize(h.size());
    hip_check(hipMemcpy(tw_[d].get(), h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice),
              "hipMemcpy(fft3 twiddles)");
  }
  modeord_ = modeord;
  batch_ = batch;
  s1_.resize(size_t(nf[0]) * nf[1] * ms[2] * size_t(batch));
}

template <typename T>
void Fft3<T>::to_modes(C *fw, C *fk, const T *const ker[3], int sign, int blk, hipStream_t s) {
  if (blk > batch_) fail(FINUFFT_ERR_INVALID_ARGUMENT, "fft3: batch exceeds setup");
  const T *tw[3] = {tw_[0].get(), tw_[1].get(), tw_[2].get()};
  const F3Passes<T> q = f3_passes<T>(1, ms_, nf_, modeord_, fw, fk, s1_.get(), ker, tw, sign);
  f3_cols<T>(q.z, blk, s);
  f3_cols<T>(q.y, blk, s);
  f3_rows<T>(q.x, blk, s);
}

template <typename T>
void Fft3<T>::to_grid(const C *fk, C *fw, const T *const ker[3], int sign, int blk,
                      hipStream_t s) {
  if (blk > batch_) fail(FINUFFT_ERR_INVALID_ARGUMENT, "fft3: batch exceeds setup");
  const T *tw[3] = {tw_[0].get(), tw_[1].get(), tw_[2].get()};
  const F3Passes<T> q =
      f3_passes<T>(0, ms_, nf_, modeord_, fw, const_cast<C *>(fk), s1_.get(), ker, tw, sign);
  f3_rows<T>(q.x, blk, s);
  f3_cols<T>(q.y, blk, s);
  f3_cols<T>(q.z, blk, s);
}

template class Fft3<float>;
template class Fft3<double>;
template size_t f3_lds_cols<float>(int);
template size_t f3_lds_cols<double>(int);
template size_t f3_lds_rows<float>(int);
template size_t f3_lds_rows<double>(int);
template F3Passes<float> f3_passes<float>(int, const int[3], const int[3], int, cplx_t<float> *,
                                          cplx_t<float> *, cplx_t<float> *,
                                          const float *const[3], const float *const[3], int);
template F3Passes<double> f3_passes<double>(int, const int[3], const int[3], int,
                                            cplx_t<double> *, cplx_t<double> *,
                                            cplx_t<double> *, const double *const[3],
                                            const double *const[3], int);

}  // namespace fa
