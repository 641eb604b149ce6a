// Pruned 3-D transforms between a plan's fine grid and its modes, one axis
// per launch, for fine grids whose sides are powers of two in [16, 4096].
//
// The reference spreads into the fine grid, runs a full 3-D FFT over it and
// then reads the ms0 x ms1 x ms2 mode cells out of the result with the
// deconvolution (src/cuda/execute.cu:105-153: spread, cufftExec, deconvolve);
// type 2 amplifies the modes into the zero-padded grid and transforms all of
// it (execute.cu:156-204). Only the mode cells of the transformed grid are
// used (type 1), or are nonzero before the transform (type 2), so the same
// numbers come from three line passes that drop what is not needed:
//   type 1: z pass  fw (nf0 nf1 nf2) -> S1 (nf0 nf1 ms2), keeping the ms2
//           mode planes of each z line;
//           y pass  S1 -> S2 (nf0 ms1 ms2), keeping ms1 of each y line;
//           x pass  S2 -> fk (ms0 ms1 ms2), keeping ms0 of each row, divided
//           by the kernel's Fourier coefficients and written in the user's
//           mode order.
//   type 2: the same passes in reverse, each starting from the zero-padded
//           mode rows / columns: x pass fk -> S2 (amplification on the way),
//           y pass S2 -> S1, z pass S1 -> fw.
// Grid traffic per transform is about 2.6 grid volumes, against an
// amplification or deconvolution pass plus a multi-pass 3-D FFT of the
// whole grid. S1 is a work buffer of the plan; S2 lives in the part of fw
// that the pass which needs it does not touch (type 1: fw is consumed by the
// z pass before the y pass writes S2 there; type 2: the y pass reads S2 from
// fw before the z pass overwrites fw).
//
// Each pass transforms lines of length N in LDS: a workgroup of 256 threads
// takes G lines (G N = 8192 complex values in fp32, 4096 in fp64, G <= 64),
// runs log8 N radix-8 passes (then one radix-4 or radix-2 pass) as
// self-sorting Stockham steps, the first reading the lines from global
// memory and the last writing the kept outputs to global memory, the steps
// between through LDS. Lines along y and z are taken G adjacent in x, so
// every global access is a run of G consecutive cells.
#pragma once

#include <cstdint>

#include "common.hpp"
#include "spreadinterp.hpp"

namespace fa {

constexpr int F3_THREADS = 256;
constexpr int F3_MIN_N = 16, F3_MAX_N = 4096;

// One strided-axis pass (y or z lines). Line a (in [0, A)) at outer index b
// and vector t holds the cells in[t in_v + b in_b + pos A + a] for pos along
// the axis: pos a cell index in [0, N) (full) or a compact mode index in
// [0, m) (in_compact), where compact index c stands for cell c (c < h) or
// N - m + c (c >= h), h = m - m / 2. The output is written the same way.
template <typename T>
struct F3Cols {
  const cplx_t<T> *in;
  cplx_t<T> *out;
  int64_t A;             // lines per outer index (cells of the faster axes)
  int64_t in_b, out_b;   // outer-index strides
  int64_t in_v, out_v;   // vector strides
  int64_t nb;            // outer indices
  int n;                 // line length N
  int nblk_a;            // workgroups along a: ceil(A / G) (set at launch)
  int m;                 // modes along the axis
  int in_compact, out_compact;
  const T *tw;           // N complex values exp(2 pi i j / N), interleaved
  T sign;                // exponent sign of the transform
};

// The x pass. Rows are indexed by the compact mode pair (cy, cz), r = cy +
// m1 cz, r < m1 m2. to_modes (type 1): row r of S2 (nf0 cells at
// s2 + t s2_v + r N) -> fk entries (kx, ky, kz) in the user's mode order,
// divided by k0[|fx|] k1[|fy|] k2[|fz|]. Otherwise (type 2) the reverse,
// with the modes multiplied by the reciprocal kernel coefficients and the
// row zero-padded.
template <typename T>
struct F3Rows {
  cplx_t<T> *s2;
  cplx_t<T> *fk;
  int64_t s2_v, fk_v;    // vector strides
  int ms[3];
  int modeord;
  int to_modes;
  int nrows;             // m1 m2
  int n;                 // row length nf0
  const T *ker[3];       // kernel Fourier coefficients, nf_d / 2 + 1 each
  const T *tw;
  T sign;
};

// The three passes of one transform of a plan (vector 0; further vectors at
// the vector strides). to_modes 1: z, y, x (type-1 direction), else x, y,
// z (type-2 direction). fw is the fine grid, s1 the work buffer of
// nf0 nf1 ms2 cells per vector, tw[d] the twiddle table of axis d.
template <typename T>
struct F3Passes {
  F3Cols<T> z, y;
  F3Rows<T> x;
};
template <typename T>
F3Passes<T> f3_passes(int to_modes, const int ms[3], const int nf[3], int modeord,
                      cplx_t<T> *fw, cplx_t<T> *fk, cplx_t<T> *s1, const T *const ker[3],
                      const T *const tw[3], int sign);

// Lines per workgroup for length N.
template <typename T>
constexpr int f3_lines(int n) {
  const int g = (sizeof(T) == 4 ? 8192 : 4096) / n;
  return g > 64 ? 64 : (g < 1 ? 1 : g);
}

// LDS bytes of a column pass / a row pass of length n (the data layout plus
// the twiddle table).
template <typename T>
size_t f3_lds_cols(int n);
template <typename T>
size_t f3_lds_rows(int n);

// True when every side is a power of two in [F3_MIN_N, F3_MAX_N] (3-D).
bool fft3_supported(int dim, const int nf[3]);

// The plan's pruned transforms (types 1 and 2 directions).
template <typename T>
class Fft3 {
 public:
  using C = cplx_t<T>;
  // Prepares the twiddle tables and the S1 work buffer for batch vectors.
  void setup(const int ms[3], const int nf[3], int batch, int modeord);
  // fw (blk vectors, consumed) -> fk, deconvolved (type-1 direction).
  void to_modes(C *fw, C *fk, const T *const ker[3], int sign, int blk, hipStream_t s);
  // fk -> fw (blk vectors), amplified (type-2 direction).
  void to_grid(const C *fk, C *fw, const T *const ker[3], int sign, int blk, hipStream_t s);

 private:
  int ms_[3] = {1, 1, 1}, nf_[3] = {1, 1, 1};
  int modeord_ = 0, batch_ = 0;
  dbuf<T> tw_[3];
  dbuf<C> s1_;
};

}  // namespace fa
