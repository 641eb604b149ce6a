// Pruned, fused amplification + FFT of 2-D type-2 fine grids, in LDS.
//
// The reference amplifies the modes into the zero-padded fine grid and runs
// a full 2-D FFT on it (src/cuda/execute.cu:105-204: deconvolve, then
// cufftExec). Before the transform only the ms0 x ms1 mode cells of the
// nf0 x nf1 grid are nonzero, so the same grid is produced by
//   * a column pass: each workgroup reads the modes of a few mode columns
//     (the amplification by the kernel's Fourier coefficients on the way),
//     transforms them along y (length nf1, zero padding implied) in LDS and
//     writes those ms0 grid columns;
//   * a row pass: each workgroup reads the ms0 mode entries of a few grid
//     rows, transforms them along x (length nf0) in LDS and writes the rows.
// The grid is written about 1.5 times and read about 0.5 times, against an
// amplification pass plus a multi-pass 2-D FFT over the whole padded grid.
// The transforms are mixed-radix (4, 2, 3, 5) self-sorting Stockham passes
// between two LDS buffers, with twiddles exp(2 pi i j / n) from two small
// tables staged in LDS.
// Blocks are renumbered so that neighbouring column / row groups, whose
// writes share cache lines, run on the same XCD (one L2).
#pragma once

#include <cstdint>
#include <vector>

#include "plan_kernels.hpp"

namespace fa {

// Radices of a length n = 2^a 3^b 5^c in pass order, 4 bits each.
struct RadixPlan {
  uint64_t code = 0;
  int npass = 0;
};
// false when n has another prime factor or needs more than 16 passes.
bool radix_plan(int n, RadixPlan &out);

constexpr int FFT_THREADS = 256;
// Complex elements a workgroup aims to transform: a workgroup takes as many
// columns (up to 16, a power of two) or rows as fit this budget, at least one.
// Small workgroups (8 KB of LDS per buffer) let many share a CU.
template <typename T>
constexpr int fft_col_elems() {
  return sizeof(T) == 4 ? 1024 : 512;
}
template <typename T>
constexpr int fft_row_elems() {
  return sizeof(T) == 4 ? 1024 : 512;
}
// Twiddles are staged in LDS as two tables, w^j for j < 64 and w^(64 j) for
// j < n / 64 (n <= 4096), and multiplied: w^i = w^(i mod 64) w^(64 (i / 64)).
constexpr int FFT_TW_LO = 64, FFT_TW_HI = 64;
constexpr int FFT_MAX_N = FFT_TW_LO * FFT_TW_HI;

// One problem of the fused launches. The column pass gives every vector ncol
// workgroups of cw (a power of two) mode columns; the row pass gives every
// vector nrow workgroups of rw grid rows. col_begin / row_begin are the first
// blocks of this problem in the respective launch. All problems of a launch
// share the grid shape.
template <typename T>
struct FusedFftItem {
  int ms[2], nf[2];
  int modeord;
  int cw, lcw, ncol;  // lcw = log2 cw
  int rw, lrw, nrow;  // lrw = log2 rw
  RadixPlan rp[2];
  const T *rker[2];  // reciprocal kernel Fourier coefficients per axis
  const T *tw[2];   // (cos, sin)(2 pi j / nf_d), j < nf_d, as pairs
  int64_t nmodes, nftot;
  int64_t col_begin, row_begin;
};

// Whether a 2-D grid can take the fused path (both lengths 2,3,5-smooth and
// at most FFT_MAX_N), its workgroup shapes and the LDS bytes of the column
// and row launches.
template <typename T>
bool fused_fft_shape(const int nf[2], int &cw, int &lcw, int &rw, int &lrw, size_t &lds_col,
                     size_t &lds_row);

// Twiddle table of length n (2n values: cos, sin of 2 pi j / n), computed in
// double.
template <typename T>
void fft_twiddles(int n, std::vector<T> &out);

// Raises the kernels' LDS limit (before any graph capture).
template <typename T>
void fused_fft_prepare();

// Column pass of a round: modes fk (vector fk_vec_offset of the round's first)
// -> the mode columns of fw, transformed along y with exponent sign `sign`.
template <typename T>
void fused_fft_cols(const FusedFftItem<T> *items, const cplx_t<T> *const *fk_base,
                    cplx_t<T> *const *fw_base, int nprob, int64_t fk_vec_offset, int sign,
                    int64_t total_blocks, size_t lds_bytes, hipStream_t stream);
// Row pass: the mode columns of every row of fw -> the whole row, transformed
// along x.
template <typename T>
void fused_fft_rows(const FusedFftItem<T> *items, cplx_t<T> *const *fw_base, int nprob, int sign,
                    int64_t total_blocks, size_t lds_bytes, hipStream_t stream);

}  // namespace fa
