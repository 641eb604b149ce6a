// Interface between the plan and the spreading / interpolation kernels.
// The kernels themselves live in spreadinterp.hip and are instantiated once
// per (precision, dimension) translation unit.
#pragma once

#include <hip/hip_runtime.h>

#include "kernel_math.hpp"

namespace fa {

template <typename T>
struct cplx_of;
template <>
struct cplx_of<float> {
  using type = float2;
};
template <>
struct cplx_of<double> {
  using type = double2;
};
template <typename T>
using cplx_t = typename cplx_of<T>::type;

// Everything a spread or interp launch needs about the points, their sorted
// order and the fine grid. All pointers are device pointers.
template <typename T>
struct SpreadCtx {
  int dim = 1;
  int ns = 2;
  // 1: point driven (GM / GM-sort), 2: LDS subproblems (SM), 3: output
  // driven (spreading only; LDS subproblems, np points staged at a time)
  int method = 1;
  int M = 0;
  int nf[3] = {1, 1, 1};
  const T *xyz[3] = {nullptr, nullptr, nullptr};
  // Coordinates permuted into the sorted order (sxyz[d][i] = xyz[d][idx[i]]);
  // null when not materialized.
  const T *sxyz[3] = {nullptr, nullptr, nullptr};
  KerParams<T> kp{};

  const int *idx = nullptr;  // point order (sorted or identity), length M

  // bin geometry of the sort (x fastest)
  int binsize[3] = {1, 1, 1};
  int nbins[3] = {1, 1, 1};
  const int *binstart = nullptr;  // nbins_tot + 1 entries, binstart[nbins_tot] = M

  // subproblems (method 2): subproblem s covers points
  // idx[binstart[b] + (s - substart[b]) * maxsub ...] of bin b = sub2bin[s]
  int maxsub = 1024;
  const int *substart = nullptr;  // nbins_tot + 1 entries, last = total count
  const int *sub2bin = nullptr;
  int sub_capacity = 0;  // launch grid (>= the device-side total)
  size_t lds_bytes = 0;   // LDS bytes of the method-2 interpolation kernels
  size_t lds_spread = 0;  // LDS bytes of the method-2 spreading kernels (fp64 tiles)
  // Spread point-driven (method-1 kernel) although the plan is method 2: set
  // when the fp64 spreading tile of the plan's bins exceeds the LDS limit
  // (interpolation plans, whose bins are sized for plan-precision tiles).
  int spread_gm = 0;
  // Method-2 spreading by whole bins in 2^dim parity colors with a plain
  // (non-atomic) tile flush; requires per dimension an even number (>= 2) of
  // equal bins of at least 2 ceil(ns/2) + 1 cells.
  int colored = 0;

  // Column-window spreading (3-D, fp32, ns <= 8): colS is the sub-column side
  // (0 = off) and colstart the sorted start of every (bin, sub-column)
  // (nbins_tot * nsub + 1 entries). Colored launches skip bins with more than
  // dense_min points; those bins are covered by n_dense_sub subproblems
  // (substart / sub2bin built over the dense bins only) with an atomic flush.
  // Interpolation by subproblems (substart / sub2bin cover every bin).
  int interp_sub = 0;

  int colS = 0;
  const int *colstart = nullptr;
  // Interpolation plans: each bin is ordered by the LDS bank class of the
  // points' first tile cells (32 classes), and colstart holds the sorted
  // start of every (bin, class) (nbins_tot * 32 + 1 entries).
  int bank = 0;
  int dense_min = 0;
  int n_dense_sub = 0;

  // Output-driven spreading (method 3): points staged per batch and the
  // kernel's LDS bytes (fp64 tile + od_stage_bytes).
  int od_np = 0;
  size_t lds_od = 0;
};

// LDS bytes after the tile of a bank-ordered interpolation unit: its class
// bounds, leftover prefix, full rounds and dealing mode (68 ints).
constexpr size_t INTERP_BANK_LDS = 512;

// Staging bytes of the output-driven spreader for np points: strengths (re
// and im arrays), kernel values (dim * ns per point) and first tile cells
// (dim ints per point).
inline constexpr size_t od_stage_bytes(int dim, int ns, int np, size_t real_bytes) {
  return size_t(np) * ((2 + size_t(dim) * ns) * real_bytes + size_t(dim) * sizeof(int));
}

// Global-memory pointer types. A pointer read from memory (a work
// descriptor) is generic to the compiler, which then uses flat instructions
// and cannot load through it with scalar loads; holding it in one of these
// types keeps every access a global (or, where uniform, scalar) one.
template <typename P>
using gptr = __attribute__((address_space(1))) P *;
template <typename P>
__device__ __forceinline__ gptr<P> to_global(P *p) {
  return (gptr<P>)p;
}
// Read-only tables (the Horner coefficients): the constant address space,
// whose loads the compiler may always make scalar.
template <typename P>
using kptr = __attribute__((address_space(4))) const P *;
template <typename P>
__device__ __forceinline__ kptr<P> to_const(const P *p) {
  return (kptr<P>)p;
}

// Column spreading geometry. A wave owns the footprint of one sub-column:
// 8 x 8 (x, y) grid columns, one per lane, for the points whose first node
// lies in an S x S square, S the largest power of two with S + ns - 1 <= 8;
// each lane holds its column's whole depth in the bin's tile (bin depth at
// most COL_ZC) in registers. The tile is fp64; every wave stages COL_BATCH
// points of COL_REC floats at a time. Eight waves with 32-point batches keep
// the workgroup's LDS at two workgroups per CU (16 waves) where four waves
// with 64-point batches gave 8 waves per CU.
constexpr int COL_MAX_NS = 8;
constexpr int COL_WAVES = 8;  // waves per workgroup
constexpr int COL_BATCH = 32; // points a wave stages at a time (<= 64)
constexpr int COL_REC = 32;   // floats staged per point
constexpr int COL_ZC = 8;     // largest bin depth (first-node planes)
inline constexpr int col_side(int ns) { return ns <= 5 ? 4 : (ns <= 7 ? 2 : 1); }
inline constexpr size_t col_stage_bytes() {
  return size_t(COL_WAVES) * COL_BATCH * COL_REC * sizeof(float);
}
// LDS bytes of a column-spreading unit with a tile of `cells` cells.
inline constexpr size_t col_lds_bytes(int64_t cells) {
  return size_t(cells) * 2 * sizeof(double) + col_stage_bytes();
}

// c (M x blk, contiguous per vector) -> fw (nf_tot x blk). fw must be zeroed
// by the caller. Dispatches on ctx.dim / ctx.ns / ctx.method.
template <typename T>
void spread(const SpreadCtx<T> &ctx, const cplx_t<T> *c, cplx_t<T> *fw, int blk,
            hipStream_t stream);

// fw (nf_tot x blk) -> c (M x blk).
template <typename T>
void interp(const SpreadCtx<T> &ctx, cplx_t<T> *c, const cplx_t<T> *fw, int blk,
            hipStream_t stream);

// fw (nf_tot x blk) -> c (M x blk, may be null) and dc (dim x blk x M): the
// interpolant and its derivatives with respect to each point coordinate,
// through the derivative of the kernel.
template <typename T>
void interp_grad(const SpreadCtx<T> &ctx, cplx_t<T> *c, cplx_t<T> *dc, const cplx_t<T> *fw,
                 int blk, hipStream_t stream);

// One problem of a fused interpolation launch over many independent type-2
// problems (bin subproblems of each problem, nvec vectors per round).
template <typename T>
struct InterpItem {
  SpreadCtx<T> ctx;   // interp_sub = 1: substart / sub2bin cover every bin
  int64_t nftot;      // fine-grid cells per vector
  int nsub;           // subproblems of the problem (its work units)
  int64_t unit_begin;  // first work unit of this problem in the launch
};

// Interpolates all problems of a launch, one workgroup per work unit: unit u
// of problem p is its subproblem u - unit_begin, over all nvec vectors of the
// round (the kernel weights of a point are computed once and applied to
// every vector). fw_base[p] is the problem's vector-0 grid of the round
// (vectors nftot apart), c_base[p] its value array (vector c_vec_offset is
// the round's first). All problems share the kernel (width ns and
// parameters kp, passed once as launch arguments).
template <typename T>
void interp_batch(const InterpItem<T> *items, int nprob, int dim, int ns, const KerParams<T> &kp,
                  const cplx_t<T> *const *fw_base, cplx_t<T> *const *c_base,
                  int64_t c_vec_offset, int nvec, int64_t total_units, size_t lds_bytes,
                  hipStream_t stream);

// Spreads all problems of a launch, one workgroup per work unit (a
// subproblem u - unit_begin of problem p; substart / sub2bin cover every
// bin), for each of the nvec vectors of the round in turn: c_base[p] holds
// the problem's strengths (vector c_vec_offset is the round's first),
// fw_base[p] its vector-0 grid of the round (vectors nftot apart), which the
// caller has cleared. lds_bytes: the largest fp64 spreading tile.
template <typename T>
void spread_batch(const InterpItem<T> *items, int nprob, int dim, int ns,
                  const cplx_t<T> *const *c_base, int64_t c_vec_offset, cplx_t<T> *const *fw_base,
                  int nvec, int64_t total_units, size_t lds_bytes, hipStream_t stream);

// Per-dimension entry points (one translation unit each).
template <typename T, int DIM>
void spread_batch_dim(const InterpItem<T> *items, int nprob, int ns,
                      const cplx_t<T> *const *c_base, int64_t c_vec_offset,
                      cplx_t<T> *const *fw_base, int nvec, int64_t total_units, size_t lds_bytes,
                      hipStream_t stream);
template <typename T, int DIM>
void spread_dim(const SpreadCtx<T> &ctx, const cplx_t<T> *c, cplx_t<T> *fw, int blk,
                hipStream_t stream);
template <typename T, int DIM>
void interp_dim(const SpreadCtx<T> &ctx, cplx_t<T> *c, const cplx_t<T> *fw, int blk,
                hipStream_t stream);
template <typename T, int DIM>
void interp_batch_dim(const InterpItem<T> *items, int nprob, int ns, const KerParams<T> &kp,
                      const cplx_t<T> *const *fw_base, cplx_t<T> *const *c_base,
                      int64_t c_vec_offset, int nvec, int64_t total_units, size_t lds_bytes,
                      hipStream_t stream);
template <typename T, int DIM>
void interp_grad_dim(const SpreadCtx<T> &ctx, cplx_t<T> *c, cplx_t<T> *dc, const cplx_t<T> *fw,
                     int blk, hipStream_t stream);


}  // namespace fa
