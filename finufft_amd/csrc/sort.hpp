// Deterministic bin sort of the nonuniform points and the subproblem table.
//
// The reference GPU sort (calc_bin_size_noghost + exclusive scan +
// calc_inverse_of_global_sort_idx, include/cufinufft/spreadinterp.hpp:299-325)
// orders points by bin but leaves the order inside a bin to atomicAdd
// arrival. This sort is a stable LSD radix sort on the same bin key, so its
// permutation is fully determined: bins in x-fastest order, points inside a
// bin by increasing original index (the order of the reference's CPU
// bin_sort_singlethread_impl, include/finufft/spread.hpp:459-584).
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"

namespace fa {

struct BinGeom {
  int dim = 1;
  int nf[3] = {1, 1, 1};
  int binsize[3] = {1, 1, 1};
  int nbins[3] = {1, 1, 1};
  // 0: the reference's bins of the folded coordinate, floor(r / binsize)
  //    (include/cufinufft/spreadinterp.hpp:214-229);
  // 1: bins of the first support node s = ceil(r - ns/2) (wrapped into
  //    [0, nf)), computed with the spreading kernels' own fold, so that a bin
  //    of size b owns exactly the nodes [off, off + b + ns - 1).
  int start_bins = 0;
  int ns = 2;
  int fold64 = 0;
  // colS > 0 (3-D start bins only) refines the key inside each bin to
  // (sub-column, local z of the first node): a sub-column is colS x colS
  // first-node positions in x and y over the bin's whole z extent. Points of a
  // sub-column are then contiguous and ordered by z (ties by index).
  int colS = 0;
  // bank = 1 (start bins, no sub-columns) refines the key inside each bin by
  // the LDS bank class of the point's first tile cell: (lx + P0 ly +
  // P0 P1 lz) mod 32 with P = binsize + ns - 1. The interpolation kernel deals
  // a bin's points to lanes with a stride, so that each half-wave reads 32
  // different bank classes.
  int bank = 0;
  int total() const { return nbins[0] * nbins[1] * nbins[2]; }
  int nsub() const { return colS ? (binsize[0] / colS) * (binsize[1] / colS) : 1; }
  int64_t nkeys() const {
    return int64_t(total()) * nsub() * (colS ? binsize[2] : 1) * (bank && !colS ? 32 : 1);
  }
};

// Values per point of the packed coordinates the sort gathers from.
__host__ __device__ constexpr int pack_stride(int dim) { return dim == 3 ? 4 : dim; }

// Scratch reused across setpts calls (grows only). `packed` holds the points'
// coordinates packed per point (pack_stride(dim) values of the plan's real
// type) while sorted copies are made.
struct SortScratch {
  dbuf<int> keys_a, keys_b, vals_b, hist, scan_tmp;
  dbuf<unsigned char> packed;
};

// idx[M]: sorted point order. binstart[nbins+1]: first sorted position of each
// bin, binstart[nbins] = M. With g.colS, colstart[nbins * nsub + 1] receives
// the first sorted position of each (bin, sub-column); with g.bank (and a
// non-null colstart), colstart[nbins * 32 + 1] that of each (bin, bank class). sorted[d] (M values
// each, d < dim; or null): the coordinates in the sorted order. Nothing is
// synchronized with the host.
template <typename T>
void bin_sort(const BinGeom &g, int M, const T *const xyz[3], int *idx, int *binstart,
              SortScratch &s, hipStream_t stream, int *colstart, T *const sorted[3]);

// idx[i] = i.
void iota(int *idx, int M, hipStream_t stream);

// substart[nbins+1] (exclusive scan of ceil(count/maxsub) over the bins with
// more than dense_min points, 0 for the others; last entry = total) and
// sub2bin[capacity], capacity = subproblem_capacity(M, nbins, maxsub).
void build_subproblems(const int *binstart, int nbins, int maxsub, int dense_min, int *substart,
                       int *sub2bin, SortScratch &s, hipStream_t stream);
inline int subproblem_capacity(int M, int nbins, int maxsub) {
  return nbins + (M + maxsub - 1) / maxsub;
}

// out[i] = in[idx[i]] for i in [0, M).
template <typename T>
void permute(const int *idx, const T *in, T *out, int M, hipStream_t stream);

// out[i] = sum_{k<i} in[k] for i in [0, n); in and out may alias.
void exclusive_scan(const int *in, int *out, int n, SortScratch &s, hipStream_t stream);

}  // namespace fa
