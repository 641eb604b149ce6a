// Deterministic bin sort (see sort.hpp).
//
// Layout of the work: the bin key of every point is computed once (4 bytes
// per point); then 1..4 passes of a stable LSD radix sort of (key, index)
// pairs with up to 8-bit digits. Each pass is two kernels over 4096-point
// tiles (256 lanes x 16 points) plus a scan of the digit-major tile
// histogram:
//   radix_count    tile histogram of the digit
//   exclusive_scan global digit offsets per tile
//   radix_scatter  stable rank inside the tile, then scatter
// Inside a wave the rank of a point among the points with the same digit
// comes from ballots over the digit bits (no atomics anywhere), and each wave
// owns a contiguous quarter of the tile, so the order is the original index
// order. Bin start offsets are read off the sorted keys.
#include "sort.hpp"

#include <type_traits>

#include "kernel_math.hpp"

namespace fa {
namespace {

constexpr int SCAN_T = 256;
constexpr int SCAN_I = 8;
constexpr int SCAN_TILE = SCAN_T * SCAN_I;

constexpr int RS_T = 256;
constexpr int RS_WAVES = RS_T / 64;
constexpr int RS_I = 16;
constexpr int RS_TILE = RS_T * RS_I;

__device__ __forceinline__ unsigned long long lanes_below() {
  return (1ull << __lane_id()) - 1ull;
}

// Exclusive scan of v over the 256-lane block; total receives the block sum.
__device__ int block_exclusive_scan(int v, int *wave_sums, int &total) {
  const int lane = __lane_id(), wave = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wave_sums[wave] = x;
  __syncthreads();
  int pre = 0;
  total = 0;
#pragma unroll
  for (int w = 0; w < SCAN_T / 64; ++w) {
    const int t = wave_sums[w];
    pre += (w < wave) ? t : 0;
    total += t;
  }
  __syncthreads();
  return pre + x - v;
}

__global__ void __launch_bounds__(SCAN_T) scan_reduce_kernel(const int *in, int n, int *sums) {
  __shared__ int ws[SCAN_T / 64];
  const int64_t base = int64_t(blockIdx.x) * SCAN_TILE + threadIdx.x * SCAN_I;
  int s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_I; ++k) s += (base + k < n) ? in[base + k] : 0;
  int total;
  block_exclusive_scan(s, ws, total);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ void __launch_bounds__(SCAN_T)
    scan_apply_kernel(const int *in, int *out, int n, const int *offs) {
  __shared__ int ws[SCAN_T / 64];
  const int64_t base = int64_t(blockIdx.x) * SCAN_TILE + threadIdx.x * SCAN_I;
  int v[SCAN_I];
  int s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_I; ++k) {
    v[k] = (base + k < n) ? in[base + k] : 0;
    s += v[k];
  }
  int total;
  int run = block_exclusive_scan(s, ws, total) + (offs ? offs[blockIdx.x] : 0);
#pragma unroll
  for (int k = 0; k < SCAN_I; ++k) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
}

void scan_rec(const int *in, int *out, int n, int *tmp, hipStream_t stream) {
  const int nb = (n + SCAN_TILE - 1) / SCAN_TILE;
  if (nb == 1) {
    scan_apply_kernel<<<1, SCAN_T, 0, stream>>>(in, out, n, nullptr);
    launch_check("scan_apply_kernel");
    return;
  }
  int *sums = tmp;
  scan_reduce_kernel<<<nb, SCAN_T, 0, stream>>>(in, n, sums);
  launch_check("scan_reduce_kernel");
  scan_rec(sums, sums, nb, tmp + nb, stream);
  scan_apply_kernel<<<nb, SCAN_T, 0, stream>>>(in, out, n, sums);
  launch_check("scan_apply_kernel");
}

// Lanes of the wave holding the same digit as this lane (valid lanes only).
__device__ __forceinline__ unsigned long long digit_peers(unsigned d, bool valid, int bits) {
  unsigned long long m = __ballot(valid);
  for (int b = 0; b < bits; ++b) {
    const bool bit = (d >> b) & 1u;
    const unsigned long long bal = __ballot(bit);
    m &= bit ? bal : ~bal;
  }
  return valid ? m : 0ull;
}

__global__ void __launch_bounds__(RS_T)
    radix_count_kernel(const int *keys, int n, int shift, int bits, int *hist, int ntiles) {
  __shared__ int cnt[RS_WAVES][256];
  for (int i = threadIdx.x; i < RS_WAVES * 256; i += RS_T) (&cnt[0][0])[i] = 0;
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = __lane_id();
  const unsigned mask = (1u << bits) - 1u;
  const int64_t base = int64_t(blockIdx.x) * RS_TILE + wave * (RS_I * 64);
  for (int k = 0; k < RS_I; ++k) {
    const int64_t i = base + k * 64 + lane;
    const bool valid = i < n;
    const unsigned d = valid ? (unsigned(keys[i]) >> shift) & mask : 0u;
    const unsigned long long m = digit_peers(d, valid, bits);
    if (valid && (m & lanes_below()) == 0ull) cnt[wave][d] += __popcll(m);
  }
  __syncthreads();
  for (int dd = threadIdx.x; dd <= int(mask); dd += RS_T) {
    int s = 0;
#pragma unroll
    for (int w = 0; w < RS_WAVES; ++w) s += cnt[w][dd];
    hist[int64_t(dd) * ntiles + blockIdx.x] = s;
  }
}

__global__ void __launch_bounds__(RS_T)
    radix_scatter_kernel(const int *keys_in, const int *vals_in, int *keys_out, int *vals_out,
                         int n, int shift, int bits, const int *offs, int ntiles) {
  __shared__ int cnt[RS_WAVES][256];
  for (int i = threadIdx.x; i < RS_WAVES * 256; i += RS_T) (&cnt[0][0])[i] = 0;
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = __lane_id();
  const unsigned mask = (1u << bits) - 1u;
  const int64_t base = int64_t(blockIdx.x) * RS_TILE + wave * (RS_I * 64);
  int key[RS_I], val[RS_I], rank[RS_I];
#pragma unroll
  for (int k = 0; k < RS_I; ++k) {
    const int64_t i = base + k * 64 + lane;
    const bool valid = i < n;
    key[k] = valid ? keys_in[i] : 0;
    val[k] = vals_in ? (valid ? vals_in[i] : 0) : int(i);
    const unsigned d = (unsigned(key[k]) >> shift) & mask;
    const unsigned long long m = digit_peers(d, valid, bits);
    const int r = __popcll(m & lanes_below());
    const int before = cnt[wave][d];
    rank[k] = before + r;
    if (valid && r == 0) cnt[wave][d] = before + __popcll(m);
  }
  __syncthreads();
  for (int dd = threadIdx.x; dd <= int(mask); dd += RS_T) {
    int run = offs[int64_t(dd) * ntiles + blockIdx.x];
#pragma unroll
    for (int w = 0; w < RS_WAVES; ++w) {
      const int t = cnt[w][dd];
      cnt[w][dd] = run;
      run += t;
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < RS_I; ++k) {
    const int64_t i = base + k * 64 + lane;
    if (i < n) {
      const unsigned d = (unsigned(key[k]) >> shift) & mask;
      const int pos = cnt[wave][d] + rank[k];
      if (keys_out) keys_out[pos] = key[k];
      vals_out[pos] = val[k];
    }
  }
}

// Folded coordinate, with non-finite input pinned to 0 so that the key stays
// inside [0, nbins) (finite inputs are unaffected).
template <typename T>
__device__ __forceinline__ T folded(T x, int n) {
  const T r = fold_rescale(x, n);
  return (r >= T(0) && r < T(n)) ? r : T(0);
}

// A point's coordinates packed as one aligned vector of pack_stride(DIM)
// values ((x, y) in 2-D, (x, y, z, 0) in 3-D), so that the sorted copies
// cost one gathered read per point instead of one per coordinate (a random
// read costs a whole cache line, however few of its bytes are used).
template <typename T, int DIM>
__device__ __forceinline__ void pack_point(T *__restrict__ p, const T (&v)[3]) {
  using V2 = typename std::conditional<std::is_same<T, float>::value, float2, double2>::type;
  if constexpr (DIM == 2) {
    *reinterpret_cast<V2 *>(p) = V2{v[0], v[1]};
  } else if constexpr (DIM == 3) {
    if constexpr (std::is_same<T, float>::value) {
      *reinterpret_cast<float4 *>(p) = make_float4(v[0], v[1], v[2], 0.f);
    } else {
      reinterpret_cast<V2 *>(p)[0] = V2{v[0], v[1]};
      reinterpret_cast<V2 *>(p)[1] = V2{v[2], T(0)};
    }
  }
}

template <typename T, int DIM>
__global__ void bin_key_kernel(int M, const T *__restrict__ x, const T *__restrict__ y,
                               const T *__restrict__ z, BinGeom g, T inv0, T inv1, T inv2,
                               int *__restrict__ keys, T *__restrict__ packed) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  T v[3] = {x[i], T(0), T(0)};
  if constexpr (DIM > 1) v[1] = y[i];
  if constexpr (DIM > 2) v[2] = z[i];
  if (DIM > 1 && packed) pack_point<T, DIM>(packed + int64_t(i) * pack_stride(DIM), v);
  if (g.start_bins) {
    int key = 0, stride = 1, l[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      int s = start_index(v[d], g.nf[d], g.ns, g.fold64);
      s = s < 0 ? s + g.nf[d] : s;
      const int b = min(max(s / g.binsize[d], 0), g.nbins[d] - 1);
      l[d] = min(max(s - b * g.binsize[d], 0), g.binsize[d] - 1);
      key += b * stride;
      stride *= g.nbins[d];
    }
    if (DIM == 3 && g.colS > 0) {
      const int nsx = g.binsize[0] / g.colS, nsy = g.binsize[1] / g.colS;
      const int sub = min(l[0] / g.colS, nsx - 1) + nsx * min(l[1] / g.colS, nsy - 1);
      key = (key * (nsx * nsy) + sub) * g.binsize[2] + l[2];
    } else if (g.bank) {
      const int P0 = g.binsize[0] + g.ns - 1, P1 = g.binsize[1] + g.ns - 1;
      const int cell = l[0] + P0 * (DIM > 1 ? l[1] + P1 * (DIM > 2 ? l[2] : 0) : 0);
      key = key * 32 + (cell & 31);
    }
    keys[i] = key;
    return;
  }
  int key = bin_of(folded(v[0], g.nf[0]), inv0, g.nbins[0]);
  if constexpr (DIM > 1) key += g.nbins[0] * bin_of(folded(v[1], g.nf[1]), inv1, g.nbins[1]);
  if constexpr (DIM > 2)
    key += g.nbins[0] * g.nbins[1] * bin_of(folded(v[2], g.nf[2]), inv2, g.nbins[2]);
  keys[i] = key;
}

// out_d[q] = coordinate d of point idx[q], from the packed points.
template <typename T, int DIM>
__global__ void gather_points_kernel(const int *__restrict__ idx, const T *__restrict__ packed,
                                     T *__restrict__ o0, T *__restrict__ o1, T *__restrict__ o2,
                                     int M) {
  using V2 = typename std::conditional<std::is_same<T, float>::value, float2, double2>::type;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= M) return;
  const T *p = packed + int64_t(idx[q]) * pack_stride(DIM);
  if constexpr (DIM == 2) {
    const V2 v = *reinterpret_cast<const V2 *>(p);
    o0[q] = v.x;
    o1[q] = v.y;
  } else {
    if constexpr (std::is_same<T, float>::value) {
      const float4 v = *reinterpret_cast<const float4 *>(p);
      o0[q] = v.x;
      o1[q] = v.y;
      o2[q] = v.z;
    } else {
      const V2 a = reinterpret_cast<const V2 *>(p)[0];
      const V2 b = reinterpret_cast<const V2 *>(p)[1];
      o0[q] = a.x;
      o1[q] = a.y;
      o2[q] = b.x;
    }
  }
}

// binstart[g] = first sorted position whose key / div is >= g, g in [0, ng].
__global__ void bin_bounds_kernel(const int *sk, int M, int ng, int div, int *binstart) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > M) return;
  const int prev = i == 0 ? -1 : sk[i - 1] / div;
  const int cur = i == M ? ng : sk[i] / div;
  for (int b = prev + 1; b <= cur; ++b) binstart[b] = i;
}

__global__ void strided_copy_kernel(const int *in, int stride, int n, int *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i <= n) out[i] = in[int64_t(i) * stride];
}

template <typename T>
__global__ void permute_kernel(const int *__restrict__ idx, const T *__restrict__ in,
                               T *__restrict__ out, int M) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < M) out[i] = in[idx[i]];
}

__global__ void iota_kernel(int *idx, int M) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < M) idx[i] = i;
}

__global__ void two_point_bounds_kernel(int *binstart, int nb, int M) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b <= nb) binstart[b] = (b == nb) ? M : 0;
}

__global__ void numsub_kernel(const int *binstart, int nb, int maxsub, int dense_min,
                              int *numsub) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nb) {
    const int cnt = binstart[b + 1] - binstart[b];
    numsub[b] = cnt > dense_min ? (cnt + maxsub - 1) / maxsub : 0;
  }
  if (b == nb) numsub[b] = 0;
}

__global__ void sub2bin_kernel(const int *substart, int nb, int *sub2bin) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  for (int s = substart[b]; s < substart[b + 1]; ++s) sub2bin[s] = b;
}

inline int blocks_for(int64_t n, int t) { return int((n + t - 1) / t); }

}  // namespace

void exclusive_scan(const int *in, int *out, int n, SortScratch &s, hipStream_t stream) {
  if (n <= 0) return;
  size_t need = 0;
  for (int64_t m = n; m > SCAN_TILE;) {
    m = (m + SCAN_TILE - 1) / SCAN_TILE;
    need += m;
  }
  s.scan_tmp.reserve_at_least(need + 1);
  scan_rec(in, out, n, s.scan_tmp.get(), stream);
}

template <typename T>
void permute(const int *idx, const T *in, T *out, int M, hipStream_t stream) {
  if (M <= 0) return;
  permute_kernel<T><<<blocks_for(M, 256), 256, 0, stream>>>(idx, in, out, M);
  launch_check("permute_kernel");
}
template void permute<float>(const int *, const float *, float *, int, hipStream_t);
template void permute<double>(const int *, const double *, double *, int, hipStream_t);

void iota(int *idx, int M, hipStream_t stream) {
  if (M <= 0) return;
  iota_kernel<<<blocks_for(M, 256), 256, 0, stream>>>(idx, M);
  launch_check("iota_kernel");
}

template <typename T>
void bin_sort(const BinGeom &g, int M, const T *const xyz[3], int *idx, int *binstart,
              SortScratch &s, hipStream_t stream, int *colstart, T *const sorted[3]) {
  const int nb = g.total();
  const bool cols = g.colS > 0 && g.dim == 3 && g.start_bins;
  const bool banks = !cols && g.bank && g.start_bins;
  if (cols && (colstart == nullptr || g.nkeys() > INT32_MAX))
    fail(FINUFFT_ERR_INVALID_ARGUMENT, "bin_sort: invalid sub-column geometry");
  const int nsub = cols ? g.nsub() : 1;
  if (M == 0 || (nb == 1 && !cols && !banks)) {
    iota(idx, M, stream);
    if (sorted)
      for (int d = 0; d < g.dim; ++d) permute<T>(idx, xyz[d], sorted[d], M, stream);
    two_point_bounds_kernel<<<blocks_for(nb + 1, 256), 256, 0, stream>>>(binstart, nb, M);
    launch_check("two_point_bounds_kernel");
    if (cols) {
      two_point_bounds_kernel<<<blocks_for(int64_t(nb) * nsub + 1, 256), 256, 0, stream>>>(
          colstart, nb * nsub, M);
      launch_check("two_point_bounds_kernel");
    }
    return;
  }
  s.keys_a.reserve_at_least(M);
  s.keys_b.reserve_at_least(M);
  s.vals_b.reserve_at_least(M);
  // 2-D and 3-D: the key pass also packs each point's coordinates, gathered
  // once per point into the sorted copies after the last radix pass
  T *packed = nullptr;
  if (sorted && g.dim > 1) {
    s.packed.reserve_at_least(size_t(M) * pack_stride(g.dim) * sizeof(T));
    packed = reinterpret_cast<T *>(s.packed.get());
  }
  const T inv0 = T(1) / T(g.binsize[0]), inv1 = T(1) / T(g.binsize[1]),
          inv2 = T(1) / T(g.binsize[2]);
  switch (g.dim) {
    case 1:
      bin_key_kernel<T, 1><<<blocks_for(M, 256), 256, 0, stream>>>(
          M, xyz[0], xyz[1], xyz[2], g, inv0, inv1, inv2, s.keys_a.get(), nullptr);
      break;
    case 2:
      bin_key_kernel<T, 2><<<blocks_for(M, 256), 256, 0, stream>>>(
          M, xyz[0], xyz[1], xyz[2], g, inv0, inv1, inv2, s.keys_a.get(), packed);
      break;
    default:
      bin_key_kernel<T, 3><<<blocks_for(M, 256), 256, 0, stream>>>(
          M, xyz[0], xyz[1], xyz[2], g, inv0, inv1, inv2, s.keys_a.get(), packed);
  }
  launch_check("bin_key_kernel");

  const int64_t nkeys = (cols || banks) ? g.nkeys() : nb;
  if (nkeys > INT32_MAX) fail(FINUFFT_ERR_INVALID_ARGUMENT, "bin_sort: too many keys");
  const int B = std::max(1, 32 - __builtin_clz(unsigned(nkeys - 1)));  // bits of the largest key
  const int passes = (B + 7) / 8;
  const int bpp = (B + passes - 1) / passes;
  const int ntiles = blocks_for(M, RS_TILE);
  s.hist.reserve_at_least(size_t(256) * ntiles);

  int *kin = s.keys_a.get();
  const int *vin = nullptr;  // identity on the first pass
  for (int p = 0; p < passes; ++p) {
    const int shift = p * bpp;
    const int bits = std::min(bpp, B - shift);
    int *kout = (p % 2 == 0) ? s.keys_b.get() : s.keys_a.get();
    int *vout = ((passes - 1 - p) % 2 == 0) ? idx : s.vals_b.get();
    radix_count_kernel<<<ntiles, RS_T, 0, stream>>>(kin, M, shift, bits, s.hist.get(), ntiles);
    launch_check("radix_count_kernel");
    exclusive_scan(s.hist.get(), s.hist.get(), (1 << bits) * ntiles, s, stream);
    radix_scatter_kernel<<<ntiles, RS_T, 0, stream>>>(kin, vin, kout, vout, M, shift, bits,
                                                      s.hist.get(), ntiles);
    launch_check("radix_scatter_kernel");
    kin = kout;
    vin = vout;
  }
  if (cols) {
    bin_bounds_kernel<<<blocks_for(int64_t(M) + 1, 256), 256, 0, stream>>>(
        kin, M, nb * nsub, g.binsize[2], colstart);
    launch_check("bin_bounds_kernel");
    strided_copy_kernel<<<blocks_for(int64_t(nb) + 1, 256), 256, 0, stream>>>(colstart, nsub, nb,
                                                                               binstart);
    launch_check("strided_copy_kernel");
  } else {
    bin_bounds_kernel<<<blocks_for(int64_t(M) + 1, 256), 256, 0, stream>>>(
        kin, M, nb, banks ? 32 : 1, binstart);
    launch_check("bin_bounds_kernel");
    if (banks && colstart) {  // the start of every (bin, bank class)
      bin_bounds_kernel<<<blocks_for(int64_t(M) + 1, 256), 256, 0, stream>>>(kin, M, nb * 32, 1,
                                                                           colstart);
      launch_check("bin_bounds_kernel");
    }
  }
  if (sorted) {
    if (g.dim == 1) {
      permute<T>(idx, xyz[0], sorted[0], M, stream);
    } else if (g.dim == 2) {
      gather_points_kernel<T, 2><<<blocks_for(M, 256), 256, 0, stream>>>(
          idx, packed, sorted[0], sorted[1], nullptr, M);
      launch_check("gather_points_kernel");
    } else {
      gather_points_kernel<T, 3><<<blocks_for(M, 256), 256, 0, stream>>>(
          idx, packed, sorted[0], sorted[1], sorted[2], M);
      launch_check("gather_points_kernel");
    }
  }
}

template void bin_sort<float>(const BinGeom &, int, const float *const[3], int *, int *,
                              SortScratch &, hipStream_t, int *, float *const[3]);
template void bin_sort<double>(const BinGeom &, int, const double *const[3], int *, int *,
                               SortScratch &, hipStream_t, int *, double *const[3]);

void build_subproblems(const int *binstart, int nb, int maxsub, int dense_min, int *substart,
                       int *sub2bin, SortScratch &s, hipStream_t stream) {
  numsub_kernel<<<blocks_for(int64_t(nb) + 1, 256), 256, 0, stream>>>(binstart, nb, maxsub,
                                                                       dense_min, substart);
  launch_check("numsub_kernel");
  exclusive_scan(substart, substart, nb + 1, s, stream);
  sub2bin_kernel<<<blocks_for(nb, 256), 256, 0, stream>>>(substart, nb, sub2bin);
  launch_check("sub2bin_kernel");
}

}  // namespace fa
