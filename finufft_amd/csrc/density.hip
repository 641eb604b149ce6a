// Point-density statistics of a set of nonuniform points: the features the
// cost model reads besides the problem sizes. Each coordinate is folded into
// [0, g) with the library's fold and the points are counted per cell of the
// g^dim coarse grid; the host turns the counts into (max / mean, 99th
// percentile / mean, occupied fraction). The reference has no counterpart
// (its method tables do not look at the points, src/cuda/heuristics.cu:
// 270-281).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "common.hpp"
#include "density.hpp"
#include "kernel_math.hpp"

namespace fa {
namespace {

constexpr int DENS_T = 256;

template <typename T>
__global__ void __launch_bounds__(DENS_T)
    density_count_kernel(int dim, int64_t M, const T *__restrict__ x, const T *__restrict__ y,
                         const T *__restrict__ z, int g, unsigned *__restrict__ counts) {
  const int64_t i = int64_t(blockIdx.x) * DENS_T + threadIdx.x;
  if (i >= M) return;
  const T *c[3] = {x, y, z};
  int cell = 0, stride = 1;
  for (int d = 0; d < dim; ++d) {
    const T r = fold_rescale(c[d][i], g);
    int b = (r >= T(0) && r < T(g)) ? int(r) : 0;  // a non-finite coordinate counts in cell 0
    b = b >= g ? g - 1 : b;
    cell += b * stride;
    stride *= g;
  }
  atomicAdd(&counts[cell], 1u);
}

}  // namespace

template <typename T>
void density_stats(int dim, int64_t M, const T *x, const T *y, const T *z, int g, double *stats,
                   hipStream_t stream) {
  if (dim < 1 || dim > 3 || g < 1) fail(FINUFFT_ERR_INVALID_ARGUMENT, "density: bad dimension");
  int64_t ncell = 1;
  for (int d = 0; d < dim; ++d) ncell *= g;
  if (ncell > (int64_t(1) << 20)) fail(FINUFFT_ERR_INVALID_ARGUMENT, "density: grid too large");
  if (M < 0 || (M > 0 && (!x || (dim > 1 && !y) || (dim > 2 && !z))))
    fail(FINUFFT_ERR_INVALID_ARGUMENT, "density: bad points");
  if (M == 0) {
    stats[0] = stats[1] = stats[2] = 0.0;
    return;
  }
  const int64_t blocks = (M + DENS_T - 1) / DENS_T;
  if (blocks > int64_t(0x7fffffff)) fail(FINUFFT_ERR_INVALID_ARGUMENT, "density: too many points");
  dbuf<unsigned> counts(static_cast<size_t>(ncell));
  hip_check(hipMemsetAsync(counts.get(), 0, size_t(ncell) * sizeof(unsigned), stream),
            "hipMemsetAsync(density)");
  density_count_kernel<T><<<unsigned(blocks), DENS_T, 0, stream>>>(dim, M, x, y, z, g,
                                                                    counts.get());
  launch_check("density_count_kernel");
  std::vector<unsigned> h(static_cast<size_t>(ncell));
  hip_check(hipMemcpyAsync(h.data(), counts.get(), h.size() * sizeof(unsigned),
                           hipMemcpyDeviceToHost, stream),
            "hipMemcpyAsync(density)");
  hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize(density)");
  const double mean = double(M) / double(ncell);
  int64_t occupied = 0;
  for (unsigned v : h) occupied += v > 0;
  const size_t k = std::min(h.size() - 1, size_t(0.99 * double(h.size())));
  std::nth_element(h.begin(), h.begin() + k, h.end());
  const double p99 = h[k];
  const double mx = *std::max_element(h.begin() + k, h.end());
  stats[0] = mx / mean;
  stats[1] = p99 / mean;
  stats[2] = double(occupied) / double(ncell);
}

template void density_stats<float>(int, int64_t, const float *, const float *, const float *,
                                   int, double *, hipStream_t);
template void density_stats<double>(int, int64_t, const double *, const double *,
                                    const double *, int, double *, hipStream_t);

}  // namespace fa

// ------------------------------------------------- allocation-free variant
//
// The same statistics with every buffer in a file-scope device array and the
// finishing done on the GPU, for callers that cannot allocate (graph
// capture, tight serving loops): no allocation, no host synchronization, the
// result written to a device array. Calls that overlap on one GPU share the
// array and race; the array keeps the last call's counts (each call clears
// what it uses).

namespace fa {
namespace {

constexpr int64_t DENS_SCRATCH_CELLS = int64_t(1) << 18;  // 1 MiB of counts
static __device__ __align__(256) unsigned scratch[DENS_SCRATCH_CELLS];

constexpr int DENS_FIN_T = 1024;

__global__ void __launch_bounds__(DENS_T) density_clear_kernel(unsigned *counts, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * DENS_T + threadIdx.x;
  if (i < n) counts[i] = 0u;
}

// One workgroup: max, occupied count and the k-th smallest count (k =
// floor(0.99 ncell), the 99th percentile as the host version takes it:
// the smallest v with at least k + 1 counts <= v, found by bisection).
__global__ void __launch_bounds__(DENS_FIN_T)
    density_finish_kernel(const unsigned *__restrict__ counts, int64_t ncell, int64_t M,
                          double *__restrict__ stats) {
  __shared__ unsigned red_u[DENS_FIN_T / 64];
  __shared__ unsigned long long red_n[DENS_FIN_T / 64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  auto block_max = [&](unsigned v) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, unsigned(__shfl_down(int(v), o)));
    if (lane == 0) red_u[wave] = v;
    __syncthreads();
    unsigned r = red_u[0];
    for (int w = 1; w < DENS_FIN_T / 64; ++w) r = max(r, red_u[w]);
    __syncthreads();
    return r;
  };
  auto block_sum = [&](unsigned long long v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
    if (lane == 0) red_n[wave] = v;
    __syncthreads();
    unsigned long long r = 0;
    for (int w = 0; w < DENS_FIN_T / 64; ++w) r += red_n[w];
    __syncthreads();
    return r;
  };
  unsigned mx = 0;
  unsigned long long occ = 0;
  for (int64_t i = tid; i < ncell; i += DENS_FIN_T) {
    const unsigned v = counts[i];
    mx = max(mx, v);
    occ += v > 0u;
  }
  mx = block_max(mx);
  occ = block_sum(occ);
  const int64_t k = min(ncell - 1, int64_t(0.99 * double(ncell)));
  unsigned lo = 0, hi = mx;  // answer in [lo, hi]
  while (lo < hi) {
    const unsigned mid = lo + (hi - lo) / 2;
    unsigned long long le = 0;
    for (int64_t i = tid; i < ncell; i += DENS_FIN_T) le += counts[i] <= mid;
    le = block_sum(le);
    if (le >= (unsigned long long)(k + 1))
      hi = mid;
    else
      lo = mid + 1;
  }
  if (tid == 0) {
    const double mean = double(M) / double(ncell);
    stats[0] = double(mx) / mean;
    stats[1] = double(lo) / mean;
    stats[2] = double(occ) / double(ncell);
  }
}

__global__ void density_zero_stats_kernel(double *stats) {
  if (threadIdx.x < 3) stats[threadIdx.x] = 0.0;
}

constexpr int DENS_MAX_DEVICES = 64;
std::mutex scratch_mu;
unsigned *scratch_dev[DENS_MAX_DEVICES] = {};  // `scratch` per device, once looked up

// The device address of `scratch` on the current device, looked up on the
// first call on each device with that device current (each GPU holds its
// own copy of the array at its own address).
unsigned *scratch_ptr() {
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= DENS_MAX_DEVICES) fail(FINUFFT_ERR_INVALID_ARGUMENT, "device id");
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_dev[dev]) {
    void *p = nullptr;
    hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    scratch_dev[dev] = static_cast<unsigned *>(p);
  }
  return scratch_dev[dev];
}

}  // namespace

int64_t density_noalloc_max_cells() { return DENS_SCRATCH_CELLS; }

template <typename T>
void density_stats_noalloc(int dim, int64_t M, const T *x, const T *y, const T *z, int g,
                           double *d_stats, hipStream_t stream) {
  if (dim < 1 || dim > 3 || g < 1) fail(FINUFFT_ERR_INVALID_ARGUMENT, "density: bad dimension");
  if (!d_stats) fail(FINUFFT_ERR_INVALID_ARGUMENT, "density: null stats");
  int64_t ncell = 1;
  for (int d = 0; d < dim; ++d) {
    ncell *= g;
    if (ncell > DENS_SCRATCH_CELLS)
      fail(FINUFFT_ERR_INVALID_ARGUMENT,
           "density (no allocation): g^dim exceeds the scratch array (2^18 cells)");
  }
  if (M < 0 || (M > 0 && (!x || (dim > 1 && !y) || (dim > 2 && !z))))
    fail(FINUFFT_ERR_INVALID_ARGUMENT, "density: bad points");
  if (M == 0) {
    density_zero_stats_kernel<<<1, 64, 0, stream>>>(d_stats);
    launch_check("density_zero_stats_kernel");
    return;
  }
  const int64_t blocks = (M + DENS_T - 1) / DENS_T;
  if (blocks > int64_t(0x7fffffff)) fail(FINUFFT_ERR_INVALID_ARGUMENT, "density: too many points");
  unsigned *counts = scratch_ptr();
  density_clear_kernel<<<unsigned((ncell + DENS_T - 1) / DENS_T), DENS_T, 0, stream>>>(counts,
                                                                                      ncell);
  launch_check("density_clear_kernel");
  density_count_kernel<T><<<unsigned(blocks), DENS_T, 0, stream>>>(dim, M, x, y, z, g, counts);
  launch_check("density_count_kernel");
  density_finish_kernel<<<1, DENS_FIN_T, 0, stream>>>(counts, ncell, M, d_stats);
  launch_check("density_finish_kernel");
}

template void density_stats_noalloc<float>(int, int64_t, const float *, const float *,
                                           const float *, int, double *, hipStream_t);
template void density_stats_noalloc<double>(int, int64_t, const double *, const double *,
                                            const double *, int, double *, hipStream_t);

}  // namespace fa
