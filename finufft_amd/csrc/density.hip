// Point-density statistics of a set of nonuniform points: the features the
// cost model reads besides the problem sizes. Each coordinate is folded into
// [0, g) with the library's fold and the points are counted per cell of the
// g^dim coarse grid; the host turns the counts into (max / mean, 99th
// percentile / mean, occupied fraction). The reference has no counterpart
// (its method tables do not look at the points, src/cuda/heuristics.cu:
// 270-281).
#include <hip/hip_runtime.h>

#include <algorithm>
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

