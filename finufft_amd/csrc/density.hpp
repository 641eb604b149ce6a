// Point-density statistics for the cost model (see density.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace fa {

// stats[0] = max / mean, stats[1] = 99th percentile / mean, stats[2] =
// occupied fraction of the counts over the g^dim cells of the folded points
// (g^dim <= 2^20); synchronizes the stream.
template <typename T>
void density_stats(int dim, int64_t M, const T *x, const T *y, const T *z, int g, double *stats,
                   hipStream_t stream);

// The same statistics with no allocation and no synchronization: counts in
// a file-scope device array of density_noalloc_max_cells() cells, the result
// written to the device array d_stats[3] on `stream`.
template <typename T>
void density_stats_noalloc(int dim, int64_t M, const T *x, const T *y, const T *z, int g,
                           double *d_stats, hipStream_t stream);
int64_t density_noalloc_max_cells();

}  // namespace fa
