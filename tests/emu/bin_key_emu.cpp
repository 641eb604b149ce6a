// Runs the sort's bin-key kernel (finufft_amd/csrc/sort.hip, compiled from
// that source for the host, FA_REAL / FA_DIM on the command line) on the
// coordinates in a binary file and writes the keys; then, in 2-D and 3-D,
// gathers the coordinates the key kernel packed into the stable key order
// with the sort's gather kernel and writes them after the keys.
//   bin_key_emu in.bin out.bin   (in: M, then geometry ints, then coordinates)
#include "hip_shim.hpp"
#include "bin_key_kernels.inc"

#include <algorithm>
#include <numeric>
#include <vector>

using namespace fa;
using T = FA_REAL;
constexpr int DIM = FA_DIM;

int main(int argc, char **argv) {
  if (argc != 3) return 2;
  FILE *f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  // header: M, start_bins, ns, fold64, colS, bank, nf[3], binsize[3]
  int64_t h[12];
  if (std::fread(h, sizeof(int64_t), 12, f) != 12) return 2;
  const int M = int(h[0]);
  BinGeom g;
  g.dim = DIM;
  g.start_bins = int(h[1]);
  g.ns = int(h[2]);
  g.fold64 = int(h[3]);
  g.colS = int(h[4]);
  g.bank = int(h[5]);
  for (int d = 0; d < 3; ++d) {
    g.nf[d] = int(h[6 + d]);
    g.binsize[d] = int(h[9 + d]);
    g.nbins[d] = (g.nf[d] + g.binsize[d] - 1) / g.binsize[d];
  }
  T *x[3] = {nullptr, nullptr, nullptr};
  for (int d = 0; d < DIM; ++d) {
    x[d] = static_cast<T *>(std::malloc(sizeof(T) * std::max(M, 1)));
    std::vector<double> v(M);
    if (std::fread(v.data(), sizeof(double), M, f) != size_t(M)) return 2;
    for (int i = 0; i < M; ++i) x[d][i] = T(v[i]);
  }
  std::fclose(f);
  int *keys = static_cast<int *>(std::malloc(sizeof(int) * std::max(M, 1)));
  // exactly the packed size (no packing in 1-D)
  T *packed = DIM > 1 ? static_cast<T *>(std::malloc(sizeof(T) * pack_stride(DIM) *
                                                     std::max(M, 1)))
                      : nullptr;
  const T inv0 = T(1) / T(g.binsize[0]), inv1 = T(1) / T(g.binsize[1]),
          inv2 = T(1) / T(g.binsize[2]);
  const int nb = (M + 255) / 256;
  emu_launch(nb, 256, 0, [&] {
    bin_key_kernel<T, DIM>(M, x[0], x[1], x[2], g, inv0, inv1, inv2, keys, packed);
  });
  FILE *o = std::fopen(argv[2], "wb");
  if (!o) return 2;
  std::fwrite(keys, sizeof(int), M, o);
  if constexpr (DIM > 1) {
    std::vector<int> idx(static_cast<size_t>(M));
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return keys[a] < keys[b]; });
    T *s[3] = {nullptr, nullptr, nullptr};
    for (int d = 0; d < DIM; ++d) s[d] = static_cast<T *>(std::malloc(sizeof(T) * M));
    emu_launch(nb, 256, 0, [&] {
      gather_points_kernel<T, DIM>(idx.data(), packed, s[0], s[1], s[2], M);
    });
    for (int d = 0; d < DIM; ++d) {
      for (int i = 0; i < M; ++i) {
        const double v = double(s[d][i]);
        std::fwrite(&v, sizeof(double), 1, o);
      }
      std::free(s[d]);
    }
  }
  std::fclose(o);
  for (int d = 0; d < DIM; ++d) std::free(x[d]);
  std::free(keys);
  std::free(packed);
  return 0;
}
