// Runs the scan and subproblem-table kernels of finufft_amd/csrc/sort.hip
// (compiled from that source for the host): the bin bounds of sorted keys,
// the multi-level exclusive scan, and the subproblem table of a bin count.
//   subproblems_emu in.bin out.bin
// in: M nb div maxsub dense_min (int64) then M sorted keys (int32).
// out: binstart[nb + 1], substart[nb + 1], sub2bin[capacity], scan of a
//      test vector of length 40 * SCAN_TILE + 17.
#include "hip_shim.hpp"
#include "subproblem_kernels.inc"

#include <vector>

using namespace fa;

namespace {
// The library's recursive scan (tile sums, scan of the sums, apply), with the
// emulated launches in place of the device ones.
void scan(const int *in, int *out, int n, int *tmp) {
  const int nb = (n + SCAN_TILE - 1) / SCAN_TILE;
  if (nb == 1) {
    emu_launch(1, SCAN_T, 0, [&] { scan_apply_kernel(in, out, n, nullptr); });
    return;
  }
  int *sums = tmp;
  emu_launch(nb, SCAN_T, 0, [&] { scan_reduce_kernel(in, n, sums); });
  scan(sums, sums, nb, tmp + nb);
  emu_launch(nb, SCAN_T, 0, [&] { scan_apply_kernel(in, out, n, sums); });
}
template <typename X>
X *alloc(size_t n) {
  return static_cast<X *>(std::malloc(sizeof(X) * std::max<size_t>(n, 1)));
}
}  // namespace

int main(int argc, char **argv) {
  if (argc != 3) return 2;
  FILE *f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  int64_t h[5];
  if (std::fread(h, sizeof(int64_t), 5, f) != 5) return 2;
  const int M = int(h[0]), nb = int(h[1]), div = int(h[2]), maxsub = int(h[3]);
  const int dense_min = int(h[4]);
  int *keys = alloc<int>(M);
  if (std::fread(keys, sizeof(int), M, f) != size_t(M)) return 2;
  std::fclose(f);

  int *binstart = alloc<int>(nb + 1);
  if (M == 0) {
    emu_launch((nb + 1 + 255) / 256, 256, 0, [&] { two_point_bounds_kernel(binstart, nb, M); });
  } else {
    emu_launch((M + 1 + 255) / 256, 256, 0, [&] { bin_bounds_kernel(keys, M, nb, div, binstart); });
  }
  // subproblem table (build_subproblems)
  int *substart = alloc<int>(nb + 1);
  const int cap = nb + (M + maxsub - 1) / maxsub;
  int *sub2bin = alloc<int>(cap);
  for (int i = 0; i < cap; ++i) sub2bin[i] = -1;
  emu_launch((nb + 1 + 255) / 256, 256, 0,
             [&] { numsub_kernel(binstart, nb, maxsub, dense_min, substart); });
  std::vector<int> tmpv;
  {
    size_t need = 0;
    for (int64_t m = nb + 1; m > SCAN_TILE;) {
      m = (m + SCAN_TILE - 1) / SCAN_TILE;
      need += m;
    }
    tmpv.resize(need + 1);
  }
  int *tmp = alloc<int>(tmpv.size());
  scan(substart, substart, nb + 1, tmp);
  emu_launch((nb + 255) / 256, 256, 0, [&] { sub2bin_kernel(substart, nb, sub2bin); });
  // a longer scan: tile sums, their scan, and the apply pass
  const int nl = 40 * SCAN_TILE + 17;
  int *a = alloc<int>(nl), *sa = alloc<int>(nl);
  for (int i = 0; i < nl; ++i) a[i] = int((int64_t(i) * 7919) % 13);
  int *tmp2 = alloc<int>(nl / SCAN_TILE + nl / SCAN_TILE / SCAN_TILE + 8);
  scan(a, sa, nl, tmp2);

  FILE *o = std::fopen(argv[2], "wb");
  if (!o) return 2;
  std::fwrite(binstart, sizeof(int), nb + 1, o);
  std::fwrite(substart, sizeof(int), nb + 1, o);
  std::fwrite(sub2bin, sizeof(int), cap, o);
  std::fwrite(sa, sizeof(int), nl, o);
  std::fclose(o);
  for (void *q : {(void *)keys, (void *)binstart, (void *)substart, (void *)sub2bin, (void *)tmp,
                  (void *)a, (void *)sa, (void *)tmp2})
    std::free(q);
  return 0;
}
