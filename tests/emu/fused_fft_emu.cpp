// Runs the batch path's fused column and row transform kernels
// (finufft_amd/csrc/fused_fft.hip, compiled from that source for the host)
// on one 2-D grid and writes their inputs and output for the test to check
// against numpy's FFT.
//   fused_fft_emu <f|d> nf0 nf1 m0 m1 modeord sign seed out.bin
// Exit codes: 0 ok, 2 usage, 3 grid not taken by the fused path.
#include "hip_shim.hpp"
#include "fused_fft_kernels.inc"

#include <random>

using namespace fa;

template <typename T>
int run(int nf0, int nf1, int m0, int m1, int modeord, int sign, int seed, const char *out) {
  const int nf[2] = {nf0, nf1};
  int cw, lcw, rw, lrw;
  size_t lds_col, lds_row;
  if (!fused_fft_shape<T>(nf, cw, lcw, rw, lrw, lds_col, lds_row)) return 3;
  std::mt19937 rng(seed);
  std::uniform_real_distribution<double> U(-1, 1), K(0.5, 1.5);
  const int64_t nmodes = int64_t(m0) * m1, nftot = int64_t(nf0) * nf1;
  // exact-size heap arrays: AddressSanitizer checks every kernel access
  auto *fk = static_cast<cplx_t<T> *>(std::malloc(sizeof(cplx_t<T>) * nmodes));
  auto *fw = static_cast<cplx_t<T> *>(std::malloc(sizeof(cplx_t<T>) * nftot));
  for (int64_t i = 0; i < nmodes; ++i) fk[i] = {T(U(rng)), T(U(rng))};
  for (int64_t i = 0; i < nftot; ++i) fw[i] = {T(NAN), T(NAN)};
  T *rk[2], *tw[2];
  FusedFftItem<T> *item = static_cast<FusedFftItem<T> *>(std::malloc(sizeof(FusedFftItem<T>)));
  *item = FusedFftItem<T>{};
  item->modeord = modeord;
  for (int d = 0; d < 2; ++d) {
    item->ms[d] = d == 0 ? m0 : m1;
    item->nf[d] = nf[d];
    rk[d] = static_cast<T *>(std::malloc(sizeof(T) * (nf[d] / 2 + 1)));
    for (int j = 0; j <= nf[d] / 2; ++j) rk[d][j] = T(K(rng));
    std::vector<T> t;
    fft_twiddles<T>(nf[d], t);
    tw[d] = static_cast<T *>(std::malloc(sizeof(T) * t.size()));
    std::copy(t.begin(), t.end(), tw[d]);
    item->rker[d] = rk[d];
    item->tw[d] = tw[d];
    radix_plan(nf[d], item->rp[d]);
  }
  item->cw = cw;
  item->lcw = lcw;
  item->rw = rw;
  item->lrw = lrw;
  item->ncol = (m0 + cw - 1) / cw;
  item->nrow = (nf1 + rw - 1) / rw;
  item->nmodes = nmodes;
  item->nftot = nftot;
  item->col_begin = 0;
  item->row_begin = 0;
  auto **fk_base = static_cast<const cplx_t<T> **>(std::malloc(sizeof(void *)));
  auto **fw_base = static_cast<cplx_t<T> **>(std::malloc(sizeof(void *)));
  fk_base[0] = fk;
  fw_base[0] = fw;
  // two blocks past the end exercise the kernels' early return
  const int64_t ncb = item->ncol, nrb = item->nrow;
  emu_launch(ncb + 2, FFT_THREADS, lds_col,
             [&] { fused_cols_kernel<T>(item, fk_base, fw_base, 1, 0, sign, ncb); });
  emu_launch(nrb + 2, FFT_THREADS, lds_row,
             [&] { fused_rows_kernel<T>(item, fw_base, 1, sign, nrb); });
  FILE *f = std::fopen(out, "wb");
  if (!f) return 2;
  auto put = [&](double v) { std::fwrite(&v, sizeof(double), 1, f); };
  for (int64_t i = 0; i < nmodes; ++i) put(fk[i].x), put(fk[i].y);
  for (int d = 0; d < 2; ++d)
    for (int j = 0; j <= nf[d] / 2; ++j) put(rk[d][j]);
  for (int64_t i = 0; i < nftot; ++i) put(fw[i].x), put(fw[i].y);
  std::fclose(f);
  std::free(fk);
  std::free(fw);
  for (int d = 0; d < 2; ++d) std::free(rk[d]), std::free(tw[d]);
  std::free(item);
  std::free(fk_base);
  std::free(fw_base);
  return 0;
}

int main(int argc, char **argv) {
  if (argc != 10) return 2;
  const int a[7] = {std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4]), std::atoi(argv[5]),
                    std::atoi(argv[6]), std::atoi(argv[7]), std::atoi(argv[8])};
  if (argv[1][0] == 'd') return run<double>(a[0], a[1], a[2], a[3], a[4], a[5], a[6], argv[9]);
  return run<float>(a[0], a[1], a[2], a[3], a[4], a[5], a[6], argv[9]);
}
