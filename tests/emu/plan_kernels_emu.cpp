// Runs the deconvolution / amplification kernels of
// finufft_amd/csrc/plan_kernels.hip (compiled from that source for the host,
// FA_REAL / FA_DIM on the command line) on seeded data and writes inputs and
// outputs for the test to check against numpy.
//   plan_kernels_emu ms0 ms1 ms2 nf0 nf1 nf2 modeord blk seed out.bin
// Exit codes: 0 ok, 2 usage, 4 the batched amplification differs from the
// per-plan one.
#include "hip_shim.hpp"
#include "plan_kernels_kernels.inc"

#include <random>

using namespace fa;
using T = FA_REAL;
constexpr int DIM = FA_DIM;
using C = cplx_t<T>;

template <typename X>
X *alloc(size_t n) {
  return static_cast<X *>(std::malloc(sizeof(X) * std::max<size_t>(n, 1)));
}

int main(int argc, char **argv) {
  if (argc != 11) return 2;
  int ms[3], nf[3];
  for (int d = 0; d < 3; ++d) ms[d] = d < DIM ? std::atoi(argv[1 + d]) : 1;
  for (int d = 0; d < 3; ++d) nf[d] = d < DIM ? std::atoi(argv[4 + d]) : 1;
  const int modeord = std::atoi(argv[7]), blk = std::atoi(argv[8]), seed = std::atoi(argv[9]);
  const char *out = argv[10];
  std::mt19937 rng(seed);
  std::uniform_real_distribution<double> U(-1, 1), K(0.5, 1.5);
  const int64_t nmodes = int64_t(ms[0]) * ms[1] * ms[2], nftot = int64_t(nf[0]) * nf[1] * nf[2];
  T *ker[3];
  for (int d = 0; d < 3; ++d) {
    const int n = nf[d] / 2 + 1;
    ker[d] = alloc<T>(n);
    for (int j = 0; j < n; ++j) ker[d][j] = T(d < DIM ? K(rng) : 1.0);
  }
  ModeMap mm{};
  mm.dim = DIM;
  mm.modeord = modeord;
  for (int d = 0; d < 3; ++d) {
    mm.ms[d] = ms[d];
    mm.nf[d] = nf[d];
    mm.rms[d] = 1.0 / ms[d];
    mm.rnf[d] = 1.0 / nf[d];
  }
  C *fw_in = alloc<C>(size_t(nftot) * blk), *fk_out = alloc<C>(size_t(nmodes) * blk);
  C *fk_in = alloc<C>(size_t(nmodes) * blk), *fw_out = alloc<C>(size_t(nftot) * blk);
  C *fw_b = alloc<C>(size_t(nftot) * blk);
  for (int64_t i = 0; i < nftot * blk; ++i) fw_in[i] = {T(U(rng)), T(U(rng))};
  for (int64_t i = 0; i < nmodes * blk; ++i) fk_in[i] = {T(U(rng)), T(U(rng))};
  for (int64_t i = 0; i < nmodes * blk; ++i) fk_out[i] = {T(NAN), T(NAN)};
  for (int64_t i = 0; i < nftot * blk; ++i) fw_out[i] = fw_b[i] = {T(NAN), T(NAN)};

  // grid-stride kernels on a short grid, so that every thread takes several
  // elements
  emu_launch(3, EW_T, 0, [&] {
    deconv_to_modes_kernel<T, DIM>(mm, ker[0], ker[1], ker[2], fw_in, fk_out, nmodes, nftot);
  }, blk);
  emu_launch(5, EW_T, 0, [&] {
    amplify_to_grid_kernel<T, DIM>(mm, ker[0], ker[1], ker[2], fw_out, fk_in, nmodes, nftot);
  }, blk);
  // the batched amplification of the same problem, twice in one launch
  // (vectors 0 .. blk - 1 of the round), must give the same grid
  AmpItem<T> *items = alloc<AmpItem<T>>(2);
  const C **fkb = alloc<const C *>(2);
  C **fwb = alloc<C *>(2);
  C *fw_b2 = alloc<C>(size_t(nftot) * blk);
  int64_t blocks = 0;
  for (int q = 0; q < 2; ++q) {
    AmpItem<T> &a = items[q];
    a = AmpItem<T>{};
    a.dim = DIM;
    a.modeord = modeord;
    for (int d = 0; d < 3; ++d) {
      a.ms[d] = ms[d];
      a.nf[d] = nf[d];
      a.rnf[d] = 1.0 / nf[d];
      a.ker[d] = ker[d];
    }
    a.nmodes = nmodes;
    a.nftot = nftot;
    a.chunks = int((nftot + AMP_CHUNK - 1) / AMP_CHUNK);
    a.blk_begin = blocks;
    blocks += int64_t(a.chunks) * blk;
    fkb[q] = fk_in;
  }
  fwb[0] = fw_b;
  fwb[1] = fw_b2;
  emu_launch(blocks + 2, EW_T, 0,
             [&] { amplify_batch_kernel<T, DIM>(items, fkb, fwb, 2, 0, blocks); });
  bool same = std::memcmp(fw_b, fw_out, sizeof(C) * size_t(nftot) * blk) == 0 &&
              std::memcmp(fw_b2, fw_out, sizeof(C) * size_t(nftot) * blk) == 0;
  // the batched deconvolution (type-1 batches) of the same grids, twice in
  // one launch, must give the modes of the per-plan deconvolution
  C *fk_b1 = alloc<C>(size_t(nmodes) * blk), *fk_b2 = alloc<C>(size_t(nmodes) * blk);
  for (int64_t i = 0; i < nmodes * blk; ++i) fk_b1[i] = fk_b2[i] = {T(NAN), T(NAN)};
  const C **fwc = alloc<const C *>(2);
  C **fkw = alloc<C *>(2);
  fwc[0] = fwc[1] = fw_in;
  fkw[0] = fk_b1;
  fkw[1] = fk_b2;
  emu_launch(blocks + 2, EW_T, 0,
             [&] { deconv_batch_kernel<T, DIM>(items, fwc, fkw, 2, 0, blocks); });
  same = same && std::memcmp(fk_b1, fk_out, sizeof(C) * size_t(nmodes) * blk) == 0 &&
         std::memcmp(fk_b2, fk_out, sizeof(C) * size_t(nmodes) * blk) == 0;
  std::free(fk_b1);
  std::free(fk_b2);
  std::free(fwc);
  std::free(fkw);

  FILE *f = std::fopen(out, "wb");
  if (!f) return 2;
  auto put = [&](double v) { std::fwrite(&v, sizeof(double), 1, f); };
  for (int d = 0; d < DIM; ++d)
    for (int j = 0; j <= nf[d] / 2; ++j) put(double(ker[d][j]));
  auto putc = [&](const C *a, int64_t n) {
    for (int64_t i = 0; i < n; ++i) put(double(a[i].x)), put(double(a[i].y));
  };
  putc(fw_in, nftot * blk);
  putc(fk_out, nmodes * blk);
  putc(fk_in, nmodes * blk);
  putc(fw_out, nftot * blk);
  std::fclose(f);
  for (int d = 0; d < 3; ++d) std::free(ker[d]);
  std::free(fw_in);
  std::free(fk_out);
  std::free(fk_in);
  std::free(fw_out);
  std::free(fw_b);
  std::free(fw_b2);
  std::free(items);
  std::free(fkb);
  std::free(fwb);
  return same ? 0 : 4;
}
