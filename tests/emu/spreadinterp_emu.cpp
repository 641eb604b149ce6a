// Runs the spreading / interpolation kernels of finufft_amd/csrc/spreadinterp.hip
// (compiled from that source for the host, FA_REAL / FA_DIM / EMU_NS set on
// the command line) on seeded points and writes inputs and outputs for the
// test to check against a direct FP64 sum. The point sort and subproblem
// table are built here with the library's own first-node rule (start_index).
//   spreadinterp_emu <gm|sub|colored|col|colsub|colmix|od|bin|binsub|mv|igm|grad>
//                    nf0 nf1 nf2 b0 b1 b2 M maxsub
//                    blk clustered seed out.bin
// (od: output-driven spreading with EMU_NP points staged per batch, default
// 64.)
// Exit codes: 0 ok, 2 usage, 3 geometry not valid for the variant, 4 the two
// problems of the multi-vector launch differ.
#include "hip_shim.hpp"
#include "spreadinterp_kernels.inc"

#include <random>
#include <string>
#include <type_traits>

using namespace fa;
using T = FA_REAL;
constexpr int DIM = FA_DIM;
constexpr int NS = EMU_NS;
using C = cplx_t<T>;
constexpr double kPi = 3.14159265358979323846;

template <typename X>
X *alloc(size_t n) {
  return static_cast<X *>(std::malloc(sizeof(X) * std::max<size_t>(n, 1)));
}

int main(int argc, char **argv) {
  if (argc != 14) return 2;
  const std::string var = argv[1];
  int nf[3], bs[3];
  for (int d = 0; d < 3; ++d) nf[d] = d < DIM ? std::atoi(argv[2 + d]) : 1;
  for (int d = 0; d < 3; ++d) bs[d] = d < DIM ? std::atoi(argv[5 + d]) : 1;
  const int M = std::atoi(argv[8]), maxsub = std::atoi(argv[9]), blk = std::atoi(argv[10]);
  const int clustered = std::atoi(argv[11]), seed = std::atoi(argv[12]);
  const char *out = argv[13];
  const bool col = var == "col" || var == "colsub" || var == "colmix";
  const bool spreads =
      var == "gm" || var == "sub" || var == "colored" || var == "od" || var == "smv" || col;
  if (col && !(DIM == 3 && sizeof(T) == 4 && NS <= COL_MAX_NS && bs[0] == 8 && bs[1] == 8))
    return 3;
  if (var == "colored" || var == "col" || var == "colmix")
    // whole-bin colors need an even number (>= 2) of equal bins
    for (int d = 0; d < DIM; ++d) {
      const int n = nf[d] / bs[d];
      if (nf[d] % bs[d] || n < 2 || n % 2 || bs[d] < NS - 1) return 3;
    }

  std::mt19937 rng(seed);
  std::uniform_real_distribution<double> U(-kPi, kPi), V(-1, 1);
  T *x[3] = {nullptr, nullptr, nullptr};
  for (int d = 0; d < DIM; ++d) {
    x[d] = alloc<T>(M);
    for (int i = 0; i < M; ++i)
      x[d][i] = T(clustered && i % 2 ? -kPi + 0.3 * (U(rng) + kPi) / (2 * kPi) : U(rng));
  }
  // first-node bins, stable order (the device sort's rule)
  int nb[3] = {1, 1, 1};
  for (int d = 0; d < DIM; ++d) nb[d] = (nf[d] + bs[d] - 1) / bs[d];
  const int nbt = nb[0] * nb[1] * nb[2];
  // Column spreading refines the key inside a bin to (sub-column of S x S
  // first nodes in x, y; local z of the first node), the library's sort key.
  const int S = col ? col_side(NS) : 1, NSX = col ? 8 / S : 1, NSUB = NSX * NSX;
  // Bank-ordered interpolation (bbin, bsub): the key inside a bin is the LDS
  // bank class of the first tile cell, (lx + P0 ly + P0 P1 lz) mod 32.
  const bool bank = var == "bbin" || var == "bsub";
  const int nz = col ? bs[2] : (bank ? 32 : 1);
  const int nkeys = nbt * NSUB * nz;
  std::vector<int> key(M), kcnt(nkeys + 1, 0);
  const int fold64 = sizeof(T) == 4 ? 1 : 0;
  for (int i = 0; i < M; ++i) {
    int k = 0, stride = 1, loc[3] = {0, 0, 0};
    for (int d = 0; d < DIM; ++d) {
      int s = start_index<T>(x[d][i], nf[d], NS, fold64);
      if (s < 0) s += nf[d];
      const int b = std::min(s / bs[d], nb[d] - 1);
      loc[d] = s - b * bs[d];
      k += b * stride;
      stride *= nb[d];
    }
    if (col) k = (k * NSUB + (loc[0] / S) + NSX * (loc[1] / S)) * nz + loc[2];
    if (bank) {
      const int P0 = bs[0] + NS - 1, P1 = DIM > 1 ? bs[1] + NS - 1 : 1;
      const int cell = loc[0] + P0 * (loc[1] + P1 * loc[2]);
      k = k * 32 + (cell & 31);
    }
    key[i] = k;
    ++kcnt[k + 1];
  }
  for (int b = 0; b < nkeys; ++b) kcnt[b + 1] += kcnt[b];
  std::vector<int> cnt(nbt + 1);
  for (int b = 0; b <= nbt; ++b) cnt[b] = kcnt[int64_t(b) * NSUB * nz];
  int *idx = alloc<int>(M), *binstart = alloc<int>(nbt + 1);
  int *colstart = alloc<int>(size_t(nbt) * (bank ? 32 : NSUB) + 1);
  {
    std::vector<int> pos(kcnt.begin(), kcnt.end() - 1);
    for (int i = 0; i < M; ++i) idx[pos[key[i]]++] = i;
    std::copy(cnt.begin(), cnt.end(), binstart);
    if (bank)
      std::copy(kcnt.begin(), kcnt.end(), colstart);  // (bin, class) starts
    else
      for (int q = 0; q <= nbt * NSUB; ++q) colstart[q] = kcnt[int64_t(q) * nz];
  }
  // Column spreading: colored launches take the bins of at most dense_min
  // points; the subproblem table covers the denser bins only.
  const int dense_min = var == "col" ? M : var == "colmix" ? maxsub : 0;
  int *substart = alloc<int>(nbt + 1);
  substart[0] = 0;
  for (int b = 0; b < nbt; ++b) {
    const int n = binstart[b + 1] - binstart[b];
    substart[b + 1] = substart[b] + (n > dense_min ? (n + maxsub - 1) / maxsub : 0);
  }
  const int nsub = substart[nbt];
  const int cap = nbt + (M + maxsub - 1) / maxsub;  // the library's launch bound
  int *sub2bin = alloc<int>(cap);
  for (int b = 0; b < nbt; ++b)
    for (int s = substart[b]; s < substart[b + 1]; ++s) sub2bin[s] = b;
  T *sx[3] = {nullptr, nullptr, nullptr};
  for (int d = 0; d < DIM; ++d) {
    sx[d] = alloc<T>(M);
    for (int i = 0; i < M; ++i) sx[d][i] = x[d][idx[i]];
  }

  SpreadCtx<T> ctx;
  ctx.dim = DIM;
  ctx.ns = NS;
  ctx.method = var == "gm" || var == "igm" ? 1 : var == "od" ? 3 : 2;
  ctx.M = M;
  for (int d = 0; d < 3; ++d) {
    ctx.nf[d] = nf[d];
    ctx.xyz[d] = x[d];
    ctx.sxyz[d] = sx[d];
    ctx.binsize[d] = bs[d];
    ctx.nbins[d] = nb[d];
  }
  ctx.kp.beta = T(2.30 * NS);
  ctx.kp.es_c = T(4.0 / (double(NS) * NS));
  ctx.kp.horner = 0;  // direct exp(sqrt()) evaluation: no coefficient table
  ctx.kp.fold64 = fold64;
  ctx.kp.coef = nullptr;
  ctx.idx = idx;
  ctx.binstart = binstart;
  ctx.maxsub = maxsub;
  ctx.substart = substart;
  ctx.sub2bin = sub2bin;
  ctx.sub_capacity = cap;
  int64_t ntile = 1;
  for (int d = 0; d < DIM; ++d) ntile *= bs[d] + NS - 1;
  ctx.lds_bytes = size_t(ntile) * 2 * sizeof(T);
  ctx.lds_spread = size_t(ntile) * 2 * sizeof(double);
  ctx.interp_sub = var == "binsub" || var == "bsub" ? 1 : 0;
  if (bank) {
    ctx.bank = 1;
    ctx.colstart = colstart;
  }
  if (var == "od") {
    const char *e = std::getenv("EMU_NP");
    ctx.od_np = e ? std::atoi(e) : 64;
    if (ctx.od_np < 1) return 2;
    ctx.lds_od = ctx.lds_spread + od_stage_bytes(DIM, NS, ctx.od_np, sizeof(T));
  }
  if (col) {
    ctx.colS = S;
    ctx.colstart = colstart;
    ctx.dense_min = dense_min;
    ctx.n_dense_sub = nsub;
    ctx.lds_spread = size_t(ntile) * 16 + col_stage_bytes();
  }

  const int64_t nftot = int64_t(nf[0]) * nf[1] * nf[2];
  C *c = alloc<C>(size_t(M) * blk);
  C *dc = nullptr;
  C *fw = alloc<C>(size_t(nftot) * blk);
  if (spreads) {
    for (int64_t i = 0; i < int64_t(M) * blk; ++i) c[i] = {T(V(rng)), T(V(rng))};
    for (int64_t i = 0; i < nftot * blk; ++i) fw[i] = {T(0), T(0)};
  } else {
    for (int64_t i = 0; i < nftot * blk; ++i) fw[i] = {T(V(rng)), T(V(rng))};
    for (int64_t i = 0; i < int64_t(M) * blk; ++i) c[i] = {T(NAN), T(NAN)};
  }

  if (var == "gm") {
    const int64_t g = (M + SPREAD_THREADS - 1) / SPREAD_THREADS;
    emu_launch(g, SPREAD_THREADS, 0, [&] { spread_gm_kernel<T, DIM, NS>(ctx, c, fw, nftot); }, blk);
  } else if (var == "od") {
    emu_launch(cap, BIN_THREADS, ctx.lds_od,
               [&] { spread_od_kernel<T, DIM, NS>(ctx, c, fw, nftot, ctx.od_np); }, blk);
  } else if (var == "sub") {
    emu_launch(cap, BIN_THREADS, ctx.lds_spread,
               [&] { spread_sub_kernel<T, DIM, NS>(ctx, c, fw, nftot); }, blk);
  } else if (var == "colored") {
    ctx.colored = 1;
    for (int color = 0; color < (1 << DIM); ++color) {
      int col[3] = {0, 0, 0};
      int64_t n = 1;
      for (int d = 0; d < DIM; ++d) {
        col[d] = (color >> d) & 1;
        n *= (nb[d] - col[d] + 1) / 2;
      }
      emu_launch(n, BIN_THREADS, ctx.lds_spread, [&] {
        spread_bin_colored_kernel<T, DIM, NS>(ctx, c, fw, nftot, col[0], col[1], col[2]);
      }, blk);
    }
  } else if (col) {
    // a generic lambda, so that the fp32 3-D kernels are only instantiated
    // in the build they exist for
    auto launch_col = [&](auto &cx, auto *cc, auto *ff) {
      using TT = std::remove_cvref_t<decltype(cx.kp.beta)>;
      if constexpr (DIM == 3 && std::is_same_v<TT, float> && NS <= COL_MAX_NS) {
        if (var != "colsub")
          for (int color = 0; color < 8; ++color) {
            const int col3[3] = {color & 1, (color >> 1) & 1, (color >> 2) & 1};
            int64_t n = 1;
            for (int d = 0; d < 3; ++d) n *= (nb[d] - col3[d] + 1) / 2;
            emu_launch(n, 64 * COL_WAVES, cx.lds_spread, [&] {
              spread_col_colored_kernel<NS>(cx, cc, ff, nftot, col3[0], col3[1], col3[2]);
            }, blk);
          }
        if (nsub > 0)
          emu_launch(nsub, 64 * COL_WAVES, cx.lds_spread,
                     [&] { spread_col_sub_kernel<NS>(cx, cc, ff, nftot); }, blk);
      }
    };
    launch_col(ctx, c, fw);
  } else if (var == "bin" || var == "binsub" || bank) {
    emu_launch(var == "bin" || var == "bbin" ? nbt : cap, BIN_THREADS,
               ctx.lds_bytes + (bank ? INTERP_BANK_LDS : 0),
               [&] { interp_bin_kernel<T, DIM, NS>(ctx, c, fw, nftot); }, blk);
  } else if (var == "mv") {
    // the batch path's multi-vector interpolation: two problems (the same
    // points and grids, separate outputs) in one launch, all blk vectors per
    // work unit; the second problem's values must equal the first's
    ctx.interp_sub = 1;
    InterpItem<T> *items = alloc<InterpItem<T>>(2);
    const C **fwb = alloc<const C *>(2);
    C **cb = alloc<C *>(2);
    C *c2 = alloc<C>(size_t(M) * blk);
    for (int64_t i = 0; i < int64_t(M) * blk; ++i) c2[i] = {T(NAN), T(NAN)};
    for (int q = 0; q < 2; ++q) {
      items[q] = InterpItem<T>{};
      items[q].ctx = ctx;
      items[q].nftot = nftot;
      items[q].nsub = nsub;
      items[q].unit_begin = int64_t(q) * nsub;
      fwb[q] = fw;
    }
    cb[0] = c;
    cb[1] = c2;
    const int64_t units = 2 * int64_t(nsub);
    emu_launch(units + 2, MV_THREADS, ctx.lds_bytes, [&] {
      interp_batch_kernel<T, DIM, NS>(items, 2, ctx.kp, fwb, cb, 0, blk, units);
    });
    const bool same = std::memcmp(c, c2, sizeof(C) * size_t(M) * blk) == 0;
    std::free(items);
    std::free(fwb);
    std::free(cb);
    std::free(c2);
    if (!same) return 4;
  } else if (var == "smv") {
    // the batch path's multi-problem spreading: two problems (the same
    // points and strengths, separate grids) in one launch, all blk vectors
    // per work unit; the second problem's grids must equal the first's
    InterpItem<T> *items = alloc<InterpItem<T>>(2);
    const C **cb = alloc<const C *>(2);
    C **fwb = alloc<C *>(2);
    C *fw2 = alloc<C>(size_t(nftot) * blk);
    for (int64_t i = 0; i < nftot * blk; ++i) fw2[i] = {T(0), T(0)};
    for (int q = 0; q < 2; ++q) {
      items[q] = InterpItem<T>{};
      items[q].ctx = ctx;
      items[q].nftot = nftot;
      items[q].nsub = nsub;
      items[q].unit_begin = int64_t(q) * nsub;
      cb[q] = c;
    }
    fwb[0] = fw;
    fwb[1] = fw2;
    const int64_t units = 2 * int64_t(nsub);
    // two blocks past the end exercise the kernel's early return
    emu_launch(units + 2, BIN_THREADS, ctx.lds_spread, [&] {
      spread_batch_kernel<T, DIM, NS>(items, 2, cb, 0, fwb, blk, units);
    });
    // equal to rounding (the LDS tile sums its atomic adds in thread order)
    double dmax = 0, amax = 0;
    for (int64_t i = 0; i < nftot * blk; ++i) {
      dmax = std::max({dmax, std::abs(double(fw[i].x - fw2[i].x)),
                       std::abs(double(fw[i].y - fw2[i].y))});
      amax = std::max({amax, std::abs(double(fw[i].x)), std::abs(double(fw[i].y))});
    }
    const bool same = dmax <= (sizeof(T) == 4 ? 1e-5 : 1e-12) * amax;
    std::free(items);
    std::free(cb);
    std::free(fwb);
    std::free(fw2);
    if (!same) return 4;
  } else if (var == "grad") {
    // values and coordinate derivatives of the interpolation (autograd)
    dc = alloc<C>(size_t(DIM) * M * blk);
    for (int64_t i = 0; i < int64_t(DIM) * M * blk; ++i) dc[i] = {T(NAN), T(NAN)};
    const int64_t g = (M + SPREAD_THREADS - 1) / SPREAD_THREADS;
    emu_launch(g, SPREAD_THREADS, 0,
               [&] { interp_grad_kernel<T, DIM, NS>(ctx, c, dc, fw, nftot, blk); });
  } else if (var == "igm") {
    const int64_t g = (M + SPREAD_THREADS - 1) / SPREAD_THREADS;
    emu_launch(g, SPREAD_THREADS, 0, [&] { interp_gm_kernel<T, DIM, NS>(ctx, c, fw, nftot, blk); });
  } else {
    return 2;
  }

  FILE *f = std::fopen(out, "wb");
  if (!f) return 2;
  auto put = [&](double v) { std::fwrite(&v, sizeof(double), 1, f); };
  put(double(nsub));
  for (int d = 0; d < DIM; ++d)
    for (int i = 0; i < M; ++i) put(double(x[d][i]));
  for (int64_t i = 0; i < int64_t(M) * blk; ++i) put(double(c[i].x)), put(double(c[i].y));
  for (int64_t i = 0; i < nftot * blk; ++i) put(double(fw[i].x)), put(double(fw[i].y));
  if (dc)
    for (int64_t i = 0; i < int64_t(DIM) * M * blk; ++i) put(double(dc[i].x)), put(double(dc[i].y));
  std::fclose(f);
  std::free(dc);
  for (int d = 0; d < DIM; ++d) std::free(x[d]), std::free(sx[d]);
  std::free(idx);
  std::free(binstart);
  std::free(substart);
  std::free(sub2bin);
  std::free(colstart);
  std::free(c);
  std::free(fw);
  return 0;
}
