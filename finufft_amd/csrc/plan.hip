// NUFFT plan implementation (see plan.hpp).
#include "plan.hpp"

#include <atomic>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <limits>
#include <string>
#include <vector>

#include "plan_kernels.hpp"

namespace fa {

// ------------------------------------------------------------ dispatch glue

template <typename T>
void spread(const SpreadCtx<T> &ctx, const cplx_t<T> *c, cplx_t<T> *fw, int blk,
            hipStream_t stream) {
  switch (ctx.dim) {
    case 1: return spread_dim<T, 1>(ctx, c, fw, blk, stream);
    case 2: return spread_dim<T, 2>(ctx, c, fw, blk, stream);
    default: return spread_dim<T, 3>(ctx, c, fw, blk, stream);
  }
}

template <typename T>
void interp(const SpreadCtx<T> &ctx, cplx_t<T> *c, const cplx_t<T> *fw, int blk,
            hipStream_t stream) {
  switch (ctx.dim) {
    case 1: return interp_dim<T, 1>(ctx, c, fw, blk, stream);
    case 2: return interp_dim<T, 2>(ctx, c, fw, blk, stream);
    default: return interp_dim<T, 3>(ctx, c, fw, blk, stream);
  }
}

template <typename T>
void spread_batch(const InterpItem<T> *items, int nprob, int dim, int ns,
                  const cplx_t<T> *const *c_base, int64_t c_vec_offset, cplx_t<T> *const *fw_base,
                  int nvec, int64_t total_units, size_t lds_bytes, hipStream_t stream) {
  switch (dim) {
    case 1:
      return spread_batch_dim<T, 1>(items, nprob, ns, c_base, c_vec_offset, fw_base, nvec,
                                    total_units, lds_bytes, stream);
    case 2:
      return spread_batch_dim<T, 2>(items, nprob, ns, c_base, c_vec_offset, fw_base, nvec,
                                    total_units, lds_bytes, stream);
    default:
      return spread_batch_dim<T, 3>(items, nprob, ns, c_base, c_vec_offset, fw_base, nvec,
                                    total_units, lds_bytes, stream);
  }
}
template void spread_batch<float>(const InterpItem<float> *, int, int, int,
                                  const cplx_t<float> *const *, int64_t, cplx_t<float> *const *,
                                  int, int64_t, size_t, hipStream_t);
template void spread_batch<double>(const InterpItem<double> *, int, int, int,
                                   const cplx_t<double> *const *, int64_t,
                                   cplx_t<double> *const *, int, int64_t, size_t, hipStream_t);

template <typename T>
void interp_batch(const InterpItem<T> *items, int nprob, int dim, int ns, const KerParams<T> &kp,
                  const cplx_t<T> *const *fw_base, cplx_t<T> *const *c_base,
                  int64_t c_vec_offset, int nvec, int64_t total_units, size_t lds_bytes,
                  hipStream_t stream) {
  switch (dim) {
    case 1:
      return interp_batch_dim<T, 1>(items, nprob, ns, kp, fw_base, c_base, c_vec_offset, nvec,
                                    total_units, lds_bytes, stream);
    case 2:
      return interp_batch_dim<T, 2>(items, nprob, ns, kp, fw_base, c_base, c_vec_offset, nvec,
                                    total_units, lds_bytes, stream);
    default:
      return interp_batch_dim<T, 3>(items, nprob, ns, kp, fw_base, c_base, c_vec_offset, nvec,
                                    total_units, lds_bytes, stream);
  }
}
template void interp_batch<float>(const InterpItem<float> *, int, int, int,
                                  const KerParams<float> &,
                                  const cplx_t<float> *const *, cplx_t<float> *const *, int64_t,
                                  int, int64_t, size_t, hipStream_t);
template void interp_batch<double>(const InterpItem<double> *, int, int, int,
                                   const KerParams<double> &,
                                   const cplx_t<double> *const *, cplx_t<double> *const *,
                                   int64_t, int, int64_t, size_t, hipStream_t);

template <typename T>
void interp_grad(const SpreadCtx<T> &ctx, cplx_t<T> *c, cplx_t<T> *dc, const cplx_t<T> *fw,
                 int blk, hipStream_t stream) {
  switch (ctx.dim) {
    case 1: return interp_grad_dim<T, 1>(ctx, c, dc, fw, blk, stream);
    case 2: return interp_grad_dim<T, 2>(ctx, c, dc, fw, blk, stream);
    default: return interp_grad_dim<T, 3>(ctx, c, dc, fw, blk, stream);
  }
}

namespace {

int device_lds_limit(int dev) {
  int v = 0;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess ||
      v <= 0)
    v = 64 * 1024;
  return std::min(v, 160 * 1024);
}

template <typename T>
constexpr int max_ns_of() {
  return std::is_same<T, float>::value ? MAX_NS_F32 : MAX_NS;
}

// Cells of the padded tile of a start-node bin: prod_d (b_d + ns - 1).
int64_t tile_cells(int dim, const int *b, int ns) {
  int64_t c = 1;
  for (int d = 0; d < dim; ++d) c *= int64_t(b[d]) + ns - 1;
  return c;
}

// True when whole-bin spreading can run in 2^dim parity colors with a plain
// tile flush: per dimension an even number (>= 2) of equal bins, each at least
// ns - 1 nodes wide (so that same-parity tiles never touch).
bool colorable(int dim, const int *nf, const int *b, int ns) {
  for (int d = 0; d < dim; ++d) {
    if (b[d] < ns - 1 || nf[d] % b[d] != 0) return false;
    const int nb = nf[d] / b[d];
    if (nb < 2 || nb % 2 != 0) return false;
  }
  return true;
}

// Bin shape for the LDS kernels: the largest bin volume whose tile fits in
// `budget` bytes (cell_bytes per tile cell) with enough bins to fill the GPU
// (min_units work units, per color when colored). With want_colored the
// search is over colorable shapes only; returns false when none fits.
bool choose_bins(int dim, const int *nf, int ns, size_t cell_bytes, size_t budget,
                 bool want_colored, int64_t min_units, int *out) {
  std::vector<int> cand[3];
  for (int d = 0; d < 3; ++d) {
    if (d >= dim) {
      cand[d] = {1};
      continue;
    }
    for (int b = 1; b <= nf[d]; ++b) {
      if (want_colored) {
        if (nf[d] % b || b < ns - 1 || (nf[d] / b) % 2 || nf[d] / b < 2) continue;
      } else if (b > 8 && nf[d] % b && (b & (b - 1))) {
        continue;  // keep the search small: divisors, powers of two and b <= 8
      }
      cand[d].push_back(b);
    }
    if (cand[d].empty()) return false;
  }
  bool found = false;
  int64_t best_vol = 0;
  double best_ratio = 1e300;
  bool best_fill = false;
  for (int bx : cand[0])
    for (int by : cand[1])
      for (int bz : cand[2]) {
        const int b[3] = {bx, by, bz};
        const int64_t cells = tile_cells(dim, b, ns);
        if (size_t(cells) * cell_bytes > budget) continue;
        int64_t vol = 1, units = 1;
        for (int d = 0; d < dim; ++d) {
          vol *= b[d];
          units *= (nf[d] + b[d] - 1) / b[d];
        }
        if (want_colored) units >>= dim;
        const bool fill = units >= min_units;
        const double ratio = double(cells) / double(vol);
        // Prefer shapes that fill the GPU, then bigger bins, then less padding.
        const bool better = !found || (fill && !best_fill) ||
                            (fill == best_fill && (vol > best_vol ||
                                                   (vol == best_vol && ratio < best_ratio)));
        if (better) {
          found = true;
          best_fill = fill;
          best_vol = vol;
          best_ratio = ratio;
          out[0] = bx;
          out[1] = by;
          out[2] = bz;
        }
      }
  return found;
}

// Bin shape of the column spreader (3-D, fp32, ns <= 8): 8 x 8 first nodes in
// (x, y) (a multiple of every sub-column side) and the tallest z extent up to
// COL_ZC (the register depth of a column) whose tile plus the staging area
// fits, in order of preference:
// colorable within `budget` (two workgroups per CU), colorable within the
// device limit, any shape within `budget`, any within the limit. Returns
// false when nothing fits.
bool choose_col_bins(const int *nf, int ns, size_t budget, size_t limit, int *out,
                     bool *colored) {
  static const int zc[] = {16, 12, 8, 6, 4, 2, 1};
  for (int pass = 0; pass < 4; ++pass) {
    const bool want_col = pass < 2;
    const int64_t cap = int64_t(pass % 2 == 0 ? budget : limit);
    for (int bz : zc) {
      if (bz > nf[2] || bz > COL_ZC) continue;  // a bin never exceeds its grid
      const int b[3] = {8, 8, bz};
      if (int64_t(col_lds_bytes(tile_cells(3, b, ns))) > cap) continue;
      const bool colr = colorable(3, nf, b, ns);
      if (want_col && !colr) continue;
      out[0] = 8;
      out[1] = 8;
      out[2] = bz;
      *colored = colr;
      return true;
    }
  }
  return false;
}

// 2-D and 3-D interpolation bins: x-elongated shapes (wide rows spread a
// wave's gathers over more LDS banks) whose padded tile fits `budget` (two
// workgroups per CU) and is at most max_overhead times the bin (tile loads
// stay a small multiple of the grid): the first in order of preference that
// gives min_units bins, else (small grids) the one with the most bins.
bool choose_interp_bins(int dim, const int *nf, int ns, size_t cell_bytes, size_t budget,
                        int64_t min_units, int *out) {
  static const int pref3[][3] = {{32, 8, 8}, {32, 8, 4}, {16, 8, 8}, {16, 8, 4},
                                 {16, 4, 4}, {8, 8, 8},  {8, 4, 4},  {4, 4, 4}};
  static const int pref2[][3] = {{64, 32, 1}, {64, 16, 1}, {32, 16, 1}, {32, 8, 1},
                                 {16, 8, 1},  {8, 8, 1},   {8, 4, 1},   {4, 4, 1}};
  const double max_overhead = dim == 3 ? 4.5 : 2.0;
  int64_t best_units = -1;
  for (int i = 0; i < 8; ++i) {
    const int *b = dim == 3 ? pref3[i] : pref2[i];
    bool fits = true;
    int64_t units = 1, vol = 1;
    for (int d = 0; d < dim; ++d) {
      fits = fits && b[d] <= nf[d];
      units *= (nf[d] + b[d] - 1) / b[d];
      vol *= b[d];
    }
    const int64_t cells = tile_cells(dim, b, ns);
    if (!fits || size_t(cells) * cell_bytes > budget) continue;
    // A tile much larger than its bin (small bins, wide kernel) costs more
    // staging than the extra workgroups gain.
    if (double(cells) > max_overhead * double(vol)) continue;
    if (units >= min_units) {
      for (int d = 0; d < 3; ++d) out[d] = d < dim ? b[d] : 1;
      return true;
    }
    if (units > best_units) {
      best_units = units;
      for (int d = 0; d < 3; ++d) out[d] = d < dim ? b[d] : 1;
    }
  }
  return best_units > 0;
}

}  // namespace

// ------------------------------------------------------------ makeplan

template <typename T>
Plan<T>::Plan(int type, int dim, const int *nmodes, int iflag, int ntransf, T tol,
              const cufinufft_opts &opts)
    : type_(type), dim_(dim), ntransf_(ntransf), iflag_(iflag >= 0 ? 1 : -1), tol_(tol),
      opts_(opts) {
  // Argument checks come first, in the order of makeplan.cu:256-329, and
  // before any device call.
  if (type_ < 1 || type_ > 3) fail(FINUFFT_ERR_TYPE_NOTVALID, "invalid type");
  if (ntransf_ < 1) fail(FINUFFT_ERR_NTRANS_NOTVALID, "invalid ntransf");

  opts_.gpu_maxbatchsize = std::max(opts_.gpu_maxbatchsize, 0);
  // The caller's bin sizes; the sizes the plan settles on live in geom_ only,
  // so that a later type-3 setpts chooses afresh for its new grid.
  user_bin_[0] = opts_.gpu_binsizex;
  user_bin_[1] = opts_.gpu_binsizey;
  user_bin_[2] = opts_.gpu_binsizez;
  if (opts_.gpu_method != 3) opts_.gpu_np = 0;
  user_np_ = opts_.gpu_np;  // as the bins: chosen afresh at every type-3 setpts
  if (type_ != 3) {
    for (int d = 0; d < 3; ++d) ms_[d] = d < dim_ ? nmodes[d] : 1;
    spread_only_ = opts_.gpu_spreadinterponly != 0;
  } else {
    spread_only_ = true;  // the outer type-3 stage only spreads
  }
  stream_ = static_cast<hipStream_t>(opts_.gpu_stream);

  if (opts_.upsampfac == 0.0) {  // auto (makeplan.cu:289-297)
    opts_.upsampfac = 2.0;
    if (type_ == 3 && tol_ >= T(1e-9)) opts_.upsampfac = 1.25;
  }
  setup_kernel();
  if (opts_.gpu_method < 0 || opts_.gpu_method > 4)
    fail(FINUFFT_ERR_INSUFFICIENT_SHMEM, "invalid gpu_method");
  if (opts_.gpu_method == 4)
    fail(FINUFFT_ERR_METHOD_NOTVALID, "block-gather (gpu_method=4) is not provided");
  // The output-driven method only spreads: the reference's interpolation
  // dispatch refuses it at execute (spreadinterp.cpp:53-61), this plan at
  // makeplan with the same code. (The adjoint of a method-3 type-1 plan
  // interpolates by the LDS method on the plan's bins.)
  if (opts_.gpu_method == 3 && type_ == 2)
    fail(FINUFFT_ERR_METHOD_NOTVALID, "gpu_method 3 (output driven) only spreads");

  device_guard dg(opts_.gpu_device_id);
  upload_coefficients();
  // Method choice. The reference picks LDS subproblems for types 1 and 3 and
  // point-driven interpolation for type 2 (makeplan.cu:308-314); on this GPU
  // staging the grid in LDS is faster for interpolation too, so the LDS
  // method is the default for every type, with the point-driven method as
  // the fallback when the tile does not fit.
  auto_method_ = opts_.gpu_method == 0;
  if (auto_method_) opts_.gpu_method = 2;

  if (type_ != 3) {
    for (int d = 0; d < dim_; ++d) {
      int64_t n = spread_only_ ? ms_[d] : fine_grid_len(ms_[d], opts_.upsampfac, ns_);
      if (n > MAX_NF) fail(FINUFFT_ERR_MAXNALLOC, "fine grid exceeds the int32 index range");
      nf_[d] = int(n);
    }
    for (int d = 0; d < dim_; ++d)
      if (nf_[d] < 2 * ns_)
        fail(FINUFFT_ERR_SPREAD_BOX_SMALL, "grid dims smaller than 2*nspread");
    nftot_ = int64_t(nf_[0]) * nf_[1] * nf_[2];
    nmodes_ = int64_t(ms_[0]) * ms_[1] * ms_[2];
    if (nftot_ > MAX_NF) fail(FINUFFT_ERR_MAXNALLOC, "fine grid exceeds the int32 index range");
  }
  // Vectors per FFT batch: the user's choice, else up to 8 vectors while the
  // batch grid stays within 2 GiB.
  if (opts_.gpu_maxbatchsize > 0) {
    batchsize_ = std::min(opts_.gpu_maxbatchsize, ntransf_);
  } else {
    const int64_t per = std::max<int64_t>(nftot_, 1) * int64_t(sizeof(C));
    batchsize_ = int(std::max<int64_t>(1, std::min<int64_t>({ntransf_, 8, (int64_t(2) << 30) / per})));
  }
  if (type_ != 3) {
    setup_bins();  // the bin choice counts the vectors of a batch
    prepare_grid();
  }
}

template <typename T>
void Plan<T>::setup_kernel() {
  // setup_spreadinterp (makeplan.cu:43-119). gpu_kerevalmeth selects the
  // kernel: 1 (the default) Horner polynomials of the PSWF kernel with the
  // reference CPU library's width law (its v2.5 kernel, which the GPU
  // library's setup_spreadinterp lists as its own to-do, makeplan.cu:52);
  // 0 the ES kernel evaluated directly and 2 the ES kernel's Horner
  // polynomials, both with the GPU library's width rule.
  const double sigma = opts_.upsampfac;
  const int meth = opts_.gpu_kerevalmeth;
  if (meth < 0 || meth > 2) fail(FINUFFT_ERR_KERFORMULA_NOTVALID, "gpu_kerevalmeth must be 0, 1 or 2");
  if (sigma != 2.0 && sigma != 1.25) {
    if (meth != 0)
      fail(FINUFFT_ERR_HORNER_WRONG_BETA,
           "nonstandard upsampfac cannot use the Horner kernel (kerevalmeth=1 or 2)");
    if (sigma <= 1.0) fail(FINUFFT_ERR_UPSAMPFAC_TOO_SMALL, "upsampfac must exceed 1");
    if (!opts_.gpu_spreadinterponly && sigma > 4.0 && opts_.debug)
      std::fprintf(stderr, "[finufft_amd] warning: upsampfac=%.3g is too large to be beneficial\n",
                   sigma);
  }
  T eps = tol_;
  const T eps_mach = std::numeric_limits<T>::epsilon();
  if (eps < eps_mach) {
    if (opts_.debug)
      std::fprintf(stderr, "[finufft_amd] warning: increasing tol=%.3g to eps_mach=%.3g\n",
                   double(eps), double(eps_mach));
    eps = eps_mach;
  }
  const KernelShape ks =
      meth == 1 ? kernel_shape_pswf(double(eps), sigma, dim_, type_, max_ns_of<T>(),
                                    std::is_same<T, float>::value)
                : kernel_shape(double(eps), double(eps / T(10)), sigma, max_ns_of<T>());
  if (ks.clipped && opts_.debug)
    std::fprintf(stderr, "[finufft_amd] warning: kernel width clipped to %d\n", ks.ns);
  ns_ = ks.ns;
  beta_ = ks.beta;
  formula_ = ks.formula;
  if (opts_.debug)
    std::printf("[finufft_amd] kerevalmeth=%d tol=%.3g sigma=%.3g: %s ns=%d beta=%.4g\n", meth,
                double(eps), sigma, formula_ == KER_PSWF ? "PSWF" : "ES", ns_, beta_);
}

template <typename T>
void Plan<T>::upload_coefficients() {
  if (opts_.gpu_kerevalmeth == 0) return;
  const int nc = horner_nc_sigma2(ns_);
  const std::vector<double> c64 = formula_ == KER_PSWF
                                      ? horner_coeffs_pswf(ns_, beta_, nc)
                                      : horner_coeffs(ns_, opts_.upsampfac, beta_, nc);
  std::vector<T> c(c64.begin(), c64.end());
  coef_.resize(c.size());
  hip_check(hipMemcpy(coef_.get(), c.data(), c.size() * sizeof(T), hipMemcpyHostToDevice),
            "hipMemcpy(coef)");
}

template <typename T>
void Plan<T>::setup_bins() {
  // Bin sizes per method (heuristics.cu:325-445, re-derived for this GPU's
  // LDS). Needs nf_. Method 1 keeps the reference's bins of the folded
  // coordinate; the LDS methods bin by first support node.
  const int lds_limit = device_lds_limit(opts_.gpu_device_id);
  int *bs = geom_.binsize;
  // The automatic method is chosen again for every grid (type 3 calls this
  // at each setpts).
  if (auto_method_) opts_.gpu_method = 2;
  // User bin sizes are taken as given up to the grid size: a start-node bin
  // wider than its grid would give a tile past the wrap range of the tile
  // kernels (two conditional subtractions of nf).
  int user[3];
  for (int d = 0; d < 3; ++d) user[d] = std::min(user_bin_[d], nf_[d]);
  const bool user_bins = user[0] > 0 || user[1] > 0 || user[2] > 0;
  auto set_ref_bins = [&]() {
    const int b = dim_ == 1 ? 1024 : dim_ == 2 ? 40 : 8;
    for (int d = 0; d < 3; ++d) bs[d] = d < dim_ ? (user[d] > 0 ? user[d] : b) : 1;
    geom_.start_bins = 0;
    colored_ = false;
    lds_bytes_ = 0;
    lds_spread_ = 0;
  };
  geom_.colS = 0;
  geom_.bank = 0;
  od_np_ = 0;
  lds_od_ = 0;
  // Output-driven spreading (method 3; plans of types 1 and 3).
  const bool od = opts_.gpu_method == 3 && type_ != 2;
  // Column spreading: 3-D fp32 spreading plans with ns <= 8. User bin sizes
  // select it when x and y are 8 (the footprint of one wave) and z is at most
  // COL_ZC.
  const bool col_ok = !od && opts_.gpu_method != 1 && type_ != 2 && dim_ == 3 &&
                      std::is_same<T, float>::value && ns_ <= COL_MAX_NS && nf_[0] >= 8 &&
                      nf_[1] >= 8 &&
                      (!user_bins ||
                       (user[0] == 8 && user[1] == 8 && user[2] > 0 && user[2] <= COL_ZC));
  bool col_done = false;
  if (col_ok) {
    const size_t budget = std::min<size_t>(size_t(lds_limit), 80 * 1024);
    bool colr = false;
    if (user_bins) {
      bs[0] = bs[1] = 8;
      bs[2] = user[2];
      colr = colorable(3, nf_, bs, ns_);
      col_done = int64_t(col_lds_bytes(tile_cells(3, bs, ns_))) <= lds_limit;
    } else {
      col_done = choose_col_bins(nf_, ns_, budget, size_t(lds_limit), bs, &colr);
    }
    if (col_done) {
      geom_.start_bins = 1;
      geom_.colS = fa::col_side(ns_);
      colored_ = colr;
      lds_spread_ = col_lds_bytes(tile_cells(3, bs, ns_));
      lds_bytes_ = size_t(tile_cells(3, bs, ns_)) * 2 * sizeof(T);
    }
  }
  if (od) {
    // Start-node bins whose fp64 tile and np staged points fit the LDS: the
    // user's np and bins, the user's bins with as many points as fit (a
    // multiple of 16, at most 256), or 64 points (or the user's np) with the
    // largest bins that fit (the cases of heuristics.cu:394-423).
    const size_t cell = 2 * sizeof(double);
    int np = user_np_;
    bool ok = np >= 0;
    if (ok && user_bins) {
      for (int d = 0; d < 3; ++d)
        bs[d] = d < dim_ ? (user[d] > 0 ? user[d] : std::min(8, nf_[d])) : 1;
      const int64_t tile = tile_cells(dim_, bs, ns_) * int64_t(cell);
      if (np == 0) {
        const int64_t room = (int64_t(lds_limit) - tile) /
                             int64_t(od_stage_bytes(dim_, ns_, 1, sizeof(T)));
        np = int(std::min<int64_t>(256, std::max<int64_t>(room, 0)) & ~int64_t(15));
      }
      ok = np >= 1 && tile + int64_t(od_stage_bytes(dim_, ns_, np, sizeof(T))) <= lds_limit;
    } else if (ok) {
      if (np == 0) np = 64;
      const int64_t stage = int64_t(od_stage_bytes(dim_, ns_, np, sizeof(T)));
      const int64_t min_units = (1024 + batchsize_ - 1) / batchsize_;
      ok = stage < lds_limit &&
           choose_bins(dim_, nf_, ns_, cell, size_t(lds_limit - stage), false, min_units, bs);
    }
    if (!ok)
      fail(FINUFFT_ERR_INSUFFICIENT_SHMEM,
           "output-driven spreading: the tile and np staged points do not fit the LDS");
    geom_.start_bins = 1;
    colored_ = false;
    lds_spread_ = size_t(tile_cells(dim_, bs, ns_)) * cell;
    lds_bytes_ = size_t(tile_cells(dim_, bs, ns_)) * 2 * sizeof(T);
    od_np_ = np;
    lds_od_ = lds_spread_ + od_stage_bytes(dim_, ns_, np, sizeof(T));
  } else if (col_done) {
    // bins chosen above
  } else if (opts_.gpu_method == 1) {
    set_ref_bins();
  } else {
    // Spreading accumulates in fp64 LDS tiles (16 bytes per complex cell);
    // interpolation stages the grid in the plan precision.
    const bool spreads = type_ != 2;
    const size_t cell = spreads ? 2 * sizeof(double) : 2 * sizeof(T);
    bool ok = true;
    if (user_bins) {
      for (int d = 0; d < 3; ++d)
        bs[d] = d < dim_ ? (user[d] > 0 ? user[d] : std::min(8, nf_[d])) : 1;
      // (interpolation tiles leave room for the bank-ordered dealing's bounds)
      ok = size_t(tile_cells(dim_, bs, ns_)) * cell + (spreads ? 0 : INTERP_BANK_LDS) <=
           size_t(lds_limit);
    } else {
      // about 1024 workgroups per launch (bins x vectors of a batch)
      const int64_t min_units = (1024 + batchsize_ - 1) / batchsize_;
      ok = !spreads && dim_ >= 2 &&
           choose_interp_bins(dim_, nf_, ns_, cell, std::min<size_t>(size_t(lds_limit), 64 * 1024),
                              min_units, bs);
      if (!ok)
        ok = spreads && choose_bins(dim_, nf_, ns_, cell, size_t(lds_limit), true, min_units, bs);
      if (!ok) ok = choose_bins(dim_, nf_, ns_, cell, size_t(lds_limit), false, min_units, bs);
    }
    if (!ok) {
      if (auto_method_) {
        opts_.gpu_method = 1;
        set_ref_bins();
      } else {
        fail(FINUFFT_ERR_INSUFFICIENT_SHMEM, "LDS tile does not fit for gpu_method 2/3");
      }
    } else {
      geom_.start_bins = 1;
      colored_ = spreads && colorable(dim_, nf_, bs, ns_);
      lds_bytes_ = size_t(tile_cells(dim_, bs, ns_)) * 2 * sizeof(T);
      lds_spread_ = size_t(tile_cells(dim_, bs, ns_)) * 2 * sizeof(double);
      // interpolation plans order each bin by LDS bank class
      geom_.bank = spreads ? 0 : 1;
    }
  }
  spread_gm_ = lds_spread_ > size_t(lds_limit);
  for (int d = 0; d < 3; ++d)
    if (bs[d] < 1) fail(FINUFFT_ERR_BINSIZE_NOTVALID, "invalid bin size");
  if (opts_.gpu_maxsubprobsize < 1) opts_.gpu_maxsubprobsize = 1024;
  geom_.dim = dim_;
  geom_.ns = ns_;
  geom_.fold64 = std::is_same<T, float>::value ? 1 : 0;
  for (int d = 0; d < 3; ++d) {
    geom_.nf[d] = nf_[d];
    geom_.nbins[d] = (nf_[d] + bs[d] - 1) / bs[d];
  }
}

namespace {
std::atomic<int> g_fft_path{0};
}  // namespace

int fft_path_setting() { return g_fft_path.load(std::memory_order_relaxed); }
void set_fft_path_setting(int path) { g_fft_path.store(path, std::memory_order_relaxed); }

template <typename T>
void Plan<T>::prepare_grid() {
  // The work grid, the FFT plan and the kernel's Fourier series.
  if (!spread_only_ || type_ == 3) {
    fw_.resize(size_t(nftot_) * batchsize_);
  }
  if (!spread_only_) {
    fft3_.reset();
    if (fft_path_setting() == 0 && fft3_supported(dim_, nf_)) {
      fft3_ = std::make_unique<Fft3<T>>();
      fft3_->setup(ms_, nf_, batchsize_, opts_.modeord);
    } else {
      fft_.setup(dim_, nf_, batchsize_, std::is_same<T, double>::value);
    }
    int q = 0;
    std::vector<double> z(3 * MAX_NQUAD, 0.0), f(3 * MAX_NQUAD, 0.0);
    const KernelFn ker(formula_, ns_, beta_);
    for (int d = 0; d < dim_; ++d)
      fourier_nodes(ker, ns_, false, &q, z.data() + d * MAX_NQUAD, f.data() + d * MAX_NQUAD);
    dbuf<double> dz(z.size()), df(f.size());
    hip_check(hipMemcpy(dz.get(), z.data(), z.size() * sizeof(double), hipMemcpyHostToDevice),
              "hipMemcpy(nodes)");
    hip_check(hipMemcpy(df.get(), f.data(), f.size() * sizeof(double), hipMemcpyHostToDevice),
              "hipMemcpy(weights)");
    T *outs[3] = {nullptr, nullptr, nullptr};
    for (int d = 0; d < dim_; ++d) {
      fwker_[d].resize(size_t(nf_[d] / 2 + 1));
      outs[d] = fwker_[d].get();
    }
    fseries<T>(dim_, nf_, q, df.get(), dz.get(), outs, stream_);
    hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize(fseries)");
  }
}

// ------------------------------------------------------------ setpts

template <typename T>
void Plan<T>::sort_points(const T *x, const T *y, const T *z) {
  xyz_[0] = x;
  xyz_[1] = dim_ > 1 ? y : nullptr;
  xyz_[2] = dim_ > 2 ? z : nullptr;
  const int nb = geom_.total();
  idx_.reserve_at_least(std::max(M_, 1));
  binstart_.reserve_at_least(size_t(nb) + 1);
  const T *pts[3] = {xyz_[0], xyz_[1], xyz_[2]};
  if (geom_.colS) colstart_.reserve_at_least(size_t(nb) * geom_.nsub() + 1);
  if (geom_.bank) colstart_.reserve_at_least(size_t(nb) * 32 + 1);
  if (opts_.gpu_method == 1 && !opts_.gpu_sort) {
    iota(idx_.get(), M_, stream_);
  } else {
    // the permutation, and the coordinates in sorted order so that
    // spread/interp read them coalesced
    T *sorted[3] = {nullptr, nullptr, nullptr};
    for (int d = 0; d < dim_; ++d) {
      sxyz_[d].reserve_at_least(std::max(M_, 1));
      sorted[d] = sxyz_[d].get();
    }
    bin_sort<T>(geom_, M_, pts, idx_.get(), binstart_.get(), scratch_, stream_,
                (geom_.colS || geom_.bank) ? colstart_.get() : nullptr, sorted);
  }
  if (opts_.gpu_method != 1) {
    substart_.reserve_at_least(size_t(nb) + 1);
    // Column spreading runs whole bins in parity colors and only the bins
    // denser than dense_min_ (all bins when the shape is not colorable) as
    // subproblems; their count is read back so that the launch is exact.
    // The limit is one maxsub: a few whole-bin workgroups of far more points
    // than the typical bin (the centre of radial spokes) would hold back
    // their color's launch.
    dense_min_ = geom_.colS && colored_ ? opts_.gpu_maxsubprobsize : 0;
    // Interpolation plans cut bins into subproblems of sub_size_ points (one
    // per bin for typical densities, several for dense bins).
    // Column spreading cuts its dense bins into subproblems of dense_min_
    // points: each subproblem adds its whole padded tile to the grid with
    // atomics, and the dense centre of radial spokes (bins of 1e5 points)
    // otherwise sends about a hundred tiles' atomics to the same cells.
    sub_size_ = type_ == 2 ? 4 * opts_.gpu_maxsubprobsize
                           : (dense_min_ > 0 ? dense_min_ : opts_.gpu_maxsubprobsize);
    sub_capacity_ = subproblem_capacity(M_, nb, sub_size_);
    sub2bin_.reserve_at_least(std::max(sub_capacity_, 1));
    build_subproblems(binstart_.get(), nb, sub_size_, dense_min_, substart_.get(),
                      sub2bin_.get(), scratch_, stream_);
    n_dense_sub_ = 0;
    if (geom_.colS) {
      hip_check(hipMemcpyAsync(&n_dense_sub_, substart_.get() + nb, sizeof(int),
                               hipMemcpyDeviceToHost, stream_),
                "hipMemcpyAsync(subproblem count)");
      hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize(subproblem count)");
    }
  }
}

template <typename T>
void Plan<T>::setpts(int M, const T *x, const T *y, const T *z, int N, const T *s, const T *t,
                     const T *u) {
  device_guard dg(opts_.gpu_device_id);
  ++setpts_gen_;
  if (M < 0) fail(FINUFFT_ERR_NUM_NU_PTS_INVALID, "negative number of points");
  if (type_ == 3) {
    // the reference's order: the targets' count, then the target arrays
    // (setpts.cu:95-119); then (beyond the reference) the source arrays
    if (N < 0) fail(FINUFFT_ERR_NUM_NU_PTS_INVALID, "type 3: negative number of targets");
    if (int64_t(N) > MAX_NF) fail(FINUFFT_ERR_NUM_NU_PTS_INVALID, "type 3: too many targets");
    if (dim_ > 0 && s == nullptr) fail(FINUFFT_ERR_INVALID_ARGUMENT, "type 3: s is null");
    if (dim_ > 1 && t == nullptr) fail(FINUFFT_ERR_INVALID_ARGUMENT, "type 3: t is null");
    if (dim_ > 2 && u == nullptr) fail(FINUFFT_ERR_INVALID_ARGUMENT, "type 3: u is null");
  }
  if (M > 0 && (x == nullptr || (dim_ > 1 && y == nullptr) || (dim_ > 2 && z == nullptr)))
    fail(FINUFFT_ERR_INVALID_ARGUMENT, "null coordinate array");
  if (type_ != 3) {
    M_ = M;
    if (M == 0) {
      // No points: nothing is sorted and execute launches nothing but the
      // clearing of a type-1 output.
      for (int d = 0; d < 3; ++d) xyz_[d] = nullptr;
      n_dense_sub_ = 0;
      return;
    }
    sort_points(x, y, z);
    return;
  }
  // ---- type 3 (setpts.cu:92-313)
  t3_set_ = false;
  M_ = M;
  Nk_ = N;
  if (M == 0 || N == 0) {
    // No sources (the output is zero) or no targets (there is no output):
    // execute clears the output or returns, with no other launch.
    inner_.reset();
    t3_set_ = true;
    return;
  }
  const T *src[3] = {x, y, z};
  const T *trg[3] = {s, t, u};
  double h[3] = {1, 1, 1}, gam[3] = {1, 1, 1};
  int64_t nf123[3] = {1, 1, 1};
  for (int d = 0; d < dim_; ++d) {
    // an empty set of sources or targets has zero width and centre
    double lo = 0, hi = 0;
    if (M > 0) minmax<T>(src[d], M, &lo, &hi, stream_);
    const auto [X, Cc] = widcen(lo, hi);
    lo = hi = 0;
    if (N > 0) minmax<T>(trg[d], N, &lo, &hi, stream_);
    const auto [S, D] = widcen(lo, hi);
    // The type-3 parameters live in the plan's precision, as in the reference.
    t3C_[d] = double(T(Cc));
    t3D_[d] = double(T(D));
    const auto [nfd, hd, gd] = nhg_type3(opts_.upsampfac, double(T(X)), double(T(S)), ns_, MAX_NF);
    nf123[d] = nfd;
    h[d] = double(T(hd));
    gam[d] = double(T(gd));
  }
  const int64_t nft = nf123[0] * nf123[1] * nf123[2];
  if (nft * std::max(batchsize_, 1) > MAX_NF)
    fail(FINUFFT_ERR_MAXNALLOC, "type 3: fine grid batch exceeds the int32 index range");
  for (int d = 0; d < 3; ++d) nf_[d] = int(nf123[d]);
  nftot_ = nft;
  setup_bins();

  for (int d = 0; d < dim_; ++d) {
    xyzp_[d].reserve_at_least(std::max(M, 1));
    stup_[d].reserve_at_least(std::max(N, 1));
    const T inv_gam = T(1) / T(gam[d]);
    affine<T>(src[d], xyzp_[d].get(), M, t3C_[d], double(inv_gam), stream_);
    const T scale = T(h[d]) * T(gam[d]);
    affine<T>(trg[d], stup_[d].get(), N, t3D_[d], double(scale), stream_);
  }
  prephase_.reserve_at_least(std::max(M, 1));
  t3_prephase<T>(dim_, src, t3D_, iflag_, M, prephase_.get(), stream_);
  // phihat at the rescaled targets and the deconvolution factor
  {
    int q = 0;
    std::vector<double> zq(MAX_NQUAD), fq(MAX_NQUAD);
    fourier_nodes(KernelFn(formula_, ns_, beta_), ns_, true, &q, zq.data(), fq.data());
    dbuf<double> dz(MAX_NQUAD), df(MAX_NQUAD);
    hip_check(hipMemcpy(dz.get(), zq.data(), MAX_NQUAD * sizeof(double), hipMemcpyHostToDevice),
              "hipMemcpy(t3 nodes)");
    hip_check(hipMemcpy(df.get(), fq.data(), MAX_NQUAD * sizeof(double), hipMemcpyHostToDevice),
              "hipMemcpy(t3 weights)");
    deconv_.reserve_at_least(std::max(N, 1));
    const T *stup[3] = {stup_[0].get(), stup_[1].get(), stup_[2].get()};
    t3_deconv<T>(dim_, trg, stup, t3C_, t3D_, iflag_, q, df.get(), dz.get(), N, deconv_.get(),
                 stream_);
    hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize(t3 deconv)");
  }
  prepare_grid();
  cpre_.reserve_at_least(size_t(std::max(M, 1)) * batchsize_);
  sort_points(xyzp_[0].get(), xyzp_[1].get(), xyzp_[2].get());

  // inner type-2 plan over the outer fine grid (setpts.cu:295-313); mode
  // order is forced to CMCL order, which the type-3 algebra assumes (the
  // reference CPU library does the same, include/finufft/setpts.hpp:298).
  cufinufft_opts t2opts = opts_;
  t2opts.gpu_spreadinterponly = 0;
  t2opts.gpu_method = 0;
  t2opts.modeord = 0;
  t2opts.gpu_binsizex = t2opts.gpu_binsizey = t2opts.gpu_binsizez = 0;
  t2opts.gpu_maxbatchsize = batchsize_;
  inner_.reset();
  inner_ = std::make_unique<Plan<T>>(2, dim_, nf_, iflag_, batchsize_, tol_, t2opts);
  inner_->setpts(N, stup_[0].get(), stup_[1].get(), stup_[2].get(), 0, nullptr, nullptr, nullptr);
  t3_set_ = true;
}

// ------------------------------------------------------------ execute

template <typename T>
SpreadCtx<T> Plan<T>::spread_ctx() const {
  SpreadCtx<T> c;
  c.dim = dim_;
  c.ns = ns_;
  c.method = opts_.gpu_method == 1 ? 1 : (opts_.gpu_method == 3 && type_ != 2 ? 3 : 2);
  c.M = M_;
  const bool sorted = !(opts_.gpu_method == 1 && !opts_.gpu_sort);
  for (int d = 0; d < 3; ++d) {
    c.nf[d] = nf_[d];
    c.xyz[d] = xyz_[d];
    c.sxyz[d] = (d < dim_ && sorted) ? sxyz_[d].get() : (d < dim_ ? xyz_[d] : nullptr);
    c.binsize[d] = geom_.binsize[d];
    c.nbins[d] = geom_.nbins[d];
  }
  c.kp.beta = T(beta_);
  c.kp.es_c = T(4.0 / (double(ns_) * ns_));
  c.kp.horner = opts_.gpu_kerevalmeth != 0 ? 1 : 0;
  c.kp.fold64 = std::is_same<T, float>::value ? 1 : 0;
  c.kp.coef = coef_.get();
  c.idx = idx_.get();
  c.binstart = binstart_.get();
  c.maxsub = sub_size_;
  c.interp_sub = type_ == 2 && opts_.gpu_method != 1 ? 1 : 0;
  c.substart = substart_.get();
  c.sub2bin = sub2bin_.get();
  c.sub_capacity = sub_capacity_;
  c.lds_bytes = lds_bytes_;
  c.lds_spread = lds_spread_;
  c.spread_gm = spread_gm_ ? 1 : 0;
  c.colored = colored_ ? 1 : 0;
  c.colS = geom_.colS;
  c.colstart = colstart_.get();
  c.bank = geom_.bank;
  c.dense_min = dense_min_;
  c.n_dense_sub = n_dense_sub_;
  c.od_np = od_np_;
  c.lds_od = lds_od_;
  return c;
}

template <typename T>
void Plan<T>::exec_type1(C *d_c, C *d_fk, int fft_sign, int ntr) {
  if (M_ == 0) {  // no points: the modes (or the spread grid) are zero
    zero<T>(d_fk, nmodes_ * ntr, stream_);
    return;
  }
  const SpreadCtx<T> ctx = spread_ctx();
  for (int i = 0; i * batchsize_ < ntr; ++i) {
    const int blk = std::min(ntr - i * batchsize_, batchsize_);
    const C *c = d_c + int64_t(i) * batchsize_ * M_;
    C *fk = d_fk + int64_t(i) * batchsize_ * nmodes_;
    C *fw = spread_only_ ? fk : fw_.get();
    timer_.begin(ST_ZERO, stream_);
    zero<T>(fw, nftot_ * blk, stream_);
    timer_.end(stream_);
    timer_.begin(ST_SPREAD, stream_);
    spread<T>(ctx, c, fw, blk, stream_);
    timer_.end(stream_);
    if (spread_only_) continue;
    grid_to_modes(fw, fk, fft_sign, blk);
  }
}

template <typename T>
void Plan<T>::grid_to_modes(C *fw, C *fk, int fft_sign, int blk) {
  const T *ker[3] = {fwker_[0].get(), fwker_[1].get(), fwker_[2].get()};
  if (fft3_) {
    timer_.begin(ST_FFT, stream_);
    fft3_->to_modes(fw, fk, ker, fft_sign, blk, stream_);
    timer_.end(stream_);
    return;
  }
  timer_.begin(ST_FFT, stream_);
  fft_.exec(fw, fft_sign, stream_);
  timer_.end(stream_);
  timer_.begin(ST_DECONV, stream_);
  deconvolve<T>(dim_, opts_.modeord, true, ms_, nf_, ker, fw, fk, blk, stream_);
  timer_.end(stream_);
}

template <typename T>
void Plan<T>::modes_to_grid(C *fk, C *fw, int fft_sign, int blk) {
  const T *ker[3] = {fwker_[0].get(), fwker_[1].get(), fwker_[2].get()};
  if (fft3_) {
    timer_.begin(ST_FFT, stream_);
    fft3_->to_grid(fk, fw, ker, fft_sign, blk, stream_);
    timer_.end(stream_);
    return;
  }
  timer_.begin(ST_DECONV, stream_);
  deconvolve<T>(dim_, opts_.modeord, false, ms_, nf_, ker, fw, fk, blk, stream_);
  timer_.end(stream_);
  timer_.begin(ST_FFT, stream_);
  fft_.exec(fw, fft_sign, stream_);
  timer_.end(stream_);
}

template <typename T>
void Plan<T>::exec_type2(C *d_c, C *d_fk, int fft_sign, int ntr) {
  if (M_ == 0) return;  // no points: no values to compute
  const SpreadCtx<T> ctx = spread_ctx();
  for (int i = 0; i * batchsize_ < ntr; ++i) {
    const int blk = std::min(ntr - i * batchsize_, batchsize_);
    C *c = d_c + int64_t(i) * batchsize_ * M_;
    C *fk = d_fk + int64_t(i) * batchsize_ * nmodes_;
    C *fw = fk;
    if (!spread_only_) {
      fw = fw_.get();
      modes_to_grid(fk, fw, fft_sign, blk);
    }
    timer_.begin(ST_INTERP, stream_);
    interp<T>(ctx, c, fw, blk, stream_);
    timer_.end(stream_);
  }
}

template <typename T>
void Plan<T>::exec_type3(C *d_c, C *d_fk) {
  if (Nk_ == 0) return;  // no targets
  if (M_ == 0) {         // no sources: the target values are zero
    zero<T>(d_fk, int64_t(Nk_) * ntransf_, stream_);
    return;
  }
  const SpreadCtx<T> ctx = spread_ctx();
  for (int i = 0; i * batchsize_ < ntransf_; ++i) {
    const int blk = std::min(ntransf_ - i * batchsize_, batchsize_);
    const C *c = d_c + int64_t(i) * batchsize_ * M_;
    C *fk = d_fk + int64_t(i) * batchsize_ * Nk_;
    cmul<T>(c, prephase_.get(), cpre_.get(), M_, blk, stream_);
    zero<T>(fw_.get(), nftot_ * blk, stream_);
    spread<T>(ctx, cpre_.get(), fw_.get(), blk, stream_);
    inner_->exec_type2(fk, fw_.get(), inner_->iflag_, blk);
    cmul<T>(fk, deconv_.get(), fk, Nk_, blk, stream_);
  }
}

template <typename T>
void Plan<T>::check_data(const C *c, const C *fk) const {
  // Null data arrays are refused where a kernel would read or write them.
  if (type_ == 3 && !t3_set_)
    fail(FINUFFT_ERR_INVALID_ARGUMENT, "type 3: setpts must come before execute");
  if (M_ > 0 && c == nullptr) fail(FINUFFT_ERR_INVALID_ARGUMENT, "null c array");
  if ((type_ != 3 || Nk_ > 0) && fk == nullptr) fail(FINUFFT_ERR_INVALID_ARGUMENT, "null fk array");
}

template <typename T>
void Plan<T>::execute(C *c, C *fk) {
  device_guard dg(opts_.gpu_device_id);
  check_data(c, fk);
  switch (type_) {
    case 1: return exec_type1(c, fk, iflag_, ntransf_);
    case 2: return exec_type2(c, fk, iflag_, ntransf_);
    default: return exec_type3(c, fk);
  }
}

template <typename T>
void Plan<T>::execute_adjoint(C *c, C *fk) {
  device_guard dg(opts_.gpu_device_id);
  check_data(c, fk);
  if (type_ == 1) return exec_type2(c, fk, -iflag_, ntransf_);
  if (type_ == 2) return exec_type1(c, fk, -iflag_, ntransf_);
  fail(FINUFFT_ERR_TYPE_NOTVALID, "adjoint is provided for types 1 and 2");
}

template <typename T>
void Plan<T>::interp_grad(C *d_c, C *d_dc, C *d_fk) {
  // The type-2 direction of the plan (the transform itself for type 2, the
  // adjoint for type 1), with the coordinate derivatives of the result.
  device_guard dg(opts_.gpu_device_id);
  if (type_ == 3) fail(FINUFFT_ERR_TYPE_NOTVALID, "interp_grad is provided for types 1 and 2");
  if (spread_only_) fail(FINUFFT_ERR_INVALID_ARGUMENT, "interp_grad needs the full transform");
  if (d_dc == nullptr) fail(FINUFFT_ERR_INVALID_ARGUMENT, "interp_grad: dc is null");
  if (d_fk == nullptr) fail(FINUFFT_ERR_INVALID_ARGUMENT, "interp_grad: fk is null");
  if (M_ == 0) return;  // no points: no values or derivatives
  const int fft_sign = type_ == 2 ? iflag_ : -iflag_;
  const SpreadCtx<T> ctx = spread_ctx();
  const int64_t mtot = int64_t(M_) * ntransf_;
  for (int i = 0; i * batchsize_ < ntransf_; ++i) {
    const int blk = std::min(ntransf_ - i * batchsize_, batchsize_);
    C *c = d_c ? d_c + int64_t(i) * batchsize_ * M_ : nullptr;
    C *fk = d_fk + int64_t(i) * batchsize_ * nmodes_;
    C *fw = fw_.get();
    modes_to_grid(fk, fw, fft_sign, blk);
    // dc is (dim, ntransf, M); this batch writes vectors [i b, i b + blk) of
    // every dimension, so the kernel's per-dimension stride must be mtot.
    if (blk == ntransf_) {
      ::fa::interp_grad<T>(ctx, c, d_dc, fw, blk, stream_);
    } else {
      for (int t = 0; t < blk; ++t) {
        grad_scratch_.reserve_at_least(size_t(dim_) * M_);
        ::fa::interp_grad<T>(ctx, c ? c + int64_t(t) * M_ : nullptr, grad_scratch_.get(),
                             fw + int64_t(t) * nftot_, 1, stream_);
        for (int d = 0; d < dim_; ++d)
          hip_check(hipMemcpyAsync(d_dc + d * mtot + (int64_t(i) * batchsize_ + t) * M_,
                                   grad_scratch_.get() + int64_t(d) * M_, size_t(M_) * sizeof(C),
                                   hipMemcpyDeviceToDevice, stream_),
                    "hipMemcpyAsync(grad)");
      }
    }
  }
}

template class Plan<float>;
template class Plan<double>;

}  // namespace fa
