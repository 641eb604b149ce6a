// Plan-time host math (see host_math.hpp for what each routine restates).
#include "host_math.hpp"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <limits>

namespace fa {

KernelShape kernel_shape(double eps, double eps_div10, double upsampfac, int max_ns) {
  KernelShape ks{};
  int ns;
  if (upsampfac == 2.0) {
    // One digit per power of ten: ns = ceil(-log10(eps/10)) (makeplan.cu:93).
    ns = int(std::ceil(-std::log10(eps_div10)));
  } else {
    // Width formula for other sigma (makeplan.cu:94-96).
    ns = int(std::ceil(-std::log(eps) / (PI * std::sqrt(1.0 - 1.0 / upsampfac))));
  }
  ns = std::max(2, ns);
  ks.clipped = ns > max_ns;
  if (ks.clipped) ns = max_ns;
  ks.ns = ns;
  double betaoverns = 2.30;
  if (ns == 2) betaoverns = 2.20;
  if (ns == 3) betaoverns = 2.26;
  if (ns == 4) betaoverns = 2.38;
  if (upsampfac != 2.0) betaoverns = 0.97 * PI * (1.0 - 1.0 / (2.0 * upsampfac));
  ks.beta = betaoverns * ns;
  ks.formula = KER_ES;
  return ks;
}

KernelShape kernel_shape_pswf(double eps, double upsampfac, int dim, int type, int max_ns,
                              bool is_float) {
  KernelShape ks{};
  // Aliasing law tol = tolfac exp(-(ns - 1) pi sqrt(1 - 1/sigma)) with the
  // reference's empirical per-dimension and type-3 factors.
  double tolfac = 0.18;
  for (int d = 1; d < dim; ++d) tolfac *= 1.4;
  if (type == 3) tolfac *= 1.4;
  int ns = int(std::ceil(std::log(tolfac / eps) / (PI * std::sqrt(1.0 - 1.0 / upsampfac)) + 1.0));
  ns = std::max(2, ns);
  ks.clipped = ns > max_ns;
  ns = std::min(ns, max_ns);
  // fp32 at low upsampling loses digits to cancellation beyond width 8
  if (is_float && upsampfac < 1.4) ns = std::min(ns, 8);
  ks.ns = ns;
  ks.beta = PI * ns * (1.0 - 1.0 / (2.0 * upsampfac)) - 0.05;
  ks.formula = KER_PSWF;
  return ks;
}

Prolate0::Prolate0(double c) {
  // Even Legendre degrees k = 2m, m < n; n covers the eigenvector's decay
  // (its coefficients fall off like (c / 2k)^k) with a wide margin.
  const int n = int(c) + 40;
  std::vector<long double> d(n), e(n, 0.0L);
  const long double cc = (long double)c * c;
  for (int m = 0; m < n; ++m) {
    const long double k = 2.0L * m;
    // x^2 P_k = A P_{k+2} + B P_k + C P_{k-2}; in the orthonormal basis the
    // operator is k(k+1) + c^2 B on the diagonal and c^2 A sqrt((2k+1)/(2k+5))
    // between k and k+2 (symmetric).
    long double B = (k + 1) * (k + 1) / ((2 * k + 1) * (2 * k + 3));
    if (m > 0) B += k * k / ((2 * k + 1) * (2 * k - 1));
    d[m] = k * (k + 1) + cc * B;
    e[m] = cc * (k + 1) * (k + 2) / ((2 * k + 3) * std::sqrt((2 * k + 1) * (2 * k + 5)));
  }
  // Inverse iteration with zero shift: the matrix is positive definite and
  // psi_0 is its lowest eigenvector; each step solves the tridiagonal system
  // by elimination without pivoting (stable for SPD matrices).
  std::vector<long double> v(n, 0.0L), w(n), cp(n), dp(n);
  v[0] = 1.0L;
  for (int it = 0; it < 300; ++it) {
    cp[0] = e[0] / d[0];
    dp[0] = v[0] / d[0];
    for (int m = 1; m < n; ++m) {
      const long double den = d[m] - e[m - 1] * cp[m - 1];
      cp[m] = e[m] / den;
      dp[m] = (v[m] - e[m - 1] * dp[m - 1]) / den;
    }
    w[n - 1] = dp[n - 1];
    for (int m = n - 2; m >= 0; --m) w[m] = dp[m] - cp[m] * w[m + 1];
    long double nrm = 0;
    for (int m = 0; m < n; ++m) nrm += w[m] * w[m];
    nrm = std::sqrt(nrm);
    for (int m = 0; m < n; ++m) v[m] = w[m] / nrm;
  }
  a_.resize(n);
  for (int m = 0; m < n; ++m) a_[m] = v[m] * std::sqrt(2.0L * m + 0.5L);
  scale_ = 1.0L / raw(0.0L);
}

long double Prolate0::raw(long double x) const {
  // sum_m a_m P_2m(x) with the three-term recurrence of P_k
  long double p0 = 1, p1 = x, s = a_[0];
  const int kmax = 2 * int(a_.size()) - 2;
  for (int k = 1; k < kmax; ++k) {
    const long double p2 = ((2 * k + 1) * x * p1 - k * p0) / (k + 1);
    p0 = p1;
    p1 = p2;
    if ((k + 1) % 2 == 0) s += a_[(k + 1) / 2] * p1;
  }
  return s;
}

double Prolate0::operator()(double x) const {
  if (std::abs(x) > 1.0) return 0.0;
  return double(raw(x) * scale_);
}

KernelFn::KernelFn(int formula, int ns, double beta) : formula_(formula), ns_(ns), beta_(beta) {
  if (formula_ == KER_PSWF) pswf_.emplace_back(beta);
}

double KernelFn::operator()(double x) const {
  if (formula_ == KER_PSWF) {
    const double z = 2.0 * x / ns_;
    return std::abs(z) >= 1.0 ? 0.0 : pswf_[0](z);
  }
  return es_kernel(x, ns_, beta_);
}

double es_kernel(double x, int ns, double beta) {
  const double z = 2.0 * x / ns;
  if (std::abs(z) >= 1.0) return 0.0;
  return std::exp(beta * (std::sqrt(1.0 - z * z) - 1.0));
}

namespace {

// Least squares min ||A c - b|| for a tall real matrix (rows x cols, row
// major) and several right-hand sides, by Householder QR in long double.
void lstsq_householder(int rows, int cols, std::vector<long double> &A, int nrhs,
                       std::vector<long double> &B, std::vector<long double> &X) {
  auto a = [&](int i, int j) -> long double & { return A[size_t(i) * cols + j]; };
  auto b = [&](int i, int j) -> long double & { return B[size_t(i) * nrhs + j]; };
  for (int k = 0; k < cols; ++k) {
    long double norm = 0;
    for (int i = k; i < rows; ++i) norm += a(i, k) * a(i, k);
    norm = std::sqrt(norm);
    const long double alpha = a(k, k) > 0 ? -norm : norm;
    std::vector<long double> v(rows - k);
    for (int i = k; i < rows; ++i) v[i - k] = a(i, k);
    v[0] -= alpha;
    long double vn = 0;
    for (auto e : v) vn += e * e;
    if (vn == 0) continue;
    for (int j = k; j < cols; ++j) {
      long double d = 0;
      for (int i = k; i < rows; ++i) d += v[i - k] * a(i, j);
      d = 2 * d / vn;
      for (int i = k; i < rows; ++i) a(i, j) -= d * v[i - k];
    }
    for (int j = 0; j < nrhs; ++j) {
      long double d = 0;
      for (int i = k; i < rows; ++i) d += v[i - k] * b(i, j);
      d = 2 * d / vn;
      for (int i = k; i < rows; ++i) b(i, j) -= d * v[i - k];
    }
  }
  X.assign(size_t(cols) * nrhs, 0);
  for (int j = 0; j < nrhs; ++j) {
    for (int k = cols - 1; k >= 0; --k) {
      long double s = b(k, j);
      for (int m = k + 1; m < cols; ++m) s -= a(k, m) * X[size_t(m) * nrhs + j];
      X[size_t(k) * nrhs + j] = s / a(k, k);
    }
  }
}

}  // namespace

std::vector<double> horner_coeffs(int ns, double upsampfac, double beta, int nc_out) {
  // Degree rule of devel/get_degree_and_beta.m.
  int degree;
  if (upsampfac == 2.0)
    degree = ns + 1 + (ns <= 7 ? 1 : 0) - (ns == 2 ? 1 : 0);
  else
    degree = int(std::ceil(0.7 * ns + 1.3));
  const int fitd = 20, m = 7;
  const int npts = 4 * m, cols = fitd + 1;
  // Collocation points on the boundary of the box [-1,1]^2 in the complex plane.
  std::vector<std::complex<long double>> zc;
  zc.reserve(npts);
  for (int side = 0; side < 4; ++side)
    for (int j = 1; j <= m; ++j) {
      const long double l = 2.0L * (j - 0.5L) / m - 1.0L;
      switch (side) {
        case 0: zc.emplace_back(l, -1.0L); break;
        case 1: zc.emplace_back(1.0L, l); break;
        case 2: zc.emplace_back(l, 1.0L); break;
        default: zc.emplace_back(-1.0L, l); break;
      }
    }
  // The point set is closed under conjugation and the kernel is real on the
  // real axis, so the least-squares coefficients are real: fit the stacked
  // real and imaginary parts.
  std::vector<long double> A(size_t(2 * npts) * cols), B(size_t(2 * npts) * ns);
  for (int p = 0; p < npts; ++p) {
    std::complex<long double> pw = 1.0L;
    for (int k = 0; k < cols; ++k) {
      A[size_t(p) * cols + k] = pw.real();
      A[size_t(p + npts) * cols + k] = pw.imag();
      pw *= zc[p];
    }
  }
  const long double h = 1.0L / ns;
  for (int i = 0; i < ns; ++i) {
    const long double center = -1.0L + h * (2 * i + 1);
    for (int p = 0; p < npts; ++p) {
      const std::complex<long double> t = center + zc[p] * h;
      const std::complex<long double> val =
          std::exp((long double)beta * (std::sqrt(1.0L - t * t) - 1.0L));
      B[size_t(p) * ns + i] = val.real();
      B[size_t(p + npts) * ns + i] = val.imag();
    }
  }
  std::vector<long double> X;
  lstsq_householder(2 * npts, cols, A, ns, B, X);
  std::vector<double> out(size_t(nc_out) * ns, 0.0);
  for (int k = 0; k <= degree && k < nc_out; ++k)
    for (int i = 0; i < ns; ++i) out[size_t(k) * ns + i] = double(X[size_t(k) * ns + i]);
  return out;
}

std::vector<double> horner_coeffs_pswf(int ns, double beta, int nc_out) {
  const Prolate0 psi(beta);
  const int npts = 4 * nc_out, cols = nc_out;
  std::vector<long double> A(size_t(npts) * cols), B(size_t(npts) * ns);
  for (int p = 0; p < npts; ++p) {
    // Chebyshev points of the first kind on [-1, 1]
    const long double t = std::cos(PI * (2 * p + 1) / (2.0 * npts));
    long double pw = 1;
    for (int k = 0; k < cols; ++k) {
      A[size_t(p) * cols + k] = pw;
      pw *= t;
    }
    for (int i = 0; i < ns; ++i) {
      // interval i of the support [-1, 1], its local variable t
      const double z = (-1.0 + (2.0 * i + 1.0) / ns) + double(t) / ns;
      B[size_t(p) * ns + i] = psi(z);
    }
  }
  std::vector<long double> X;
  lstsq_householder(npts, cols, A, ns, B, X);
  std::vector<double> out(size_t(nc_out) * ns, 0.0);
  for (int k = 0; k < nc_out; ++k)
    for (int i = 0; i < ns; ++i) out[size_t(k) * ns + i] = double(X[size_t(k) * ns + i]);
  return out;
}

int64_t next235(int64_t n, int64_t required_factor) {
  n = std::max<int64_t>(n, 1);
  required_factor = std::max<int64_t>(required_factor, 1);
  const int64_t target = (n + required_factor - 1) / required_factor;
  // Smallest 5-smooth integer >= target, by enumerating 2^a 3^b 5^c.
  int64_t best = std::numeric_limits<int64_t>::max();
  for (int64_t p5 = 1;; p5 *= 5) {
    for (int64_t p35 = p5;; p35 *= 3) {
      int64_t v = p35;
      while (v < target) v *= 2;
      best = std::min(best, v);
      if (p35 >= target) break;
    }
    if (p5 >= target) break;
  }
  return best * required_factor;
}

void gaussquad(int n, double *x, double *w) {
  // Newton iteration on the Legendre three-term recurrence from the
  // asymptotic node guesses cos(pi (i + 3/4) / (n + 1/2)).
  for (int i = 0; i < (n + 1) / 2; ++i) {
    long double t = std::cos(PI * (i + 0.75) / (n + 0.5));
    long double dp = 1;
    for (int it = 0; it < 100; ++it) {
      long double p0 = 1, p1 = t;
      for (int k = 2; k <= n; ++k) {
        const long double p2 = ((2 * k - 1) * t * p1 - (k - 1) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      if (n == 1) p0 = 1;
      // derivative of P_n at t
      dp = n * (t * p1 - p0) / (t * t - 1);
      const long double dt = p1 / dp;
      t -= dt;
      if (std::fabs((double)dt) < 1e-18) break;
    }
    // recompute the derivative at the converged node for the weight
    long double p0 = 1, p1 = t;
    for (int k = 2; k <= n; ++k) {
      const long double p2 = ((2 * k - 1) * t * p1 - (k - 1) * p0) / k;
      p0 = p1;
      p1 = p2;
    }
    if (n == 1) p0 = 1;
    dp = n * (t * p1 - p0) / (t * t - 1);
    const double wt = double(2.0L / ((1 - t * t) * dp * dp));
    // ascending order: negative nodes first
    x[i] = -double(t);
    x[n - 1 - i] = double(t);
    w[i] = wt;
    w[n - 1 - i] = wt;
  }
  if (n % 2 == 1) x[n / 2] = 0.0;
}

int64_t fine_grid_len(int64_t ms, double upsampfac, int ns) {
  int64_t nf = int64_t(std::ceil(upsampfac * double(ms)));
  if (nf < 2 * ns) nf = 2 * ns;
  if (nf < MAX_NF) nf = next235(nf, 2);
  return nf;
}

void fourier_nodes(const KernelFn &ker, int ns, bool type3, int *q, double *z, double *f) {
  const double J2 = ns / 2.0;
  const int qq = type3 ? int(2 + 2.0 * J2) : int(2 + 3.0 * J2);
  std::vector<double> xg(2 * qq), wg(2 * qq);
  gaussquad(2 * qq, xg.data(), wg.data());
  for (int n = 0; n < qq; ++n) {
    z[n] = J2 * xg[n];  // nodes in (-ns/2, 0); the kernel is even
    f[n] = J2 * wg[n] * ker(z[n]);
  }
  *q = qq;
}

std::tuple<int64_t, double, double> nhg_type3(double sigma, double X, double S, int ns,
                                             int64_t max_nf) {
  double Xs = X, Ss = S;
  if (Xs == 0.0) {
    if (Ss == 0.0)
      Xs = Ss = 1.0;
    else
      Xs = 1.0 / Ss;
  } else {
    Ss = std::max(Ss, 1.0 / Xs);
  }
  double nfd = 2.0 * sigma * Ss * Xs / PI + (ns + 1);
  if (!std::isfinite(nfd)) nfd = 0.0;
  int64_t nf = std::max<int64_t>(int64_t(nfd), 2 * ns);
  if (nf < max_nf) nf = next235(nf, 2);
  return {nf, 2.0 * PI / double(nf), double(nf) / (2.0 * sigma * Ss)};
}

std::tuple<double, double> widcen(double lo, double hi) {
  double w = (hi - lo) / 2, c = (hi + lo) / 2;
  if (std::abs(c) < ARRAYWIDCEN_GROWFRAC * w) {
    w += std::abs(c);
    c = 0.0;
  }
  return {w, c};
}

}  // namespace fa
