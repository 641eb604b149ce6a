// Plan-time host math: kernel width / shape rules, Horner coefficient
// generation, FFT size rounding, Gauss-Legendre quadrature and the type-3
// grid rule. Pure C++ (no HIP), unit-testable on the CPU.
#pragma once

#include <cstdint>
#include <tuple>
#include <vector>

namespace fa {

constexpr double PI = 3.141592653589793238462643383279502884;
constexpr int MAX_NQUAD = 100;
constexpr double ARRAYWIDCEN_GROWFRAC = 0.1;
constexpr int64_t MAX_NF = 2147483647;  // fine grid sizes are int32 on the GPU path

// Kernel families. ES is exp(beta (sqrt(1 - z^2) - 1)), the reference GPU
// library's kernel (include/cufinufft/spreadinterp.hpp:91-134); PSWF is the
// prolate spheroidal wave function psi_0^c of bandwidth c = beta, the
// reference CPU library's default kernel (src/common/kernel.cpp:41-57,
// kerformula 8, include/finufft/makeplan.hpp:150). Both are normalised to 1
// at z = 0 and cut off at |z| = 1.
enum KernelFormula : int { KER_ES = 0, KER_PSWF = 1 };

// Kernel width, shape parameter and family for a tolerance and upsampling
// factor. eps must already be clamped to the working precision's epsilon;
// eps_div10 is eps/10 formed in the plan's precision.
//  * ES (the reference GPU plan, src/cuda/makeplan.cu:43-119): at sigma = 2
//    ns = ceil(-log10(eps/10)) (log10 taken in double of the plan-precision
//    eps/10), else ceil(-log(eps) / (pi sqrt(1 - 1/sigma))); beta = 2.30 ns
//    (2.20, 2.26, 2.38 for ns = 2, 3, 4) at sigma = 2, else
//    0.97 pi (1 - 1/(2 sigma)) ns.
//  * PSWF (the reference CPU plan, src/common/kernel.cpp:60-96 and :98-150):
//    ns = ceil(log(tolfac / eps) / (pi sqrt(1 - 1/sigma)) + 1) with
//    tolfac = 0.18 * 1.4^(dim-1) (* 1.4 for type 3); fp32 plans below
//    sigma = 1.4 keep ns <= 8 (constants.h FLOAT_MAX_NS_CC);
//    beta = pi ns (1 - 1/(2 sigma)) - 0.05.
// Both clip ns to [2, max_ns] (max_ns is 12 for fp32, 16 for fp64).
struct KernelShape {
  int ns;
  double beta;
  bool clipped;  // the width was clipped to max_ns
  int formula;   // KernelFormula
};
KernelShape kernel_shape(double eps, double eps_div10, double upsampfac, int max_ns);
KernelShape kernel_shape_pswf(double eps, double upsampfac, int dim, int type, int max_ns,
                              bool is_float);

// psi_0^c on [-1, 1], normalised to psi(0) = 1 (0 outside [-1, 1]). psi_0^c
// is the eigenfunction of lowest eigenvalue of the prolate operator
// -d/dx (1 - x^2) d/dx + c^2 x^2; in the orthonormal even Legendre basis
// that operator is a symmetric positive definite tridiagonal matrix, whose
// lowest eigenvector is found here by inverse iteration in long double
// (Bouwkamp's method; Xiao, Rokhlin and Yarvin, Inverse Problems 17, 2001).
class Prolate0 {
 public:
  explicit Prolate0(double c);
  double operator()(double x) const;

 private:
  std::vector<long double> a_;  // coefficients of sqrt(2k + 1/2) P_2k
  long double scale_ = 1;       // 1 / psi(0)
  long double raw(long double x) const;
};

// Kernel value phi(2x/ns) for |x| < ns/2 (x in grid units) of either family.
class KernelFn {
 public:
  KernelFn(int formula, int ns, double beta);
  double operator()(double x) const;

 private:
  int formula_, ns_;
  double beta_;
  std::vector<Prolate0> pswf_;  // empty for ES
};

// Horner coefficients of the kernel on the ns unit intervals of its support
// (row k = coefficient of z^k, z in [-1, 1] across each interval), nc_out
// rows. ES at sigma = 2 or 1.25 restates the collocation least-squares fit
// of devel/ker_ppval_coeff_mat.m (degree-20 fit at 7 points per side of the
// box [-1,1]^2 around each interval, then truncation to the degree rule of
// get_degree_and_beta.m), so it reproduces the reference's tables; PSWF is a
// least-squares fit of degree nc_out - 1 on 4 nc_out Chebyshev points per
// interval (the reference CPU library interpolates at nc Chebyshev points,
// include/finufft_common/kernel.h:19-64).
std::vector<double> horner_coeffs(int ns, double upsampfac, double beta, int nc_out);
std::vector<double> horner_coeffs_pswf(int ns, double beta, int nc_out);

// Direct ES kernel phi(2x/ns) for |x| < ns/2 (spreadinterp.hpp:29-39).
double es_kernel(double x, int ns, double beta);

// Smallest 2,3,5-smooth multiple of required_factor that is >= n
// (src/common/utils.cpp:84-121).
int64_t next235(int64_t n, int64_t required_factor);

// n-node Gauss-Legendre rule on [-1, 1] (src/common/utils.cpp:18-58).
void gaussquad(int n, double *x, double *w);

// Fine grid length for types 1/2: next235(max(ceil(sigma*m), 2 ns), 2)
// (src/cuda/fft.cu:166-179).
int64_t fine_grid_len(int64_t ms, double upsampfac, int ns);

// Quadrature nodes z_n and weights f_n for the kernel's Fourier transform:
// q = floor(2 + 1.5 ns) nodes for the type-1/2 series (fft.cu:185-202) or
// floor(2 + ns) for type-3 targets (fft.cu:144-157).
void fourier_nodes(const KernelFn &ker, int ns, bool type3, int *q, double *z, double *f);

// Type-3 grid: (nf, h, gamma) from half-widths X (sources) and S (targets)
// (include/finufft_common/kernel.h:165-181).
std::tuple<int64_t, double, double> nhg_type3(double sigma, double X, double S, int ns,
                                             int64_t max_nf);

// Half-width and center of an interval holding [lo, hi]; the center is used
// only if it shrinks the width by at least ARRAYWIDCEN_GROWFRAC
// (include/cufinufft/utils.hpp:121-131).
std::tuple<double, double> widcen(double lo, double hi);

}  // namespace fa
