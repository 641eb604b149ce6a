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

// Kernel width and ES beta for a tolerance and upsampling factor; the rules of
// the reference GPU plan (src/cuda/makeplan.cu:43-119). eps must already be
// clamped to the working precision's epsilon; eps_div10 is eps/10 formed in the
// plan's precision (the sigma = 2 rule takes log10 of it, in double).
// max_ns is 12 (fp32) or 16 (fp64).
struct KernelShape {
  int ns;
  double beta;
  bool clipped;  // the width was clipped to max_ns
};
KernelShape kernel_shape(double eps, double eps_div10, double upsampfac, int max_ns);

// Horner coefficients of the ES kernel on the ns unit intervals of its
// support, for sigma = 2 or 1.25 (the only two the polynomial path serves).
// Restates the collocation least-squares fit of devel/ker_ppval_coeff_mat.m
// (degree-20 fit at 7 points per side of the box [-1,1]^2 around each
// interval, then truncation to the degree rule of get_degree_and_beta.m).
// Returns nc_out x ns values, row k = coefficient of z^k for each interval,
// with rows beyond the sigma's own degree set to zero so that the table is
// always nc_out = horner_nc_sigma2(ns) rows tall.
std::vector<double> horner_coeffs(int ns, double upsampfac, double beta, int nc_out);

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
void fourier_nodes(int ns, double beta, bool type3, int *q, double *z, double *f);

// Type-3 grid: (nf, h, gamma) from half-widths X (sources) and S (targets)
// (include/finufft_common/kernel.h:165-181).
std::tuple<int64_t, double, double> nhg_type3(double sigma, double X, double S, int ns,
                                             int64_t max_nf);

// Half-width and center of an interval holding [lo, hi]; the center is used
// only if it shrinks the width by at least ARRAYWIDCEN_GROWFRAC
// (include/cufinufft/utils.hpp:121-131).
std::tuple<double, double> widcen(double lo, double hi);

}  // namespace fa
