// C entry points over the reference's own kernel-selection sources, for the
// oracle (test infrastructure only). Built by `make -C oracle ref` from
// src/common/pswf.cpp, src/common/kernel.cpp and src/common/utils.cpp as they
// lie in the reference tree, into oracle/_ref/libref_kernel.so; nothing of
// the reference is copied here, this file only calls it.
#include <cstdint>

#include <finufft_common/kernel.h>
#include <finufft_common/pswf.h>
#include <finufft_common/spread_opts.h>
#include <finufft_common/utils.h>

extern "C" {

// psi_0^c at n points (finufft::common::PSWF0, include/finufft_common/pswf.h).
void ref_pswf(double c, const double *z, int n, double *out) {
  const finufft::common::PSWF0 psi(c);
  for (int i = 0; i < n; ++i) out[i] = psi(z[i]);
}

// Width before clipping (theoretical_kernel_ns, src/common/kernel.cpp:79-96).
int ref_theoretical_ns(double tol, int dim, int type, double sigma) {
  finufft_spread_opts so;
  so.upsampfac = sigma;
  so.kerformula = 0;
  return finufft::kernel::theoretical_kernel_ns(tol, dim, type, so);
}

// Width after the shared clamp (clamp_kernel_ns, kernel.h).
int ref_clamp_ns(int ns, double sigma, int is_float) {
  return finufft::kernel::clamp_kernel_ns(ns, sigma, is_float ? 12 : 16, is_float != 0);
}

// Shape parameter for a kernel formula (set_kernel_shape_given_ns,
// src/common/kernel.cpp:98-150).
double ref_beta(int ns, double sigma, int kerformula) {
  finufft_spread_opts so;
  so.nspread = ns;
  so.upsampfac = sigma;
  so.kerformula = kerformula;
  finufft::kernel::set_kernel_shape_given_ns(so, 0);
  return so.beta;
}

// The reference CPU library's Horner coefficients of one kernel interval
// (poly_fit at n Chebyshev points, include/finufft_common/kernel.h:19-64),
// highest degree first, for the PSWF of shape beta, interval j of width ns.
void ref_pswf_panel_coeffs(double beta, int ns, int j, int n, double *out) {
  const finufft::common::PSWF0 psi(beta);
  const double xshift = double(2 * j + 1 - ns);
  auto f = [&](double x) { return psi((x + xshift) / double(ns)); };
  const auto c = finufft::kernel::poly_fit<double>(f, n);
  for (int k = 0; k < n; ++k) out[k] = c[k];
}

int64_t ref_next235(int64_t n, int64_t f) { return finufft::common::next235(n, f); }

void ref_gaussquad(int n, double *x, double *w) { finufft::common::gaussquad(n, x, w); }
}
