// Host+device math of the spreading kernel and the coordinate fold.
//
//  * fold_rescale: x in R -> [0, nf) with the reference GPU's rounding
//    (fma round-to-nearest, then subtract and multiply rounded DOWN;
//    include/cufinufft/spreadinterp.hpp:61-89). The round-down operations are
//    computed exactly from round-to-nearest results plus their error terms
//    (TwoSum / fma), so the same code gives bit-identical results on host and
//    device without touching the FP rounding mode.
//  * ES kernel phi(z) = exp(beta (sqrt(1 - z^2) - 1)), evaluated directly
//    (spreadinterp.hpp:91-111) or by per-interval Horner polynomials whose
//    coefficients are generated at plan time (horner.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>

namespace fa {

#define FA_HD __host__ __device__ inline

// ----- exact directed rounding built from round-to-nearest arithmetic -----

FA_HD float next_down(float s) {
  if (s == 0.0f) return -1.40129846e-45f;  // -denorm_min
  uint32_t b;
  memcpy(&b, &s, 4);
  b = (s > 0.0f) ? b - 1u : b + 1u;
  float r;
  memcpy(&r, &b, 4);
  return r;
}
FA_HD double next_down(double s) {
  if (s == 0.0) return -4.9406564584124654e-324;
  uint64_t b;
  memcpy(&b, &s, 8);
  b = (s > 0.0) ? b - 1u : b + 1u;
  double r;
  memcpy(&r, &b, 8);
  return r;
}

// a - b rounded toward -inf. TwoSum gives the exact error e of s = RN(a - b)
// (a - b = s + e exactly); when e < 0 the exact value lies below s. An exact
// zero difference is -0 when rounding toward -inf (IEEE 754 6.3), except
// +0 - (-0) = +0.
template <typename T>
FA_HD T sub_rd(T a, T b) {
  const T nb = -b;
  const T s = a + nb;
  if (s == T(0)) return (a == T(0) && !std::signbit(a) && std::signbit(b)) ? T(0) : T(-0.0);
  const T bb = s - a;
  const T e = (a - (s - bb)) + (nb - bb);
  return (e < T(0)) ? next_down(s) : s;
}

// a * b rounded toward -inf; fma(a, b, -p) is the exact product error.
template <typename T>
FA_HD T mul_rd(T a, T b) {
  const T p = a * b;
  const T e = std::fma(a, b, -p);
  return (e < T(0)) ? next_down(p) : p;
}

// Folds x (any real; [-pi, pi) is the nominal box) periodically and rescales
// to [0, n). Matches the reference device fold_rescale bit for bit.
template <typename T>
FA_HD T fold_rescale(T x, int n) {
  const T x2pi = T(0.159154943091895345554011992339482617);
  T r = std::fma(x, x2pi, T(0.5));
  r = sub_rd(r, std::floor(r));
  return mul_rd(r, static_cast<T>(n));
}

// Bin of a folded coordinate, clamped to [0, nbins) as the reference's
// compute_bin_index (spreadinterp.hpp:214-229).
template <typename T>
FA_HD int bin_of(T rescaled, T inv_binsize, int nbins) {
  int b = int(std::floor(rescaled * inv_binsize));
  b = b >= nbins ? b - 1 : b;
  return b < 0 ? 0 : b;
}

// ----- polynomial degree table -----

// Number of Horner coefficients for width w at sigma = 2 (degree
// w + 1 + (w<=7) - (w==2), devel/get_degree_and_beta.m:24-25) and at
// sigma = 1.25 (degree ceil(0.7 w + 1.3), :31). The sigma = 2 count is the
// larger, so device code is compiled for it and pads sigma = 1.25 with
// leading zeros (an exact no-op in Horner's rule).
constexpr int horner_nc_sigma2(int w) { return w + 2 + (w <= 7 ? 1 : 0) - (w == 2 ? 1 : 0); }
inline int horner_nc_sigma125(int w) { return int(std::ceil(0.7 * w + 1.3)) + 1; }

constexpr int MAX_NS = 16;                       // largest kernel width (fp64)
constexpr int MAX_NS_F32 = 12;                   // largest kernel width (fp32)
constexpr int MAX_NC = horner_nc_sigma2(MAX_NS);  // 18

// Kernel evaluation parameters shared by all spread/interp kernels.
template <typename T>
struct KerParams {
  T beta;         // ES shape parameter
  T es_c;         // (2/ns)^2
  int horner;     // 1: Horner polynomial, 0: direct exp(sqrt())
  int fold64;     // fp32 plans: fold the coordinate in fp64 (see start_and_kernel)
  const T *coef;  // device: coef[k*ns + i], k = 0..NC(ns)-1, ascending powers
};

// Intervals I and NS-1-I of the Horner table at z (w = z^2) from the even
// and odd parts of interval I's polynomial, then the pairs further out, down
// to I = 0 (compile-time indices throughout, so that ker stays in
// registers; the last value written is ker[NS-1], as in the direct
// evaluation, so that the two branches of eval_kernel_c end alike and no
// store is merged across them into a run-time index).
template <typename T, int NS, int I, typename CP>
__device__ __forceinline__ void horner_pairs(T *ker, T z, T w, CP coef) {
  if constexpr (I >= 0) {
    constexpr int NC = horner_nc_sigma2(NS);
    constexpr int NE = (NC + 1) / 2, NO = NC / 2;  // even / odd coefficients
    T e = coef[2 * (NE - 1) * NS + I];
#pragma unroll
    for (int j = NE - 2; j >= 0; --j) e = std::fma(e, w, coef[2 * j * NS + I]);
    T o = coef[(2 * NO - 1) * NS + I];
#pragma unroll
    for (int j = NO - 2; j >= 0; --j) o = std::fma(o, w, coef[(2 * j + 1) * NS + I]);
    ker[I] = std::fma(z, o, e);
    if constexpr (NS - 1 - I != I) ker[NS - 1 - I] = std::fma(-z, o, e);
    horner_pairs<T, NS, I - 1>(ker, z, w, coef);
  }
}

// Fills ker[i] = phi at the ns grid offsets for a point whose first grid node
// sits at xl = s - x_rescaled in [-ns/2, -ns/2 + 1]; z = 2 xl + ns - 1 is the
// same offset scaled to [-1, 1] (the Horner variable).
// coef: the Horner table (kp.coef, or the same pointer with a known address
// space).
template <typename T, int NS, typename CP>
__device__ __forceinline__ void eval_kernel_c(T *ker, T xl, T z, const KerParams<T> &kp,
                                              CP coef) {
  if (kp.horner) {
    // The kernel is even, so interval NS-1-i's polynomial is interval i's
    // reflected: p_{NS-1-i}(z) = p_i(-z) (the tables are symmetric to
    // rounding). With p_i(z) = E_i(z^2) + z O_i(z^2), one evaluation of the
    // even and odd parts gives both values of a pair: about NC FMAs per pair
    // instead of 2 (NC - 1).
    horner_pairs<T, NS, (NS - 1) / 2>(ker, z, z * z, coef);
  } else {
#pragma clang loop unroll(full)
    for (int i = 0; i < NS; ++i) {
      const T a = xl + T(i);
      const T zsq = kp.es_c * a * a;
      ker[i] = (zsq < T(1)) ? std::exp(kp.beta * (std::sqrt(T(1) - zsq) - T(1))) : T(0);
    }
  }
}

template <typename T, int NS>
__device__ __forceinline__ void eval_kernel(T *ker, T xl, T z, const KerParams<T> &kp) {
  eval_kernel_c<T, NS>(ker, xl, z, kp, kp.coef);
}

// Start node and kernel values for one coordinate: s = ceil(x - ns/2)
// (spreadinterp.hpp:154-196).
//
// fp32 plans with kp.fold64 fold the coordinate in fp64 and round only the
// in-cell offset to fp32. The fp32 fold of the reference keeps ~6e-8 of the
// period, a phase error of about 2 pi k 6e-8 at mode k (1e-5 at k = 128),
// which alone exceeds a 1e-6 tolerance on large grids; in fp64 the error is
// set by the fp32 offset instead (~1e-7 grid cells).
// First support node s = ceil(x_folded - ns/2) (may be negative) with exactly
// the arithmetic of start_and_kernel, for a runtime width. Used by the sort to
// bin points by their first node.
template <typename T>
__device__ __forceinline__ int start_index(T x, int nf, int ns, int fold64) {
  if constexpr (sizeof(T) == 4) {
    if (fold64) {
      double r = fold_rescale(double(x), nf);
      r = (r >= 0.0 && r < double(nf)) ? r : 0.0;
      return int(std::ceil(r - ns * 0.5));
    }
  }
  T r = fold_rescale(x, nf);
  r = (r >= T(0) && r < T(nf)) ? r : T(0);
  return int(std::ceil(r - T(ns * 0.5)));
}

template <typename T, int NS, typename CP>
__device__ __forceinline__ int start_and_kernel_c(T x, int nf, T *ker, const KerParams<T> &kp,
                                                  CP coef) {
  if constexpr (sizeof(T) == 4) {
    if (kp.fold64) {
      double r = fold_rescale(double(x), nf);
      r = (r >= 0.0 && r < double(nf)) ? r : 0.0;
      const int s = int(std::ceil(r - NS * 0.5));
      const double xl = double(s) - r;
      eval_kernel_c<T, NS>(ker, T(xl), T(std::fma(2.0, xl, double(NS - 1))), kp, coef);
      return s;
    }
  }
  T r = fold_rescale(x, nf);
  // A non-finite coordinate folds to NaN; pin it to the grid so that every
  // index stays in bounds (its contribution is then meaningless, as it is in
  // the reference, but harmless).
  r = (r >= T(0) && r < T(nf)) ? r : T(0);
  const int s = int(std::ceil(r - T(NS * 0.5)));
  const T xl = T(s) - r;
  eval_kernel_c<T, NS>(ker, xl, std::fma(T(2), xl, T(NS - 1)), kp, coef);
  return s;
}

template <typename T, int NS>
__device__ __forceinline__ int start_and_kernel(T x, int nf, T *ker, const KerParams<T> &kp) {
  return start_and_kernel_c<T, NS>(x, nf, ker, kp, kp.coef);
}

// Kernel values and their derivatives with respect to the offset xl: the
// Horner polynomial p(z), z = 2 xl + ns - 1, gives d/dxl = 2 p'(z); the direct
// form phi(a) = exp(beta (sqrt(1 - c a^2) - 1)), a = xl + i, gives
// phi'(a) = -phi(a) beta c a / sqrt(1 - c a^2) (0 outside the support).
template <typename T, int NS>
__device__ __forceinline__ void eval_kernel_d(T *ker, T *dker, T xl, T z,
                                              const KerParams<T> &kp) {
  if (kp.horner) {
    constexpr int NC = horner_nc_sigma2(NS);
#pragma unroll
    for (int i = 0; i < NS; ++i) {
      ker[i] = kp.coef[(NC - 1) * NS + i];
      dker[i] = T(0);
    }
#pragma unroll
    for (int k = NC - 2; k >= 0; --k) {
#pragma unroll
      for (int i = 0; i < NS; ++i) {
        dker[i] = std::fma(dker[i], z, ker[i]);
        ker[i] = std::fma(ker[i], z, kp.coef[k * NS + i]);
      }
    }
#pragma unroll
    for (int i = 0; i < NS; ++i) dker[i] *= T(2);
  } else {
#pragma unroll
    for (int i = 0; i < NS; ++i) {
      const T a = xl + T(i);
      const T zsq = kp.es_c * a * a;
      if (zsq < T(1)) {
        const T root = std::sqrt(T(1) - zsq);
        ker[i] = std::exp(kp.beta * (root - T(1)));
        dker[i] = -ker[i] * kp.beta * kp.es_c * a / root;
      } else {
        ker[i] = T(0);
        dker[i] = T(0);
      }
    }
  }
}

// start_and_kernel with the derivatives d ker / d xl (xl = s - x_rescaled).
template <typename T, int NS>
__device__ __forceinline__ int start_and_kernel_d(T x, int nf, T *ker, T *dker,
                                                  const KerParams<T> &kp) {
  if constexpr (sizeof(T) == 4) {
    if (kp.fold64) {
      double r = fold_rescale(double(x), nf);
      r = (r >= 0.0 && r < double(nf)) ? r : 0.0;
      const int s = int(std::ceil(r - NS * 0.5));
      const double xl = double(s) - r;
      eval_kernel_d<T, NS>(ker, dker, T(xl), T(std::fma(2.0, xl, double(NS - 1))), kp);
      return s;
    }
  }
  T r = fold_rescale(x, nf);
  r = (r >= T(0) && r < T(nf)) ? r : T(0);
  const int s = int(std::ceil(r - T(NS * 0.5)));
  const T xl = T(s) - r;
  eval_kernel_d<T, NS>(ker, dker, xl, std::fma(T(2), xl, T(NS - 1)), kp);
  return s;
}

}  // namespace fa
