// Runs the pruned 3-D line-transform kernels (finufft_amd/csrc/fft3.hip,
// compiled from that source for the host) on one grid, with the pass
// descriptors built by the library's own f3_passes.
//   fft3_emu <f|d> nf0 nf1 nf2 m0 m1 m2 modeord sign to_modes nvec in.bin out.bin
// in.bin (doubles): the input vectors (fw when to_modes, else fk; complex
// interleaved), then the kernel coefficients of each axis (nf_d / 2 + 1).
// out.bin (doubles): the output vectors (fk when to_modes, else fw).
// Exit codes: 0 ok, 2 usage or i/o, 3 unsupported line length.
#include "hip_shim.hpp"
#include "fft3_kernels.inc"

#include <cmath>
#include <vector>

using namespace fa;

template <typename T, int N>
void cols_n(const F3Cols<T> &p, int nvec) {
  F3Cols<T> q = p;
  const int g = f3_lines<T>(N);
  q.nblk_a = int((q.A + g - 1) / g);
  emu_launch(int64_t(q.nblk_a) * q.nb, F3_THREADS, f3_lds_cols<T>(N),
             [&] { fft3_cols_kernel<T, N>(q); }, nvec);
}

template <typename T, int N>
void rows_n(const F3Rows<T> &p, int nvec) {
  const int g = f3_lines<T>(N);
  emu_launch((p.nrows + g - 1) / g, F3_THREADS, f3_lds_rows<T>(N),
             [&] { fft3_rows_kernel<T, N>(p); }, nvec);
}

#define EMU_DISPATCH(n, F, ...)                   \
  switch (n) {                                    \
    case 16: F<T, 16>(__VA_ARGS__); break;        \
    case 32: F<T, 32>(__VA_ARGS__); break;        \
    case 64: F<T, 64>(__VA_ARGS__); break;        \
    case 128: F<T, 128>(__VA_ARGS__); break;      \
    case 256: F<T, 256>(__VA_ARGS__); break;      \
    case 512: F<T, 512>(__VA_ARGS__); break;      \
    case 1024: F<T, 1024>(__VA_ARGS__); break;    \
    default: return 3;                            \
  }

template <typename T>
int run(const int nf[3], const int ms[3], int modeord, int sign, int to_modes, int nvec,
        const char *inp, const char *outp) {
  using C = cplx_t<T>;
  const int64_t nftot = int64_t(nf[0]) * nf[1] * nf[2];
  const int64_t nmodes = int64_t(ms[0]) * ms[1] * ms[2];
  const int64_t s1n = int64_t(nf[0]) * nf[1] * ms[2];
  FILE *f = std::fopen(inp, "rb");
  if (!f) return 2;
  auto get = [&](double &v) { return std::fread(&v, sizeof(double), 1, f) == 1; };
  // exact-size heap arrays: AddressSanitizer checks every kernel access
  auto *fw = static_cast<C *>(std::malloc(sizeof(C) * nftot * nvec));
  auto *fk = static_cast<C *>(std::malloc(sizeof(C) * nmodes * nvec));
  auto *s1 = static_cast<C *>(std::malloc(sizeof(C) * s1n * nvec));
  for (int64_t i = 0; i < nftot * nvec; ++i) fw[i] = {T(NAN), T(NAN)};
  for (int64_t i = 0; i < nmodes * nvec; ++i) fk[i] = {T(NAN), T(NAN)};
  for (int64_t i = 0; i < s1n * nvec; ++i) s1[i] = {T(NAN), T(NAN)};
  C *src = to_modes ? fw : fk;
  const int64_t nsrc = (to_modes ? nftot : nmodes) * nvec;
  for (int64_t i = 0; i < nsrc; ++i) {
    double re, im;
    if (!get(re) || !get(im)) return 2;
    src[i] = {T(re), T(im)};
  }
  T *ker[3], *tw[3];
  for (int d = 0; d < 3; ++d) {
    ker[d] = static_cast<T *>(std::malloc(sizeof(T) * (nf[d] / 2 + 1)));
    for (int j = 0; j <= nf[d] / 2; ++j) {
      double v;
      if (!get(v)) return 2;
      ker[d][j] = T(v);
    }
    tw[d] = static_cast<T *>(std::malloc(sizeof(T) * 2 * nf[d]));
    for (int j = 0; j < nf[d]; ++j) {
      const double a = 2.0 * PI * double(j) / double(nf[d]);
      tw[d][2 * j] = T(std::cos(a));
      tw[d][2 * j + 1] = T(std::sin(a));
    }
  }
  std::fclose(f);
  const T *kc[3] = {ker[0], ker[1], ker[2]};
  const T *tc[3] = {tw[0], tw[1], tw[2]};
  const F3Passes<T> q = f3_passes<T>(to_modes, ms, nf, modeord, fw, fk, s1, kc, tc, sign);
  if (to_modes) {
    EMU_DISPATCH(nf[2], cols_n, q.z, nvec)
    EMU_DISPATCH(nf[1], cols_n, q.y, nvec)
    EMU_DISPATCH(nf[0], rows_n, q.x, nvec)
  } else {
    EMU_DISPATCH(nf[0], rows_n, q.x, nvec)
    EMU_DISPATCH(nf[1], cols_n, q.y, nvec)
    EMU_DISPATCH(nf[2], cols_n, q.z, nvec)
  }
  FILE *o = std::fopen(outp, "wb");
  if (!o) return 2;
  const C *dst = to_modes ? fk : fw;
  const int64_t ndst = (to_modes ? nmodes : nftot) * nvec;
  for (int64_t i = 0; i < ndst; ++i) {
    const double v[2] = {double(dst[i].x), double(dst[i].y)};
    std::fwrite(v, sizeof(double), 2, o);
  }
  std::fclose(o);
  std::free(fw);
  std::free(fk);
  std::free(s1);
  for (int d = 0; d < 3; ++d) std::free(ker[d]), std::free(tw[d]);
  return 0;
}

int main(int argc, char **argv) {
  if (argc != 14) return 2;
  int a[10];
  for (int i = 0; i < 10; ++i) a[i] = std::atoi(argv[2 + i]);
  const int nf[3] = {a[0], a[1], a[2]}, ms[3] = {a[3], a[4], a[5]};
  const int modeord = a[6], sign = a[7], to_modes = a[8], nvec = a[9];
  if (argv[1][0] == 'd') return run<double>(nf, ms, modeord, sign, to_modes, nvec, argv[12], argv[13]);
  return run<float>(nf, ms, modeord, sign, to_modes, nvec, argv[12], argv[13]);
}
