// Runs the cost model's network kernels (finufft_amd/csrc/mlp.hip, compiled
// from that source for the host) on seeded data and writes inputs and outputs
// for the test to check against the FP64 numpy restatement.
//   mlp_emu n seed out.bin
#include "hip_shim.hpp"
#include "mlp_kernels.inc"

#include <random>

using namespace fa;

int main(int argc, char **argv) {
  if (argc != 4) return 2;
  const int n = std::atoi(argv[1]), seed = std::atoi(argv[2]);
  std::mt19937 rng(seed);
  std::normal_distribution<float> N(0.f, 0.5f);
  std::uniform_real_distribution<float> U(0.f, 1.f);
  auto vec = [](size_t k) { return static_cast<float *>(std::malloc(sizeof(float) * k)); };
  float *p = vec(NPARAM), *x = vec(size_t(n) * IN), *t = vec(size_t(n) * OUT);
  float *m = vec(size_t(n) * OUT), *y = vec(size_t(n) * OUT), *g = vec(NPARAM);
  float *loss = vec(1), *m1 = vec(NPARAM), *m2 = vec(NPARAM), *p2 = vec(NPARAM);
  for (int i = 0; i < NPARAM; ++i) p[i] = N(rng);
  for (int i = 0; i < n * IN; ++i) x[i] = 2 * N(rng);
  for (int i = 0; i < n * OUT; ++i) t[i] = N(rng), m[i] = U(rng) > 0.25f ? 1.f : 0.f;
  for (int i = 0; i < NPARAM; ++i) g[i] = 0.f, m1[i] = 0.1f * N(rng), m2[i] = 0.01f * U(rng);
  loss[0] = 0.f;
  const int nb = (n + MT - 1) / MT;
  emu_launch(nb, MT, 0, [&] { mlp_forward_kernel(p, n, x, y); });
  emu_launch(nb, MT, 0, [&] { mlp_grad_kernel(p, n, x, t, m, g, loss); });
  // one Adam step at step 3 (bias corrections as the library's launcher)
  const float b1 = 0.9f, b2 = 0.999f, lr = 1e-2f;
  const float c1 = 1.f - std::pow(b1, 3.f), c2 = 1.f - std::pow(b2, 3.f);
  std::copy(p, p + NPARAM, p2);
  float *m1_in = vec(NPARAM), *m2_in = vec(NPARAM);
  std::copy(m1, m1 + NPARAM, m1_in);
  std::copy(m2, m2 + NPARAM, m2_in);
  emu_launch((NPARAM + 255) / 256, 256, 0,
             [&] { mlp_adam_kernel(p2, g, m1, m2, lr, b1, b2, c1, c2, 1e-8f); });
  FILE *f = std::fopen(argv[3], "wb");
  if (!f) return 2;
  auto put = [&](const float *a, size_t k) {
    for (size_t i = 0; i < k; ++i) {
      const double v = a[i];
      std::fwrite(&v, sizeof(double), 1, f);
    }
  };
  put(p, NPARAM);
  put(x, size_t(n) * IN);
  put(t, size_t(n) * OUT);
  put(m, size_t(n) * OUT);
  put(y, size_t(n) * OUT);
  put(g, NPARAM);
  put(loss, 1);
  put(m1_in, NPARAM);
  put(m2_in, NPARAM);
  put(p2, NPARAM);
  std::fclose(f);
  for (float *a : {p, x, t, m, y, g, loss, m1, m2, p2, m1_in, m2_in}) std::free(a);
  return 0;
}
