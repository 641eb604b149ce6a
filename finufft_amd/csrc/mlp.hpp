// The cost model's network kernels (see mlp.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace fa {

void mlp_forward(const float *p, int n, const float *x, float *y, hipStream_t s);
void mlp_grad(const float *p, int n, const float *x, const float *t, const float *m, float *g,
              float *loss, hipStream_t s);
void mlp_adam(float *p, const float *g, float *m1, float *m2, int step, float lr, hipStream_t s);

}  // namespace fa
