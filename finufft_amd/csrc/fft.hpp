// Batched in-place complex FFT of the fine grid through rocFFT.
//
// Replaces the reference's cufftPlanMany / cufftExec{C2C,Z2Z}
// (src/cuda/makeplan.cu:364-377, include/cufinufft/types.hpp:36-43): the grid
// is x-fastest, unnormalized, and sign +1 means exp(+i k x) (rocFFT's
// "inverse", cuFFT's CUFFT_INVERSE).
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"

struct rocfft_plan_t;
struct rocfft_execution_info_t;

namespace fa {

class FftPlan {
 public:
  FftPlan() = default;
  FftPlan(const FftPlan &) = delete;
  FftPlan &operator=(const FftPlan &) = delete;
  ~FftPlan();

  // Grid nf[0] x nf[1] x nf[2] (x fastest), batch vectors of nf_tot each.
  void setup(int dim, const int nf[3], int batch, bool fp64);
  // In place on batch contiguous vectors; the plan for each sign is created
  // on first use.
  void exec(void *data, int sign, hipStream_t stream);
  // Creates the plan of one sign and its work buffer now (before a stream
  // capture, where plan creation must not happen).
  void prepare(int sign);
  bool ready() const { return dim_ > 0; }

 private:
  void release();
  rocfft_plan_t *plan_for(int sign);

  int dim_ = 0;
  size_t len_[3] = {1, 1, 1};
  int batch_ = 1;
  bool fp64_ = false;
  rocfft_plan_t *plans_[2] = {nullptr, nullptr};  // [0]: sign -1, [1]: sign +1
  rocfft_execution_info_t *info_ = nullptr;
  dbuf<char> work_;
};

}  // namespace fa
