// rocFFT wrapper (see fft.hpp).
#include "fft.hpp"

#include <rocfft/rocfft.h>

#include <mutex>
#include <string>

namespace fa {
namespace {

void rocfft_check(rocfft_status st, const char *what) {
  if (st != rocfft_status_success)
    fail(FINUFFT_ERR_CUDA_FAILURE,
         std::string("[finufft_amd] rocFFT failure in ") + what + " (status " +
             std::to_string(int(st)) + ")");
}

void rocfft_init_once() {
  static std::once_flag flag;
  std::call_once(flag, [] { rocfft_check(rocfft_setup(), "rocfft_setup"); });
}

}  // namespace

FftPlan::~FftPlan() { release(); }

void FftPlan::release() {
  for (auto &p : plans_) {
    if (p) rocfft_plan_destroy(p);
    p = nullptr;
  }
  if (info_) rocfft_execution_info_destroy(info_);
  info_ = nullptr;
  work_.release();
  dim_ = 0;
}

void FftPlan::setup(int dim, const int nf[3], int batch, bool fp64) {
  release();
  rocfft_init_once();
  dim_ = dim;
  for (int d = 0; d < 3; ++d) len_[d] = d < dim ? size_t(nf[d]) : 1;
  batch_ = batch;
  fp64_ = fp64;
  rocfft_check(rocfft_execution_info_create(&info_), "rocfft_execution_info_create");
}

rocfft_plan_t *FftPlan::plan_for(int sign) {
  rocfft_plan_t *&p = plans_[sign > 0 ? 1 : 0];
  if (p) return p;
  rocfft_check(rocfft_plan_create(&p, rocfft_placement_inplace,
                                  sign > 0 ? rocfft_transform_type_complex_inverse
                                           : rocfft_transform_type_complex_forward,
                                  fp64_ ? rocfft_precision_double : rocfft_precision_single,
                                  size_t(dim_), len_, size_t(batch_), nullptr),
               "rocfft_plan_create");
  size_t wbytes = 0;
  rocfft_check(rocfft_plan_get_work_buffer_size(p, &wbytes), "rocfft_plan_get_work_buffer_size");
  if (wbytes > work_.size()) {
    work_.resize(wbytes);
  }
  return p;
}

void FftPlan::prepare(int sign) {
  if (!ready()) fail(FINUFFT_ERR_CUDA_FAILURE, "[finufft_amd] FFT used before setup");
  plan_for(sign);
}

void FftPlan::exec(void *data, int sign, hipStream_t stream) {
  if (!ready()) fail(FINUFFT_ERR_CUDA_FAILURE, "[finufft_amd] FFT used before setup");
  rocfft_plan_t *p = plan_for(sign);
  if (work_.size())
    rocfft_check(rocfft_execution_info_set_work_buffer(info_, work_.get(), work_.size()),
                 "rocfft_execution_info_set_work_buffer");
  rocfft_check(rocfft_execution_info_set_stream(info_, stream), "rocfft_execution_info_set_stream");
  void *buf[1] = {data};
  rocfft_check(rocfft_execute(p, buf, nullptr, info_), "rocfft_execute");
}

}  // namespace fa
