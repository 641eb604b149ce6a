// Shared host-side plumbing: error codes carried by exceptions, the C-ABI
// exception barrier, HIP error checks and a small owning device buffer.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <new>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include <finufft_errors.h>

namespace fa {

// Exception carrying one of the FINUFFT_ERR_* codes. Thrown inside the library
// and turned back into the integer code at the C boundary (same contract as
// the reference's finufft::exception + safe_finufft_call,
// include/finufft_common/safe_call.h:13-81).
class nufft_error : public std::runtime_error {
 public:
  nufft_error(int code, const std::string &what) : std::runtime_error(what), code_(code) {}
  int code() const noexcept { return code_; }

 private:
  int code_;
};

[[noreturn]] inline void fail(int code, const std::string &what) {
  throw nufft_error(code, what);
}

// Runs f(); maps exceptions to codes: nufft_error -> its code, bad_alloc ->
// FINUFFT_ERR_ALLOC, anything else -> FINUFFT_ERR_UNKNOWN_EXCEPTION.
template <class F>
int safe_call(F &&f) noexcept {
  try {
    f();
    return 0;
  } catch (const nufft_error &e) {
    return e.code();
  } catch (const std::bad_alloc &) {
    return FINUFFT_ERR_ALLOC;
  } catch (...) {
    return FINUFFT_ERR_UNKNOWN_EXCEPTION;
  }
}

inline void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) {
    std::string msg = std::string("[finufft_amd] HIP failure in ") + what + ": " +
                      hipGetErrorString(e);
    fail(e == hipErrorOutOfMemory ? FINUFFT_ERR_ALLOC : FINUFFT_ERR_CUDA_FAILURE, msg);
  }
}

// Checks the last kernel launch.
inline void launch_check(const char *what) { hip_check(hipGetLastError(), what); }

// Owning device array (hipMalloc'd, freed on destruction). Resizing discards
// the contents.
template <typename E>
class dbuf {
 public:
  dbuf() = default;
  explicit dbuf(size_t n) { resize(n); }
  dbuf(const dbuf &) = delete;
  dbuf &operator=(const dbuf &) = delete;
  dbuf(dbuf &&o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
  dbuf &operator=(dbuf &&o) noexcept {
    if (this != &o) {
      release();
      p_ = o.p_; n_ = o.n_;
      o.p_ = nullptr; o.n_ = 0;
    }
    return *this;
  }
  ~dbuf() { release(); }

  void resize(size_t n) {
    if (n == n_) return;
    release();
    if (n) {
      void *q = nullptr;
      hip_check(hipMalloc(&q, n * sizeof(E)), "hipMalloc");
      p_ = static_cast<E *>(q);
    }
    n_ = n;
  }
  // Grows only; keeps the allocation when it is already large enough.
  void reserve_at_least(size_t n) {
    if (n > n_) resize(n);
  }
  void release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  E *get() const { return p_; }
  size_t size() const { return n_; }

 private:
  E *p_ = nullptr;
  size_t n_ = 0;
};

// Makes opts.gpu_device_id current for the lifetime of the object and
// restores the caller's device afterwards (the reference's DeviceSwitcher,
// include/cufinufft/cufinufft_plan_t.hpp:115-137).
class device_guard {
 public:
  explicit device_guard(int dev) {
    hip_check(hipGetDevice(&prev_), "hipGetDevice");
    if (prev_ != dev) hip_check(hipSetDevice(dev), "hipSetDevice");
    dev_ = dev;
  }
  ~device_guard() {
    if (prev_ != dev_) (void)hipSetDevice(prev_);
  }

 private:
  int prev_ = 0, dev_ = 0;
};

// Per-stage GPU timing with HIP events recorded on the plan's stream.
// Disabled by default; when enabled, every stage of execute is bracketed by a
// pair of events, and collect() waits for them and adds up the milliseconds.
enum Stage { ST_ZERO = 0, ST_SPREAD, ST_FFT, ST_DECONV, ST_INTERP, ST_T3, ST_COUNT = 8 };

class StageTimer {
 public:
  StageTimer() = default;
  StageTimer(const StageTimer &) = delete;
  StageTimer &operator=(const StageTimer &) = delete;
  ~StageTimer() {
    for (auto &r : recs_) {
      (void)hipEventDestroy(r.a);
      (void)hipEventDestroy(r.b);
    }
    for (auto e : pool_) (void)hipEventDestroy(e);
  }
  void enable(bool on) { on_ = on; }
  bool enabled() const { return on_; }
  void begin(int stage, hipStream_t s) {
    if (!on_) return;
    cur_ = Rec{stage, take(), take()};
    hip_check(hipEventRecord(cur_.a, s), "hipEventRecord");
  }
  void end(hipStream_t s) {
    if (!on_) return;
    hip_check(hipEventRecord(cur_.b, s), "hipEventRecord");
    recs_.push_back(cur_);
  }
  // Adds the elapsed time of every recorded stage into ms[ST_COUNT] and
  // recycles the events.
  void collect(double *ms) {
    for (int i = 0; i < ST_COUNT; ++i) ms[i] = 0;
    for (auto &r : recs_) {
      hip_check(hipEventSynchronize(r.b), "hipEventSynchronize");
      float t = 0;
      hip_check(hipEventElapsedTime(&t, r.a, r.b), "hipEventElapsedTime");
      ms[r.stage] += t;
      pool_.push_back(r.a);
      pool_.push_back(r.b);
    }
    recs_.clear();
  }

 private:
  struct Rec {
    int stage;
    hipEvent_t a, b;
  };
  hipEvent_t take() {
    if (!pool_.empty()) {
      hipEvent_t e = pool_.back();
      pool_.pop_back();
      return e;
    }
    hipEvent_t e;
    hip_check(hipEventCreate(&e), "hipEventCreate");
    return e;
  }
  bool on_ = false;
  Rec cur_{};
  std::vector<Rec> recs_;
  std::vector<hipEvent_t> pool_;
};

}  // namespace fa
