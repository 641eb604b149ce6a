// Host stand-ins for the HIP device constructs that the kernel sources use,
// so that a kernel's own source compiles as ordinary C++ (clang, for the
// vector types) and runs with one OS thread per lane. A workgroup's threads
// meet at __syncthreads through one std::barrier; LDS is a heap block of
// exactly the launch's dynamic size, so AddressSanitizer checks every access
// to it as well as to the global arrays. Test infrastructure only.
#pragma once

#include <algorithm>
#include <barrier>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <thread>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __align__(x)
#define __builtin_amdgcn_readfirstlane(x) (x)

struct emu_dim3 {
  unsigned x = 0, y = 0, z = 0;
};
inline thread_local emu_dim3 threadIdx, blockIdx;
inline std::barrier<> *emu_block_barrier = nullptr;
inline unsigned char *emu_lds = nullptr;
inline void __syncthreads() { emu_block_barrier->arrive_and_wait(); }

using std::max;
using std::min;

struct float2 {
  float x, y;
};
struct double2 {
  double x, y;
};
using hipStream_t = void *;

namespace fa {
template <typename T>
struct cplx_of;
template <>
struct cplx_of<float> {
  using type = float2;
};
template <>
struct cplx_of<double> {
  using type = double2;
};
template <typename T>
using cplx_t = typename cplx_of<T>::type;
template <typename P>
using gptr = P *;
template <typename P>
inline P *to_global(P *p) {
  return p;
}
constexpr double PI = 3.14159265358979323846;
}  // namespace fa

// Runs blocks [0, nblocks) of a kernel one after another, nthreads OS threads
// each, with a fresh LDS block of lds_bytes filled with 0xff (so that a read
// before a write shows up as NaN).
inline void emu_launch(int64_t nblocks, int nthreads, size_t lds_bytes,
                       const std::function<void()> &body) {
  for (int64_t b = 0; b < nblocks; ++b) {
    std::unique_ptr<unsigned char[]> lds(new unsigned char[std::max<size_t>(lds_bytes, 1)]);
    std::memset(lds.get(), 0xff, lds_bytes);
    emu_lds = lds.get();
    std::barrier<> bar(nthreads);
    emu_block_barrier = &bar;
    std::vector<std::thread> th;
    th.reserve(nthreads);
    for (int t = 0; t < nthreads; ++t)
      th.emplace_back([&, t] {
        threadIdx.x = unsigned(t);
        blockIdx.x = unsigned(b);
        body();
      });
    for (auto &x : th) x.join();
    emu_lds = nullptr;
  }
}
