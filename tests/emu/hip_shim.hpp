// Host stand-ins for the HIP device constructs that the kernel sources use,
// so that a kernel's own source compiles as ordinary C++ (clang, for the
// vector types) and runs with one OS thread per lane. A workgroup's threads
// meet at __syncthreads through one std::barrier; LDS is a heap block of
// exactly the launch's dynamic size, so AddressSanitizer checks every access
// to it as well as to the global arrays. Test infrastructure only.
#pragma once

#include <algorithm>
#include <atomic>
#include <barrier>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <thread>
#include <type_traits>
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
struct dim3 {
  unsigned x = 1, y = 1, z = 1;
  dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
inline thread_local emu_dim3 threadIdx, blockIdx, blockDim, gridDim;
// static shared arrays of a kernel: one copy, used by one block at a time
#define __shared__ static
inline std::barrier<> *emu_block_barrier = nullptr;
inline std::barrier<> *emu_wave_barriers[16] = {};
inline unsigned char *emu_lds = nullptr;
inline void __syncthreads() { emu_block_barrier->arrive_and_wait(); }

// A wave's lanes run as separate threads: the points where the device code
// relies on the wave executing together (wave barriers, cross-lane reads)
// become barriers over the wave's 64 threads. Fences have no work left to do.
inline void __builtin_amdgcn_wave_barrier() {
  emu_wave_barriers[threadIdx.x >> 6]->arrive_and_wait();
}
#define __builtin_amdgcn_fence(order, scope) ((void)0)
inline uint64_t emu_lane_slots[16][64];
template <typename V>
inline V __builtin_amdgcn_readlane(V v, int k) {
  static_assert(sizeof(V) <= 8, "readlane of a 32- or 64-bit value");
  const unsigned w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  uint64_t bits = 0;
  std::memcpy(&bits, &v, sizeof(V));
  emu_wave_barriers[w]->arrive_and_wait();
  emu_lane_slots[w][lane] = bits;
  emu_wave_barriers[w]->arrive_and_wait();
  bits = emu_lane_slots[w][k];
  emu_wave_barriers[w]->arrive_and_wait();
  V r;
  std::memcpy(&r, &bits, sizeof(V));
  return r;
}

// Lane index in the wave, and the up-shuffle of a wave (lanes below the
// offset keep their own value), as a collective over the wave's threads.
inline unsigned __lane_id() { return threadIdx.x & 63; }
template <typename V>
inline V __shfl_up(V v, unsigned o, int width = 64) {
  static_assert(sizeof(V) <= 8, "shuffle of a 32- or 64-bit value");
  (void)width;
  const unsigned w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  uint64_t bits = 0;
  std::memcpy(&bits, &v, sizeof(V));
  emu_wave_barriers[w]->arrive_and_wait();
  emu_lane_slots[w][lane] = bits;
  emu_wave_barriers[w]->arrive_and_wait();
  if (lane >= o) bits = emu_lane_slots[w][lane - o];
  emu_wave_barriers[w]->arrive_and_wait();
  V r;
  std::memcpy(&r, &bits, sizeof(V));
  return r;
}

// Ballot of a predicate over the wave (a collective), and the popcount.
inline unsigned long long __ballot(int pred) {
  const unsigned w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  emu_wave_barriers[w]->arrive_and_wait();
  emu_lane_slots[w][lane] = pred ? 1 : 0;
  emu_wave_barriers[w]->arrive_and_wait();
  unsigned long long m = 0;
  for (int k = 0; k < 64; ++k)
    if (emu_lane_slots[w][k]) m |= 1ull << k;
  emu_wave_barriers[w]->arrive_and_wait();
  return m;
}
inline int __popcll(unsigned long long v) { return __builtin_popcountll(v); }

// Atomic adds (LDS or global) as host atomics.
template <typename V>
inline V atomicAdd(V *p, V v) {
  return std::atomic_ref<V>(*p).fetch_add(v);
}
template <typename V>
inline V unsafeAtomicAdd(V *p, V v) {
  return std::atomic_ref<V>(*p).fetch_add(v);
}

using std::max;
using std::min;

struct float2 {
  float x, y;
};
struct alignas(16) float4 {
  float x, y, z, w;
};
inline float4 make_float4(float x, float y, float z, float w) { return float4{x, y, z, w}; }
inline float2 make_float2(float x, float y) { return float2{x, y}; }
inline int __float_as_int(float f) {
  int i;
  std::memcpy(&i, &f, 4);
  return i;
}
inline float __int_as_float(int i) {
  float f;
  std::memcpy(&f, &i, 4);
  return f;
}
struct double2 {
  double x, y;
};
using hipStream_t = void *;

#ifndef EMU_NO_FA_TYPES  // the sources that define these themselves set it
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
#endif

// Runs the blocks (x, y) of a grid nblocks x nblocks_y one after another on
// one team of nthreads OS threads (whole waves when nthreads >= 64). The
// dynamic LDS block is exactly lds_bytes (so ASan bounds it) and is refilled
// with 0xff before every block, so that a read before a write within a block
// shows up as NaN.
inline void emu_launch(int64_t nblocks, int nthreads, size_t lds_bytes,
                       const std::function<void()> &body, int nblocks_y = 1) {
  const int64_t total = nblocks * nblocks_y;
  if (total <= 0) return;
  std::unique_ptr<unsigned char[]> lds(new unsigned char[std::max<size_t>(lds_bytes, 1)]);
  unsigned char *base = lds.get();
  auto refill = [base, lds_bytes]() noexcept { std::memset(base, 0xff, lds_bytes); };
  refill();
  emu_lds = base;
  std::barrier<> bar(nthreads);
  emu_block_barrier = &bar;
  std::barrier<decltype(refill)> next_block(nthreads, refill);
  std::vector<std::unique_ptr<std::barrier<>>> waves;
  for (int w = 0; w * 64 < nthreads; ++w) {
    waves.emplace_back(new std::barrier<>(std::min(64, nthreads - 64 * w)));
    emu_wave_barriers[w] = waves.back().get();
  }
  std::vector<std::thread> th;
  th.reserve(nthreads);
  for (int t = 0; t < nthreads; ++t)
    th.emplace_back([&, t] {
      threadIdx.x = unsigned(t);
      blockDim.x = unsigned(nthreads);
      gridDim.x = unsigned(nblocks);
      gridDim.y = unsigned(nblocks_y);
      for (int64_t g = 0; g < total; ++g) {
        blockIdx.x = unsigned(g % nblocks);
        blockIdx.y = unsigned(g / nblocks);
        body();
        next_block.arrive_and_wait();
      }
    });
  for (auto &x : th) x.join();
  emu_lds = nullptr;
}
