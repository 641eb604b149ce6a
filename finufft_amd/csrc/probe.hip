// LDS and VALU throughput probes (measurement helpers for the bench).
//
// The spreading kernels keep their accumulators in LDS; which LDS operations
// they use (float or double atomic adds, plain read-modify-write, uniform
// broadcast reads) is decided by the throughputs measured here, with the
// library's own compile flags. Each probe launches many workgroups of 256
// lanes; every lane runs `reps` iterations of 8 operations of one kind at
// addresses that are distinct within a wave (or uniform, for the broadcast
// probes), so bank conflicts do not enter. The plain atomic probes revisit
// each lane's 8 addresses every iteration; the rotating ones spread a lane's
// updates over 64 (fp32) or 32 (fp64) addresses. The host side turns the elapsed
// time into wave instructions per clock per CU.
#include <hip/hip_runtime.h>

#include "common.hpp"

namespace fa {
namespace {

constexpr int PROBE_T = 256;
constexpr int PROBE_WORDS = 16384;  // 64 KiB of LDS

enum ProbeMode {
  P_ADD_F32 = 0,      // ds_add_f32, lane-distinct dwords
  P_ADD_F64 = 1,      // ds_add_f64, lane-distinct qwords
  P_ADD_U32 = 2,      // ds_add_u32, lane-distinct dwords
  P_RMW_F32 = 3,      // ds_read_b32 + v_add + ds_write_b32 (wave-private cells)
  P_READ_B32 = 4,     // ds_read_b32, lane-distinct
  P_READ_B32_U = 5,   // ds_read_b32, one address per wave (broadcast)
  P_READ_B128 = 6,    // ds_read_b128, lane-distinct
  P_READ_B128_U = 7,  // ds_read_b128, one address per wave (broadcast)
  P_FMA_F32 = 8,      // v_fma_f32 (VALU issue rate)
  P_ADD_F32_ROT = 9,  // ds_add_f32, each lane cycling through 64 addresses
  P_ADD_F64_ROT = 10, // ds_add_f64, each lane cycling through 32 addresses
  // ds_read_b64 with lane l of a wave at qword l * S (+ a wave-uniform
  // shift): bank conflicts appear at the strides that map several lanes of
  // one issue group to one bank
  P_READ_B64_S1 = 11,
  P_READ_B64_S2 = 12,
  P_READ_B64_S16 = 13,
  P_READ_B64_S32 = 14,
  P_READ_B64_S64 = 15,
  // ds_add_f32 as P_ADD_F32, with the wave's fp32 denormals flushed (MODE
  // register FP_DENORM bits [5:4] cleared at kernel start; the library's
  // kernels run with fp32 denormals kept)
  P_ADD_F32_FTZ = 16,
  P_COUNT = 17
};

// s_setreg operand: hardware register MODE (1), bit offset 4, 2 bits.
constexpr int HWREG_MODE_FP32_DENORM = ((2 - 1) << 11) | (4 << 6) | 1;

constexpr int b64_stride(int mode) {
  return mode == P_READ_B64_S1    ? 1
         : mode == P_READ_B64_S2  ? 2
         : mode == P_READ_B64_S16 ? 16
         : mode == P_READ_B64_S32 ? 32
                                  : 64;
}

template <int MODE>
__global__ void __launch_bounds__(PROBE_T) lds_probe_kernel(float *out, int reps) {
  __shared__ __align__(16) float t[PROBE_WORDS];
  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  for (int i = tid; i < PROBE_WORDS; i += PROBE_T) t[i] = float(i & 7);
  __syncthreads();
  float acc = 0.f;
  if constexpr (MODE == P_ADD_F32_FTZ) {
    __builtin_amdgcn_s_setreg(HWREG_MODE_FP32_DENORM, 0);
    for (int r = 0; r < reps; ++r)
#pragma unroll
      for (int k = 0; k < 8; ++k) atomicAdd(&t[(tid + k * PROBE_T) & (PROBE_WORDS - 1)], 1.0f);
  } else if constexpr (MODE == P_ADD_F32) {
    for (int r = 0; r < reps; ++r)
#pragma unroll
      for (int k = 0; k < 8; ++k) atomicAdd(&t[(tid + k * PROBE_T) & (PROBE_WORDS - 1)], 1.0f);
  } else if constexpr (MODE == P_ADD_F32_ROT) {
    for (int r = 0; r < reps; ++r)
#pragma unroll
      for (int k = 0; k < 8; ++k)
        atomicAdd(&t[(tid + k * PROBE_T + (r & 7) * 8 * PROBE_T) & (PROBE_WORDS - 1)], 1.0f);
  } else if constexpr (MODE == P_ADD_F64_ROT) {
    double *d = reinterpret_cast<double *>(t);
    for (int r = 0; r < reps; ++r)
#pragma unroll
      for (int k = 0; k < 8; ++k)
        atomicAdd(&d[(tid + k * PROBE_T + (r & 3) * 8 * PROBE_T) & (PROBE_WORDS / 2 - 1)], 1.0);
  } else if constexpr (MODE == P_ADD_F64) {
    double *d = reinterpret_cast<double *>(t);
    for (int r = 0; r < reps; ++r)
#pragma unroll
      for (int k = 0; k < 8; ++k)
        atomicAdd(&d[(tid + k * PROBE_T) & (PROBE_WORDS / 2 - 1)], 1.0);
  } else if constexpr (MODE == P_ADD_U32) {
    unsigned *u = reinterpret_cast<unsigned *>(t);
    for (int r = 0; r < reps; ++r)
#pragma unroll
      for (int k = 0; k < 8; ++k) atomicAdd(&u[(tid + k * PROBE_T) & (PROBE_WORDS - 1)], 1u);
  } else if constexpr (MODE == P_RMW_F32) {
    for (int r = 0; r < reps; ++r)
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        float *p = &t[(tid + k * PROBE_T) & (PROBE_WORDS - 1)];
        *p = *p + 1.0f;
      }
  } else if constexpr (MODE == P_READ_B32 || MODE == P_READ_B32_U) {
    for (int r = 0; r < reps; ++r)
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int a = MODE == P_READ_B32 ? tid + k * PROBE_T : wave * 64 + k * 8 + (r & 7);
        acc += t[(a + r) & (PROBE_WORDS - 1)];
      }
  } else if constexpr (MODE >= P_READ_B64_S1 && MODE <= P_READ_B64_S64) {
    // PROBE_WORDS / 2 qwords; lane * 64 + k + r stays below it after the mask
    const double *v = reinterpret_cast<const double *>(t);
    const int lane = tid & 63;
    for (int r = 0; r < reps; ++r)
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const double q = v[(lane * b64_stride(MODE) + k + (r & 7) + wave * 8) &
                           (PROBE_WORDS / 2 - 1)];
        acc += float(q);
      }
  } else if constexpr (MODE == P_READ_B128 || MODE == P_READ_B128_U) {
    const float4 *v = reinterpret_cast<const float4 *>(t);
    for (int r = 0; r < reps; ++r)
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int a = MODE == P_READ_B128 ? tid + k * PROBE_T : wave * 64 + k * 8 + (r & 7);
        const float4 q = v[(a + r) & (PROBE_WORDS / 4 - 1)];
        acc += (q.x + q.y) + (q.z + q.w);
      }
  } else {
    float x = float(tid), y = 1.0001f;
    for (int r = 0; r < reps; ++r)
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        x = std::fma(x, y, 0.5f);
        y = std::fma(y, x, -0.25f);
      }
    acc = x + y;
  }
  __syncthreads();
  acc += t[tid];
  out[blockIdx.x * PROBE_T + tid] = acc;
}

template <int MODE>
void launch_probe(float *out, int blocks, int reps, hipStream_t s) {
  lds_probe_kernel<MODE><<<blocks, PROBE_T, 0, s>>>(out, reps);
  launch_check("lds_probe_kernel");
}

}  // namespace

// Runs probe `mode` over `blocks` workgroups of 256 lanes with `reps`
// iterations of 8 operations each; returns the elapsed milliseconds and the
// number of wave-level operations issued (the FMA probe counts two per
// iteration step).
void lds_probe(int mode, int blocks, int reps, double *ms, double *wave_ops) {
  if (mode < 0 || mode >= P_COUNT) fail(FINUFFT_ERR_INVALID_ARGUMENT, "lds_probe: bad mode");
  if (blocks < 1 || reps < 1) fail(FINUFFT_ERR_INVALID_ARGUMENT, "lds_probe: bad size");
  dbuf<float> out(size_t(blocks) * PROBE_T);
  hipEvent_t a, b;
  hip_check(hipEventCreate(&a), "hipEventCreate");
  hip_check(hipEventCreate(&b), "hipEventCreate");
  auto run = [&](int r) {
    switch (mode) {
      case P_ADD_F32: return launch_probe<P_ADD_F32>(out.get(), blocks, r, nullptr);
      case P_ADD_F64: return launch_probe<P_ADD_F64>(out.get(), blocks, r, nullptr);
      case P_ADD_U32: return launch_probe<P_ADD_U32>(out.get(), blocks, r, nullptr);
      case P_RMW_F32: return launch_probe<P_RMW_F32>(out.get(), blocks, r, nullptr);
      case P_READ_B32: return launch_probe<P_READ_B32>(out.get(), blocks, r, nullptr);
      case P_READ_B32_U: return launch_probe<P_READ_B32_U>(out.get(), blocks, r, nullptr);
      case P_READ_B128: return launch_probe<P_READ_B128>(out.get(), blocks, r, nullptr);
      case P_READ_B128_U: return launch_probe<P_READ_B128_U>(out.get(), blocks, r, nullptr);
      case P_ADD_F32_ROT: return launch_probe<P_ADD_F32_ROT>(out.get(), blocks, r, nullptr);
      case P_ADD_F64_ROT: return launch_probe<P_ADD_F64_ROT>(out.get(), blocks, r, nullptr);
      case P_READ_B64_S1: return launch_probe<P_READ_B64_S1>(out.get(), blocks, r, nullptr);
      case P_READ_B64_S2: return launch_probe<P_READ_B64_S2>(out.get(), blocks, r, nullptr);
      case P_READ_B64_S16: return launch_probe<P_READ_B64_S16>(out.get(), blocks, r, nullptr);
      case P_READ_B64_S32: return launch_probe<P_READ_B64_S32>(out.get(), blocks, r, nullptr);
      case P_READ_B64_S64: return launch_probe<P_READ_B64_S64>(out.get(), blocks, r, nullptr);
      case P_ADD_F32_FTZ: return launch_probe<P_ADD_F32_FTZ>(out.get(), blocks, r, nullptr);
      default: return launch_probe<P_FMA_F32>(out.get(), blocks, r, nullptr);
    }
  };
  run(1);  // warm-up
  // the loop-free part (LDS fill, final store) is timed with reps = 1 and
  // subtracted
  float t1 = 0, tr = 0;
  hip_check(hipEventRecord(a, nullptr), "hipEventRecord");
  run(1);
  hip_check(hipEventRecord(b, nullptr), "hipEventRecord");
  hip_check(hipEventSynchronize(b), "hipEventSynchronize");
  hip_check(hipEventElapsedTime(&t1, a, b), "hipEventElapsedTime");
  hip_check(hipEventRecord(a, nullptr), "hipEventRecord");
  run(reps);
  hip_check(hipEventRecord(b, nullptr), "hipEventRecord");
  hip_check(hipEventSynchronize(b), "hipEventSynchronize");
  hip_check(hipEventElapsedTime(&tr, a, b), "hipEventElapsedTime");
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  *ms = double(tr) - double(t1);
  const double per_it = mode == P_FMA_F32 ? 16.0 : 8.0;
  *wave_ops = double(blocks) * (PROBE_T / 64) * double(reps - 1) * per_it;
}

}  // namespace fa
