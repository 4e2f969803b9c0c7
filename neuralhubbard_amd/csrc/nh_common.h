// Shared device/host helpers of the gfx950 backflow library.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "../../include/nh.h"
#include "nh_modes.h"

#define NH_HIP(call)                                  \
  do {                                                \
    hipError_t e_ = (call);                           \
    if (e_ != hipSuccess) return NH_ERR_HIP;          \
  } while (0)

#define NH_LAUNCH_CHECK() NH_HIP(hipGetLastError())

namespace nh {

constexpr int kWave = 64;
constexpr int kGroup = 8;

__device__ inline int lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0)); }

__device__ inline int lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0));
}

__device__ inline float gelu_tanh(float x) {
  const float k0 = 0.7978845608028654f;  // sqrt(2 / pi)
  float u = k0 * fmaf(0.044715f * x, x * x, x);
  return 0.5f * x * (1.0f + tanhf(u));
}

}  // namespace nh
