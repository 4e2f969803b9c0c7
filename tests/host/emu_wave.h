// Workgroup emulation for the host builds of the kernels' device routines (tests/host): one OS
// thread per GPU thread, __syncthreads as a barrier of all of them, and the wave-level builtins
// (readfirstlane, ballot, the wave max, shfl_xor and the 16x16x4 FP32 MFMA) as exchanges through
// a per-wave buffer between barriers of its 64 lanes.  readfirstlane checks that its argument is
// wave-uniform.
#pragma once

#include <algorithm>
#include <barrier>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

#define __host__
#define __device__
#define __forceinline__ inline

constexpr int kWave = 64;
struct alignas(16) float4 {
  float x, y, z, w;
};
struct alignas(16) int4 {
  int x, y, z, w;
};
struct alignas(16) uint4 {
  uint32_t x, y, z, w;
};
struct alignas(8) float2 {
  float x, y;
};
inline float2 make_float2(float a, float b) { return {a, b}; }
inline float4 make_float4(float a, float b, float c, float d) { return {a, b, c, d}; }
inline uint32_t __float_as_uint(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return u;
}
inline float __uint_as_float(uint32_t u) {
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}
using std::isfinite;
using std::max;
using std::min;
typedef float nh_f32x4 __attribute__((vector_size(16)));

struct EmuIdx {
  int x;
};
thread_local EmuIdx threadIdx;

struct EmuWave {
  std::barrier<> bar{kWave};
  uint64_t u[kWave];
  double d[kWave];
  float a[kWave], b[kWave];
};
static std::barrier<>* g_block;
static EmuWave* g_waves;

inline void __syncthreads() { g_block->arrive_and_wait(); }
static EmuWave& my_wave() { return g_waves[threadIdx.x / kWave]; }
static int my_lane() { return threadIdx.x % kWave; }

// readfirstlane: lane 0's value; the routine applies it to wave-uniform values only, which the
// emulation checks
template <class T>
T emu_readfirstlane(T v) {
  EmuWave& w = my_wave();
  uint64_t bits = 0;
  std::memcpy(&bits, &v, sizeof(T));
  w.u[my_lane()] = bits;
  w.bar.arrive_and_wait();
  const uint64_t r = w.u[0];
  for (int l = 0; l < kWave; ++l)
    if (w.u[l] != r) {
      std::fprintf(stderr, "readfirstlane of a value that is not wave-uniform\n");
      std::abort();
    }
  w.bar.arrive_and_wait();
  T out;
  std::memcpy(&out, &r, sizeof(T));
  return out;
}

inline uint32_t wave_max_u32(uint32_t v) {
  EmuWave& w = my_wave();
  w.u[my_lane()] = v;
  w.bar.arrive_and_wait();
  uint64_t m = 0;
  for (int l = 0; l < kWave; ++l) m = std::max(m, w.u[l]);
  w.bar.arrive_and_wait();
  return (uint32_t)m;
}

// row16_sum_f32: lane 0's row of 16 summed in the grouping of the device's DPP sequence
// (quad_perm [1,0,3,2], [2,3,0,1], row_ror:4, row_ror:8, then lane 0): ((Q0 + Q3) + (Q2 + Q1)).
inline float row16_sum_f32(float v) {
  EmuWave& w = my_wave();
  w.d[my_lane()] = v;
  w.bar.arrive_and_wait();
  float q[4];
  for (int k = 0; k < 4; ++k) {
    const float a = (float)w.d[4 * k], b = (float)w.d[4 * k + 1], c = (float)w.d[4 * k + 2],
                d = (float)w.d[4 * k + 3];
    q[k] = (a + b) + (c + d);
  }
  const float r = (q[0] + q[3]) + (q[2] + q[1]);
  w.bar.arrive_and_wait();
  return r;
}

inline double __shfl_xor(double x, int o, int width) {
  EmuWave& w = my_wave();
  const int lane = my_lane();
  w.d[lane] = x;
  w.bar.arrive_and_wait();
  const double r = w.d[(lane ^ o) % width];
  w.bar.arrive_and_wait();
  return r;
}

// v_mfma_f32_16x16x4f32: lane l holds A[l % 16][l / 16] and B[l / 16][l % 16]; its four results
// are D[4 (l / 16) + r][l % 16]
inline nh_f32x4 emu_mfma(float a, float b, nh_f32x4 c) {
  EmuWave& w = my_wave();
  const int lane = my_lane();
  w.a[lane] = a;
  w.b[lane] = b;
  w.bar.arrive_and_wait();
  const int j = lane & 15;
  for (int r = 0; r < 4; ++r) {
    const int i = 4 * (lane >> 4) + r;
    float acc = c[r];
    for (int k = 0; k < 4; ++k) acc = std::fma(w.a[16 * k + i], w.b[16 * k + j], acc);
    c[r] = acc;
  }
  w.bar.arrive_and_wait();
  return c;
}

inline uint64_t __ballot(bool p) {
  EmuWave& w = my_wave();
  w.u[my_lane()] = p ? 1u : 0u;
  w.bar.arrive_and_wait();
  uint64_t m = 0;
  for (int l = 0; l < kWave; ++l) m |= (w.u[l] ? 1ull : 0ull) << l;
  w.bar.arrive_and_wait();
  return m;
}
inline int lanes_below(uint64_t mask) {
  const int l = my_lane();
  return l == 0 ? 0 : __builtin_popcountll(mask & (~0ull >> (64 - l)));
}
inline int __popcll(uint64_t m) { return __builtin_popcountll(m); }

#define __builtin_amdgcn_readfirstlane(x) emu_readfirstlane(x)
#define __builtin_amdgcn_rcpf(x) (1.f / (x))
#define __builtin_amdgcn_sched_barrier(x) ((void)0)
#define NH_OPAQUE_S(x) asm volatile("" : "+r"(x))
#define NH_OPAQUE_V(x) asm volatile("" : "+r"(x))
#define __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, cbsz, abid, blgp) emu_mfma(a, b, c)

