// Host emulation of the blocked Pfaffian pf_blocked2 (neuralhubbard_amd/csrc/nh_pf_blocked.h) and
// the blocked LU lu_blocked (nh_lu_blocked.h):
// one OS thread per GPU thread of the workgroup, __syncthreads as a barrier of all of them, and
// the wave-level builtins the routine uses (readfirstlane, the wave max, shfl_xor and the
// 16x16x4 FP32 MFMA) as exchanges through a per-wave buffer between barriers of its 64 lanes.
// The triangle and the scratch are separate heap blocks of exactly the sizes the kernels give
// them, so a build with AddressSanitizer reports any access outside them; the scratch starts as
// NaN, so a result that depends on scratch never written comes out NaN.
//
//   pf_blocked_host NT P n tri.bin   ->  prints "sign logabs" of the packed strict lower triangle
//                                        (n (n - 1) / 2 float32, row i at i (i - 1) / 2)
//   pf_blocked_host lu NT n mat.bin   ->  "sign logabs" of det of the row-major n x n float32
//                                        matrix, held as the BCS kernels hold it (leading
//                                        dimension n + 1, 16-column panels, 9-bit row keys)

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

constexpr int kWave = 64;
struct alignas(16) float4 {
  float x, y, z, w;
};
struct alignas(16) uint4 {
  uint32_t x, y, z, w;
};
struct alignas(8) float2 {
  float x, y;
};
inline float2 make_float2(float a, float b) { return {a, b}; }
inline uint32_t __float_as_uint(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return u;
}
using std::isfinite;
using std::max;
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

#define __builtin_amdgcn_readfirstlane(x) emu_readfirstlane(x)
#define __builtin_amdgcn_rcpf(x) (1.f / (x))
#define __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, cbsz, abid, blgp) emu_mfma(a, b, c)

inline int toff(int i, int j) { return ((i * (i - 1)) >> 1) + j; }

#include "nh_pf_blocked.h"

#include "nh_lu_blocked.h"

template <int NT>
static void run_lu(int n, const std::vector<float>& in, float& sg, double& la) {
  const int ld = n + 1;
  float* M = static_cast<float*>(std::malloc(sizeof(float) * ((size_t)n * ld + kWave)));
  for (size_t i = 0; i < (size_t)n * ld + kWave; ++i) M[i] = 0.f;  // padding column and dummies
  for (int r = 0; r < n; ++r) std::memcpy(M + (size_t)r * ld, in.data() + (size_t)r * n, sizeof(float) * n);
  const int nscr = lu_scratch_words(n);
  uint32_t* scr = static_cast<uint32_t*>(std::malloc(sizeof(uint32_t) * nscr));
  for (int i = 0; i < nscr; ++i) scr[i] = 0x7fc00000u;  // NaN
  std::barrier<> block(NT);
  std::vector<EmuWave> waves(NT / kWave);
  g_block = &block;
  g_waves = waves.data();
  std::vector<std::thread> th;
  for (int t = 0; t < NT; ++t)
    th.emplace_back([&, t] {
      threadIdx.x = t;
      float s;
      double l;
      lu_blocked<NT, 16, 9>(M, n, ld, scr, s, l);
      if (t == 0) {
        sg = s;
        la = l;
      }
    });
  for (auto& x : th) x.join();
  std::free(scr);
  std::free(M);
}

template <int NT, int P, int KB>
static void run(int n, const std::vector<float>& in, float& sg, double& la) {
  const size_t ntri = (size_t)n * (n - 1) / 2;
  float* tri = static_cast<float*>(std::malloc(sizeof(float) * (ntri + kWave)));
  std::memcpy(tri, in.data(), sizeof(float) * ntri);
  for (int l = 0; l < kWave; ++l) tri[ntri + l] = 0.f;  // the dummy floats after the triangle
  const int nscr = pf2_scratch_floats<P>(n);
  float* scr = static_cast<float*>(std::malloc(sizeof(float) * nscr));
  for (int i = 0; i < nscr; ++i) scr[i] = std::numeric_limits<float>::quiet_NaN();
  std::barrier<> block(NT);
  std::vector<EmuWave> waves(NT / kWave);
  g_block = &block;
  g_waves = waves.data();
  std::vector<std::thread> th;
  for (int t = 0; t < NT; ++t)
    th.emplace_back([&, t] {
      threadIdx.x = t;
      float s;
      double l;
      pf_blocked2<NT, P, KB>(tri, n, scr, s, l);
      if (t == 0) {
        sg = s;
        la = l;
      }
    });
  for (auto& x : th) x.join();
  std::free(scr);
  std::free(tri);
}

static int main_lu(int argc, char** argv) {
  const int NT = std::atoi(argv[2]), n = std::atoi(argv[3]);
  if (n < 1 || n > NT || n > 512) return 2;
  std::vector<float> in((size_t)n * n);
  FILE* f = std::fopen(argv[4], "rb");
  if (!f || std::fread(in.data(), sizeof(float), in.size(), f) != in.size()) return 2;
  std::fclose(f);
  float sg = 0.f;
  double la = 0.0;
  if (NT == 256) run_lu<256>(n, in, sg, la);
  else if (NT == 512) run_lu<512>(n, in, sg, la);
  else return 2;
  std::printf("%.1f %.17g\n", sg, la);
  return 0;
}

int main(int argc, char** argv) {
  if (argc == 5 && std::strcmp(argv[1], "lu") == 0) return main_lu(argc, argv);
  if (argc != 5) {
    std::fprintf(stderr, "usage: %s NT P n tri.bin\n", argv[0]);
    return 2;
  }
  const int NT = std::atoi(argv[1]), P = std::atoi(argv[2]), n = std::atoi(argv[3]);
  if (n < 2 || n % 2 || n > 1024) return 2;
  std::vector<float> in((size_t)n * (n - 1) / 2);
  FILE* f = std::fopen(argv[4], "rb");
  if (!f || std::fread(in.data(), sizeof(float), in.size(), f) != in.size()) return 2;
  std::fclose(f);
  float sg = 0.f;
  double la = 0.0;
  // 256 threads: the image kernels' global-triangle form (10-bit row keys); 1024 threads: the
  // standalone nh_slogpf form (8-bit row keys, n <= 256)
  if (NT == 1024 && n > 256) return 2;
  if (NT == 256 && P == 4) run<256, 4, 10>(n, in, sg, la);
  else if (NT == 256 && P == 8) run<256, 8, 10>(n, in, sg, la);
  else if (NT == 256 && P == 12) run<256, 12, 10>(n, in, sg, la);
  else if (NT == 256 && P == 16) run<256, 16, 10>(n, in, sg, la);
  else if (NT == 1024 && P == 4) run<1024, 4, 8>(n, in, sg, la);
  else if (NT == 1024 && P == 8) run<1024, 8, 8>(n, in, sg, la);
  else return 2;
  std::printf("%.1f %.17g\n", sg, la);
  return 0;
}
