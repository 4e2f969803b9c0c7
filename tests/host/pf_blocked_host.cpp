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
//                                        dimension n + 1, 16-column panels, 9-bit row keys: the
//                                        workspace form of the BCS image kernel)
//   pf_blocked_host lu8 NT n mat.bin  ->  the same with 8-bit row keys (the LDS form of the BCS
//                                        image kernel and nh_slogdet, n <= 256)
//   pf_blocked_host lazy NT n tri.bin ->  "sign logabs" of the packed strict lower triangle by
//                                        pf_lazy (nh_pf_lazy.h) on the column-major skew square
//                                        that nh_slogpf_ws's column form builds (16-step panels;
//                                        256 threads with 8-bit row keys, 512 with 10-bit ones)
//   pf_blocked_host lazy12 NT n tri.bin -> the same with 12-step panels (512 threads, 256 < n <= 352)

#include "emu_wave.h"

inline int toff(int i, int j) { return ((i * (i - 1)) >> 1) + j; }

#include "nh_pf_blocked.h"

#include "nh_lu_blocked.h"

#include "nh_pf_lazy.h"

// the column form of nh_slogpf_ws: M[j * ld + i] = A[i][j] below the diagonal, -A[j][i] above,
// zero on the diagonal and in the padding rows, then room for the first compacted complement
// (NaN, so a value read before it is written shows)
template <int NT, int KB, int P>
static void run_lazy(int n, const std::vector<float>& in, float& sg, double& la) {
  const int ld = pfl_ld(n);
  const size_t nm = pfl_matrix_floats<P>(n);
  float* M = static_cast<float*>(std::malloc(sizeof(float) * nm));
  for (size_t i = 0; i < nm; ++i) M[i] = std::numeric_limits<float>::quiet_NaN();
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < ld; ++i)
      M[(size_t)j * ld + i] = i >= n || i == j ? 0.f : (i > j ? in[toff(i, j)] : -in[toff(j, i)]);
  const int nscr = pfl_scratch_floats<NT, P>(ld);
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
      pf_lazy<NT, P, KB>(M, n, scr, s, l);
      if (t == 0) {
        sg = s;
        la = l;
      }
    });
  for (auto& x : th) x.join();
  std::free(scr);
  std::free(M);
}

template <int NT, int KB>
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
      lu_blocked<NT, 16, KB>(M, n, ld, scr, s, l);
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
  const bool k8 = std::strcmp(argv[1], "lu8") == 0;
  if (n < 1 || n > NT || n > (k8 ? 256 : 512)) return 2;
  std::vector<float> in((size_t)n * n);
  FILE* f = std::fopen(argv[4], "rb");
  if (!f || std::fread(in.data(), sizeof(float), in.size(), f) != in.size()) return 2;
  std::fclose(f);
  float sg = 0.f;
  double la = 0.0;
  if (NT == 256 && k8) run_lu<256, 8>(n, in, sg, la);
  else if (NT == 512 && k8) run_lu<512, 8>(n, in, sg, la);
  else if (NT == 256) run_lu<256, 9>(n, in, sg, la);
  else if (NT == 512) run_lu<512, 9>(n, in, sg, la);
  else return 2;
  std::printf("%.1f %.17g\n", sg, la);
  return 0;
}

static int main_lazy(char** argv) {
  const int NT = std::atoi(argv[2]), n = std::atoi(argv[3]);
  if (n < 2 || n % 2 || n > NT) return 2;
  std::vector<float> in((size_t)n * (n - 1) / 2);
  FILE* f = std::fopen(argv[4], "rb");
  if (!f || std::fread(in.data(), sizeof(float), in.size(), f) != in.size()) return 2;
  std::fclose(f);
  float sg = 0.f;
  double la = 0.0;
  const bool p12 = std::strcmp(argv[1], "lazy12") == 0;
  if (NT == 256 && !p12) run_lazy<256, 8, 16>(n, in, sg, la);
  else if (NT == 512 && !p12) run_lazy<512, 10, 16>(n, in, sg, la);
  else if (NT == 512) run_lazy<512, 10, 12>(n, in, sg, la);
  else return 2;
  std::printf("%.1f %.17g\n", sg, la);
  return 0;
}

int main(int argc, char** argv) {
  if (argc == 5 && (std::strcmp(argv[1], "lu") == 0 || std::strcmp(argv[1], "lu8") == 0)) return main_lu(argc, argv);
  if (argc == 5 && (std::strcmp(argv[1], "lazy") == 0 || std::strcmp(argv[1], "lazy12") == 0)) return main_lazy(argv);
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
  // standalone nh_slogpf form (8-bit row keys, n <= 256) and, above n = 256, the global-triangle
  // form of the 32x32 image kernel and of nh_slogpf_ws (10-bit row keys)
  if (NT == 256 && P == 4) run<256, 4, 10>(n, in, sg, la);
  else if (NT == 256 && P == 8) run<256, 8, 10>(n, in, sg, la);
  else if (NT == 256 && P == 12) run<256, 12, 10>(n, in, sg, la);
  else if (NT == 256 && P == 16) run<256, 16, 10>(n, in, sg, la);
  else if (NT == 1024 && P == 4 && n <= 256) run<1024, 4, 8>(n, in, sg, la);
  else if (NT == 1024 && P == 8 && n <= 256) run<1024, 8, 8>(n, in, sg, la);
  else if (NT == 1024 && P == 8) run<1024, 8, 10>(n, in, sg, la);
  else return 2;
  std::printf("%.1f %.17g\n", sg, la);
  return 0;
}
