// Host emulation of one workgroup of the Pfaffian image kernel's column-major form
// (image_pf_cm in neuralhubbard_amd/csrc/nh_image_pf.h: occupied-mode gather, feature staging,
// the pairing square built into the workspace, and pf_lazy): one OS thread per GPU thread,
// __syncthreads as a barrier of all of them, the wave builtins as exchanges (emu_wave.h).
// Every buffer is a separate heap block of exactly the size the library gives it (the image's
// workspace slice, the dynamic LDS of the launch, the feature rows of one image), so a build
// with AddressSanitizer reports any access outside them; the workspace and the LDS start as
// NaN, so a result that depends on memory never written comes out NaN.
//
//   image_pf_host case.bin  ->  prints "sign logabs" of the image
//   case.bin: int32 N, LX, LY, n, n_up, D, g, hop, has_F0, P; int8 s[2N]; float32 Y[2N][D]
//             (features of the image configuration, by mode), Wt[4][N][D], F0[2N][2N] if has_F0.

#include "emu_wave.h"

constexpr int kGroup = 8;

#include "nh_modes.h"
using namespace nh;

inline int toff(int i, int j) { return ((i * (i - 1)) >> 1) + j; }

#include "nh_pf_blocked.h"

#include "nh_pf_lazy.h"

#include "nh_image_pf.h"

template <class T>
static T* exact_alloc(size_t count, T fill) {
  T* p = static_cast<T*>(std::malloc(sizeof(T) * (count ? count : 1)));
  for (size_t i = 0; i < count; ++i) p[i] = fill;
  return p;
}

template <int D, int P, int NT, int KB>
static int run(const int* h, FILE* f) {
  const int N = h[0], LX = h[1], LY = h[2], n = h[3], n_up = h[4], g = h[6], hop = h[7], has_f0 = h[8];
  const int two_n = 2 * N;
  if (n <= 64 || n > NT || n % 2 || n_up % 4 || N != LX * LY || g < 0 || g > 7) return 2;
  int8_t* s = exact_alloc<int8_t>(two_n, 0);
  float* Y = exact_alloc<float>((size_t)two_n * D, 0.f);
  float* wt = exact_alloc<float>((size_t)4 * N * D, 0.f);
  float* F0 = has_f0 ? exact_alloc<float>((size_t)two_n * two_n, 0.f) : nullptr;
  bool ok = std::fread(s, 1, two_n, f) == (size_t)two_n &&
            std::fread(Y, sizeof(float), (size_t)two_n * D, f) == (size_t)two_n * D &&
            std::fread(wt, sizeof(float), (size_t)4 * N * D, f) == (size_t)4 * N * D;
  if (ok && F0) ok = std::fread(F0, sizeof(float), (size_t)two_n * two_n, f) == (size_t)two_n * two_n;
  if (!ok) return 2;
  const float qnan = std::numeric_limits<float>::quiet_NaN();
  const size_t gstride = pfl_matrix_floats<P>(n);  // as image_global_floats
  float* gmat = exact_alloc<float>(gstride, qnan);
  const size_t lds_bytes = pf_cm_lds_bytes<D, NT, P>(n, N);
  float* lds = exact_alloc<float>(lds_bytes / sizeof(float), qnan);
  int32_t* item_hop = exact_alloc<int32_t>(1, hop);
  float* img_sign = exact_alloc<float>(kGroup, qnan);
  double* img_logabs = exact_alloc<double>(kGroup, std::numeric_limits<double>::quiet_NaN());
  std::barrier<> block(NT);
  std::vector<EmuWave> waves(NT / kWave);
  g_block = &block;
  g_waves = waves.data();
  std::vector<std::thread> th;
  // image g of item 0 in a launch that starts at image g (one workgroup)
  for (int t = 0; t < NT; ++t)
    th.emplace_back([&, t] {
      threadIdx.x = t;
      image_pf_cm<D, NT, P, KB>(g, N, LX, LY, n, n_up, s, nullptr, item_hop, g, kGroup, Y, wt, F0, img_sign,
                                img_logabs, gmat, gstride, lds);
    });
  for (auto& x : th) x.join();
  std::printf("%.1f %.17g\n", img_sign[g], img_logabs[g]);
  std::free(s);
  std::free(Y);
  std::free(wt);
  std::free(F0);
  std::free(gmat);
  std::free(lds);
  std::free(item_hop);
  std::free(img_sign);
  std::free(img_logabs);
  return 0;
}

int main(int argc, char** argv) {
  if (argc != 2) {
    std::fprintf(stderr, "usage: %s case.bin\n", argv[0]);
    return 2;
  }
  FILE* f = std::fopen(argv[1], "rb");
  int h[10];
  if (!f || std::fread(h, sizeof(int), 10, f) != 10) return 2;
  const int D = h[5], P = h[9];
  // 256 threads, 8-bit row keys, as the library launches the kernel (n <= 256)
  int rc = 2;
  if (D == 16 && P == 16) rc = run<16, 16, 256, 8>(h, f);
  else if (D == 16 && P == 8) rc = run<16, 8, 256, 8>(h, f);
  else if (D == 32 && P == 16) rc = run<32, 16, 256, 8>(h, f);
  else if (D == 8 && P == 16) rc = run<8, 16, 256, 8>(h, f);
  std::fclose(f);
  return rc;
}
