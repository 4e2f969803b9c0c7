// Test-only probe: checks the gfx950 MFMA operand/result lane layouts and the DPP row
// rotation direction that the feature-network kernel relies on.  One wave, every access
// inside the 16 x K / K x 16 / 16 x 16 buffers the host allocates.
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef short bf16x8 __attribute__((ext_vector_type(8)));

__device__ inline short to_bf16(float x) {  // round to nearest even
  uint32_t u = __float_as_uint(x);
  u += 0x7fff + ((u >> 16) & 1);
  return (short)(u >> 16);
}

// D[16][16] = A[16][4] B[4][16]; assumed: A lane l -> (row l%16, k l/16); B lane l -> (k l/16,
// col l%16); D lane l, reg i -> (row 4*(l/16)+i, col l%16).
__global__ void k_f32(const float* A, const float* B, float* D) {
  int l = threadIdx.x;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  acc = __builtin_amdgcn_mfma_f32_16x16x4f32(A[(l % 16) * 4 + l / 16], B[(l / 16) * 16 + l % 16], acc, 0, 0, 0);
  for (int i = 0; i < 4; ++i) D[(4 * (l / 16) + i) * 16 + l % 16] = acc[i];
}

// D[16][16] = A[16][32] B[32][16]; assumed: A lane l -> row l%16, k 8*(l/16)+j (j<8);
// B lane l -> col l%16, k 8*(l/16)+j; D as above.
__global__ void k_bf16(const float* A, const float* B, float* D) {
  int l = threadIdx.x;
  bf16x8 a, b;
  for (int j = 0; j < 8; ++j) {
    a[j] = to_bf16(A[(l % 16) * 32 + 8 * (l / 16) + j]);
    b[j] = to_bf16(B[(8 * (l / 16) + j) * 16 + l % 16]);
  }
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
  for (int i = 0; i < 4; ++i) D[(4 * (l / 16) + i) * 16 + l % 16] = acc[i];
}

// out[l] = DPP row_ror:1 of in (0x121), and row_ror:15 (0x12F)
__global__ void k_dpp(const float* in, float* out) {
  int l = threadIdx.x;
  float v = in[l];
  out[l] = __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x121, 0xf, 0xf, false));
  out[64 + l] = __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x12F, 0xf, 0xf, false));
}

extern "C" int probe_run(int which, const float* A, const float* B, float* D) {
  if (which == 0) hipLaunchKernelGGL(k_f32, dim3(1), dim3(64), 0, 0, A, B, D);
  else if (which == 1) hipLaunchKernelGGL(k_bf16, dim3(1), dim3(64), 0, 0, A, B, D);
  else hipLaunchKernelGGL(k_dpp, dim3(1), dim3(64), 0, 0, A, D);
  return hipDeviceSynchronize() == hipSuccess ? 0 : 1;
}
