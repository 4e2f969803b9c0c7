// Backflow feature network on MFMA (gfx950) for lattices with LY == 16 sites per row and
// C == 32 channels: FP32 (v_mfma_f32_16x16x4_f32) and split-FP16 (v_mfma_f32_16x16x32_f16,
// FP32 accumulation) variants.  Split FP16 carries every operand as hi + lo FP16 parts and
// accumulates lo*hi + hi*lo + hi*hi: 22 significant bits per operand, features within ~2e-7 of
// their largest magnitude (split BF16 at the same MFMA cost gave ~6e-6 and |d log psi| up to
// 3.6e-3 at 16x16; plain BF16 ~4e-3 per feature).  The weights are packed pre-scaled by
// kLpScale = 2^6, so that their lo parts stay FP16 normals; the accumulator starts at the scaled
// bias and the layer output is multiplied by 2^-6, both exact.
//
// A lattice row of 16 sites is the M = 16 dimension of one MFMA tile: lane l of an operand
// fragment holds site y = l % 16, so the y-shift of a 3x3 tap is a DPP rotation inside each
// row of 16 lanes (periodic for free) and the x-shift is the fragment of the neighbouring
// lattice row.  One block (LX / 4 waves) evaluates one image: wave w produces lattice rows
// 4w .. 4w+3 for all 32 output channels (2 N-halves of 16), holding the residual stream h in
// the MFMA result layout (lane l, register r -> site y = 4 (l / 16) + r, channel 16 half + l % 16).
// Each layer: gelu(h) -> LDS as [x][y][c] -> operand fragments of rows 4w-1 .. 4w+4 -> 9 taps
// x 4 rows x 2 halves MFMA chains -> h += result.  The embedding (2 input channels) runs as an
// ordinary layer with zero-padded input channels.  Weights arrive pre-packed in fragment order
// (neuralhubbard_amd/mixed.py), so every lane loads 16 contiguous bytes per fragment.

#include "nh_common.h"
#include "nh_internal.h"

namespace nh {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

constexpr float kLpScale = 64.f;  // weight pre-scale of the split-FP16 packing (a power of two)

constexpr int kMfC = 32;   // channels
constexpr int kMfLY = 16;  // sites per lattice row


// DPP rotation within rows of 16 lanes: row_ror:1 (0x121) gives lane y the value of lane y - 1,
// row_ror:15 (0x12F) that of lane y + 1 (periodic in y).
template <int CTRL>
__device__ inline int rot(int v) { return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xf, 0xf, false); }
template <int DYI>
__device__ inline int rot_dy(int v) { return DYI == 0 ? rot<0x121>(v) : rot<0x12F>(v); }

// gelu in the sigmoid form x * sigmoid(2 u), u = sqrt(2/pi) (x + 0.044715 x^3): identical to
// 0.5 x (1 + tanh u); the fast form uses the hardware exp2 / reciprocal (split-FP16 path).
__device__ inline float gelu_fast(float x) {
  const float u2 = -2.0f * 0.7978845608028654f * 1.4426950408889634f * fmaf(0.044715f * x, x * x, x);
  return x * __builtin_amdgcn_rcpf(1.0f + exp2f(u2));
}

template <bool LP>
__global__ void __launch_bounds__(256) k_cnn_mfma(int LX, int n_layers, const void* __restrict__ Wp,
                                                  const float* __restrict__ Bias, const int8_t* __restrict__ s,
                                                  const int32_t* __restrict__ item_b,
                                                  const int32_t* __restrict__ item_hop, int img0, int img_end,
                                                  int ipi, float* __restrict__ Y) {
  // activation tile [LX][16][32]: split fp16 (hi tile, then lo tile; channel order natural) or
  // fp32 (channel c stored at (c % 4) * 8 + c / 4 so that a lane's 8 K-values of the 16x16x4
  // fragments are contiguous)
  extern __shared__ float lds_act[];
  const int img = img0 + blockIdx.x;
  if (img >= img_end) return;
  const int item = img / ipi;
  const int grp = (ipi == kGroup) ? (img & 7) : 0;
  const int32_t hop = item_hop ? item_hop[item] : kHopNone;
  if (hop == kHopInvalid) return;
  const int b = item_b ? item_b[item] : item;
  const int LY = kMfLY, N = LX * LY, two_n = 2 * N;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int x0 = 4 * wave;

  // layer-0 input: the 2-channel +-1 occupancy of the image, channels 2..31 zero
  {
    const int8_t* sb = s + (size_t)b * two_n;
    for (int site = t; site < N; site += blockDim.x) {
      const float o0 = (float)item_occ(sb, hop, group_map(site, grp, N, LX, LY), N);
      const float o1 = (float)item_occ(sb, hop, group_map(N + site, grp, N, LX, LY), N);
      if (LP) {  // +-1 is exact in fp16: lo part zero
        _Float16* g = reinterpret_cast<_Float16*>(lds_act) + site * kMfC;
        _Float16* gl = g + N * kMfC;
        const int4 z = make_int4(0, 0, 0, 0);
        reinterpret_cast<int4*>(g)[0] = z;
        g[0] = (_Float16)o0;
        g[1] = (_Float16)o1;
        reinterpret_cast<int4*>(g)[1] = z;
        reinterpret_cast<int4*>(g)[2] = z;
        reinterpret_cast<int4*>(g)[3] = z;
#pragma unroll
        for (int e = 0; e < 4; ++e) reinterpret_cast<int4*>(gl)[e] = z;
      } else {
        float* g = lds_act + site * kMfC;
        for (int c = 0; c < kMfC; ++c) g[c] = 0.f;
        g[0] = o0;  // channel 0 -> position 0
        g[8] = o1;  // channel 1 -> position (1 % 4) * 8 + 1 / 4 = 8
      }
    }
  }
  __syncthreads();

  f32x4 h[4][2];  // residual stream: [row][half]
  for (int l = 0; l <= n_layers; ++l) {
    f32x4 acc[4][2];
    const float* bl = Bias + l * kMfC;
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int hf = 0; hf < 2; ++hf) {
        const float bv = bl[hf * 16 + (lane & 15)] * (LP ? kLpScale : 1.f);
        acc[r][hf] = f32x4{bv, bv, bv, bv};
      }
    if (LP) {
      // operand fragments (hi, lo) of rows x0-1 .. x0+4, 8 fp16 per lane each
      f16x8 fh[6], fl[6];
#pragma unroll
      for (int u = 0; u < 6; ++u) {
        int xr = x0 - 1 + u;
        xr += xr < 0 ? LX : 0;
        xr -= xr >= LX ? LX : 0;
        const _Float16* g = reinterpret_cast<const _Float16*>(lds_act) + ((xr * LY + (lane & 15)) * kMfC + 8 * (lane >> 4));
        fh[u] = *reinterpret_cast<const f16x8*>(g);
        fl[u] = *reinterpret_cast<const f16x8*>(g + N * kMfC);
      }
      // weights: hi fragments of all layers, then lo fragments
      const f16x8* wh = reinterpret_cast<const f16x8*>(Wp) + (size_t)l * 9 * 2 * 64;
      const f16x8* wlo = wh + (size_t)(n_layers + 1) * 9 * 2 * 64;
#pragma unroll
      for (int dyi = 0; dyi < 3; ++dyi) {
        f16x8 sh[6], sl[6];
#pragma unroll
        for (int u = 0; u < 6; ++u) {
          if (dyi == 1) {
            sh[u] = fh[u];
            sl[u] = fl[u];
          } else {  // dy = -1: lane y <- y-1; dy = +1: y <- y+1
            int4 v = *reinterpret_cast<int4*>(&fh[u]);
            int4 w = *reinterpret_cast<int4*>(&fl[u]);
            if (dyi == 0) {
              v.x = rot_dy<0>(v.x); v.y = rot_dy<0>(v.y); v.z = rot_dy<0>(v.z); v.w = rot_dy<0>(v.w);
              w.x = rot_dy<0>(w.x); w.y = rot_dy<0>(w.y); w.z = rot_dy<0>(w.z); w.w = rot_dy<0>(w.w);
            } else {
              v.x = rot_dy<2>(v.x); v.y = rot_dy<2>(v.y); v.z = rot_dy<2>(v.z); v.w = rot_dy<2>(v.w);
              w.x = rot_dy<2>(w.x); w.y = rot_dy<2>(w.y); w.z = rot_dy<2>(w.z); w.w = rot_dy<2>(w.w);
            }
            sh[u] = *reinterpret_cast<f16x8*>(&v);
            sl[u] = *reinterpret_cast<f16x8*>(&w);
          }
        }
#pragma unroll
        for (int dxi = 0; dxi < 3; ++dxi) {
          const int tap = dxi * 3 + dyi;
          const f16x8 w0 = wh[(tap * 2 + 0) * 64 + lane], w1 = wh[(tap * 2 + 1) * 64 + lane];
          const f16x8 v0 = wlo[(tap * 2 + 0) * 64 + lane], v1 = wlo[(tap * 2 + 1) * 64 + lane];
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            acc[r][0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(sl[r + dxi], w0, acc[r][0], 0, 0, 0);
            acc[r][1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(sl[r + dxi], w1, acc[r][1], 0, 0, 0);
            acc[r][0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(sh[r + dxi], v0, acc[r][0], 0, 0, 0);
            acc[r][1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(sh[r + dxi], v1, acc[r][1], 0, 0, 0);
            acc[r][0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(sh[r + dxi], w0, acc[r][0], 0, 0, 0);
            acc[r][1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(sh[r + dxi], w1, acc[r][1], 0, 0, 0);
          }
        }
      }
    } else {
      const float* wl = reinterpret_cast<const float*>(Wp) + (size_t)l * 9 * 2 * 8 * 64;
#pragma unroll 1
      for (int dyi = 0; dyi < 3; ++dyi) {
        float fs[6][8];
#pragma unroll
        for (int u = 0; u < 6; ++u) {
          int xr = x0 - 1 + u;
          xr += xr < 0 ? LX : 0;
          xr -= xr >= LX ? LX : 0;
          const float4* g = reinterpret_cast<const float4*>(lds_act + (xr * LY + (lane & 15)) * kMfC + 8 * (lane >> 4));
          const float4 p0 = g[0], p1 = g[1];
          float v[8] = {p0.x, p0.y, p0.z, p0.w, p1.x, p1.y, p1.z, p1.w};
#pragma unroll
          for (int kk = 0; kk < 8; ++kk) {
            if (dyi == 1) fs[u][kk] = v[kk];
            else if (dyi == 0) fs[u][kk] = __builtin_bit_cast(float, rot_dy<0>(__builtin_bit_cast(int, v[kk])));
            else fs[u][kk] = __builtin_bit_cast(float, rot_dy<2>(__builtin_bit_cast(int, v[kk])));
          }
        }
#pragma unroll
        for (int dxi = 0; dxi < 3; ++dxi) {
          const int tap = dxi * 3 + dyi;
#pragma unroll
          for (int hf = 0; hf < 2; ++hf) {
            float wv[8];
#pragma unroll
            for (int kk = 0; kk < 8; ++kk) wv[kk] = wl[((tap * 2 + hf) * 8 + kk) * 64 + lane];
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
              for (int kk = 0; kk < 8; ++kk)
                acc[r][hf] = __builtin_amdgcn_mfma_f32_16x16x4f32(fs[r + dxi][kk], wv[kk], acc[r][hf], 0, 0, 0);
          }
        }
      }
    }
    // residual stream
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int hf = 0; hf < 2; ++hf) {
        if (LP) acc[r][hf] *= 1.f / kLpScale;
        h[r][hf] = (l == 0) ? acc[r][hf] : h[r][hf] + acc[r][hf];
      }
    if (l == n_layers) break;
    __syncthreads();  // every wave is done reading the previous activations
    // next layer input gelu(h) -> LDS
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int hf = 0; hf < 2; ++hf) {
        const int co = hf * 16 + (lane & 15);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int y = 4 * (lane >> 4) + q;
          const int site = (x0 + r) * LY + y;
          if (LP) {
            const float gv = gelu_fast(h[r][hf][q]);
            const _Float16 hi = (_Float16)gv;
            reinterpret_cast<_Float16*>(lds_act)[site * kMfC + co] = hi;
            reinterpret_cast<_Float16*>(lds_act)[N * kMfC + site * kMfC + co] = (_Float16)(gv - (float)hi);
          } else {
            lds_act[site * kMfC + (co & 3) * 8 + (co >> 2)] = gelu_tanh(h[r][hf][q]);
          }
        }
      }
    __syncthreads();
  }
  // features: y[spin = half][site][a = lane % 16] = h
  float* yo = Y + (size_t)(img - img0) * (two_n * 16);
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int hf = 0; hf < 2; ++hf)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int site = (x0 + r) * LY + 4 * (lane >> 4) + q;
        yo[((size_t)hf * N + site) * 16 + (lane & 15)] = h[r][hf][q];
      }
}

bool cnn_mfma_supported(const nh_model* m) { return m->LY == kMfLY && m->channels == kMfC && m->LX % 4 == 0 && m->LX <= 64; }

static int launch_mf(hipStream_t st, const nh_model* m, bool lp, const void* wp, const int8_t* s,
                     const int32_t* item_b, const int32_t* item_hop, int img0, int img_end, int ipi, float* Y) {
  const int nimg = img_end - img0;
  if (nimg <= 0) return NH_OK;
  const int threads = 64 * (m->LX / 4);
  const size_t lds = (size_t)m->LX * kMfLY * kMfC * 4;  // fp32, or fp16 hi + lo
  if (threads > 256) return NH_ERR_SHAPE;
  if (lp)
    hipLaunchKernelGGL(k_cnn_mfma<true>, dim3(nimg), dim3(threads), lds, st, m->LX, m->n_layers, wp, m->conv_b, s,
                       item_b, item_hop, img0, img_end, ipi, Y);
  else
    hipLaunchKernelGGL(k_cnn_mfma<false>, dim3(nimg), dim3(threads), lds, st, m->LX, m->n_layers, wp, m->conv_b, s,
                       item_b, item_hop, img0, img_end, ipi, Y);
  NH_LAUNCH_CHECK();
  return NH_OK;
}

int launch_cnn_mfma(hipStream_t st, const nh_model* m, bool lp, const int8_t* s, const int32_t* item_b,
                    const int32_t* item_hop, int img0, int img_end, int ipi, float* Y) {
  const void* wp = lp ? m->conv_w_lp : m->conv_w_mfma;
  if (!wp || !cnn_mfma_supported(m)) return NH_ERR_UNSUPPORTED;
  return launch_mf(st, m, lp, wp, s, item_b, item_hop, img0, img_end, ipi, Y);
}

}  // namespace nh
