// Backflow feature network on MFMA (gfx950) for lattices with LY == 16 or 32 sites per row and
// C == 32 channels: FP32 (v_mfma_f32_16x16x4_f32) and split-FP16 (v_mfma_f32_16x16x32_f16,
// FP32 accumulation) variants.  Split FP16 carries every operand as hi + lo FP16 parts and
// accumulates lo*hi + hi*lo + hi*hi: 22 significant bits per operand, features within ~2e-7 of
// their largest magnitude (split BF16 at the same MFMA cost gave ~6e-6 and |d log psi| up to
// 3.6e-3 at 16x16; plain BF16 ~4e-3 per feature).  The weights are packed pre-scaled by
// kLpScale = 2^6, so that their lo parts stay FP16 normals; the accumulator starts at the scaled
// bias and the layer output is multiplied by 2^-6, both exact.
//
// A segment of 16 sites of a lattice row (SEG = LY / 16 segments per row) is the M = 16
// dimension of one MFMA tile: lane l of an operand fragment holds site y = 16 seg + l % 16, so
// the y-shift of a 3x3 tap is a DPP rotation inside each row of 16 lanes, periodic for free when
// SEG = 1; with SEG = 2 the lane at the segment edge takes the rotated value of the row's other
// segment instead, which is its periodic neighbour.  The x-shift is the fragment of the
// neighbouring lattice row.  One block (LX SEG / 4 waves) evaluates one image: wave w produces
// 4 segment-rows (lattice rows RW w .. RW w + RW - 1, RW = 4 / SEG, all their segments) for all 32
// output channels (2 N-halves of 16), holding the residual stream h in the MFMA result layout
// (lane l, register r -> site y = 16 seg + 4 (l / 16) + r, channel 16 half + l % 16).  Each layer:
// gelu(h) -> LDS as [x][y][c] -> operand fragments of rows RW w - 1 .. RW w + RW -> 9 taps x 4
// segment-rows x 2 halves MFMA chains -> h += result.  The embedding (2 input channels) runs as
// an ordinary layer with zero-padded input channels.  Weights arrive pre-packed in fragment
// order (neuralhubbard_amd/mixed.py), so every lane loads 16 contiguous bytes per fragment.

#include "nh_common.h"
#include "nh_internal.h"

namespace nh {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

constexpr float kLpScale = 64.f;  // weight pre-scale of the split-FP16 packing (a power of two)

constexpr int kMfC = 32;   // channels


// DPP rotation within rows of 16 lanes: row_ror:1 (0x121) gives lane y the value of lane y - 1,
// row_ror:15 (0x12F) that of lane y + 1 (periodic in y).
template <int CTRL>
__device__ inline int rot(int v) { return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xf, 0xf, false); }
template <int DYI>
__device__ inline int rot_dy(int v) { return DYI == 0 ? rot<0x121>(v) : rot<0x12F>(v); }

// y-shift of a 16-lane segment of a row with SEG segments: the rotation inside the segment, and
// at the segment edge (lane 0 for y - 1, lane 15 for y + 1) the rotated value of the other segment.
template <int SEG, int DYI>
__device__ inline int shift_y(int mine, int other, int lane) {
  const int r = rot_dy<DYI>(mine);
  if constexpr (SEG == 1) {
    return r;
  } else {
    const int o = rot_dy<DYI>(other);
    return (lane & 15) == (DYI == 0 ? 0 : 15) ? o : r;
  }
}

// gelu in the sigmoid form x * sigmoid(2 u), u = sqrt(2/pi) (x + 0.044715 x^3): identical to
// 0.5 x (1 + tanh u); the fast form uses the hardware exp2 / reciprocal (split-FP16 path).
__device__ inline float gelu_fast(float x) {
  const float u2 = -2.0f * 0.7978845608028654f * 1.4426950408889634f * fmaf(0.044715f * x, x * x, x);
  return x * __builtin_amdgcn_rcpf(1.0f + exp2f(u2));
}

template <bool LP, int SEG>
__global__ void __launch_bounds__(SEG == 1 ? 256 : 1024) k_cnn_mfma(int LX, int n_layers, const void* __restrict__ Wp,
                                                        const float* __restrict__ Bias, const int8_t* __restrict__ s,
                                                        const int32_t* __restrict__ item_b,
                                                        const int32_t* __restrict__ item_hop, int img0, int img_end,
                                                        int ipi, float* __restrict__ Y) {
  // activation tile [LX][LY][32]: split fp16 (hi tile, then lo tile; channel order natural) or
  // fp32 (channel c stored at (c % 4) * 8 + c / 4 so that a lane's 8 K-values of the 16x16x4
  // fragments are contiguous)
  extern __shared__ float lds_act[];
  constexpr int RW = 4 / SEG;       // lattice rows per wave
  constexpr int NF = (RW + 2) * SEG;  // operand fragments: rows x0 - 1 .. x0 + RW, every segment
  const int img = img0 + blockIdx.x;
  if (img >= img_end) return;
  const int item = img / ipi;
  const int grp = (ipi == kGroup) ? (img & 7) : 0;
  const int32_t hop = item_hop ? item_hop[item] : kHopNone;
  if (hop == kHopInvalid) return;
  const int b = item_b ? item_b[item] : item;
  const int LY = 16 * SEG, N = LX * LY, two_n = 2 * N;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int x0 = RW * wave;

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

  // segment-row v = r * SEG + seg of this wave: lattice row x0 + r, sites 16 seg .. 16 seg + 15
  f32x4 h[4][2];  // residual stream: [segment-row][half]
  for (int l = 0; l <= n_layers; ++l) {
    f32x4 acc[4][2];
    const float* bl = Bias + l * kMfC;
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
      for (int hf = 0; hf < 2; ++hf) {
        const float bv = bl[hf * 16 + (lane & 15)] * (LP ? kLpScale : 1.f);
        acc[v][hf] = f32x4{bv, bv, bv, bv};
      }
    if constexpr (LP && SEG == 2) {
      // 32-site rows under 1024 threads (128 VGPRs): per lattice row of the wave (unrolled, so
      // the accumulators stay in registers) the taps run as loops, one row's two segments
      // loaded per x-tap and each segment shifted just before its MFMAs; h, acc, one fragment
      // pair and the tap's weights are all that is live
      const f16x8* wh = reinterpret_cast<const f16x8*>(Wp) + (size_t)l * 9 * 2 * 64;
      const f16x8* wlo = wh + (size_t)(n_layers + 1) * 9 * 2 * 64;
#pragma unroll
      for (int r = 0; r < RW; ++r) {
#pragma unroll 1
        for (int dxi = 0; dxi < 3; ++dxi) {
          int xr = x0 - 1 + r + dxi;
          xr += xr < 0 ? LX : 0;
          xr -= xr >= LX ? LX : 0;
          f16x8 fh[2], fl[2];
#pragma unroll
          for (int sg = 0; sg < 2; ++sg) {
            const _Float16* g = reinterpret_cast<const _Float16*>(lds_act) +
                                ((xr * LY + 16 * sg + (lane & 15)) * kMfC + 8 * (lane >> 4));
            fh[sg] = *reinterpret_cast<const f16x8*>(g);
            fl[sg] = *reinterpret_cast<const f16x8*>(g + N * kMfC);
          }
#pragma unroll 1
          for (int dyi = 0; dyi < 3; ++dyi) {
            const int tap = dxi * 3 + dyi;
            const f16x8 w0 = wh[(tap * 2 + 0) * 64 + lane], w1 = wh[(tap * 2 + 1) * 64 + lane];
            const f16x8 v0 = wlo[(tap * 2 + 0) * 64 + lane], v1 = wlo[(tap * 2 + 1) * 64 + lane];
#pragma unroll
            for (int sg = 0; sg < 2; ++sg) {
              f16x8 sh = fh[sg], sl = fl[sg];
              if (dyi != 1) {  // wave-uniform
                const int4 v = *reinterpret_cast<int4*>(&fh[sg]), vo = *reinterpret_cast<int4*>(&fh[1 - sg]);
                const int4 w = *reinterpret_cast<int4*>(&fl[sg]), wo = *reinterpret_cast<int4*>(&fl[1 - sg]);
                int4 a, c;
                if (dyi == 0) {
                  a = make_int4(shift_y<2, 0>(v.x, vo.x, lane), shift_y<2, 0>(v.y, vo.y, lane),
                                shift_y<2, 0>(v.z, vo.z, lane), shift_y<2, 0>(v.w, vo.w, lane));
                  c = make_int4(shift_y<2, 0>(w.x, wo.x, lane), shift_y<2, 0>(w.y, wo.y, lane),
                                shift_y<2, 0>(w.z, wo.z, lane), shift_y<2, 0>(w.w, wo.w, lane));
                } else {
                  a = make_int4(shift_y<2, 2>(v.x, vo.x, lane), shift_y<2, 2>(v.y, vo.y, lane),
                                shift_y<2, 2>(v.z, vo.z, lane), shift_y<2, 2>(v.w, vo.w, lane));
                  c = make_int4(shift_y<2, 2>(w.x, wo.x, lane), shift_y<2, 2>(w.y, wo.y, lane),
                                shift_y<2, 2>(w.z, wo.z, lane), shift_y<2, 2>(w.w, wo.w, lane));
                }
                sh = *reinterpret_cast<f16x8*>(&a);
                sl = *reinterpret_cast<f16x8*>(&c);
              }
              const int v = r * 2 + sg;
              acc[v][0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(sl, w0, acc[v][0], 0, 0, 0);
              acc[v][1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(sl, w1, acc[v][1], 0, 0, 0);
              acc[v][0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(sh, v0, acc[v][0], 0, 0, 0);
              acc[v][1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(sh, v1, acc[v][1], 0, 0, 0);
              acc[v][0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(sh, w0, acc[v][0], 0, 0, 0);
              acc[v][1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(sh, w1, acc[v][1], 0, 0, 0);
            }
          }
        }
      }
    } else if constexpr (LP) {
      // operand fragments (hi, lo) of rows x0-1 .. x0+RW, every segment, 8 fp16 per lane each
      f16x8 fh[NF], fl[NF];
#pragma unroll
      for (int u = 0; u < RW + 2; ++u) {
        int xr = x0 - 1 + u;
        xr += xr < 0 ? LX : 0;
        xr -= xr >= LX ? LX : 0;
#pragma unroll
        for (int sg = 0; sg < SEG; ++sg) {
          const _Float16* g = reinterpret_cast<const _Float16*>(lds_act) +
                              ((xr * LY + 16 * sg + (lane & 15)) * kMfC + 8 * (lane >> 4));
          fh[u * SEG + sg] = *reinterpret_cast<const f16x8*>(g);
          fl[u * SEG + sg] = *reinterpret_cast<const f16x8*>(g + N * kMfC);
        }
      }
      // weights: hi fragments of all layers, then lo fragments
      const f16x8* wh = reinterpret_cast<const f16x8*>(Wp) + (size_t)l * 9 * 2 * 64;
      const f16x8* wlo = wh + (size_t)(n_layers + 1) * 9 * 2 * 64;
#pragma unroll
      for (int dyi = 0; dyi < 3; ++dyi) {
        f16x8 sh[NF], sl[NF];
#pragma unroll
        for (int u = 0; u < NF; ++u) {
          if (dyi == 1) {
            sh[u] = fh[u];
            sl[u] = fl[u];
          } else {  // dy = -1: lane y <- y-1; dy = +1: y <- y+1
            const int ou = (SEG == 1) ? u : (u ^ 1);  // the other segment of the same row
            const int4 v = *reinterpret_cast<int4*>(&fh[u]), vo = *reinterpret_cast<int4*>(&fh[ou]);
            const int4 w = *reinterpret_cast<int4*>(&fl[u]), wo = *reinterpret_cast<int4*>(&fl[ou]);
            int4 a, c;
            if (dyi == 0) {
              a = make_int4(shift_y<SEG, 0>(v.x, vo.x, lane), shift_y<SEG, 0>(v.y, vo.y, lane),
                            shift_y<SEG, 0>(v.z, vo.z, lane), shift_y<SEG, 0>(v.w, vo.w, lane));
              c = make_int4(shift_y<SEG, 0>(w.x, wo.x, lane), shift_y<SEG, 0>(w.y, wo.y, lane),
                            shift_y<SEG, 0>(w.z, wo.z, lane), shift_y<SEG, 0>(w.w, wo.w, lane));
            } else {
              a = make_int4(shift_y<SEG, 2>(v.x, vo.x, lane), shift_y<SEG, 2>(v.y, vo.y, lane),
                            shift_y<SEG, 2>(v.z, vo.z, lane), shift_y<SEG, 2>(v.w, vo.w, lane));
              c = make_int4(shift_y<SEG, 2>(w.x, wo.x, lane), shift_y<SEG, 2>(w.y, wo.y, lane),
                            shift_y<SEG, 2>(w.z, wo.z, lane), shift_y<SEG, 2>(w.w, wo.w, lane));
            }
            sh[u] = *reinterpret_cast<f16x8*>(&a);
            sl[u] = *reinterpret_cast<f16x8*>(&c);
          }
        }
#pragma unroll
        for (int dxi = 0; dxi < 3; ++dxi) {
          const int tap = dxi * 3 + dyi;
          const f16x8 w0 = wh[(tap * 2 + 0) * 64 + lane], w1 = wh[(tap * 2 + 1) * 64 + lane];
          const f16x8 v0 = wlo[(tap * 2 + 0) * 64 + lane], v1 = wlo[(tap * 2 + 1) * 64 + lane];
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int f = (v / SEG + dxi) * SEG + v % SEG;  // row r + dxi, same segment
            acc[v][0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(sl[f], w0, acc[v][0], 0, 0, 0);
            acc[v][1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(sl[f], w1, acc[v][1], 0, 0, 0);
            acc[v][0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(sh[f], v0, acc[v][0], 0, 0, 0);
            acc[v][1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(sh[f], v1, acc[v][1], 0, 0, 0);
            acc[v][0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(sh[f], w0, acc[v][0], 0, 0, 0);
            acc[v][1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(sh[f], w1, acc[v][1], 0, 0, 0);
          }
        }
      }
    } else if constexpr (SEG == 1) {  // the FP32 form runs on 16-site rows only
      const float* wl = reinterpret_cast<const float*>(Wp) + (size_t)l * 9 * 2 * 8 * 64;
#pragma unroll 1
      for (int dyi = 0; dyi < 3; ++dyi) {
        float fs[NF][8];
#pragma unroll
        for (int u = 0; u < RW + 2; ++u) {
          int xr = x0 - 1 + u;
          xr += xr < 0 ? LX : 0;
          xr -= xr >= LX ? LX : 0;
          float vr[SEG][8];
#pragma unroll
          for (int sg = 0; sg < SEG; ++sg) {
            const float4* g =
                reinterpret_cast<const float4*>(lds_act + (xr * LY + 16 * sg + (lane & 15)) * kMfC + 8 * (lane >> 4));
            const float4 p0 = g[0], p1 = g[1];
            vr[sg][0] = p0.x; vr[sg][1] = p0.y; vr[sg][2] = p0.z; vr[sg][3] = p0.w;
            vr[sg][4] = p1.x; vr[sg][5] = p1.y; vr[sg][6] = p1.z; vr[sg][7] = p1.w;
          }
#pragma unroll
          for (int sg = 0; sg < SEG; ++sg)
#pragma unroll
            for (int kk = 0; kk < 8; ++kk) {
              const int mine = __builtin_bit_cast(int, vr[sg][kk]);
              const int oth = __builtin_bit_cast(int, vr[SEG - 1 - sg][kk]);
              float& o = fs[u * SEG + sg][kk];
              if (dyi == 1) o = vr[sg][kk];
              else if (dyi == 0) o = __builtin_bit_cast(float, shift_y<SEG, 0>(mine, oth, lane));
              else o = __builtin_bit_cast(float, shift_y<SEG, 2>(mine, oth, lane));
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
            for (int v = 0; v < 4; ++v) {
              const int f = (v / SEG + dxi) * SEG + v % SEG;
#pragma unroll
              for (int kk = 0; kk < 8; ++kk)
                acc[v][hf] = __builtin_amdgcn_mfma_f32_16x16x4f32(fs[f][kk], wv[kk], acc[v][hf], 0, 0, 0);
            }
          }
        }
      }
    }
    // residual stream
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
      for (int hf = 0; hf < 2; ++hf) {
        if (LP) acc[v][hf] *= 1.f / kLpScale;
        h[v][hf] = (l == 0) ? acc[v][hf] : h[v][hf] + acc[v][hf];
      }
    if (l == n_layers) break;
    __syncthreads();  // every wave is done reading the previous activations
    // next layer input gelu(h) -> LDS
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
      for (int hf = 0; hf < 2; ++hf) {
        const int co = hf * 16 + (lane & 15);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int y = 16 * (v % SEG) + 4 * (lane >> 4) + q;
          const int site = (x0 + v / SEG) * LY + y;
          if (LP) {
            const float gv = gelu_fast(h[v][hf][q]);
            const _Float16 hi = (_Float16)gv;
            reinterpret_cast<_Float16*>(lds_act)[site * kMfC + co] = hi;
            reinterpret_cast<_Float16*>(lds_act)[N * kMfC + site * kMfC + co] = (_Float16)(gv - (float)hi);
          } else {
            lds_act[site * kMfC + (co & 3) * 8 + (co >> 2)] = gelu_tanh(h[v][hf][q]);
          }
        }
      }
    __syncthreads();
  }
  // features: y[spin = half][site][a = lane % 16] = h
  float* yo = Y + (size_t)(img - img0) * (two_n * 16);
#pragma unroll
  for (int v = 0; v < 4; ++v)
#pragma unroll
    for (int hf = 0; hf < 2; ++hf)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int site = (x0 + v / SEG) * LY + 16 * (v % SEG) + 4 * (lane >> 4) + q;
        yo[((size_t)hf * N + site) * 16 + (lane & 15)] = h[v][hf][q];
      }
}

// LY = 16: LX % 4 == 0, up to 16 rows (256 threads), FP32 and split FP16; LY = 32 (split FP16
// only): LX % 2 == 0, up to 32 rows (1024 threads, 128 KB of activations in LDS).
bool cnn_mfma_supported(const nh_model* m) {
  if (m->channels != kMfC) return false;
  if (m->LY == 16) return m->LX % 4 == 0 && m->LX >= 4 && m->LX <= 16;
  if (m->LY == 32) return m->LX % 2 == 0 && m->LX >= 4 && m->LX <= 32;
  return false;
}

static int launch_mf(hipStream_t st, const nh_model* m, bool lp, const void* wp, const int8_t* s,
                     const int32_t* item_b, const int32_t* item_hop, int img0, int img_end, int ipi, float* Y) {
  const int nimg = img_end - img0;
  if (nimg <= 0) return NH_OK;
  const int seg = m->LY / 16, rw = 4 / seg;
  const int threads = 64 * (m->LX / rw);
  const size_t lds = (size_t)m->LX * m->LY * kMfC * 4;  // fp32, or fp16 hi + lo
  if (threads > (seg == 1 ? 256 : 1024) || lds > 160 * 1024) return NH_ERR_SHAPE;
#define NH_MF_LAUNCH(LPV, SEGV)                                                                              \
  hipLaunchKernelGGL((k_cnn_mfma<LPV, SEGV>), dim3(nimg), dim3(threads), lds, st, m->LX, m->n_layers, wp, m->conv_b, \
                     s, item_b, item_hop, img0, img_end, ipi, Y)
  if (seg == 1) {
    if (lp) NH_MF_LAUNCH(true, 1);
    else NH_MF_LAUNCH(false, 1);
  } else {
    if (!lp) return NH_ERR_UNSUPPORTED;  // FP32 features at LY = 32 use the VALU network
    NH_MF_LAUNCH(true, 2);
  }
#undef NH_MF_LAUNCH
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
