// Backflow feature network on gfx950: FP32 path.
//
// One block evaluates IPB whole images (IPB * N slots, one slot = one site of one image per
// thread).  The activation of the previous layer, gelu(h), sits in LDS as [C][slots]; each
// thread keeps its own h[C] and accumulator in VGPRs and reads the 9 neighbouring slots.
// Weights are wave-uniform, so they stream through the scalar cache.  All layers run in one
// launch, so the only HBM traffic is the int8 configuration in and the features out.
//
// Image `img` of an item list is (item = img / ipi, g = img % 8 when ipi == 8): the
// configuration s[item_b[item]] with hop item_hop[item] applied, then mapped by group
// element g, x_img[m] = s_item[g(m)] (g is an involution).

#include "nh_common.h"
#include "nh_internal.h"

namespace nh {

template <int C, int MAXT>
__global__ void __launch_bounds__(MAXT) k_cnn_fp32(int N, int LX, int LY, int n_layers,
                                                   const float* __restrict__ W,
                                                   const float* __restrict__ Bias,
                                                   const int8_t* __restrict__ s, int two_n,
                                                   const int32_t* __restrict__ item_b,
                                                   const int32_t* __restrict__ item_hop,
                                                   int img0, int img_end, int ipi,
                                                   float* __restrict__ Y) {
  extern __shared__ float g_act[];  // [C][S]
  constexpr int D = C / 2;
  const int S = blockDim.x;
  const int IPB = S / N;
  const int t = threadIdx.x;
  const int li = t / N;
  const int site = t - li * N;
  const int img = img0 + blockIdx.x * IPB + li;
  bool active = img < img_end;
  const int item = img / ipi;
  const int grp = (ipi == kGroup) ? (img & 7) : 0;
  int32_t hop = kHopInvalid;
  int b = 0;
  if (active) {
    hop = item_hop ? item_hop[item] : kHopNone;
    b = item_b ? item_b[item] : item;
  }
  if (hop == kHopInvalid) active = false;
  if (__syncthreads_or(active) == 0) return;

  const int x = site / LY, y = site - (site / LY) * LY;
  int nb[9];
#pragma unroll
  for (int dx = -1; dx <= 1; ++dx)
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy) {
      int xx = x + dx, yy = y + dy;
      xx += (xx < 0) ? LX : 0;
      xx -= (xx >= LX) ? LX : 0;
      yy += (yy < 0) ? LY : 0;
      yy -= (yy >= LY) ? LY : 0;
      nb[(dx + 1) * 3 + (dy + 1)] = li * N + xx * LY + yy;
    }

  // embedding layer: the 2-channel +-1 occupancy of the image
  float xin0 = 0.f, xin1 = 0.f;
  if (active) {
    const int8_t* sb = s + (size_t)b * two_n;
    xin0 = (float)item_occ(sb, hop, group_map(site, grp, N, LX, LY), N);
    xin1 = (float)item_occ(sb, hop, group_map(N + site, grp, N, LX, LY), N);
  }
  g_act[t] = xin0;
  g_act[S + t] = xin1;
  __syncthreads();

  float h[C];
#pragma unroll
  for (int co = 0; co < C; ++co) h[co] = Bias[co];
  for (int tap = 0; tap < 9; ++tap) {
    const float v0 = g_act[nb[tap]];
    const float v1 = g_act[S + nb[tap]];
    const float* w = W + tap * 2 * C;
#pragma unroll
    for (int co = 0; co < C; ++co) h[co] = fmaf(w[co], v0, fmaf(w[C + co], v1, h[co]));
  }

  const float* Wl = W + 9 * 2 * C;
  for (int l = 1; l <= n_layers; ++l, Wl += 9 * C * C) {
    __syncthreads();
#pragma unroll
    for (int c = 0; c < C; ++c) g_act[c * S + t] = gelu_tanh(h[c]);
    __syncthreads();
    float acc[C];
    const float* bl = Bias + l * C;
#pragma unroll
    for (int co = 0; co < C; ++co) acc[co] = bl[co];
    for (int tap = 0; tap < 9; ++tap) {
      const int ns = nb[tap];
      const float* wt = Wl + tap * C * C;
#pragma unroll 2
      for (int ci = 0; ci < C; ++ci) {
        const float v = g_act[ci * S + ns];
        const float* w = wt + ci * C;
#pragma unroll
        for (int co = 0; co < C; ++co) acc[co] = fmaf(w[co], v, acc[co]);
      }
    }
#pragma unroll
    for (int co = 0; co < C; ++co) h[co] += acc[co];
  }

  if (!active) return;
  float* yo = Y + (size_t)(img - img0) * (2 * N * D);
  float4* up = reinterpret_cast<float4*>(yo + (size_t)site * D);
  float4* dn = reinterpret_cast<float4*>(yo + (size_t)(N + site) * D);
#pragma unroll
  for (int a = 0; a < D / 4; ++a) {
    up[a] = make_float4(h[4 * a], h[4 * a + 1], h[4 * a + 2], h[4 * a + 3]);
    dn[a] = make_float4(h[D + 4 * a], h[D + 4 * a + 1], h[D + 4 * a + 2], h[D + 4 * a + 3]);
  }
}

int cnn_slots(int N) { return N <= 256 ? (256 / N) * N : N; }

template <int C>
static int launch_c(hipStream_t st, const nh_model* m, const int8_t* s, const int32_t* item_b,
                    const int32_t* item_hop, int img0, int img_end, int ipi, float* Y) {
  const int N = m->LX * m->LY;
  const int S = cnn_slots(N);
  const int IPB = S / N;
  const int nimg = img_end - img0;
  if (nimg <= 0) return NH_OK;
  const int blocks = (nimg + IPB - 1) / IPB;
  const size_t lds = (size_t)C * S * sizeof(float);
  if (S <= 256) {
    hipLaunchKernelGGL((k_cnn_fp32<C, 256>), dim3(blocks), dim3(S), lds, st, N, m->LX, m->LY,
                       m->n_layers, m->conv_w, m->conv_b, s, 2 * N, item_b, item_hop, img0,
                       img_end, ipi, Y);
  } else {
    hipLaunchKernelGGL((k_cnn_fp32<C, 1024>), dim3(blocks), dim3(S), lds, st, N, m->LX, m->LY,
                       m->n_layers, m->conv_w, m->conv_b, s, 2 * N, item_b, item_hop, img0,
                       img_end, ipi, Y);
  }
  NH_LAUNCH_CHECK();
  return NH_OK;
}

int launch_cnn(hipStream_t st, const nh_model* m, const int8_t* s, const int32_t* item_b,
               const int32_t* item_hop, int img0, int img_end, int ipi, float* Y, float* scratch) {
  if (m->network == NH_NET_RESCONV || m->network == NH_NET_RESCONV_HEAD)
    return launch_resconv(st, m, s, item_b, item_hop, img0, img_end, ipi, Y, scratch);
  if (m->precision == NH_PREC_MIXED)
    return launch_cnn_mfma(st, m, true, s, item_b, item_hop, img0, img_end, ipi, Y);
  if (m->conv_w_mfma && m->LY == 16 && cnn_mfma_supported(m))
    return launch_cnn_mfma(st, m, false, s, item_b, item_hop, img0, img_end, ipi, Y);
  switch (m->channels) {
    case 8: return launch_c<8>(st, m, s, item_b, item_hop, img0, img_end, ipi, Y);
    case 16: return launch_c<16>(st, m, s, item_b, item_hop, img0, img_end, ipi, Y);
    case 32: return launch_c<32>(st, m, s, item_b, item_hop, img0, img_end, ipi, Y);
    default: return NH_ERR_SHAPE;
  }
}

}  // namespace nh

extern "C" int nh_features(void* stream, const nh_model* model, int32_t batch, const int8_t* s,
                           float* y) {
  if (!model || !s || !y || batch < 0) return NH_ERR_ARG;
  int rc = nh::check_model(model);
  if (rc != NH_OK) return rc;
  return nh::launch_cnn((hipStream_t)stream, model, s, nullptr, nullptr, 0, batch, 1, y);
}

extern "C" size_t nh_features_workspace_size(const nh_model* model, int32_t batch) {
  if (!model || batch < 0 || nh::check_model(model) != NH_OK) return 0;
  return sizeof(float) * nh::resconv_scratch_floats(model, batch) + 256;
}

extern "C" int nh_features_ws(void* stream, const nh_model* model, int32_t batch, const int8_t* s, float* y,
                              void* workspace, size_t workspace_bytes) {
  if (!model || !s || !y || batch < 0) return NH_ERR_ARG;
  int rc = nh::check_model(model);
  if (rc != NH_OK) return rc;
  const size_t need = sizeof(float) * nh::resconv_scratch_floats(model, batch);
  if (need && (!workspace || workspace_bytes < need)) return NH_ERR_WORKSPACE;
  return nh::launch_cnn((hipStream_t)stream, model, s, nullptr, nullptr, 0, batch, 1, y,
                        need ? static_cast<float*>(workspace) : nullptr);
}
