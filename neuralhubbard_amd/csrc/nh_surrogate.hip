// Surrogate log-amplitude network and its training kernels (gfx950, FP32).
//
//   z1 = conv3x3(x0) + b1          x0: 2-channel +-1 occupancy, periodic 3x3
//   a1 = gelu(z1)                   Cs channels
//   z2 = conv3x3(a1) + b2, a2 = gelu(z2)
//   z3 = conv3x3(a2) + b3           1 channel
//   f(s) = sum_sites z3             the surrogate log|psi|
// Parameters, flattened: w1 [9][2][Cs], b1 [Cs], w2 [9][Cs][Cs], b2 [Cs], w3 [9][Cs], b3 [1].
//
// One block evaluates IPB whole configurations (one slot = one site per thread; 256 slots up to
// 256 sites, one configuration of up to 1024 sites per 1024-thread block above); activations of
// neighbouring sites are exchanged through LDS.  Training computes per-configuration
// gradients in one block (forward, then the backward pass through the same LDS tiles) and a
// second kernel sums them over the batch in a fixed order, so results are deterministic.

#include "nh_common.h"
#include "nh_internal.h"

namespace nh {

__device__ inline float gelu_grad(float x) {
  // d/dx [0.5 x (1 + tanh u)], u = k0 (x + 0.044715 x^3)
  const float k0 = 0.7978845608028654f;
  const float u = k0 * fmaf(0.044715f * x, x * x, x);
  const float th = tanhf(u);
  const float du = k0 * fmaf(3.0f * 0.044715f, x * x, 1.0f);
  return 0.5f * (1.0f + th) + 0.5f * x * (1.0f - th * th) * du;
}

struct SurParams {
  const float *w1, *b1, *w2, *b2, *w3, *b3;
};

template <int CS>
__device__ inline SurParams sur_params(const float* p) {
  SurParams q;
  q.w1 = p;
  q.b1 = q.w1 + 9 * 2 * CS;
  q.w2 = q.b1 + CS;
  q.b2 = q.w2 + 9 * CS * CS;
  q.w3 = q.b2 + CS;
  q.b3 = q.w3 + 9 * CS;
  return q;
}

__host__ __device__ inline int sur_nparams(int cs) { return 9 * 2 * cs + cs + 9 * cs * cs + cs + 9 * cs + 1; }

__device__ inline void neighbour_slots(int li, int site, int N, int LX, int LY, int* nb) {
  const int x = site / LY, y = site - (site / LY) * LY;
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
}

// Forward of configurations item -> s[item_b[item]] with item_hop applied (no group map).
template <int CS, int NT>
__global__ void __launch_bounds__(NT) k_sur_forward(int N, int LX, int LY, const float* __restrict__ P,
                                                     const int8_t* __restrict__ s, int two_n,
                                                     const int32_t* __restrict__ item_b,
                                                     const int32_t* __restrict__ item_hop, int n_items,
                                                     float* __restrict__ out) {
  extern __shared__ float sh[];  // act [CS][S], x0 [2][S], red [S]
  const int S = blockDim.x, IPB = S / N;
  float* act = sh;
  float* x0 = act + CS * S;
  float* red = x0 + 2 * S;
  const int t = threadIdx.x, li = t / N, site = t - li * N;
  const int item = blockIdx.x * IPB + li;
  bool active = item < n_items;
  int32_t hop = kHopInvalid;
  int b = 0;
  if (active) {
    hop = item_hop ? item_hop[item] : kHopNone;
    b = item_b ? item_b[item] : item;
  }
  if (hop == kHopInvalid) active = false;
  if (__syncthreads_or(active) == 0) return;
  const SurParams q = sur_params<CS>(P);
  int nb[9];
  neighbour_slots(li, site, N, LX, LY, nb);
  const int8_t* sb = s + (size_t)b * two_n;
  x0[t] = active ? (float)item_occ(sb, hop, site, N) : 0.f;
  x0[S + t] = active ? (float)item_occ(sb, hop, N + site, N) : 0.f;
  __syncthreads();
  float z[CS];
#pragma unroll
  for (int c = 0; c < CS; ++c) z[c] = q.b1[c];
  for (int tap = 0; tap < 9; ++tap) {
    const float v0 = x0[nb[tap]], v1 = x0[S + nb[tap]];
#pragma unroll
    for (int c = 0; c < CS; ++c) z[c] = fmaf(q.w1[tap * 2 * CS + c], v0, fmaf(q.w1[tap * 2 * CS + CS + c], v1, z[c]));
  }
#pragma unroll
  for (int c = 0; c < CS; ++c) act[c * S + t] = gelu_tanh(z[c]);
  __syncthreads();
#pragma unroll
  for (int c = 0; c < CS; ++c) z[c] = q.b2[c];
  for (int tap = 0; tap < 9; ++tap)
    for (int ci = 0; ci < CS; ++ci) {
      const float v = act[ci * S + nb[tap]];
      const float* w = q.w2 + (tap * CS + ci) * CS;
#pragma unroll
      for (int c = 0; c < CS; ++c) z[c] = fmaf(w[c], v, z[c]);
    }
  __syncthreads();
#pragma unroll
  for (int c = 0; c < CS; ++c) act[c * S + t] = gelu_tanh(z[c]);
  __syncthreads();
  float z3 = q.b3[0];
  for (int tap = 0; tap < 9; ++tap)
    for (int ci = 0; ci < CS; ++ci) z3 = fmaf(q.w3[tap * CS + ci], act[ci * S + nb[tap]], z3);
  red[t] = z3;
  __syncthreads();
  if (site == 0 && active) {
    float acc = 0.f;
    for (int i = 0; i < N; ++i) acc += red[li * N + i];
    out[item] = acc;
  }
}

// Per-configuration gradient of 0.5 * wgt * (f(s) - target)^2 into G[cfg][nparams]; loss into
// L[cfg].  One configuration per block (slots = N).
template <int CS, int NT>
__global__ void __launch_bounds__(NT) k_sur_grad(int N, int LX, int LY, const float* __restrict__ P,
                                                  const int8_t* __restrict__ s, int two_n,
                                                  const float* __restrict__ target, float wgt, int batch,
                                                  float* __restrict__ G, float* __restrict__ L) {
  extern __shared__ float sh[];  // a1 [CS][N], dz [CS][N], x0 [2][N], red [N], s2 [CS]
  const int cfg = blockIdx.x;
  if (cfg >= batch) return;
  const int t = threadIdx.x;
  float* a1 = sh;
  float* dz = a1 + CS * N;
  float* x0 = dz + CS * N;
  float* red = x0 + 2 * N;
  float* s2 = red + N;
  const int np = sur_nparams(CS);
  float* g = G + (size_t)cfg * np;
  const SurParams q = sur_params<CS>(P);
  SurParams gq = sur_params<CS>(g);
  const bool own = t < N;
  const int site = own ? t : 0;
  int nb[9];
  neighbour_slots(0, site, N, LX, LY, nb);
  const int8_t* sb = s + (size_t)cfg * two_n;
  if (own) {
    x0[t] = (float)sb[site];
    x0[N + t] = (float)sb[N + site];
  }
  __syncthreads();
  // forward, keeping z1 and z2 of this site in registers
  float z1[CS], z2[CS];
#pragma unroll
  for (int c = 0; c < CS; ++c) z1[c] = q.b1[c];
  for (int tap = 0; tap < 9; ++tap) {
    const float v0 = x0[nb[tap]], v1 = x0[N + nb[tap]];
#pragma unroll
    for (int c = 0; c < CS; ++c) z1[c] = fmaf(q.w1[tap * 2 * CS + c], v0, fmaf(q.w1[tap * 2 * CS + CS + c], v1, z1[c]));
  }
  if (own)
#pragma unroll
    for (int c = 0; c < CS; ++c) a1[c * N + t] = gelu_tanh(z1[c]);
  __syncthreads();
#pragma unroll
  for (int c = 0; c < CS; ++c) z2[c] = q.b2[c];
  for (int tap = 0; tap < 9; ++tap)
    for (int ci = 0; ci < CS; ++ci) {
      const float v = a1[ci * N + nb[tap]];
      const float* w = q.w2 + (tap * CS + ci) * CS;
#pragma unroll
      for (int c = 0; c < CS; ++c) z2[c] = fmaf(w[c], v, z2[c]);
    }
  // a2 goes to dz (free until the backward pass); z3 and the channel sums of a2
  if (own)
#pragma unroll
    for (int c = 0; c < CS; ++c) dz[c * N + t] = gelu_tanh(z2[c]);
  __syncthreads();
  float z3 = q.b3[0];
  for (int tap = 0; tap < 9; ++tap)
    for (int ci = 0; ci < CS; ++ci) z3 = fmaf(q.w3[tap * CS + ci], dz[ci * N + nb[tap]], z3);
  if (own) red[t] = z3;
  if (t < CS) {
    float acc = 0.f;
    for (int i = 0; i < N; ++i) acc += dz[t * N + i];
    s2[t] = acc;
  }
  __syncthreads();
  float f = 0.f;
  for (int i = 0; i < N; ++i) f += red[i];
  const float r = f - target[cfg];
  const float dout = wgt * r;  // d/df of 0.5 wgt r^2
  if (t == 0) L[cfg] = 0.5f * wgt * r * r;
  // layer 3: dz3 = dout at every site
  for (int e = t; e < 9 * CS; e += blockDim.x) const_cast<float*>(gq.w3)[e] = dout * s2[e % CS];
  if (t == 0) const_cast<float*>(gq.b3)[0] = dout * (float)N;
  // da2[c] = dout * sum_tap w3[tap][c] (constant over sites); dz2 = da2 * gelu'(z2)
  __syncthreads();  // everyone is done reading a2 from dz
  if (own)
#pragma unroll
    for (int c = 0; c < CS; ++c) {
      float w3s = 0.f;
#pragma unroll
      for (int tap = 0; tap < 9; ++tap) w3s += q.w3[tap * CS + c];
      dz[c * N + t] = dout * w3s * gelu_grad(z2[c]);
    }
  __syncthreads();
  // dW2[tap][ci][co] = sum_site a1[ci][nb(site, tap)] dz2[co][site]; db2[co] = sum_site dz2[co]
  for (int e = t; e < 9 * CS * CS; e += blockDim.x) {
    const int co = e % CS, ci = (e / CS) % CS, tap = e / (CS * CS);
    const int dx = tap / 3 - 1, dy = tap % 3 - 1;
    float acc = 0.f;
    for (int x = 0; x < LX; ++x) {
      int xx = x + dx;
      xx += (xx < 0) ? LX : 0;
      xx -= (xx >= LX) ? LX : 0;
      for (int y = 0; y < LY; ++y) {
        int yy = y + dy;
        yy += (yy < 0) ? LY : 0;
        yy -= (yy >= LY) ? LY : 0;
        acc = fmaf(a1[ci * N + xx * LY + yy], dz[co * N + x * LY + y], acc);
      }
    }
    const_cast<float*>(gq.w2)[e] = acc;
  }
  for (int c = t; c < CS; c += blockDim.x) {
    float acc = 0.f;
    for (int i = 0; i < N; ++i) acc += dz[c * N + i];
    const_cast<float*>(gq.b2)[c] = acc;
  }
  // da1[ci][site] = sum_tap sum_co w2[tap][ci][co] dz2[co][site - d(tap)]
  float da1[CS];
#pragma unroll
  for (int c = 0; c < CS; ++c) da1[c] = 0.f;
  for (int tap = 0; tap < 9; ++tap) {
    const int src = nb[8 - tap];  // site - d(tap) is the neighbour at the mirrored tap
    for (int co = 0; co < CS; ++co) {
      const float d = dz[co * N + src];
#pragma unroll
      for (int ci = 0; ci < CS; ++ci) da1[ci] = fmaf(q.w2[(tap * CS + ci) * CS + co], d, da1[ci]);
    }
  }
  __syncthreads();  // dz2 reads done
  if (own)
#pragma unroll
    for (int c = 0; c < CS; ++c) dz[c * N + t] = da1[c] * gelu_grad(z1[c]);
  __syncthreads();
  // dW1[tap][ci][co] = sum_site x0[ci][nb(site, tap)] dz1[co][site]; db1
  for (int e = t; e < 9 * 2 * CS; e += blockDim.x) {
    const int co = e % CS, ci = (e / CS) % 2, tap = e / (2 * CS);
    const int dx = tap / 3 - 1, dy = tap % 3 - 1;
    float acc = 0.f;
    for (int x = 0; x < LX; ++x) {
      int xx = x + dx;
      xx += (xx < 0) ? LX : 0;
      xx -= (xx >= LX) ? LX : 0;
      for (int y = 0; y < LY; ++y) {
        int yy = y + dy;
        yy += (yy < 0) ? LY : 0;
        yy -= (yy >= LY) ? LY : 0;
        acc = fmaf(x0[ci * N + xx * LY + yy], dz[co * N + x * LY + y], acc);
      }
    }
    const_cast<float*>(gq.w1)[e] = acc;
  }
  for (int c = t; c < CS; c += blockDim.x) {
    float acc = 0.f;
    for (int i = 0; i < N; ++i) acc += dz[c * N + i];
    const_cast<float*>(gq.b1)[c] = acc;
  }
}

// out[p] = sum_cfg G[cfg][p] (fixed order); loss = sum_cfg L[cfg] (thread 0)
__global__ void k_sur_reduce(int np, int batch, const float* __restrict__ G, const float* __restrict__ L,
                             float* __restrict__ grad, float* __restrict__ loss) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < np) {
    float acc = 0.f;
    for (int c = 0; c < batch; ++c) acc += G[(size_t)c * np + p];
    grad[p] = acc;
  }
  if (p == 0 && loss) {
    float acc = 0.f;
    for (int c = 0; c < batch; ++c) acc += L[c];
    loss[0] = acc;
  }
}

// Adam (Kingma & Ba) with bias correction at step t >= 1.
__global__ void k_adam(int np, float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const float gi = g[p];
  const float mi = b1 * m[p] + (1.f - b1) * gi;
  const float vi = b2 * v[p] + (1.f - b2) * gi * gi;
  m[p] = mi;
  v[p] = vi;
  w[p] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

static int sur_slots(int N) { return N <= 256 ? (256 / N) * N : N; }

// Up to 256 sites: 16 or 32 channels; up to 1024 sites (one configuration per 1024-thread block):
// 16 channels, whose gradient activations (2 x 16 x 1024 floats) fit LDS.
int check_sur(const nh_surrogate* q) {
  if (!q || !q->params) return NH_ERR_ARG;
  if (q->LX < 3 || q->LY < 3 || q->LX * q->LY > 1024) return NH_ERR_SHAPE;
  if (q->channels != 16 && q->channels != 32) return NH_ERR_SHAPE;
  if (q->LX * q->LY > 256 && q->channels != 16) return NH_ERR_SHAPE;
  return NH_OK;
}

int launch_sur_forward(hipStream_t st, const nh_surrogate* q, const int8_t* s, const int32_t* item_b,
                       const int32_t* item_hop, int n_items, float* out) {
  if (n_items <= 0) return NH_OK;
  const int N = q->LX * q->LY, S = sur_slots(N), IPB = S / N;
  const int blocks = (n_items + IPB - 1) / IPB;
  const size_t lds = sizeof(float) * (size_t)(q->channels * S + 3 * S);
  if (S > 256)
    hipLaunchKernelGGL((k_sur_forward<16, 1024>), dim3(blocks), dim3(S), lds, st, N, q->LX, q->LY, q->params, s,
                       2 * N, item_b, item_hop, n_items, out);
  else if (q->channels == 32)
    hipLaunchKernelGGL((k_sur_forward<32, 256>), dim3(blocks), dim3(S), lds, st, N, q->LX, q->LY, q->params, s,
                       2 * N, item_b, item_hop, n_items, out);
  else
    hipLaunchKernelGGL((k_sur_forward<16, 256>), dim3(blocks), dim3(S), lds, st, N, q->LX, q->LY, q->params, s,
                       2 * N, item_b, item_hop, n_items, out);
  NH_LAUNCH_CHECK();
  return NH_OK;
}

}  // namespace nh

using namespace nh;

extern "C" int32_t nh_sur_nparams(int32_t channels) { return sur_nparams(channels); }

extern "C" int nh_sur_forward(void* stream, const nh_surrogate* q, int32_t batch, const int8_t* s, float* out) {
  int rc = check_sur(q);
  if (rc != NH_OK) return rc;
  if (batch < 0) return NH_ERR_ARG;
  if (batch == 0) return NH_OK;
  if (!s || !out) return NH_ERR_ARG;
  return launch_sur_forward((hipStream_t)stream, q, s, nullptr, nullptr, batch, out);
}

extern "C" size_t nh_sur_grad_workspace_size(const nh_surrogate* q, int32_t batch) {
  if (check_sur(q) != NH_OK || batch < 0) return 0;
  return sizeof(float) * ((size_t)batch * sur_nparams(q->channels) + (size_t)batch) + 512;
}

extern "C" int nh_sur_grad(void* stream, const nh_surrogate* q, int32_t batch, const int8_t* s, const float* target,
                           float weight, float* grad, float* loss, void* workspace, size_t workspace_bytes) {
  int rc = check_sur(q);
  if (rc != NH_OK) return rc;
  if (batch <= 0 || !s || !target || !grad || !workspace) return NH_ERR_ARG;
  if (workspace_bytes < nh_sur_grad_workspace_size(q, batch)) return NH_ERR_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  const int N = q->LX * q->LY, np = sur_nparams(q->channels);
  float* G = (float*)(((size_t)workspace + 255) & ~(size_t)255);
  float* L = G + (size_t)batch * np;
  const size_t lds = sizeof(float) * (size_t)(2 * q->channels * N + 3 * N + q->channels);
  const int threads = N < 64 ? 64 : ((N + 63) / 64) * 64;
  if (threads > 256)
    hipLaunchKernelGGL((k_sur_grad<16, 1024>), dim3(batch), dim3(threads), lds, st, N, q->LX, q->LY, q->params, s,
                       2 * N, target, weight, batch, G, L);
  else if (q->channels == 32)
    hipLaunchKernelGGL((k_sur_grad<32, 256>), dim3(batch), dim3(threads), lds, st, N, q->LX, q->LY, q->params, s,
                       2 * N, target, weight, batch, G, L);
  else
    hipLaunchKernelGGL((k_sur_grad<16, 256>), dim3(batch), dim3(threads), lds, st, N, q->LX, q->LY, q->params, s,
                       2 * N, target, weight, batch, G, L);
  NH_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_sur_reduce, dim3((np + 255) / 256), dim3(256), 0, st, np, batch, G, L, grad, loss);
  NH_LAUNCH_CHECK();
  return NH_OK;
}

extern "C" int nh_adam(void* stream, int32_t n, float* w, const float* g, float* m, float* v, float lr, float beta1,
                       float beta2, float eps, int32_t step) {
  if (n < 0 || step < 1 || (n > 0 && (!w || !g || !m || !v))) return NH_ERR_ARG;
  if (n == 0) return NH_OK;
  const float c1 = 1.f - powf(beta1, (float)step), c2 = 1.f - powf(beta2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, n, w, g, m, v, lr, beta1, beta2,
                     eps, c1, c2);
  NH_LAUNCH_CHECK();
  return NH_OK;
}
