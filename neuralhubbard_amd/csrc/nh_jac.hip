// Per-sample Jacobian of log|psi_P(s)| with respect to every parameter of a backflow state: the
// reverse mode that quantax's AdamSR builds through the pairing block's custom_vjp
// (code/backflow/pf_kernel.py:354-365) and consumes at code/train_pfaffian_16x16_tpm02.py:100-101.
//
// For one configuration, d log|psi_P| = sum_g w_g d log|Pf F_g| (det for BCS) with
// w_g = psi_g / psi_P, so every image's cotangent is w_g times the Pfaffian (determinant) VJP.
// The pass, on a batch of configurations and their `ipi` images (8 with the projection):
//   1. per-image amplitudes through the product path (FP32 weights), giving w_g and psi_P;
//   2. the feature network again, layer by layer, keeping every layer's pre-activations;
//   3. F_g from those features (nh_pairing_occ / nh_pairing_ud plus the base table),
//      G_g = w_g d log|Pf F_g| / dF_g (nh_slogpf_vjp / nh_slogdet_vjp), then the pairing VJP
//      (nh_pairing_occ_vjp / nh_pairing_ud_vjp) for dY_g and dWt_g;
//   4. dY_g scattered onto the last layer's channels and carried back layer by layer (the
//      transposed convolutions, the GELU and RMS-norm derivatives), each convolution's weight
//      and bias gradient summed over a configuration's images straight into its Jacobian row;
//   5. dWt summed over images and mapped to (W, theta) through build_wtilde / build_wt_ud
//      (code/backflow/pf_kernel.py:107-124, bcs_kernel.py:63-70), the base table's gradient
//      scattered from the G_g.
// Activations are [image][channel][site] (site = x * LY + y).  Everything runs in FP32 with
// FP64 reductions where a sum runs over a whole image.

#include <algorithm>

#include "nh_common.h"
#include "nh_internal.h"

namespace nh {
namespace {

constexpr int kJT = 256;
constexpr float kRmsEps = 1e-5f;  // eqx.nn.RMSNorm default, as in the residual CNN forward

// Site (x + sgn dx, y + sgn dy) of tap (dx + 1) * 3 + (dy + 1), periodic.
__device__ inline int tap_site(int site, int tap, int sgn, int LX, int LY) {
  int x = site / LY + sgn * (tap / 3 - 1), y = site % LY + sgn * (tap % 3 - 1);
  x += x < 0 ? LX : 0;
  x -= x >= LX ? LX : 0;
  y += y < 0 ? LY : 0;
  y -= y >= LY ? LY : 0;
  return x * LY + y;
}

__device__ inline float gelu_grad(float x) {
  const float k0 = 0.7978845608028654f;  // sqrt(2 / pi)
  const float x2 = x * x;
  const float th = tanhf(k0 * fmaf(0.044715f * x, x2, x));
  return 0.5f * (1.f + th) + 0.5f * x * (1.f - th * th) * k0 * fmaf(3.f * 0.044715f, x2, 1.f);
}

// Block-wide FP64 sum, valid in every thread.
__device__ inline double block_sum(double v, double* red) {
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  __syncthreads();
  if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = v;
  __syncthreads();
  double t = 0.0;
  for (int w = 0; w < kJT / kWave; ++w) t += red[w];
  return t;
}

// Image configuration x_img[m] = s_b[g(m)] (g an involution), b = img / ipi, g = img % ipi.
__global__ void k_jac_images(int N, int LX, int LY, int ipi, const int8_t* __restrict__ s, int8_t* __restrict__ xs) {
  const int img = blockIdx.y;
  const int m = blockIdx.x * kJT + threadIdx.x;
  if (m >= 2 * N) return;
  const int b = img / ipi, g = img % ipi;
  xs[(size_t)img * 2 * N + m] = s[(size_t)b * 2 * N + group_map(m, g, N, LX, LY)];
}

// The occupied modes of s_b in [lo, lo + cnt), ascending, mapped by g and shifted by -sub, into
// out[0 .. cap).  With the wrong particle number (whose amplitude is NaN) entries past cap are
// dropped and missing ones are 0, so every later index stays in range.
__device__ void mapped_list(const int8_t* sb, int lo, int cnt, int g, int N, int LX, int LY, int sub, int32_t* out,
                            int cap, int* wsum) {
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  int base = 0;
  for (int c0 = 0; c0 < cnt; c0 += kJT) {
    const int i = c0 + threadIdx.x;
    const bool occ = i < cnt && sb[lo + i] > 0;
    const uint64_t mask = __ballot(occ);
    if (lane == 0) wsum[wave] = __popcll(mask);
    __syncthreads();
    int off = base, tot = 0;
    for (int w = 0; w < kJT / kWave; ++w) {
      off += w < wave ? wsum[w] : 0;
      tot += wsum[w];
    }
    const int pos = off + lanes_below(mask);
    if (occ && pos < cap) out[pos] = group_map(lo + i, g, N, LX, LY) - sub;
    __syncthreads();
    base += tot;
  }
  for (int p = base + threadIdx.x; p < cap; p += kJT) out[p] = 0;
}

// Pfaffian: modes[img][p] = g(p-th occupied mode of s_b).  BCS: iu / idn[img][p] = the sites of
// the image's up / down electrons, each list in the ascending order of the source modes of s_b.
__global__ void __launch_bounds__(kJT) k_jac_lists(int N, int LX, int LY, int ipi, int bcs, int n_up, int n_dn,
                                                   const int8_t* __restrict__ s, int32_t* __restrict__ l0,
                                                   int32_t* __restrict__ l1) {
  __shared__ int wsum[kJT / kWave];
  const int img = blockIdx.x;
  const int b = img / ipi, g = img % ipi;
  const int8_t* sb = s + (size_t)b * 2 * N;
  if (!bcs) {
    const int n = n_up + n_dn;
    mapped_list(sb, 0, 2 * N, g, N, LX, LY, 0, l0 + (size_t)img * n, n, wsum);
    return;
  }
  const int inv = (g >> 2) & 1;  // spin inversion: the image's up electrons come from s's down block
  mapped_list(sb, inv * N, N, g, N, LX, LY, 0, l0 + (size_t)img * n_up, n_up, wsum);
  mapped_list(sb, (1 - inv) * N, N, g, N, LX, LY, N, l1 + (size_t)img * n_dn, n_dn, wsum);
}

// ---- forward, layer by layer

// First convolution of the synthetic network on the 2-channel occupancy of the image.
__global__ void k_jconv_in(int N, int LX, int LY, int C, const int8_t* __restrict__ xs, const float* __restrict__ W,
                           const float* __restrict__ bias, float* __restrict__ H) {
  const int img = blockIdx.z, o = blockIdx.y, site = blockIdx.x * kJT + threadIdx.x;
  if (site >= N) return;
  const int8_t* x = xs + (size_t)img * 2 * N;
  float acc = bias[o];
  for (int tap = 0; tap < 9; ++tap) {
    const int ns = tap_site(site, tap, 1, LX, LY);
    acc = fmaf(W[(tap * 2 + 0) * C + o], (float)x[ns], acc);
    acc = fmaf(W[(tap * 2 + 1) * C + o], (float)x[N + ns], acc);
  }
  H[((size_t)img * C + o) * N + site] = acc;
}

// Residual CNN: H[img][c][site] = emb[occupation class][c].
__global__ void k_jembed(int N, int C, const int8_t* __restrict__ xs, const float* __restrict__ emb,
                         float* __restrict__ H) {
  const int img = blockIdx.z, c = blockIdx.y, site = blockIdx.x * kJT + threadIdx.x;
  if (site >= N) return;
  const int8_t* x = xs + (size_t)img * 2 * N;
  const int cls = (x[site] > 0) + 2 * (x[N + site] > 0);
  H[((size_t)img * C + c) * N + site] = emb[cls * C + c];
}

// rinv[img] = 1 / sqrt(mean(H^2) + eps) over the image's channels and sites.
__global__ void __launch_bounds__(kJT) k_jrinv(int per, const float* __restrict__ H, float* __restrict__ rinv) {
  __shared__ double red[kJT / kWave];
  const float* h = H + (size_t)blockIdx.x * per;
  double acc = 0.0;
  for (int i = threadIdx.x; i < per; i += kJT) acc += (double)h[i] * h[i];
  const double t = block_sum(acc, red);
  if (threadIdx.x == 0) rinv[blockIdx.x] = (float)(1.0 / sqrt(t / per + (double)kRmsEps));
}

// A = act(P): mode 1 gelu(P), mode 2 gelu(P * rinv[img]) (the RMS-normed input of a residual
// block), mode 3 the image's 2-channel occupancy (P unused; per = 2N).
__global__ void k_jact(int per, int nimg, int mode, const float* __restrict__ P, const float* __restrict__ rinv,
                       const int8_t* __restrict__ xs, float* __restrict__ A) {
  const size_t e = (size_t)blockIdx.x * kJT + threadIdx.x;
  if (e >= (size_t)per * nimg) return;
  if (mode == 3) {
    A[e] = (float)xs[e];
    return;
  }
  const float v = P[e];
  A[e] = gelu_tanh(mode == 2 ? v * rinv[e / per] : v);
}

// out[img][o][site] = res + bias[o] + sum_tap sum_c W[tap][c][o] A[img][c][site + tap].
__global__ void k_jconv(int N, int LX, int LY, int Cin, int Cout, const float* __restrict__ A,
                        const float* __restrict__ W, const float* __restrict__ bias, const float* __restrict__ res,
                        float* __restrict__ out) {
  const int img = blockIdx.z, o = blockIdx.y, site = blockIdx.x * kJT + threadIdx.x;
  if (site >= N) return;
  const float* a = A + (size_t)img * Cin * N;
  float acc = bias[o];
  for (int tap = 0; tap < 9; ++tap) {
    const int ns = tap_site(site, tap, 1, LX, LY);
    const float* w = W + (size_t)tap * Cin * Cout + o;
    for (int c = 0; c < Cin; ++c) acc = fmaf(w[(size_t)c * Cout], a[(size_t)c * N + ns], acc);
  }
  const size_t e = ((size_t)img * Cout + o) * N + site;
  out[e] = (res ? res[e] : 0.f) + acc;
}

// ---- pairing side

// Features of the listed modes: Pfaffian Y[img][p][a] = scale H[img][spin d + a][site] of mode
// l0[p]; BCS Yu / Yd from the up / down site lists, through the linear head when `head` is set
// (y_k[site][a] = sum_c head[k][c][a] scale H[c][site]).
__global__ void k_jgather(int N, int C, int d, int bcs, int n0, int n1, float scale, const float* __restrict__ H,
                          const float* __restrict__ head, const int32_t* __restrict__ l0,
                          const int32_t* __restrict__ l1, float* __restrict__ Y0, float* __restrict__ Y1) {
  const int img = blockIdx.y;
  const int e = blockIdx.x * kJT + threadIdx.x;
  if (e >= (n0 + n1) * d) return;
  const int p = e / d, a = e % d;
  const float* h = H + (size_t)img * C * N;
  if (!bcs) {
    const int m = l0[(size_t)img * n0 + p];
    const int spin = m >= N;
    Y0[(size_t)img * n0 * d + e] = scale * h[(size_t)(spin * d + a) * N + (m - spin * N)];
    return;
  }
  const int k = p >= n0;
  const int q = k ? p - n0 : p;
  const int site = k ? l1[(size_t)img * n1 + q] : l0[(size_t)img * n0 + q];
  float v;
  if (head) {
    const float* hk = head + (size_t)k * C * d + a;
    float acc = 0.f;
    for (int c = 0; c < C; ++c) acc = fmaf(hk[(size_t)c * d], h[(size_t)c * N + site], acc);
    v = scale * acc;
  } else {
    v = scale * h[(size_t)(k * d + a) * N + site];
  }
  float* Y = k ? Y1 + (size_t)img * n1 * d : Y0 + (size_t)img * n0 * d;
  Y[(size_t)q * d + a] = v;
}

// Base table of the image matrix: Pfaffian F[p][q] += F0[l0[p]][l0[q]] (F0 [2N][2N]); BCS
// F[p][q] += phi0[r(up_p - dn_q)].
__global__ void k_jbase_add(int N, int LX, int LY, int bcs, int n0, int n1, const float* __restrict__ base,
                            const int32_t* __restrict__ l0, const int32_t* __restrict__ l1, float* __restrict__ F) {
  const int img = blockIdx.y;
  const int e = blockIdx.x * kJT + threadIdx.x;
  const int cols = bcs ? n1 : n0;
  if (e >= n0 * cols) return;
  const int p = e / cols, q = e % cols;
  float* f = F + (size_t)img * n0 * cols;
  if (!bcs) {
    const int32_t* md = l0 + (size_t)img * n0;
    f[e] += base[(size_t)md[p] * 2 * N + md[q]];
    return;
  }
  const int i = l0[(size_t)img * n0 + p], j = l1[(size_t)img * n1 + q];
  const int dx = (i / LY - j / LY + LX) % LX, dy = (i % LY - j % LY + LY) % LY;
  f[e] += base[dx * LY + dy];
}

// Image weights w_g = psi_g / psi_P (NaN where psi_P is zero or not finite).
__global__ void k_jweights(int nimg, int ipi, const float* __restrict__ img_sign, const double* __restrict__ img_log,
                           const float* __restrict__ sign, const double* __restrict__ logabs, float* __restrict__ w) {
  const int img = blockIdx.x * kJT + threadIdx.x;
  if (img >= nimg) return;
  const int b = img / ipi;
  const double lp = logabs[b];
  const float sp = sign[b];
  w[img] = (sp == 0.f || !isfinite(lp)) ? __builtin_nanf("")
                                         : (float)((double)img_sign[img] * sp * exp(img_log[img] - lp));
}

// dY onto the last layer's channels: dH[img][c][site] = scale (the feature of (c, site) in the
// listed modes' dY, through the head's transpose for BCS with a head), zero where no listed mode
// sits.  One block per image; the inverse lists in LDS.
__global__ void __launch_bounds__(kJT) k_jscatter(int N, int C, int d, int bcs, int n0, int n1, float scale,
                                                  const float* __restrict__ head, const int32_t* __restrict__ l0,
                                                  const int32_t* __restrict__ l1, const float* __restrict__ dY0,
                                                  const float* __restrict__ dY1, float* __restrict__ dH) {
  extern __shared__ int inv[];  // Pfaffian [2N] by mode; BCS [N] up, [N] down by site
  const int img = blockIdx.x;
  for (int i = threadIdx.x; i < 2 * N; i += kJT) inv[i] = -1;
  __syncthreads();
  if (!bcs) {
    for (int p = threadIdx.x; p < n0; p += kJT) inv[l0[(size_t)img * n0 + p]] = p;
  } else {
    for (int p = threadIdx.x; p < n0; p += kJT) inv[l0[(size_t)img * n0 + p]] = p;
    for (int q = threadIdx.x; q < n1; q += kJT) inv[N + l1[(size_t)img * n1 + q]] = q;
  }
  __syncthreads();
  const float* y0 = dY0 + (size_t)img * n0 * d;
  const float* y1 = bcs ? dY1 + (size_t)img * n1 * d : nullptr;
  float* out = dH + (size_t)img * C * N;
  for (int e = threadIdx.x; e < C * N; e += kJT) {
    const int c = e / N, site = e % N;
    float v = 0.f;
    if (!bcs) {
      const int spin = c / d, a = c % d;
      const int p = inv[spin * N + site];
      v = p >= 0 ? y0[(size_t)p * d + a] : 0.f;
    } else if (!head) {
      const int k = c / d, a = c % d;
      const int p = inv[k * N + site];
      v = p >= 0 ? (k ? y1 : y0)[(size_t)p * d + a] : 0.f;
    } else {
      const int pu = inv[site], pd = inv[N + site];
      if (pu >= 0)
        for (int a = 0; a < d; ++a) v = fmaf(head[(size_t)c * d + a], y0[(size_t)pu * d + a], v);
      if (pd >= 0)
        for (int a = 0; a < d; ++a) v = fmaf(head[(size_t)(C + c) * d + a], y1[(size_t)pd * d + a], v);
    }
    out[e] = scale * v;
  }
}

// ---- backward, layer by layer

// dA[img][c][site] = sum_tap sum_o W[tap][c][o] dOut[img][o][site - tap]; then mode 1:
// dIn = dres + dA gelu'(P); mode 2: dIn = dA gelu'(P rinv) (the gradient of the RMS-normed
// input, finished by k_jrms_back).
__global__ void k_jconv_back(int N, int LX, int LY, int Cin, int Cout, const float* __restrict__ dOut,
                             const float* __restrict__ W, int mode, const float* __restrict__ P,
                             const float* __restrict__ rinv, const float* __restrict__ dres, float* __restrict__ dIn) {
  const int img = blockIdx.z, c = blockIdx.y, site = blockIdx.x * kJT + threadIdx.x;
  if (site >= N) return;
  const float* dg = dOut + (size_t)img * Cout * N;
  float acc = 0.f;
  for (int tap = 0; tap < 9; ++tap) {
    const int ns = tap_site(site, tap, -1, LX, LY);
    const float* w = W + ((size_t)tap * Cin + c) * Cout;
    for (int o = 0; o < Cout; ++o) acc = fmaf(w[o], dg[(size_t)o * N + ns], acc);
  }
  const size_t e = ((size_t)img * Cin + c) * N + site;
  if (mode == 2) {
    dIn[e] = acc * gelu_grad(P[e] * rinv[img]);
  } else {
    dIn[e] = (dres ? dres[e] : 0.f) + acc * gelu_grad(P[e]);
  }
}

// RMS norm backward: R = P r, r = (mean(P^2) + eps)^(-1/2), so
// dP = dres + r dR - (r^3 / M) P sum(dR P), M = the image's channels x sites.
__global__ void __launch_bounds__(kJT) k_jrms_back(int per, const float* __restrict__ P, const float* __restrict__ rinv,
                                                   const float* __restrict__ dR, const float* __restrict__ dres,
                                                   float* __restrict__ dP) {
  __shared__ double red[kJT / kWave];
  const size_t o = (size_t)blockIdx.x * per;
  double acc = 0.0;
  for (int i = threadIdx.x; i < per; i += kJT) acc += (double)dR[o + i] * P[o + i];
  const double dot = block_sum(acc, red);
  const float r = rinv[blockIdx.x];
  const float k = (float)((double)r * r * r * dot / per);
  for (int i = threadIdx.x; i < per; i += kJT) dP[o + i] = dres[o + i] + r * dR[o + i] - k * P[o + i];
}

// Weight and bias gradients of one convolution, summed over the ipi images of configuration b:
// J[b][w_off + (tap Cin + c) Cout + o] = sum_img sum_site dOut[o][site] A[c][site + tap],
// J[b][b_off + o] = sum_img sum_site dOut[o][site].  Block (o, b); dOut's rows staged in LDS.
__global__ void __launch_bounds__(kJT) k_jwgrad(int N, int LX, int LY, int Cin, int Cout, int ipi,
                                                const float* __restrict__ dOut, const float* __restrict__ A,
                                                float* __restrict__ J, size_t ldJ, size_t w_off, size_t b_off) {
  extern __shared__ float dl[];  // [ipi][N]
  __shared__ double red[kJT / kWave];
  const int o = blockIdx.x, b = blockIdx.y;
  double bsum = 0.0;
  for (int i = threadIdx.x; i < ipi * N; i += kJT) {
    const int g = i / N, site = i % N;
    const float v = dOut[(((size_t)b * ipi + g) * Cout + o) * N + site];
    dl[i] = v;
    bsum += v;
  }
  const double bt = block_sum(bsum, red);  // its barriers also publish dl
  float* row = J + (size_t)b * ldJ;
  if (threadIdx.x == 0) row[b_off + o] = (float)bt;
  for (int tc = threadIdx.x; tc < 9 * Cin; tc += kJT) {
    const int tap = tc / Cin, c = tc % Cin;
    const int dx = tap / 3 - 1, dy = tap % 3 - 1;
    double acc = 0.0;
    for (int g = 0; g < ipi; ++g) {
      const float* a = A + (((size_t)b * ipi + g) * Cin + c) * N;
      const float* dg = dl + g * N;
      float part = 0.f;
      for (int x = 0; x < LX; ++x) {  // the lattice row by row: neighbour (x + dx, y + dy), periodic
        int xx = x + dx;
        xx += xx < 0 ? LX : 0;
        xx -= xx >= LX ? LX : 0;
        const float* ar = a + xx * LY;
        const float* dr = dg + x * LY;
        for (int y = 0; y < LY; ++y) {
          int yy = y + dy;
          yy += yy < 0 ? LY : 0;
          yy -= yy >= LY ? LY : 0;
          part = fmaf(dr[y], ar[yy], part);
        }
      }
      acc += part;
    }
    row[w_off + (size_t)tc * Cout + o] = (float)acc;
  }
}

// Residual CNN embedding: J[b][cls C + c] = sum_img sum over sites of class cls of dH0[c][site].
__global__ void __launch_bounds__(kJT) k_jemb(int N, int C, int ipi, const int8_t* __restrict__ xs,
                                              const float* __restrict__ dH, float* __restrict__ J, size_t ldJ) {
  const int b = blockIdx.x;
  for (int e = threadIdx.x; e < 4 * C; e += kJT) {
    const int cls = e / C, c = e % C;
    double acc = 0.0;
    for (int g = 0; g < ipi; ++g) {
      const int img = b * ipi + g;
      const int8_t* x = xs + (size_t)img * 2 * N;
      const float* dh = dH + ((size_t)img * C + c) * N;
      for (int site = 0; site < N; ++site)
        if ((x[site] > 0) + 2 * (x[N + site] > 0) == cls) acc += dh[site];
    }
    J[(size_t)b * ldJ + e] = (float)acc;
  }
}

// BCS linear head: J[b][h_off + (k C + c) d + a] = sum_img sum_p dY_k[p][a] scale H[c][site_k(p)].
__global__ void __launch_bounds__(kJT) k_jhead(int N, int C, int d, int ipi, int n0, int n1, float scale,
                                               const float* __restrict__ H, const int32_t* __restrict__ l0,
                                               const int32_t* __restrict__ l1, const float* __restrict__ dY0,
                                               const float* __restrict__ dY1, float* __restrict__ J, size_t ldJ,
                                               size_t h_off) {
  const int b = blockIdx.x;
  for (int e = threadIdx.x; e < 2 * C * d; e += kJT) {
    const int k = e / (C * d), c = (e / d) % C, a = e % d;
    const int nk = k ? n1 : n0;
    double acc = 0.0;
    for (int g = 0; g < ipi; ++g) {
      const int img = b * ipi + g;
      const int32_t* sites = (k ? l1 + (size_t)img * n1 : l0 + (size_t)img * n0);
      const float* dy = (k ? dY1 + (size_t)img * n1 * d : dY0 + (size_t)img * n0 * d);
      const float* h = H + ((size_t)img * C + c) * N;
      float part = 0.f;
      for (int p = 0; p < nk; ++p) part = fmaf(dy[(size_t)p * d + a], h[sites[p]], part);
      acc += part;
    }
    J[(size_t)b * ldJ + h_off + e] = (float)(scale * acc);
  }
}

// (W, theta) from dWt summed over the images of configuration b.  With pref_r =
// exp(-lam D_r) / (N Z), lam = softplus(theta): Pfaffian (Wt = [uu, ud, du, dd] pref, uu =
// (W1 - W1[neg]) / 2, ud = W0, du = -W0[neg], dd = (W2 - W2[neg]) / 2)
//   dW0[r] = dWt1[r] pref_r - dWt2[neg r] pref_neg r,
//   dW1[r] = (dWt0[r] pref_r - dWt0[neg r] pref_neg r) / 2, dW2 likewise from dWt3;
// BCS (Wt = W pref) dW[r] = dWt[r] pref_r; both
//   dtheta = sigmoid(theta) sum dWt Wt (<D> - D_r),  <D> = sum_r D_r exp(-lam D_r) / Z.
__global__ void __launch_bounds__(kJT) k_jwt(int N, int LX, int LY, int d, int ipi, int bcs, double theta,
                                             const float* __restrict__ Wt, const float* __restrict__ dWt,
                                             float* __restrict__ J, size_t ldJ, size_t w_off) {
  __shared__ double red[kJT / kWave];
  const int b = blockIdx.x;
  const double lam = theta > 0 ? theta + log1p(exp(-theta)) : log1p(exp(theta));
  const double sig = 1.0 / (1.0 + exp(-theta));
  auto dist = [&](int r) {
    const int x = r / LY, y = r % LY;
    const int mx = x < LX - x ? x : LX - x, my = y < LY - y ? y : LY - y;
    return sqrt((double)(mx * mx + my * my));
  };
  double z = 0.0, zd = 0.0;
  for (int r = threadIdx.x; r < N; r += kJT) {
    const double e = exp(-lam * dist(r));
    z += e;
    zd += dist(r) * e;
  }
  const double Z = block_sum(z, red);
  const double ZD = block_sum(zd, red);
  const double mean_d = ZD / Z;
  auto pref = [&](int r) { return exp(-lam * dist(r)) / ((double)N * Z); };
  auto neg = [&](int r) {
    const int x = r / LY, y = r % LY;
    return ((LX - x) % LX) * LY + (LY - y) % LY;
  };
  const int K = bcs ? 1 : 4;
  const size_t per = (size_t)K * N * d;
  auto dsum = [&](int k, int r, int a) {
    float v = 0.f;
    for (int g = 0; g < ipi; ++g) v += dWt[((size_t)b * ipi + g) * per + ((size_t)k * N + r) * d + a];
    return v;
  };
  float* row = J + (size_t)b * ldJ + w_off;
  double th = 0.0;
  for (int e = threadIdx.x; e < N * d; e += kJT) {
    const int r = e / d, a = e % d;
    const double pr = pref(r), dr = mean_d - dist(r);
    if (bcs) {
      const float g0 = dsum(0, r, a);
      row[e] = (float)(g0 * pr);
      th += g0 * (double)Wt[e] * dr;
      continue;
    }
    const int rn = neg(r);
    const double pn = pref(rn);
    const float g0 = dsum(0, r, a), g1 = dsum(1, r, a), g2 = dsum(2, r, a), g3 = dsum(3, r, a);
    row[e] = (float)(g1 * pr - dsum(2, rn, a) * pn);                          // W0 (ud)
    row[(size_t)N * d + e] = (float)(0.5 * (g0 * pr - dsum(0, rn, a) * pn));  // W1 (uu source)
    row[(size_t)2 * N * d + e] = (float)(0.5 * (g3 * pr - dsum(3, rn, a) * pn));  // W2 (dd source)
    for (int k = 0; k < 4; ++k) {
      const float gk = k == 0 ? g0 : k == 1 ? g1 : k == 2 ? g2 : g3;
      th += gk * (double)Wt[((size_t)k * N + r) * d + a] * dr;
    }
  }
  const double tt = block_sum(th, red);
  if (threadIdx.x == 0) row[(size_t)(bcs ? 1 : 3) * N * d] = (float)(sig * tt);  // theta follows W
}

// Base table: Pfaffian J[b][off + m 2N + m'] = sum_img sum over (p, q) with l0[p] = m,
// l0[q] = m' of G[p][q] (every entry of F0 independent); BCS J[b][off + r] = sum_img sum over
// (p, q) with r(up_p - dn_q) = r of G[p][q].
__global__ void __launch_bounds__(kJT) k_jbase(int N, int LX, int LY, int ipi, int bcs, int n0, int n1,
                                               const int32_t* __restrict__ l0, const int32_t* __restrict__ l1,
                                               const float* __restrict__ G, float* __restrict__ J, size_t ldJ,
                                               size_t off) {
  extern __shared__ float acc[];  // BCS: [N]
  const int b = blockIdx.x;
  float* row = J + (size_t)b * ldJ + off;
  if (!bcs) {
    const size_t tot = (size_t)4 * N * N;
    for (size_t e = threadIdx.x; e < tot; e += kJT) row[e] = 0.f;
    for (int g = 0; g < ipi; ++g) {
      __syncthreads();  // the previous image's (or the zeroing) writes are done
      const int img = b * ipi + g;
      const int32_t* md = l0 + (size_t)img * n0;
      const float* gm = G + (size_t)img * n0 * n0;
      for (int e = threadIdx.x; e < n0 * n0; e += kJT) {
        const int p = e / n0, q = e % n0;
        row[(size_t)md[p] * 2 * N + md[q]] += gm[e];
      }
    }
    return;
  }
  for (int r = threadIdx.x; r < N; r += kJT) acc[r] = 0.f;
  __syncthreads();
  for (int g = 0; g < ipi; ++g) {
    const int img = b * ipi + g;
    const float* gm = G + (size_t)img * n0 * n1;
    for (int e = threadIdx.x; e < n0 * n1; e += kJT) {
      const int p = e / n1, q = e % n1;
      const int i = l0[(size_t)img * n0 + p], j = l1[(size_t)img * n1 + q];
      const int dx = (i / LY - j / LY + LX) % LX, dy = (i % LY - j % LY + LY) % LY;
      atomicAdd(&acc[dx * LY + dy], gm[e]);
    }
  }
  __syncthreads();
  for (int r = threadIdx.x; r < N; r += kJT) row[r] = acc[r];
}

// ---- sizes

struct JacShape {
  int N, C, d, L, n0, n1, ipi;
  bool bcs, res, head;
  size_t conv_w, conv_b, W, base, P;  // Jacobian column counts; P the total
  size_t w_off(int l) const {  // packed conv weights of convolution l
    if (res) return (size_t)4 * C + (size_t)l * 9 * C * C;
    return l == 0 ? 0 : (size_t)18 * C + (size_t)(l - 1) * 9 * C * C;
  }
};

JacShape jac_shape(const nh_model* m) {
  JacShape s;
  s.N = m->LX * m->LY;
  s.C = m->channels;
  s.d = feature_dim(m);
  s.bcs = m->kind == NH_KIND_BCS;
  s.res = m->network == NH_NET_RESCONV || m->network == NH_NET_RESCONV_HEAD;
  s.head = m->network == NH_NET_RESCONV_HEAD;
  s.L = m->n_layers;
  s.n0 = s.bcs ? m->n_up : m->n_up + m->n_dn;
  s.n1 = s.bcs ? m->n_dn : 0;
  s.ipi = m->project ? kGroup : 1;
  const size_t C = s.C;
  if (s.res)
    s.conv_w = 4 * C + (size_t)2 * s.L * 9 * C * C + (s.head ? 2 * C * s.d : 0);
  else
    s.conv_w = 18 * C + (size_t)s.L * 9 * C * C;
  s.conv_b = (s.res ? (size_t)2 * s.L : (size_t)s.L + 1) * C;
  s.W = (size_t)(s.bcs ? 1 : 3) * s.N * s.d;
  s.base = !m->base ? 0 : s.bcs ? (size_t)s.N : (size_t)4 * s.N * s.N;
  s.P = s.conv_w + s.conv_b + s.W + 1 + s.base;
  return s;
}

int jac_check(const nh_model* m) {
  int rc = check_model(m);
  if (rc != NH_OK) return rc;
  const JacShape s = jac_shape(m);
  if (s.d < 1 || s.d > 64 || (kJT % s.d) != 0) return NH_ERR_SHAPE;  // the pairing VJP's limits
  if (s.bcs && (s.n0 == 0 || s.n1 == 0)) return NH_ERR_SHAPE;
  if (!s.bcs && s.n0 == 0) return NH_ERR_SHAPE;
  return NH_OK;
}

struct JacBuffers {
  int8_t* xs;
  int32_t *l0, *l1;
  float *Hs, *Zs, *rinv, *A, *dH0, *dH1, *dT, *Y0, *Y1, *dY0, *dY1, *F, *G, *dWt, *w;
  void* prod;  // the per-image amplitude pass
  float* img_sign;
  double* img_log;
};

struct JCarver {
  char* base;
  size_t off = 0;
  template <class T>
  T* take(size_t n) {
    T* p = base ? (T*)(base + off) : nullptr;
    off += (n * sizeof(T) + 255) & ~(size_t)255;
    return p;
  }
};

JacBuffers jac_carve(JCarver& c, const nh_model* m, const nh_model* mf, int64_t batch) {
  const JacShape s = jac_shape(m);
  const size_t nimg = (size_t)batch * s.ipi, N = s.N, CN = (size_t)s.C * N;
  JacBuffers b;
  b.xs = c.take<int8_t>(nimg * 2 * N);
  b.l0 = c.take<int32_t>(nimg * s.n0);
  b.l1 = s.bcs ? c.take<int32_t>(nimg * s.n1) : nullptr;
  b.Hs = c.take<float>(nimg * CN * (s.L + 1));  // every layer's (block's) output H
  b.Zs = s.res ? c.take<float>(nimg * CN * s.L) : nullptr;
  b.rinv = s.res ? c.take<float>(nimg * s.L) : nullptr;
  b.A = c.take<float>(nimg * std::max<size_t>(CN, 2 * N));
  b.dH0 = c.take<float>(nimg * CN);
  b.dH1 = c.take<float>(nimg * CN);
  b.dT = c.take<float>(nimg * CN);
  b.Y0 = c.take<float>(nimg * s.n0 * s.d);
  b.Y1 = s.bcs ? c.take<float>(nimg * s.n1 * s.d) : nullptr;
  b.dY0 = c.take<float>(nimg * s.n0 * s.d);
  b.dY1 = s.bcs ? c.take<float>(nimg * s.n1 * s.d) : nullptr;
  const size_t fsz = (size_t)s.n0 * (s.bcs ? s.n1 : s.n0);
  b.F = c.take<float>(nimg * fsz);
  b.G = c.take<float>(nimg * fsz);
  b.dWt = c.take<float>(nimg * (s.bcs ? 1 : 4) * N * s.d);
  b.w = c.take<float>(nimg);
  const size_t pb = image_amplitudes_bytes(mf, batch);
  b.prod = c.take<char>(pb);
  return b;
}

}  // namespace
}  // namespace nh

using namespace nh;

extern "C" size_t nh_log_psi_jacobian_size(const nh_model* model) {
  if (jac_check(model) != NH_OK) return 0;
  return jac_shape(model).P;
}

extern "C" size_t nh_log_psi_jacobian_workspace_size(const nh_model* model, int32_t batch) {
  if (jac_check(model) != NH_OK || batch < 0) return 0;
  nh_model mf = *model;
  mf.precision = NH_PREC_FP32;
  mf.conv_w_lp = nullptr;
  JCarver c{nullptr};
  jac_carve(c, model, &mf, batch);
  return c.off + 256;
}

extern "C" int nh_log_psi_jacobian(void* stream, const nh_model* model, double theta, int32_t batch,
                                   const int8_t* s, float* sign, double* logabs, float* J, void* workspace,
                                   size_t workspace_bytes) {
  int rc = jac_check(model);
  if (rc != NH_OK) return rc;
  if (batch < 0) return NH_ERR_ARG;
  if (batch == 0) return NH_OK;
  if (!s || !sign || !logabs || !J || !workspace || !isfinite(theta)) return NH_ERR_ARG;
  const JacShape sh = jac_shape(model);
  const int64_t nimg64 = (int64_t)batch * sh.ipi;
  if (nimg64 > 65535) return NH_ERR_SHAPE;  // one grid row per image in the pairing kernels
  if (workspace_bytes < nh_log_psi_jacobian_workspace_size(model, batch)) return NH_ERR_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  const int nimg = (int)nimg64, N = sh.N, C = sh.C, d = sh.d, ipi = sh.ipi;
  const size_t CN = (size_t)C * N;
  nh_model mf = *model;  // the amplitudes that weight the images: FP32 weights
  mf.precision = NH_PREC_FP32;
  mf.conv_w_lp = nullptr;
  JCarver cv{(char*)(((size_t)workspace + 255) & ~(size_t)255)};
  const JacBuffers bf = jac_carve(cv, model, &mf, batch);

  // 1. psi_P and the image amplitudes
  float* img_sign = nullptr;
  double* img_log = nullptr;
  rc = image_amplitudes(st, &mf, s, batch, bf.prod, &img_sign, &img_log, sign, logabs);
  if (rc != NH_OK) return rc;
  hipLaunchKernelGGL(k_jweights, dim3((nimg + kJT - 1) / kJT), dim3(kJT), 0, st, nimg, ipi, img_sign, img_log, sign,
                     logabs, bf.w);
  NH_LAUNCH_CHECK();

  // 2. image configurations, mode lists, forward with the pre-activations kept
  hipLaunchKernelGGL(k_jac_images, dim3((2 * N + kJT - 1) / kJT, nimg), dim3(kJT), 0, st, N, model->LX, model->LY,
                     ipi, s, bf.xs);
  hipLaunchKernelGGL(k_jac_lists, dim3(nimg), dim3(kJT), 0, st, N, model->LX, model->LY, ipi, sh.bcs ? 1 : 0,
                     model->n_up, model->n_dn, s, bf.l0, bf.l1);
  NH_LAUNCH_CHECK();
  const dim3 gsite((N + kJT - 1) / kJT, C, nimg);
  const unsigned gel = (unsigned)((nimg * CN + kJT - 1) / kJT);
  auto H = [&](int k) { return bf.Hs + (size_t)k * nimg * CN; };
  auto Z = [&](int k) { return bf.Zs + (size_t)k * nimg * CN; };
  auto Wc = [&](int l) { return model->conv_w + sh.w_off(l); };
  if (!sh.res) {
    hipLaunchKernelGGL(k_jconv_in, gsite, dim3(kJT), 0, st, N, model->LX, model->LY, C, bf.xs, Wc(0), model->conv_b,
                       H(0));
    for (int l = 1; l <= sh.L; ++l) {
      hipLaunchKernelGGL(k_jact, dim3(gel), dim3(kJT), 0, st, (int)CN, nimg, 1, H(l - 1), nullptr, nullptr, bf.A);
      hipLaunchKernelGGL(k_jconv, gsite, dim3(kJT), 0, st, N, model->LX, model->LY, C, C, bf.A, Wc(l),
                         model->conv_b + (size_t)l * C, H(l - 1), H(l));
    }
  } else {
    hipLaunchKernelGGL(k_jembed, gsite, dim3(kJT), 0, st, N, C, bf.xs, model->conv_w, H(0));
    for (int k = 0; k < sh.L; ++k) {
      float* rk = bf.rinv + (size_t)k * nimg;
      hipLaunchKernelGGL(k_jrinv, dim3(nimg), dim3(kJT), 0, st, (int)CN, H(k), rk);
      hipLaunchKernelGGL(k_jact, dim3(gel), dim3(kJT), 0, st, (int)CN, nimg, 2, H(k), rk, nullptr, bf.A);
      hipLaunchKernelGGL(k_jconv, gsite, dim3(kJT), 0, st, N, model->LX, model->LY, C, C, bf.A, Wc(2 * k),
                         model->conv_b + (size_t)(2 * k) * C, nullptr, Z(k));
      hipLaunchKernelGGL(k_jact, dim3(gel), dim3(kJT), 0, st, (int)CN, nimg, 1, Z(k), nullptr, nullptr, bf.A);
      hipLaunchKernelGGL(k_jconv, gsite, dim3(kJT), 0, st, N, model->LX, model->LY, C, C, bf.A, Wc(2 * k + 1),
                         model->conv_b + (size_t)(2 * k + 1) * C, H(k), H(k + 1));
    }
  }
  NH_LAUNCH_CHECK();
  const float scale = sh.res ? (float)(1.0 / sqrt((double)sh.L + 1.0)) : 1.f;
  const float* head = sh.head ? model->conv_w + sh.w_off(2 * sh.L) : nullptr;
  float* HL = H(sh.L);

  // 3. image matrices, their cotangents and the pairing VJP
  const int rows = sh.n0 + sh.n1;
  hipLaunchKernelGGL(k_jgather, dim3((rows * d + kJT - 1) / kJT, nimg), dim3(kJT), 0, st, N, C, d, sh.bcs ? 1 : 0,
                     sh.n0, sh.n1, scale, HL, head, bf.l0, bf.l1, bf.Y0, bf.Y1);
  NH_LAUNCH_CHECK();
  if (!sh.bcs) {
    rc = nh_pairing_occ(stream, model->LX, model->LY, d, nimg, sh.n0, bf.Y0, bf.l0, model->wt, bf.F);
  } else {
    rc = nh_pairing_ud(stream, model->LX, model->LY, d, nimg, sh.n0, sh.n1, bf.Y0, bf.Y1, bf.l0, bf.l1, model->wt,
                       bf.F);
  }
  if (rc != NH_OK) return rc;
  const int fcols = sh.bcs ? sh.n1 : sh.n0;
  if (model->base) {
    hipLaunchKernelGGL(k_jbase_add, dim3((sh.n0 * fcols + kJT - 1) / kJT, nimg), dim3(kJT), 0, st, N, model->LX,
                       model->LY, sh.bcs ? 1 : 0, sh.n0, sh.n1, model->base, bf.l0, bf.l1, bf.F);
    NH_LAUNCH_CHECK();
  }
  rc = sh.bcs ? nh_slogdet_vjp(stream, nimg, sh.n0, bf.F, bf.w, bf.G) : nh_slogpf_vjp(stream, nimg, sh.n0, bf.F, bf.w, bf.G);
  if (rc != NH_OK) return rc;
  if (!sh.bcs) {
    rc = nh_pairing_occ_vjp(stream, model->LX, model->LY, d, nimg, sh.n0, bf.Y0, bf.l0, model->wt, bf.G, bf.dY0,
                            bf.dWt);
  } else {
    rc = nh_pairing_ud_vjp(stream, model->LX, model->LY, d, nimg, sh.n0, sh.n1, bf.Y0, bf.Y1, bf.l0, bf.l1,
                           model->wt, bf.G, bf.dY0, bf.dY1, bf.dWt);
  }
  if (rc != NH_OK) return rc;

  // 4. back through the network
  const size_t ldJ = sh.P;
  const size_t b_base = sh.conv_w;  // conv biases follow the packed weights
  hipLaunchKernelGGL(k_jscatter, dim3(nimg), dim3(kJT), sizeof(int) * 2 * N, st, N, C, d, sh.bcs ? 1 : 0, sh.n0,
                     sh.n1, scale, head, bf.l0, bf.l1, bf.dY0, bf.dY1, bf.dH0);
  NH_LAUNCH_CHECK();
  if (sh.head) {
    hipLaunchKernelGGL(k_jhead, dim3(batch), dim3(kJT), 0, st, N, C, d, ipi, sh.n0, sh.n1, scale, HL, bf.l0, bf.l1,
                       bf.dY0, bf.dY1, J, ldJ, sh.w_off(2 * sh.L));
    NH_LAUNCH_CHECK();
  }
  float* dH = bf.dH0;  // gradient at the current layer's output
  float* dHn = bf.dH1;
  const size_t wl = sizeof(float) * (size_t)ipi * N;
  const dim3 gw(C, batch);
  if (!sh.res) {
    for (int l = sh.L; l >= 1; --l) {
      hipLaunchKernelGGL(k_jact, dim3(gel), dim3(kJT), 0, st, (int)CN, nimg, 1, H(l - 1), nullptr, nullptr, bf.A);
      hipLaunchKernelGGL(k_jwgrad, gw, dim3(kJT), wl, st, N, model->LX, model->LY, C, C, ipi, dH, bf.A, J, ldJ,
                         sh.w_off(l), b_base + (size_t)l * C);
      hipLaunchKernelGGL(k_jconv_back, gsite, dim3(kJT), 0, st, N, model->LX, model->LY, C, C, dH, Wc(l), 1,
                         H(l - 1), nullptr, dH, dHn);
      std::swap(dH, dHn);
    }
    hipLaunchKernelGGL(k_jact, dim3((unsigned)((nimg * 2 * (size_t)N + kJT - 1) / kJT)), dim3(kJT), 0, st, 2 * N,
                       nimg, 3, nullptr, nullptr, bf.xs, bf.A);
    hipLaunchKernelGGL(k_jwgrad, gw, dim3(kJT), wl, st, N, model->LX, model->LY, 2, C, ipi, dH, bf.A, J, ldJ,
                       sh.w_off(0), b_base);
  } else {
    for (int k = sh.L - 1; k >= 0; --k) {
      const float* rk = bf.rinv + (size_t)k * nimg;
      // H(k+1) = H(k) + conv_{2k+1}(gelu(Z(k))), Z(k) = conv_{2k}(gelu(rmsnorm(H(k))))
      hipLaunchKernelGGL(k_jact, dim3(gel), dim3(kJT), 0, st, (int)CN, nimg, 1, Z(k), nullptr, nullptr, bf.A);
      hipLaunchKernelGGL(k_jwgrad, gw, dim3(kJT), wl, st, N, model->LX, model->LY, C, C, ipi, dH, bf.A, J, ldJ,
                         sh.w_off(2 * k + 1), b_base + (size_t)(2 * k + 1) * C);
      hipLaunchKernelGGL(k_jconv_back, gsite, dim3(kJT), 0, st, N, model->LX, model->LY, C, C, dH, Wc(2 * k + 1), 1,
                         Z(k), nullptr, nullptr, bf.dT);  // dZ
      hipLaunchKernelGGL(k_jact, dim3(gel), dim3(kJT), 0, st, (int)CN, nimg, 2, H(k), rk, nullptr, bf.A);
      hipLaunchKernelGGL(k_jwgrad, gw, dim3(kJT), wl, st, N, model->LX, model->LY, C, C, ipi, bf.dT, bf.A, J, ldJ,
                         sh.w_off(2 * k), b_base + (size_t)(2 * k) * C);
      hipLaunchKernelGGL(k_jconv_back, gsite, dim3(kJT), 0, st, N, model->LX, model->LY, C, C, bf.dT, Wc(2 * k), 2,
                         H(k), rk, nullptr, dHn);  // dR
      hipLaunchKernelGGL(k_jrms_back, dim3(nimg), dim3(kJT), 0, st, (int)CN, H(k), rk, dHn, dH, dHn);
      std::swap(dH, dHn);
    }
    hipLaunchKernelGGL(k_jemb, dim3(batch), dim3(kJT), 0, st, N, C, ipi, bf.xs, dH, J, ldJ);
  }
  NH_LAUNCH_CHECK();

  // 5. (W, theta) and the base table
  const size_t wt_off = sh.conv_w + sh.conv_b;
  hipLaunchKernelGGL(k_jwt, dim3(batch), dim3(kJT), 0, st, N, model->LX, model->LY, d, ipi, sh.bcs ? 1 : 0, theta,
                     model->wt, bf.dWt, J, ldJ, wt_off);
  NH_LAUNCH_CHECK();
  if (model->base) {
    hipLaunchKernelGGL(k_jbase, dim3(batch), dim3(kJT), sh.bcs ? sizeof(float) * N : 0, st, N, model->LX, model->LY,
                       ipi, sh.bcs ? 1 : 0, sh.n0, sh.n1, bf.l0, bf.l1, bf.G, J, ldJ, wt_off + sh.W + 1);
    NH_LAUNCH_CHECK();
  }
  return NH_OK;
}
