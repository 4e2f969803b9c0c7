// Pairing blocks, skew Pfaffian and determinant on gfx950 (FP32 arithmetic, FP64 log-sum).
//
// One workgroup owns one matrix.  A Pfaffian matrix lives in LDS as its packed strict lower
// triangle (row i at i(i-1)/2), so n = 224 takes 98 KiB; a determinant matrix lives as a
// dense row-major square.  Elimination step k:
//   wave 0 : pivot search on column k (wave argmax, smallest index on ties), the symmetric
//            row/column swap, the pivot's log and sign, and the two update vectors;
//   all    : the rank-2 (Pfaffian) or rank-1 (LU) update of the trailing part, rows spread
//            over waves, columns over lanes, so LDS accesses are unit-stride.
// Parlett-Reid for the Pfaffian follows Wimmer, ACM TOMS 38 (2012) 30 (pfaffian_LTL): with
// pivot p = A[k+1][k], Pf *= -p, tau_i = A[i][k] / p, v_i = A[i][k+1] and
// A[i][j] += tau_i v_j - v_i tau_j for i > j >= k + 2.

#include "nh_common.h"
#include "nh_internal.h"

namespace nh {

__device__ inline int toff(int i, int j) { return ((i * (i - 1)) >> 1) + j; }

__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

__device__ inline void wave_argmax(float& best, int& bi) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    float ob = __shfl_xor(best, off);
    int oi = __shfl_xor(bi, off);
    if (ob > best || (ob == best && oi < bi)) {
      best = ob;
      bi = oi;
    }
  }
}

// Parlett-Reid on the packed lower triangle `tri` (n even). Result valid in thread 0.
template <int NT>
__device__ void pf_eliminate(float* tri, int n, float* tau, float* vv, int* flag, float& sign_out,
                             double& log_out) {
  constexpr int NW = NT / kWave;
  const int t = threadIdx.x, lane = t & (kWave - 1), wave = t / kWave;
  double logabs = 0.0;
  float sign = 1.f;
  for (int k = 0; k + 1 < n; k += 2) {
    if (wave == 0) {
      float best = -1.f;
      int bi = n;
      for (int i = k + 1 + lane; i < n; i += kWave) {
        float v = fabsf(tri[toff(i, k)]);
        if (v > best) {
          best = v;
          bi = i;
        }
      }
      wave_argmax(best, bi);
      const int a = k + 1, b = bi;
      if (b != a) {
        if (lane == 0) {
          float t0 = tri[toff(a, k)];
          tri[toff(a, k)] = tri[toff(b, k)];
          tri[toff(b, k)] = t0;
          tri[toff(b, a)] = -tri[toff(b, a)];
        }
        for (int i = a + 1 + lane; i < b; i += kWave) {
          float t1 = tri[toff(i, a)], t2 = tri[toff(b, i)];
          tri[toff(i, a)] = -t2;
          tri[toff(b, i)] = -t1;
        }
        for (int i = b + 1 + lane; i < n; i += kWave) {
          float t1 = tri[toff(i, a)];
          tri[toff(i, a)] = tri[toff(i, b)];
          tri[toff(i, b)] = t1;
        }
      }
      wave_sync();
      const float piv = tri[toff(a, k)];
      const bool bad = !(piv != 0.f) || !isfinite(piv);
      if (!bad)
        for (int i = k + 2 + lane; i < n; i += kWave) {
          tau[i] = tri[toff(i, k)] / piv;
          vv[i] = tri[toff(i, a)];
        }
      if (lane == 0) {
        if (b != a) sign = -sign;
        if (bad) {
          flag[0] = 1;
        } else {
          sign = (piv < 0.f) ? sign : -sign;  // Pf *= A[k][k+1] = -piv
          logabs += log((double)fabsf(piv));
        }
      }
    }
    __syncthreads();
    if (flag[0]) break;
    for (int i = k + 3 + wave; i < n; i += NW) {
      const float ti = tau[i], vi = vv[i];
      float* row = tri + toff(i, 0);
      for (int j = k + 2 + lane; j < i; j += kWave) row[j] = fmaf(ti, vv[j], fmaf(-vi, tau[j], row[j]));
    }
    __syncthreads();
  }
  if (flag[0]) {
    sign_out = 0.f;
    log_out = -INFINITY;
  } else {
    sign_out = sign;
    log_out = logabs;
  }
}

// LU with partial pivoting on the dense row-major M (leading dimension ld). Thread 0 result.
template <int NT>
__device__ void lu_eliminate(float* M, int n, int ld, float* lcol, float* urow, int* flag,
                             float& sign_out, double& log_out) {
  constexpr int NW = NT / kWave;
  const int t = threadIdx.x, lane = t & (kWave - 1), wave = t / kWave;
  double logabs = 0.0;
  float sign = 1.f;
  for (int k = 0; k < n; ++k) {
    if (wave == 0) {
      float best = -1.f;
      int bi = n;
      for (int i = k + lane; i < n; i += kWave) {
        float v = fabsf(M[i * ld + k]);
        if (v > best) {
          best = v;
          bi = i;
        }
      }
      wave_argmax(best, bi);
      if (bi != k)
        for (int j = k + lane; j < n; j += kWave) {
          float t1 = M[k * ld + j];
          M[k * ld + j] = M[bi * ld + j];
          M[bi * ld + j] = t1;
        }
      wave_sync();
      const float piv = M[k * ld + k];
      const bool bad = !(piv != 0.f) || !isfinite(piv);
      if (!bad) {
        for (int i = k + 1 + lane; i < n; i += kWave) lcol[i] = M[i * ld + k] / piv;
        for (int j = k + 1 + lane; j < n; j += kWave) urow[j] = M[k * ld + j];
      }
      if (lane == 0) {
        if (bi != k) sign = -sign;
        if (bad) {
          flag[0] = 1;
        } else {
          sign = (piv < 0.f) ? -sign : sign;
          logabs += log((double)fabsf(piv));
        }
      }
    }
    __syncthreads();
    if (flag[0]) break;
    for (int i = k + 1 + wave; i < n; i += NW) {
      const float li = lcol[i];
      float* row = M + i * ld;
      for (int j = k + 1 + lane; j < n; j += kWave) row[j] = fmaf(-li, urow[j], row[j]);
    }
    __syncthreads();
  }
  if (flag[0]) {
    sign_out = 0.f;
    log_out = -INFINITY;
  } else {
    sign_out = sign;
    log_out = logabs;
  }
}

// Ascending occupied modes of the item configuration, mapped by group element g, in the
// order of the item's own occupied list.  Wave 0 only; returns the count in lane 0's view.
__device__ inline int occupied_modes(const int8_t* sb, int32_t hop, int two_n, int N, int LX, int LY,
                                     int g, int* modes, int cap) {
  const int lane = threadIdx.x & (kWave - 1);
  int cnt = 0;
  for (int m0 = 0; m0 < two_n; m0 += kWave) {
    const int m = m0 + lane;
    const bool occ = (m < two_n) && item_occ(sb, hop, m) > 0;
    const uint64_t mask = __ballot(occ);
    const int pos = cnt + lanes_below(mask);
    if (occ && pos < cap) modes[pos] = group_map(m, g, N, LX, LY);
    cnt += __popcll(mask);
  }
  return cnt;
}

__device__ inline float parity_sign(long long k) { return (k & 1) ? -1.f : 1.f; }

// ------------------------------------------------------------------ Pfaffian image kernel

template <int D, int NT>
__global__ void __launch_bounds__(NT) k_image_pf(int N, int LX, int LY, int n, const int8_t* __restrict__ s,
                                                 const int32_t* __restrict__ item_b,
                                                 const int32_t* __restrict__ item_hop, int img0,
                                                 int img_end, int ipi, const float* __restrict__ Y,
                                                 const float* __restrict__ wt,
                                                 const float* __restrict__ F0,
                                                 float* __restrict__ img_sign,
                                                 double* __restrict__ img_logabs) {
  extern __shared__ float lds[];
  const int img = img0 + blockIdx.x;
  if (img >= img_end) return;
  const int item = img / ipi;
  const int g = (ipi == kGroup) ? (img & 7) : 0;
  const int32_t hop = item_hop ? item_hop[item] : kHopNone;
  if (hop == kHopInvalid) return;
  const int b = item_b ? item_b[item] : item;
  const int two_n = 2 * N;

  const int ntri = (n * (n - 1)) / 2;
  float* tri = lds;
  float* yocc = tri + ((ntri + 3) & ~3);
  float* tau = yocc + n * D;
  float* vv = tau + n;
  int* modes = reinterpret_cast<int*>(vv + n);
  int* flag = modes + n;

  const int t = threadIdx.x, lane = t & (kWave - 1), wave = t / kWave;
  if (wave == 0) {
    int cnt = occupied_modes(s + (size_t)b * two_n, hop, two_n, N, LX, LY, g, modes, n);
    if (lane == 0) flag[0] = (cnt != n) ? 2 : 0;
  }
  __syncthreads();
  if (flag[0]) {
    if (t == 0) {
      img_sign[img] = 0.f;
      img_logabs[img] = NAN;
    }
    return;
  }
  const float* Yi = Y + (size_t)(img - img0) * (two_n * D);
  for (int e = t; e < n * (D / 4); e += NT) {
    const int p = e / (D / 4), a4 = e - p * (D / 4);
    reinterpret_cast<float4*>(yocc)[e] = reinterpret_cast<const float4*>(Yi + (size_t)modes[p] * D)[a4];
  }
  __syncthreads();

  constexpr int NW = NT / kWave;
  for (int p = 1 + wave; p < n; p += NW) {
    float yp[D];
#pragma unroll
    for (int a = 0; a < D; ++a) yp[a] = yocc[p * D + a];
    const int mp = modes[p];
    const int sp = mp >= N;
    const int sitep = mp - sp * N;
    float* row = tri + toff(p, 0);
    for (int q = lane; q < p; q += kWave) {
      const int mq = modes[q];
      const int sq = mq >= N;
      const int r = disp_index(sitep, mq - sq * N, LX, LY);
      const float4* w4 = reinterpret_cast<const float4*>(wt + (size_t)((2 * sp + sq) * N + r) * D);
      const float4* y4 = reinterpret_cast<const float4*>(yocc + q * D);
      float acc = 0.f;
#pragma unroll
      for (int a4 = 0; a4 < D / 4; ++a4) {
        const float4 w = w4[a4];
        const float4 yq = y4[a4];
        acc = fmaf(yp[4 * a4] * yq.x, w.x, acc);
        acc = fmaf(yp[4 * a4 + 1] * yq.y, w.y, acc);
        acc = fmaf(yp[4 * a4 + 2] * yq.z, w.z, acc);
        acc = fmaf(yp[4 * a4 + 3] * yq.w, w.w, acc);
      }
      if (F0) acc += F0[(size_t)mp * two_n + mq];
      row[q] = acc;
    }
  }
  if (t == 0) flag[0] = 0;
  __syncthreads();

  float sg;
  double la;
  pf_eliminate<NT>(tri, n, tau, vv, flag, sg, la);
  if (t == 0) {
    // fermion_inverse_sign: reversal of the Ne creation operators
    sg *= parity_sign((long long)n * (n - 1) / 2);
    img_sign[img] = sg;
    img_logabs[img] = la;
  }
}

// ------------------------------------------------------------------ BCS image kernel

template <int D, int NT>
__global__ void __launch_bounds__(NT) k_image_bcs(int N, int LX, int LY, int n_up, int n_dn,
                                                  const int8_t* __restrict__ s,
                                                  const int32_t* __restrict__ item_b,
                                                  const int32_t* __restrict__ item_hop, int img0,
                                                  int img_end, int ipi, const float* __restrict__ Y,
                                                  const float* __restrict__ wt,
                                                  const float* __restrict__ phi0,
                                                  float* __restrict__ img_sign,
                                                  double* __restrict__ img_logabs) {
  extern __shared__ float lds[];
  const int img = img0 + blockIdx.x;
  if (img >= img_end) return;
  const int item = img / ipi;
  const int g = (ipi == kGroup) ? (img & 7) : 0;
  const int32_t hop = item_hop ? item_hop[item] : kHopNone;
  if (hop == kHopInvalid) return;
  const int b = item_b ? item_b[item] : item;
  const int two_n = 2 * N;
  const int n = n_up;  // square: n_up == n_dn checked on the host
  const int ne = n_up + n_dn;
  const int ld = n + 1;

  float* M = lds;
  float* yu = M + ((n * ld + 3) & ~3);
  float* yd = yu + n * D;
  float* lcol = yd + n * D;
  float* urow = lcol + n;
  int* modes = reinterpret_cast<int*>(urow + n);
  int* flag = modes + ne;

  const int t = threadIdx.x, lane = t & (kWave - 1), wave = t / kWave;
  if (wave == 0) {
    int cnt = occupied_modes(s + (size_t)b * two_n, hop, two_n, N, LX, LY, g, modes, ne);
    if (lane == 0) flag[0] = (cnt != ne) ? 2 : 0;
  }
  __syncthreads();
  if (flag[0]) {
    if (t == 0) {
      img_sign[img] = 0.f;
      img_logabs[img] = NAN;
    }
    return;
  }
  // rows: the image's spin-up modes in order; columns: its spin-down modes in order.
  const bool si = (g >> 2) & 1;
  const int* rows = si ? modes + n_up : modes;  // up modes (sites < N)
  const int* cols = si ? modes : modes + n_up;  // down modes (>= N)
  const float* Yi = Y + (size_t)(img - img0) * (two_n * D);
  for (int e = t; e < n * (D / 4); e += NT) {
    const int p = e / (D / 4), a4 = e - p * (D / 4);
    reinterpret_cast<float4*>(yu)[e] = reinterpret_cast<const float4*>(Yi + (size_t)rows[p] * D)[a4];
    reinterpret_cast<float4*>(yd)[e] = reinterpret_cast<const float4*>(Yi + (size_t)cols[p] * D)[a4];
  }
  __syncthreads();
  constexpr int NW = NT / kWave;
  for (int p = wave; p < n; p += NW) {
    float yp[D];
#pragma unroll
    for (int a = 0; a < D; ++a) yp[a] = yu[p * D + a];
    const int sitep = rows[p];
    for (int q = lane; q < n; q += kWave) {
      const int siteq = cols[q] - N;
      const int r = disp_index(sitep, siteq, LX, LY);
      const float4* w4 = reinterpret_cast<const float4*>(wt + (size_t)r * D);
      const float4* y4 = reinterpret_cast<const float4*>(yd + q * D);
      float acc = 0.f;
#pragma unroll
      for (int a4 = 0; a4 < D / 4; ++a4) {
        const float4 w = w4[a4];
        const float4 yq = y4[a4];
        acc = fmaf(yp[4 * a4] * w.x, yq.x, acc);
        acc = fmaf(yp[4 * a4 + 1] * w.y, yq.y, acc);
        acc = fmaf(yp[4 * a4 + 2] * w.z, yq.z, acc);
        acc = fmaf(yp[4 * a4 + 3] * w.w, yq.w, acc);
      }
      if (phi0) acc += phi0[r];
      M[p * ld + q] = acc;
    }
  }
  if (t == 0) flag[0] = 0;
  __syncthreads();
  float sg;
  double la;
  lu_eliminate<NT>(M, n, ld, lcol, urow, flag, sg, la);
  if (t == 0) {
    sg *= parity_sign((long long)n * (n - 1) / 2);  // Pf [[0,Phi],[-Phi^T,0]] = (-1)^{n(n-1)/2} det
    if (si) sg *= parity_sign(n);                   // spin-down-first image: det(-Phi^T)
    sg *= parity_sign((long long)ne * (ne - 1) / 2);
    img_sign[img] = sg;
    img_logabs[img] = la;
  }
}

static size_t pf_lds_bytes(int n, int D) {
  const int ntri = (n * (n - 1)) / 2;
  return sizeof(float) * (size_t)(((ntri + 3) & ~3) + n * D + 2 * n) + sizeof(int) * (size_t)(n + 4);
}

static size_t bcs_lds_bytes(int n, int ne, int D) {
  return sizeof(float) * (size_t)(((n * (n + 1) + 3) & ~3) + 2 * n * D + 2 * n) + sizeof(int) * (size_t)(ne + 4);
}

constexpr size_t kMaxLds = 160 * 1024;

template <int D>
static int launch_image_d(hipStream_t st, const nh_model* m, const int8_t* s, const int32_t* item_b,
                          const int32_t* item_hop, int img0, int img_end, int ipi, const float* Y,
                          float* img_sign, double* img_logabs) {
  const int N = m->LX * m->LY;
  const int nimg = img_end - img0;
  if (nimg <= 0) return NH_OK;
  if (m->kind == NH_KIND_PFAFFIAN) {
    const int n = m->n_up + m->n_dn;
    const size_t lds = pf_lds_bytes(n, D);
    if (lds > kMaxLds) return NH_ERR_SHAPE;
    if (n <= 64) {
      hipLaunchKernelGGL((k_image_pf<D, 64>), dim3(nimg), dim3(64), lds, st, N, m->LX, m->LY, n, s,
                         item_b, item_hop, img0, img_end, ipi, Y, m->wt, m->base, img_sign, img_logabs);
    } else {
      hipLaunchKernelGGL((k_image_pf<D, 512>), dim3(nimg), dim3(512), lds, st, N, m->LX, m->LY, n, s,
                         item_b, item_hop, img0, img_end, ipi, Y, m->wt, m->base, img_sign, img_logabs);
    }
  } else {
    const int n = m->n_up;
    const size_t lds = bcs_lds_bytes(n, m->n_up + m->n_dn, D);
    if (lds > kMaxLds) return NH_ERR_SHAPE;
    if (n <= 64) {
      hipLaunchKernelGGL((k_image_bcs<D, 64>), dim3(nimg), dim3(64), lds, st, N, m->LX, m->LY, m->n_up,
                         m->n_dn, s, item_b, item_hop, img0, img_end, ipi, Y, m->wt, m->base, img_sign,
                         img_logabs);
    } else {
      hipLaunchKernelGGL((k_image_bcs<D, 512>), dim3(nimg), dim3(512), lds, st, N, m->LX, m->LY, m->n_up,
                         m->n_dn, s, item_b, item_hop, img0, img_end, ipi, Y, m->wt, m->base, img_sign,
                         img_logabs);
    }
  }
  NH_LAUNCH_CHECK();
  return NH_OK;
}

int launch_image_amplitude(hipStream_t st, const nh_model* m, const int8_t* s, const int32_t* item_b,
                           const int32_t* item_hop, int img0, int img_end, int ipi, const float* Y,
                           float* img_sign, double* img_logabs) {
  switch (m->channels / 2) {
    case 4: return launch_image_d<4>(st, m, s, item_b, item_hop, img0, img_end, ipi, Y, img_sign, img_logabs);
    case 8: return launch_image_d<8>(st, m, s, item_b, item_hop, img0, img_end, ipi, Y, img_sign, img_logabs);
    case 16: return launch_image_d<16>(st, m, s, item_b, item_hop, img0, img_end, ipi, Y, img_sign, img_logabs);
    default: return NH_ERR_SHAPE;
  }
}

// ------------------------------------------------------------------ standalone entry kernels

template <int NT>
__global__ void __launch_bounds__(NT) k_slogpf(int n, const float* __restrict__ A, float* sign,
                                               double* logabs) {
  extern __shared__ float lds[];
  const int ntri = (n * (n - 1)) / 2;
  float* tri = lds;
  float* tau = tri + ((ntri + 3) & ~3);
  float* vv = tau + n;
  int* flag = reinterpret_cast<int*>(vv + n);
  const float* Ab = A + (size_t)blockIdx.x * n * n;
  for (int i = 1 + threadIdx.x / kWave; i < n; i += NT / kWave)
    for (int j = threadIdx.x & (kWave - 1); j < i; j += kWave) tri[toff(i, j)] = Ab[(size_t)i * n + j];
  if (threadIdx.x == 0) flag[0] = 0;
  __syncthreads();
  float sg;
  double la;
  pf_eliminate<NT>(tri, n, tau, vv, flag, sg, la);
  if (threadIdx.x == 0) {
    sign[blockIdx.x] = sg;
    logabs[blockIdx.x] = la;
  }
}

template <int NT>
__global__ void __launch_bounds__(NT) k_slogdet(int n, const float* __restrict__ A, float* sign,
                                                double* logabs) {
  extern __shared__ float lds[];
  const int ld = n + 1;
  float* M = lds;
  float* lcol = M + ((n * ld + 3) & ~3);
  float* urow = lcol + n;
  int* flag = reinterpret_cast<int*>(urow + n);
  const float* Ab = A + (size_t)blockIdx.x * n * n;
  for (int e = threadIdx.x; e < n * n; e += NT) M[(e / n) * ld + (e % n)] = Ab[e];
  if (threadIdx.x == 0) flag[0] = 0;
  __syncthreads();
  float sg;
  double la;
  lu_eliminate<NT>(M, n, ld, lcol, urow, flag, sg, la);
  if (threadIdx.x == 0) {
    sign[blockIdx.x] = sg;
    logabs[blockIdx.x] = la;
  }
}

__global__ void k_pairing_occ(int LX, int LY, int d, int n, const float* __restrict__ Y,
                              const int32_t* __restrict__ idx, const float* __restrict__ Wt,
                              float* __restrict__ F) {
  const int N = LX * LY;
  const size_t b = blockIdx.y;
  const float* Yb = Y + b * n * d;
  const int32_t* ib = idx + b * n;
  float* Fb = F + b * (size_t)n * n;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * n) return;
  const int p = e / n, q = e - (e / n) * n;
  if (p == q) {
    Fb[e] = 0.f;
    return;
  }
  // evaluate the lower-triangle entry and negate it for the upper one: exact antisymmetry
  const int hi = p > q ? p : q, lo = p > q ? q : p;
  const int mp = ib[hi], mq = ib[lo];
  const int sp = mp >= N, sq = mq >= N;
  const int r = disp_index(mp - sp * N, mq - sq * N, LX, LY);
  const float* w = Wt + (size_t)((2 * sp + sq) * N + r) * d;
  float acc = 0.f;
  for (int a = 0; a < d; ++a) acc = fmaf(Yb[hi * d + a] * Yb[lo * d + a], w[a], acc);
  Fb[e] = (p > q) ? acc : -acc;
}

__global__ void k_pairing_ud(int LX, int LY, int d, int n_up, int n_dn, const float* __restrict__ Yu,
                             const float* __restrict__ Yd, const int32_t* __restrict__ iu,
                             const int32_t* __restrict__ idn, const float* __restrict__ Wt,
                             float* __restrict__ F) {
  const size_t b = blockIdx.y;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_up * n_dn) return;
  const int p = e / n_dn, q = e - (e / n_dn) * n_dn;
  const int r = disp_index(iu[b * n_up + p], idn[b * n_dn + q], LX, LY);
  const float* w = Wt + (size_t)r * d;
  const float* yu = Yu + (b * n_up + p) * d;
  const float* yd = Yd + (b * n_dn + q) * d;
  float acc = 0.f;
  for (int a = 0; a < d; ++a) acc = fmaf(yu[a] * w[a], yd[a], acc);
  F[b * (size_t)n_up * n_dn + e] = acc;
}

}  // namespace nh

using namespace nh;

extern "C" int nh_slogpf(void* stream, int32_t batch, int32_t n, const float* A, float* sign,
                         double* logabs) {
  if (batch < 0 || n < 0 || (batch > 0 && (!A || !sign || !logabs))) return NH_ERR_ARG;
  if (n % 2 || n > 256) return NH_ERR_SHAPE;
  if (batch == 0) return NH_OK;
  hipStream_t st = (hipStream_t)stream;
  if (n == 0) return NH_ERR_SHAPE;
  const int ntri = (n * (n - 1)) / 2;
  const size_t lds = sizeof(float) * (size_t)(((ntri + 3) & ~3) + 2 * n) + 16;
  if (n <= 64)
    hipLaunchKernelGGL(k_slogpf<64>, dim3(batch), dim3(64), lds, st, n, A, sign, logabs);
  else
    hipLaunchKernelGGL(k_slogpf<512>, dim3(batch), dim3(512), lds, st, n, A, sign, logabs);
  NH_LAUNCH_CHECK();
  return NH_OK;
}

extern "C" int nh_slogdet(void* stream, int32_t batch, int32_t n, const float* A, float* sign,
                          double* logabs) {
  if (batch < 0 || n <= 0 || (batch > 0 && (!A || !sign || !logabs))) return NH_ERR_ARG;
  if (n > 192) return NH_ERR_SHAPE;
  if (batch == 0) return NH_OK;
  hipStream_t st = (hipStream_t)stream;
  const size_t lds = sizeof(float) * (size_t)(((n * (n + 1) + 3) & ~3) + 2 * n) + 16;
  if (n <= 64)
    hipLaunchKernelGGL(k_slogdet<64>, dim3(batch), dim3(64), lds, st, n, A, sign, logabs);
  else
    hipLaunchKernelGGL(k_slogdet<512>, dim3(batch), dim3(512), lds, st, n, A, sign, logabs);
  NH_LAUNCH_CHECK();
  return NH_OK;
}

extern "C" int nh_pairing_occ(void* stream, int32_t LX, int32_t LY, int32_t d, int32_t batch, int32_t n,
                              const float* Y, const int32_t* idx, const float* Wt, float* F) {
  if (batch < 0 || n < 0 || d <= 0 || LX < 1 || LY < 1) return NH_ERR_ARG;
  if (batch == 0 || n == 0) return NH_OK;
  if (!Y || !idx || !Wt || !F) return NH_ERR_ARG;
  if (batch > 65535) return NH_ERR_SHAPE;
  const int threads = 256;
  dim3 grid((n * n + threads - 1) / threads, batch);
  hipLaunchKernelGGL(k_pairing_occ, grid, dim3(threads), 0, (hipStream_t)stream, LX, LY, d, n, Y, idx, Wt, F);
  NH_LAUNCH_CHECK();
  return NH_OK;
}

extern "C" int nh_pairing_ud(void* stream, int32_t LX, int32_t LY, int32_t d, int32_t batch, int32_t n_up,
                             int32_t n_dn, const float* Yu, const float* Yd, const int32_t* iu,
                             const int32_t* idn, const float* Wt, float* F) {
  if (batch < 0 || n_up < 0 || n_dn < 0 || d <= 0 || LX < 1 || LY < 1) return NH_ERR_ARG;
  if (batch == 0 || n_up == 0 || n_dn == 0) return NH_OK;
  if (!Yu || !Yd || !iu || !idn || !Wt || !F) return NH_ERR_ARG;
  if (batch > 65535) return NH_ERR_SHAPE;
  const int threads = 256;
  dim3 grid((n_up * n_dn + threads - 1) / threads, batch);
  hipLaunchKernelGGL(k_pairing_ud, grid, dim3(threads), 0, (hipStream_t)stream, LX, LY, d, n_up, n_dn, Yu,
                     Yd, iu, idn, Wt, F);
  NH_LAUNCH_CHECK();
  return NH_OK;
}
