// Pairing blocks, skew Pfaffian and determinant on gfx950 (FP32 arithmetic, FP64 log-sum).
//
// One workgroup owns one matrix.  The Pfaffian image kernel for 100 <= n <= 256 (k_image_pf_cm)
// builds each image as a full skew square, column-major, in a global workspace and eliminates it
// with pf_lazy (lazy pivoting, compacted Schur complements; nh_pf_lazy.h), and so does the
// workspace entry nh_slogpf_ws for 100 <= n <= 352 (k_slogpf_cm).  nh_slogpf, the other sizes of
// nh_slogpf_ws and the other Pfaffian image kernels keep the packed strict lower triangle (row i
// at i(i-1)/2), in LDS up to n = 256 and in a global workspace above (pf_blocked2 works on
// either).  A determinant matrix lives in LDS as a dense row-major square.
//   n <= 64:  pf_eliminate / lu_eliminate, one wave, rank-2 (rank-1) updates per step.
//   n > 64:   pf_blocked2 (blocked Parlett-Reid: left-looking panels, packed pivot keys, two
//             barriers per step, FP32 MFMA trailing updates) and lu_blocked (blocked LU with
//             rows in place, one barrier per step, MFMA trailing updates).  The BCS image kernel
//             keeps its matrix in LDS up to n = 192 (unless that leaves one workgroup per CU) and
//             in a global workspace otherwise (n <= 512).
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

// Wave-wide max / min of a 32-bit value: DPP within rows of 16 lanes, readlane across rows.
#define NH_DPP(v, ctrl) ((uint32_t)__builtin_amdgcn_update_dpp(0, (int)(v), ctrl, 0xf, 0xf, false))
__device__ inline uint32_t wave_max_u32(uint32_t v) {
  v = max(v, NH_DPP(v, 0xB1));   // quad_perm [1,0,3,2]
  v = max(v, NH_DPP(v, 0x4E));   // quad_perm [2,3,0,1]
  v = max(v, NH_DPP(v, 0x124));  // row_ror:4
  v = max(v, NH_DPP(v, 0x128));  // row_ror:8
  uint32_t a = __builtin_amdgcn_readlane(v, 0), b = __builtin_amdgcn_readlane(v, 16);
  uint32_t c = __builtin_amdgcn_readlane(v, 32), d = __builtin_amdgcn_readlane(v, 48);
  return max(max(a, b), max(c, d));
}
// Sum over each row of 16 lanes of a float, lane 0's row taken for the whole wave (uniform):
// ((Q0 + Q3) + (Q2 + Q1)) of the quad sums Q_k = (v_4k + v_4k+1) + (v_4k+2 + v_4k+3).
#define NH_DPPF(v, ctrl) __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), ctrl, 0xf, 0xf, false))
__device__ inline float row16_sum_f32(float v) {
  v += NH_DPPF(v, 0xB1);   // quad_perm [1,0,3,2]
  v += NH_DPPF(v, 0x4E);   // quad_perm [2,3,0,1]
  v += NH_DPPF(v, 0x124);  // row_ror:4
  v += NH_DPPF(v, 0x128);  // row_ror:8
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 0));
}
__device__ inline int wave_min_i32(int v) {
  v = min(v, (int)NH_DPP(v, 0xB1));
  v = min(v, (int)NH_DPP(v, 0x4E));
  v = min(v, (int)NH_DPP(v, 0x124));
  v = min(v, (int)NH_DPP(v, 0x128));
  int a = __builtin_amdgcn_readlane(v, 0), b = __builtin_amdgcn_readlane(v, 16);
  int c = __builtin_amdgcn_readlane(v, 32), d = __builtin_amdgcn_readlane(v, 48);
  return min(min(a, b), min(c, d));
}

// argmax_i |col(i)| over i in [lo, n), smallest index on ties; wave-uniform result (n if empty).
template <class ColFn>
__device__ inline int wave_pivot(int lo, int n, ColFn col) {
  const int lane = threadIdx.x & (kWave - 1);
  uint32_t best = 0;
  int bi = n;
  for (int i = lo + lane; i < n; i += kWave) {
    const uint32_t u = __float_as_uint(fabsf(col(i)));
    if (bi == n || u > best) {
      best = u;
      bi = i;
    }
  }
  const uint32_t m = wave_max_u32(bi < n ? best : 0u);
  return wave_min_i32((bi < n && best == m) ? bi : n);
}

// Running product of |pivots| as an FP64 mantissa in [0.5, 1) and an integer exponent.
struct LogAcc {
  double mant = 1.0;
  int expo = 0;
  __device__ void mul(float x) {
    int e;
    mant = frexp(mant * (double)fabsf(x), &e);
    expo += e;
  }
  __device__ double log_value() const { return log(mant) + (double)expo * 0.6931471805599453; }
};

constexpr int kMaxColChunks = 4;  // n <= 256 = 4 x 64 columns per lane
constexpr int kRowBatch = 8;      // rows whose loads are in flight together

// Parlett-Reid on the packed lower triangle `tri` (n even, n <= 256). Result valid in thread 0.
// Parlett-Reid on the packed lower triangle `tri` (n even, n <= 256).  `tri` must be followed
// by kWave dummy floats, tau / vv must hold n + kWave floats, and `part` 2 * (NT / kWave) words.
// Addresses of lanes with no element to touch are redirected to the dummy / slack entries,
// so the step code has no divergent branches (each divergent branch costs an LDS drain).
// Result valid in thread 0.
//
// Step k (pivot column k, a = k + 1), two barrier-separated phases on all waves:
//   T: every wave reduces the per-wave partial maxima of |A[i][k]| left by the previous
//      phase U to the same pivot row b; one lane per row i >= k + 2 reads the column pair of
//      the permuted matrix A' = P A P (P the transposition (a b)) straight from the
//      unpermuted triangle, writes tau_i = A'[i][k] / A'[a][k], v_i = A'[i][a] and stores the
//      exchanged entry of row/column b.
//   U: A[i][j] += tau_i v_j - v_i tau_j for i > j >= k + 2; wave w owns rows i = w (mod NW),
//      a lane the columns k + 2 + lane + 64 c.  The owner of row i also holds the new A[i][k+2]
//      and keeps the running argmax of the next pivot column.
template <int NT>
__device__ void pf_eliminate(float* tri, int n, float* tau, float* vv, uint32_t* part, float& sign_out,
                             double& log_out) {
  constexpr int NW = NT / kWave;
  constexpr int RB = 4;  // rows per group in phase U
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int ntri = (n * (n - 1)) / 2;
  const int dummy = ntri + lane;
  {  // partial maxima of column 0 over rows >= 1
    uint32_t best = 0;
    int bi = n;
    for (int i = 1 + wave; i < n; i += NW) {
      const uint32_t u = __float_as_uint(fabsf(tri[toff(i, 0)]));
      if (bi == n || u > best) {
        best = u;
        bi = i;
      }
    }
    if (lane == 0) {
      part[2 * wave] = best;
      part[2 * wave + 1] = (uint32_t)bi;
    }
  }
  __syncthreads();
  LogAcc acc;
  float sign = 1.f;
  bool singular = false;
  for (int k = 0; k + 1 < n; k += 2) {
    const int a = k + 1;
    // ---- phase T
    const uint32_t pv = lane < NW ? part[2 * lane] : 0u;
    const int pidx = lane < NW ? (int)part[2 * lane + 1] : n;
    const uint32_t m = wave_max_u32(pidx < n ? pv : 0u);
    const int b0 = wave_min_i32((pidx < n && pv == m) ? pidx : n);
    const int bb = b0 < n ? b0 : a;
    const float piv = tri[toff(bb, k)];  // A'[a][k] = A[b][k]
    if (!(piv != 0.f) || !isfinite(piv)) {
      singular = true;
      break;
    }
    if (threadIdx.x == 0) {
      if (bb != a) sign = -sign;
      sign = (piv < 0.f) ? sign : -sign;  // Pf *= A[k][k+1] = -piv
      acc.mul(piv);
    }
    const float rp = 1.0f / piv;
    for (int t = threadIdx.x; t < n - k - 2; t += NT) {
      const int i = k + 2 + t;                    // i >= k + 2 > a
      const int pi = i == bb ? a : i;            // P(i)
      const int hi = pi > bb ? pi : bb, lo = pi > bb ? bb : pi;
      const float ck = tri[toff(pi, k)];          // A'[i][k] = A[P i][k]
      const float cav = tri[toff(hi, lo)];        // A'[i][a] = A[P i][b]
      const float ra = tri[toff(i, a)];           // old A[i][a]
      tau[i] = ck * rp;
      vv[i] = pi > bb ? cav : -cav;
      const bool live = bb != a && i != bb;       // new A[i][b] = old A[i][a]
      const int ad = live ? (i > bb ? toff(i, bb) : toff(bb, i)) : dummy;
      tri[ad] = i > bb ? ra : -ra;
    }
    __syncthreads();
    // ---- phase U
    float tj[kMaxColChunks], vj[kMaxColChunks];
#pragma unroll
    for (int c = 0; c < kMaxColChunks; ++c) {
      const int j = k + 2 + lane + c * kWave;
      const int jc = j < n ? j : n + lane;
      tj[c] = tau[jc];
      vj[c] = vv[jc];
    }
    uint32_t best = 0;
    int bi = n;
    for (int i0 = k + 3 + wave; i0 < n; i0 += NW * RB) {
      float ti[RB], vi[RB];
      int base[RB], len[RB];
#pragma unroll
      for (int r = 0; r < RB; ++r) {
        const int i = i0 + r * NW;
        const int ic = i < n ? i : n;  // row n: past the end, len 0 -> dummy only
        ti[r] = tau[ic < n ? ic : n + lane];
        vi[r] = vv[ic < n ? ic : n + lane];
        base[r] = toff(ic < n ? ic : 0, k + 2) + lane;
        len[r] = ic < n ? ic - (k + 2) : 0;
      }
      int lmax = len[0];
#pragma unroll
      for (int r = 1; r < RB; ++r) lmax = len[r] > lmax ? len[r] : lmax;
#pragma unroll
      for (int c = 0; c < kMaxColChunks; ++c) {
        if (c * kWave >= lmax) break;  // uniform
        int ad[RB];
        float x[RB];
#pragma unroll
        for (int r = 0; r < RB; ++r) {
          ad[r] = (c * kWave + lane < len[r]) ? base[r] + c * kWave : dummy;
          x[r] = tri[ad[r]];
        }
#pragma unroll
        for (int r = 0; r < RB; ++r) {
          x[r] = fmaf(ti[r], vj[c], fmaf(-vi[r], tj[c], x[r]));
          tri[ad[r]] = x[r];
        }
        if (c == 0) {
          // lane 0 of chunk 0 holds the new A[i][k+2]: next pivot column
#pragma unroll
          for (int r = 0; r < RB; ++r) {
            const int i = i0 + r * NW;
            const uint32_t u = __float_as_uint(fabsf(x[r]));
            if (lane == 0 && i < n && (bi == n || u > best)) {
              best = u;
              bi = i;
            }
          }
        }
      }
    }
    if (lane == 0) {
      part[2 * wave] = best;
      part[2 * wave + 1] = (uint32_t)bi;
    }
    __syncthreads();
  }
  if (singular) {
    sign_out = 0.f;
    log_out = -INFINITY;
  } else {
    sign_out = sign;
    log_out = acc.log_value();
  }
}

typedef float nh_f32x4 __attribute__((ext_vector_type(4)));

#include "nh_pf_blocked.h"

// LU with partial pivoting on the dense row-major M (leading dimension ld, n <= 256); M must be
// followed by kWave dummy floats at n * ld.
template <int NT>
__device__ void lu_eliminate(float* M, int n, int ld, float* lcol, float* urow, int* flag,
                             float& sign_out, double& log_out) {
  constexpr int NW = NT / kWave;
  const int t = threadIdx.x, lane = t & (kWave - 1), wave = t / kWave;
  LogAcc acc;
  float sign = 1.f;
  for (int k = 0; k < n; ++k) {
    if (wave == 0) {
      const int bi = wave_pivot(k, n, [&](int i) { return M[i * ld + k]; });
      if (bi != k && bi < n)
        for (int j = k + lane; j < n; j += kWave) {
          float t1 = M[k * ld + j];
          M[k * ld + j] = M[bi * ld + j];
          M[bi * ld + j] = t1;
        }
      wave_sync();
      const float piv = M[k * ld + k];
      const bool bad = !(piv != 0.f) || !isfinite(piv);
      const float rp = 1.0f / piv;
      if (!bad) {
        for (int i = k + 1 + lane; i < n; i += kWave) lcol[i] = M[i * ld + k] * rp;
        for (int j = k + 1 + lane; j < n; j += kWave) urow[j] = M[k * ld + j];
      }
      if (lane == 0) {
        if (bi != k) sign = -sign;
        if (bad) {
          flag[0] = 1;
        } else {
          sign = (piv < 0.f) ? -sign : sign;
          acc.mul(piv);
        }
      }
    }
    __syncthreads();
    if (flag[0]) break;
    float uj[kMaxColChunks];
#pragma unroll
    for (int c = 0; c < kMaxColChunks; ++c) {
      const int j = k + 1 + lane + c * kWave;
      uj[c] = (j < n) ? urow[j] : 0.f;
    }
    const int dummy = n * ld + lane;  // branch-free stores, as in pf_eliminate
    for (int i0 = k + 1 + wave; i0 < n; i0 += NW * kRowBatch) {
      float x[kRowBatch][kMaxColChunks], li[kRowBatch];
      int ad[kRowBatch][kMaxColChunks];
#pragma unroll
      for (int r = 0; r < kRowBatch; ++r) {
        const int i = i0 + r * NW;
        li[r] = lcol[i < n ? i : n - 1];
#pragma unroll
        for (int c = 0; c < kMaxColChunks; ++c) {
          const int j = k + 1 + lane + c * kWave;
          ad[r][c] = (i < n && j < n) ? i * ld + j : dummy;
          x[r][c] = M[ad[r][c]];
        }
      }
#pragma unroll
      for (int r = 0; r < kRowBatch; ++r)
#pragma unroll
        for (int c = 0; c < kMaxColChunks; ++c) M[ad[r][c]] = fmaf(-li[r], uj[c], x[r][c]);
    }
    __syncthreads();
  }
  if (flag[0]) {
    sign_out = 0.f;
    log_out = -INFINITY;
  } else {
    sign_out = sign;
    log_out = acc.log_value();
  }
}

#include "nh_lu_blocked.h"

#include "nh_pf_lazy.h"

#include "nh_image_pf.h"

// ------------------------------------------------------------------ Pfaffian image kernel

constexpr int kPanel = 8;  // elimination steps per panel of the blocked Pfaffian

// Standalone nh_slogpf keeps the triangle in LDS: 8-step panels, or 4 above n = 240 where the
// triangle leaves less LDS for the panel vectors.
__host__ __device__ inline int pf_panel_lds(int n) { return n > 240 ? 4 : kPanel; }
__host__ __device__ inline int pf_elim_floats_lds(int n) {
  return pf_panel_lds(n) == kPanel ? pf2_scratch_floats<kPanel>(n) : pf2_scratch_floats<4>(n);
}

// Image kernels above n = 64 keep the triangle in a global workspace: LDS then holds only the
// pairing build, the panel vectors and the mode lists, so four 256-thread workgroups share a CU
// and hide each other's pivot-step latency (measured 0.44 us per 224 x 224 image against
// 0.58 us with the triangle in LDS and one 1024-thread workgroup per CU).
constexpr int kLdsTriMaxN = 64;
// BCS image matrices (n x (n + 1) FP32) stay in LDS up to n = 192 (145 KB); above, up to
// kGlobalDetMaxN (one LU thread per row of a 512-thread workgroup), in the global workspace.
constexpr int kLdsDetMaxN = 192;
constexpr int kGlobalDetMaxN = 512;

// floats of the region used first by the pairing build (features of the occupied modes, one
// Wt spin block) and then by the blocked elimination (panel vectors, column buffer, pivots)
__host__ __device__ inline int pf_scratch_floats(int n, int N, int D) {
  const int build = n * (D + 4) + N * D;
  const int elim = pf2_scratch_floats<kPanel>(n);
  return ((build > elim ? build : elim) + 3) & ~3;
}

// GTRI (n > kLdsTriMaxN): the triangle of image img sits in the global workspace gtri at
// (img - img0) * gstride floats (followed by kWave dummy floats).
template <int D, int NT, bool GTRI>
__global__ void __launch_bounds__(NT) k_image_pf(int N, int LX, int LY, int n, int n_up,
                                                 const int8_t* __restrict__ s,
                                                 const int32_t* __restrict__ item_b,
                                                 const int32_t* __restrict__ item_hop, int img0,
                                                 int img_end, int ipi, const float* __restrict__ Y,
                                                 const float* __restrict__ wt,
                                                 const float* __restrict__ F0,
                                                 float* __restrict__ img_sign,
                                                 double* __restrict__ img_logabs,
                                                 float* __restrict__ gtri, size_t gstride) {
  extern __shared__ float lds[];
  constexpr int YS = D + 4;  // padded row stride of the occupied features
  const int img = img0 + blockIdx.x;
  if (img >= img_end) return;
  const int item = img / ipi;
  const int g = (ipi == kGroup) ? (img & 7) : 0;
  const int32_t hop = item_hop ? item_hop[item] : kHopNone;
  if (hop == kHopInvalid) return;
  const int b = item_b ? item_b[item] : item;
  const int two_n = 2 * N;

  const int ntri = (n * (n - 1)) / 2;
  float* tri;
  float* scr;  // pairing build, then blocked elimination
  if constexpr (GTRI) {
    tri = gtri + (size_t)(img - img0) * gstride;
    scr = lds;
  } else {
    tri = lds;
    scr = tri + ((ntri + kWave + 3) & ~3);
  }
  float* yocc = scr;
  float* wtl = yocc + n * YS;
  float* tau = scr + pf_scratch_floats(n, N, D);
  float* vv = tau + (GTRI ? 0 : n + kWave);  // pf_eliminate only (n <= 64)
  int* modes = reinterpret_cast<int*>(vv + (GTRI ? 0 : n + kWave));
  int* coord = modes + n;  // (x << 16) | y of each mode's site
  int* flag = coord + n;
  uint32_t* part = reinterpret_cast<uint32_t*>(flag + 4);  // 2 * 16 + 2 words

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
    *reinterpret_cast<float4*>(yocc + p * YS + 4 * a4) =
        reinterpret_cast<const float4*>(Yi + (size_t)modes[p] * D)[a4];
  }
  for (int p = t; p < n; p += NT) {
    const int m = modes[p];
    const int site = m - (m >= N ? N : 0);
    coord[p] = ((site / LY) << 16) | (site % LY);
  }

  // Three regions of the lower triangle, each with one spin block of Wt staged in LDS:
  // the first n_up modes carry spin s0 (flipped by SpinInverse), the rest spin s1.
  const int s0 = (g >> 2) & 1, s1 = s0 ^ 1;
  constexpr int NW = NT / kWave;
  for (int region = 0; region < 3; ++region) {
    const int blk = region == 0 ? 3 * s0 : (region == 1 ? 2 * s1 + s0 : 3 * s1);
    const int p0 = region == 0 ? 1 : (region == 1 ? n_up : n_up + 1);
    const int p1 = region == 0 ? n_up : n;
    const int q0 = region == 2 ? n_up : 0;
    __syncthreads();  // previous region done with wtl
    const float4* src = reinterpret_cast<const float4*>(wt + (size_t)blk * N * D);
    for (int e = t; e < N * D / 4; e += NT) reinterpret_cast<float4*>(wtl)[e] = src[e];
    __syncthreads();
    for (int p = p0 + wave; p < p1; p += NW) {
      float yp[D];
#pragma unroll
      for (int a = 0; a < D; ++a) yp[a] = yocc[p * YS + a];
      const int mp = modes[p];
      const int sxp = coord[p] >> 16, syp = coord[p] & 0xffff;
      float* row = tri + toff(p, 0);
      const int qend = region == 1 ? n_up : p;
      // groups of up to 256 columns; within a group all chunks are computed before the first
      // store (loads overlap)
      for (int cq = q0; cq < qend; cq += kMaxColChunks * kWave) {
        float accs[kMaxColChunks];
#pragma unroll
        for (int c = 0; c < kMaxColChunks; ++c) {
          const int q = cq + lane + c * kWave;
          float acc = 0.f;
          if (q < qend) {
            const int mq = modes[q];
            int dx = sxp - (coord[q] >> 16), dy = syp - (coord[q] & 0xffff);
            dx += dx < 0 ? LX : 0;
            dy += dy < 0 ? LY : 0;
            const int r = dx * LY + dy;
            const float4* w4 = reinterpret_cast<const float4*>(wtl + r * D);
            const float4* y4 = reinterpret_cast<const float4*>(yocc + q * YS);
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
          }
          accs[c] = acc;
        }
#pragma unroll
        for (int c = 0; c < kMaxColChunks; ++c) {
          const int q = cq + lane + c * kWave;
          if (q < qend) row[q] = accs[c];
        }
      }
    }
  }
  __syncthreads();

  float sg;
  double la;
  static_assert(GTRI || NT == kWave, "the LDS triangle is for one-wave eliminations (n <= 64)");
  if constexpr (GTRI)
    pf_blocked2<NT, kPanel, 10>(tri, n, scr, sg, la);
  else
    pf_eliminate<NT>(tri, n, tau, vv, part, sg, la);
  if (t == 0) {
    // fermion_inverse_sign: reversal of the Ne creation operators
    sg *= parity_sign((long long)n * (n - 1) / 2);
    img_sign[img] = sg;
    img_logabs[img] = la;
  }
}

// Column-major full-square form (any 64 < n <= 256 with n_up % 4 == 0; dispatched from
// n = kCmMinN on): pf_lazy with 16-step panels, the square in the global workspace (200 KB per
// image at n = 224).
constexpr int kPanelCm = 16;
// smallest n of the column forms (image kernel and nh_slogpf_ws): per matrix of nh_slogpf_ws
// 0.096 / 0.116 us against 0.111 / 0.150 for the packed workspace triangle at n = 114 / 130, but
// 0.083 against 0.077 at n = 98; per 16x16-style configuration (8 images) 0.928 against 0.996 us
// at n = 128 and 0.685 against 0.694 at n = 96
constexpr int kCmMinN = 100;

template <int D>
__global__ void __launch_bounds__(256, 4) k_image_pf_cm(int N, int LX, int LY, int n, int n_up,
                                                     const int8_t* __restrict__ s,
                                                     const int32_t* __restrict__ item_b,
                                                     const int32_t* __restrict__ item_hop, int img0,
                                                     int img_end, int ipi, const float* __restrict__ Y,
                                                     const float* __restrict__ wt,
                                                     const float* __restrict__ F0,
                                                     float* __restrict__ img_sign,
                                                     double* __restrict__ img_logabs,
                                                     float* __restrict__ gmat, size_t gstride) {
  extern __shared__ float lds[];
  const int img = img0 + blockIdx.x;
  if (img >= img_end) return;
  image_pf_cm<D, 256, kPanelCm, 8>(img, N, LX, LY, n, n_up, s, item_b, item_hop, img0, ipi, Y, wt, F0,
                                   img_sign, img_logabs, gmat, gstride, lds);
}

// ------------------------------------------------------------------ BCS image kernel

// GM (bcs_global_matrix): the image's n x (n + 1) matrix sits in the global workspace gmat at
// (img - img0) * gstride floats (followed by kWave dummy floats); LDS holds the LU scratch, the
// features of the down modes (each row's up-mode features are read from Y), the mode list and a
// flag.
template <int D, int NT, bool GM>
__global__ void __launch_bounds__(NT) k_image_bcs(int N, int LX, int LY, int n_up, int n_dn,
                                                  const int8_t* __restrict__ s,
                                                  const int32_t* __restrict__ item_b,
                                                  const int32_t* __restrict__ item_hop, int img0,
                                                  int img_end, int ipi, const float* __restrict__ Y,
                                                  const float* __restrict__ wt,
                                                  const float* __restrict__ phi0,
                                                  float* __restrict__ img_sign,
                                                  double* __restrict__ img_logabs,
                                                  float* __restrict__ gmat, size_t gstride) {
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

  float* M;
  float* yu;   // staged features of the up modes (LDS form; GM reads them from Y)
  float* scr;  // blocked-LU scratch
  float* yd;
  if constexpr (GM) {
    M = gmat + (size_t)(img - img0) * gstride;
    yu = nullptr;
    scr = lds;
    yd = lds + ((lu_scratch_words(n) + 3) & ~3);
  } else {
    M = lds;
    yu = M + ((n * ld + kWave + 3) & ~3);
    scr = yu;  // the features are dead once the matrix is built
    yd = yu + n * D;
  }
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
    if constexpr (!GM)
      reinterpret_cast<float4*>(yu)[e] = reinterpret_cast<const float4*>(Yi + (size_t)rows[p] * D)[a4];
    reinterpret_cast<float4*>(yd)[e] = reinterpret_cast<const float4*>(Yi + (size_t)cols[p] * D)[a4];
  }
  __syncthreads();
  constexpr int NW = NT / kWave;
  for (int p = wave; p < n; p += NW) {
    float yp[D];
#pragma unroll
    for (int a = 0; a < D; ++a) yp[a] = GM ? Yi[(size_t)rows[p] * D + a] : yu[p * D + a];
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
  if constexpr (GM)
    lu_blocked<NT, kLuPanel, 9>(M, n, ld, reinterpret_cast<uint32_t*>(scr), sg, la);
  else if constexpr (NT > kWave)
    lu_blocked<NT, kLuPanel>(M, n, ld, reinterpret_cast<uint32_t*>(scr), sg, la);
  else
    lu_eliminate<NT>(M, n, ld, lcol, urow, flag, sg, la);
  if (t == 0) {
    sg *= parity_sign((long long)n * (n - 1) / 2);  // Pf [[0,Phi],[-Phi^T,0]] = (-1)^{n(n-1)/2} det
    if (si) sg *= parity_sign(n);                   // spin-down-first image: det(-Phi^T)
    sg *= parity_sign((long long)ne * (ne - 1) / 2);
    img_sign[img] = sg;
    img_logabs[img] = la;
  }
}

static size_t pf_lds_bytes(int n, int N, int D) {
  const int ntri = (n * (n - 1)) / 2;
  if (n > kLdsTriMaxN)
    return sizeof(float) * (size_t)pf_scratch_floats(n, N, D) + sizeof(int) * (size_t)(2 * n + 4 + 2 * 16 + 2);
  return sizeof(float) * (size_t)(((ntri + kWave + 3) & ~3) + pf_scratch_floats(n, N, D) + 2 * (n + kWave)) +
         sizeof(int) * (size_t)(2 * n + 4 + 2 * 16 + 2);
}

static size_t bcs_lds_bytes(int n, int ne, int D, bool gm) {
  // GM: LU scratch + the down modes' features; otherwise the matrix + both spins' features
  const size_t front = gm ? (size_t)((lu_scratch_words(n) + 3) & ~3) + (size_t)n * D
                          : (size_t)((n * (n + 1) + kWave + 3) & ~3) + (size_t)2 * n * D;
  return sizeof(float) * (front + (size_t)(2 * n)) + sizeof(int) * (size_t)(ne + 4);
}

// BCS determinant matrix in the global workspace rather than in LDS: above n = 192, and where the
// LDS-resident form would leave room for only one workgroup per CU (the d = 64 features of the
// released BCS state: 2.75 us per 16x16 image with the matrix in the workspace against 3.02 us;
// at d = 16 the LDS form keeps two workgroups per CU and is faster, 0.86 against 1.02 us).
static bool bcs_global_matrix(const nh_model* m) {
  const int n = m->n_up;
  if (n > kLdsDetMaxN) return true;
  return n > 64 && 2 * bcs_lds_bytes(n, m->n_up + m->n_dn, feature_dim(m), false) > 160 * 1024;
}

int64_t image_chunk(const nh_model* m) {
  if (m->kind == NH_KIND_BCS)
    return !bcs_global_matrix(m) || m->n_up <= 256 ? 32768 : 1024;  // 0.8 MB per image at n = 448
  const int n = m->n_up + m->n_dn;
  if (n <= kLdsTriMaxN) return 32768;
  // each image of a chunk also keeps its matrix (the square and its first compacted complement,
  // 347 KB at n = 224; the packed triangle, 1.6 MB, at n = 896): 32768 images per launch at
  // n <= 256 (11.4 GB of workspace at n = 224)
  return n <= 256 ? 32768 : 1024;
}

constexpr size_t kMaxLds = 160 * 1024;

static size_t pf_cm_lds(int n, int N, int D) {
  switch (D) {
    case 4: return pf_cm_lds_bytes<4, 256, kPanelCm>(n, N);
    case 8: return pf_cm_lds_bytes<8, 256, kPanelCm>(n, N);
    case 16: return pf_cm_lds_bytes<16, 256, kPanelCm>(n, N);
    case 32: return pf_cm_lds_bytes<32, 256, kPanelCm>(n, N);
    case 64: return pf_cm_lds_bytes<64, 256, kPanelCm>(n, N);
    default: return kMaxLds + 1;
  }
}

// Pfaffian images that run k_image_pf_cm; the others above n = 64 run k_image_pf on the packed
// triangle.  The column form is the faster one from n = 160 on (per image: 0.190 against 0.195
// us at n = 160, 0.239 against 0.283 at 192, 0.337 against 0.408 at 224) and the slower one up to
// n = 144 (0.080 against 0.070 us at n = 80, 0.163 against 0.159 at 144).  Above n = 256 (the
// 32x32 lattice, n = 896) a 1024-thread column form measured no faster than the packed triangle
// (20.3 against 20.1 us per image) for 4x its workspace.
static bool pf_column_form(const nh_model* m) {
  const int n = m->n_up + m->n_dn;
  return m->kind == NH_KIND_PFAFFIAN && n >= kCmMinN && n <= 256 && n % 2 == 0 && m->n_up % 4 == 0 &&
         pf_cm_lds(n, m->LX * m->LY, feature_dim(m)) <= kMaxLds;
}

size_t image_global_floats(const nh_model* m) {
  if (m->kind == NH_KIND_BCS) {
    const size_t n = (size_t)m->n_up;
    return !bcs_global_matrix(m) ? 0 : ((n * (n + 1) + kWave + 63) & ~(size_t)63);
  }
  const int n = m->n_up + m->n_dn;
  if (n <= kLdsTriMaxN) return 0;
  if (pf_column_form(m)) return pfl_matrix_floats<kPanelCm>(n);
  return (size_t)((((size_t)n * (n - 1)) / 2 + kWave + 63) & ~(size_t)63);
}

template <int D>
static int launch_image_d(hipStream_t st, const nh_model* m, const int8_t* s, const int32_t* item_b,
                          const int32_t* item_hop, int img0, int img_end, int ipi, const float* Y,
                          float* img_sign, double* img_logabs, float* gtri) {
  const int N = m->LX * m->LY;
  const int nimg = img_end - img0;
  if (nimg <= 0) return NH_OK;
  if (m->kind == NH_KIND_PFAFFIAN) {
    const int n = m->n_up + m->n_dn;
    const size_t lds = pf_lds_bytes(n, N, D);
    if (lds > kMaxLds) return NH_ERR_SHAPE;
    if (n <= 64) {
      hipLaunchKernelGGL((k_image_pf<D, 64, false>), dim3(nimg), dim3(64), lds, st, N, m->LX, m->LY, n, m->n_up,
                         s, item_b, item_hop, img0, img_end, ipi, Y, m->wt, m->base, img_sign, img_logabs,
                         nullptr, (size_t)0);
    } else if (!gtri) {
      return NH_ERR_ARG;
    } else if (pf_column_form(m)) {
      hipLaunchKernelGGL((k_image_pf_cm<D>), dim3(nimg), dim3(256), pf_cm_lds(n, N, D), st, N, m->LX, m->LY, n,
                         m->n_up, s, item_b, item_hop, img0, img_end, ipi, Y, m->wt, m->base, img_sign,
                         img_logabs, gtri, image_global_floats(m));
    } else if (n <= 256) {
      hipLaunchKernelGGL((k_image_pf<D, 256, true>), dim3(nimg), dim3(256), lds, st, N, m->LX, m->LY, n,
                         m->n_up, s, item_b, item_hop, img0, img_end, ipi, Y, m->wt, m->base, img_sign,
                         img_logabs, gtri, image_global_floats(m));
    } else {
      hipLaunchKernelGGL((k_image_pf<D, 1024, true>), dim3(nimg), dim3(1024), lds, st, N, m->LX, m->LY, n,
                         m->n_up, s, item_b, item_hop, img0, img_end, ipi, Y, m->wt, m->base, img_sign,
                         img_logabs, gtri, image_global_floats(m));
    }
  } else {
    const int n = m->n_up;
    const bool gm = bcs_global_matrix(m);
    const size_t lds = bcs_lds_bytes(n, m->n_up + m->n_dn, D, gm);
    if (lds > kMaxLds) return NH_ERR_SHAPE;
    if (n <= 64) {
      hipLaunchKernelGGL((k_image_bcs<D, 64, false>), dim3(nimg), dim3(64), lds, st, N, m->LX, m->LY, m->n_up,
                         m->n_dn, s, item_b, item_hop, img0, img_end, ipi, Y, m->wt, m->base, img_sign,
                         img_logabs, nullptr, (size_t)0);
    } else if (!gm) {
      hipLaunchKernelGGL((k_image_bcs<D, 512, false>), dim3(nimg), dim3(512), lds, st, N, m->LX, m->LY, m->n_up,
                         m->n_dn, s, item_b, item_hop, img0, img_end, ipi, Y, m->wt, m->base, img_sign,
                         img_logabs, nullptr, (size_t)0);
    } else if (n > kGlobalDetMaxN) {
      return NH_ERR_SHAPE;
    } else if (!gtri) {
      return NH_ERR_ARG;
    } else {
      hipLaunchKernelGGL((k_image_bcs<D, 512, true>), dim3(nimg), dim3(512), lds, st, N, m->LX, m->LY, m->n_up,
                         m->n_dn, s, item_b, item_hop, img0, img_end, ipi, Y, m->wt, m->base, img_sign,
                         img_logabs, gtri, image_global_floats(m));
    }
  }
  NH_LAUNCH_CHECK();
  return NH_OK;
}

int launch_image_amplitude(hipStream_t st, const nh_model* m, const int8_t* s, const int32_t* item_b,
                           const int32_t* item_hop, int img0, int img_end, int ipi, const float* Y,
                           float* img_sign, double* img_logabs, float* gtri) {
  switch (feature_dim(m)) {
    case 4: return launch_image_d<4>(st, m, s, item_b, item_hop, img0, img_end, ipi, Y, img_sign, img_logabs, gtri);
    case 8: return launch_image_d<8>(st, m, s, item_b, item_hop, img0, img_end, ipi, Y, img_sign, img_logabs, gtri);
    case 16:
      return launch_image_d<16>(st, m, s, item_b, item_hop, img0, img_end, ipi, Y, img_sign, img_logabs, gtri);
    case 32:
      return launch_image_d<32>(st, m, s, item_b, item_hop, img0, img_end, ipi, Y, img_sign, img_logabs, gtri);
    case 64:
      return launch_image_d<64>(st, m, s, item_b, item_hop, img0, img_end, ipi, Y, img_sign, img_logabs, gtri);
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
  float* tau = tri + ((ntri + kWave + 3) & ~3);
  float* vv = tau + n + kWave;
  uint32_t* part = reinterpret_cast<uint32_t*>(vv + n + kWave);
  const float* Ab = A + (size_t)blockIdx.x * n * n;
  for (int i = 1 + threadIdx.x / kWave; i < n; i += NT / kWave)
    for (int j = threadIdx.x & (kWave - 1); j < i; j += kWave) tri[toff(i, j)] = Ab[(size_t)i * n + j];
  __syncthreads();
  float sg;
  double la;
  if constexpr (NT > kWave) {
    if (pf_panel_lds(n) == kPanel)
      pf_blocked2<NT, kPanel>(tri, n, tau, sg, la);
    else
      pf_blocked2<NT, 4>(tri, n, tau, sg, la);
  } else {
    pf_eliminate<NT>(tri, n, tau, vv, part, sg, la);
  }
  if (threadIdx.x == 0) {
    sign[blockIdx.x] = sg;
    logabs[blockIdx.x] = la;
  }
}

// Workspace form (64 < n <= NT): matrix b0 + blockIdx.x's strict lower triangle is packed into
// its slice of the workspace (gstride floats, the triangle then kWave dummy floats) and
// eliminated there; LDS holds only the panel scratch, so up to n = 256 four 256-thread
// workgroups share a CU.
template <int NT>
__global__ void __launch_bounds__(NT) k_slogpf_ws(int n, int b0, const float* __restrict__ A, float* sign,
                                                  double* logabs, float* __restrict__ gtri, size_t gstride) {
  extern __shared__ float lds[];
  const int b = b0 + blockIdx.x;
  float* tri = gtri + (size_t)blockIdx.x * gstride;
  const float* Ab = A + (size_t)b * n * n;
  for (int i = 1 + threadIdx.x / kWave; i < n; i += NT / kWave)
    for (int j = threadIdx.x & (kWave - 1); j < i; j += kWave) tri[toff(i, j)] = Ab[(size_t)i * n + j];
  __syncthreads();
  float sg;
  double la;
  pf_blocked2<NT, kPanel, 10>(tri, n, lds, sg, la);
  if (threadIdx.x == 0) {
    sign[b] = sg;
    logabs[b] = la;
  }
}

// Column form of the workspace entry (100 <= n <= 352; per matrix 0.172 against 0.257 us at
// n = 162, 0.297 against 0.536 at 224, 0.427 against 0.753 at 256 for the packed workspace
// triangle): matrix b0 + blockIdx.x's strict lower
// triangle becomes a column-major skew square in its workspace slice (M[j * ld + i] = A[i][j]
// for i > j, -A[j][i] for i < j, zero on the diagonal and in the padding rows n .. ld - 1;
// gstride >= pfl_matrix_floats<P>(n) floats), which pf_lazy eliminates.  The lower
// triangle is read once, in 64 x 64 tiles with all 16 loads of a thread in flight together:
// each tile is stored negated, as it stands, into the upper half (column i of M holds row i of
// the tile) and transposed through LDS into the lower half.  LDS: the larger of the tile
// (64 x 65 floats) and pf_lazy's scratch.
template <int NT, int KB, int P>
__global__ void __launch_bounds__(NT, NT == 256 ? 4 : 2) k_slogpf_cm(int n, int b0, const float* __restrict__ A,
                                                                     float* sign, double* logabs,
                                                                     float* __restrict__ gmat, size_t gstride) {
  extern __shared__ float lds[];
  constexpr int TS = kWave + 1, PER = kWave * kWave / NT;
  const int b = b0 + blockIdx.x;
  const int ld = pfl_ld(n);
  float* M = gmat + (size_t)blockIdx.x * gstride;
  const float* Ab = A + (size_t)b * n * n;
  const int c = threadIdx.x & (kWave - 1), r0 = threadIdx.x / kWave;
  for (int j0 = 0; j0 < n; j0 += kWave)
    for (int i0 = j0; i0 < n; i0 += kWave) {
      float v[PER];
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const int i = i0 + r0 + k * (NT / kWave), j = j0 + c;
        v[k] = i < n && j < i ? Ab[i * n + j] : 0.f;  // 32-bit offsets (n <= 512): one register each
      }
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const int r = r0 + k * (NT / kWave), i = i0 + r, j = j0 + c;
        if (i < n && j < i) M[i * ld + j] = -v[k];
        lds[r * TS + c] = v[k];
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const int jr = r0 + k * (NT / kWave), i = i0 + c, j = j0 + jr;
        if (i < n && j < i) M[j * ld + i] = lds[c * TS + jr];
      }
      __syncthreads();
    }
  for (int j = threadIdx.x; j < n; j += NT) {
    M[j * ld + j] = 0.f;
    for (int i = n; i < ld; ++i) M[j * ld + i] = 0.f;
  }
  __syncthreads();
  float sg;
  double la;
  pf_lazy<NT, P, KB>(M, n, lds, sg, la);
  if (threadIdx.x == 0) {
    sign[b] = sg;
    logabs[b] = la;
  }
}

template <int NT, int P>
static size_t slogpf_cm_lds(int n) {
  const size_t tile = (size_t)kWave * (kWave + 1), scr = (size_t)pfl_scratch_floats<NT, P>(pfl_ld(n));
  return sizeof(float) * (tile > scr ? tile : scr);
}

template <int NT>
__global__ void __launch_bounds__(NT) k_slogdet(int n, const float* __restrict__ A, float* sign,
                                                double* logabs) {
  extern __shared__ float lds[];
  const int ld = n + 1;
  float* M = lds;
  float* lcol = M + ((n * ld + kWave + 3) & ~3);
  float* urow = lcol + n;
  int* flag = reinterpret_cast<int*>(urow + n);
  const float* Ab = A + (size_t)blockIdx.x * n * n;
  for (int e = threadIdx.x; e < n * n; e += NT) M[(e / n) * ld + (e % n)] = Ab[e];
  if (threadIdx.x == 0) flag[0] = 0;
  __syncthreads();
  float sg;
  double la;
  if constexpr (NT > kWave)
    lu_blocked<NT, kLuPanel>(M, n, ld, reinterpret_cast<uint32_t*>(lcol), sg, la);
  else
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

// Tiled form for d in {4, 8, 16, 32, 64}: a block owns kPairRows rows p of one sample with all of
// the sample's occupied features (padded rows) and site coordinates in LDS; its threads sweep q
// so that every F row is written with consecutive addresses, and the Wt row of a pair comes from
// L2 as float4s.  Same arithmetic order as k_pairing_occ (exactly antisymmetric output).
constexpr int kPairRows = 16;

template <int D>
__global__ void __launch_bounds__(256) k_pairing_occ_t(int LX, int LY, int n, const float* __restrict__ Y,
                                                       const int32_t* __restrict__ idx,
                                                       const float* __restrict__ Wt, float* __restrict__ F) {
  extern __shared__ float pt_lds[];
  constexpr int YS = D + 4;
  float* ys = pt_lds;                                   // [n][YS]
  int* coord = reinterpret_cast<int*>(ys + n * YS);     // [n]: spin << 30 | x << 15 | y, -1 if invalid
  const int N = LX * LY;
  const size_t b = blockIdx.y;
  const int t = threadIdx.x;
  const float4* Y4 = reinterpret_cast<const float4*>(Y + b * n * D);
  for (int e = t; e < n * (D / 4); e += 256) {
    const int p = e / (D / 4), a4 = e - p * (D / 4);
    *reinterpret_cast<float4*>(ys + p * YS + 4 * a4) = Y4[e];
  }
  const int32_t* ib = idx + b * n;
  for (int q = t; q < n; q += 256) {
    const int m = ib[q];
    const int sp = m >= N, site = m - sp * N;
    coord[q] = (m < 0 || m >= 2 * N) ? -1 : ((sp << 30) | ((site / LY) << 15) | (site % LY));
  }
  __syncthreads();
  float* Fb = F + b * (size_t)n * n;
  const int p0 = blockIdx.x * kPairRows;
  for (int pr = 0; pr < kPairRows; ++pr) {
    const int p = p0 + pr;
    if (p >= n) break;
    for (int q = t; q < n; q += 256) {
      float val = 0.f;
      if (p != q) {
        const int hi = p > q ? p : q, lo = p > q ? q : p;
        const int ch = coord[hi], cl = coord[lo];
        if (ch < 0 || cl < 0) {
          val = NAN;
        } else {
          int dx = ((ch >> 15) & 0x7fff) - ((cl >> 15) & 0x7fff), dy = (ch & 0x7fff) - (cl & 0x7fff);
          dx += dx < 0 ? LX : 0;
          dy += dy < 0 ? LY : 0;
          const float4* w4 = reinterpret_cast<const float4*>(Wt + ((size_t)(2 * (ch >> 30) + (cl >> 30)) * N +
                                                                   dx * LY + dy) * D);
          const float4* yh = reinterpret_cast<const float4*>(ys + hi * YS);
          const float4* yl = reinterpret_cast<const float4*>(ys + lo * YS);
          float acc = 0.f;
#pragma unroll
          for (int a4 = 0; a4 < D / 4; ++a4) {
            const float4 w = w4[a4], u = yh[a4], v = yl[a4];
            acc = fmaf(u.x * v.x, w.x, acc);
            acc = fmaf(u.y * v.y, w.y, acc);
            acc = fmaf(u.z * v.z, w.z, acc);
            acc = fmaf(u.w * v.w, w.w, acc);
          }
          val = p > q ? acc : -acc;
        }
      }
      Fb[(size_t)p * n + q] = val;
    }
  }
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
  const int small = 2 * (n + kWave) + 2 * 16 + 4;  // pf_eliminate: tau, vv, part
  const int big = pf_elim_floats_lds(n);
  const size_t lds = sizeof(float) * (size_t)(((ntri + kWave + 3) & ~3) + (n <= 64 ? small : big));
  // 256 threads while two or more triangles share a CU (per matrix 0.032 / 0.068 / 0.128 / 0.253
  // us at n = 66 / 98 / 130 / 162 against 0.120 / 0.184 / 0.255 / 0.336 with 1024 threads), one
  // 1024-thread workgroup per CU above (0.436 against 0.660 us at n = 194)
  if (n <= 64)
    hipLaunchKernelGGL(k_slogpf<64>, dim3(batch), dim3(64), lds, st, n, A, sign, logabs);
  else if (n <= 176)
    hipLaunchKernelGGL(k_slogpf<256>, dim3(batch), dim3(256), lds, st, n, A, sign, logabs);
  else
    hipLaunchKernelGGL(k_slogpf<1024>, dim3(batch), dim3(1024), lds, st, n, A, sign, logabs);
  NH_LAUNCH_CHECK();
  return NH_OK;
}

// Workspace floats per matrix: the column-major square and its first compacted complement for
// 100 <= n <= 352 (348 KB at n = 224), the packed triangle otherwise.  Above n = 256 the column
// form runs 512 threads with 12-step panels (109 VGPRs, two workgroups per CU; per matrix 0.619 /
// 0.984 / 1.332 us at n = 258 / 320 / 352 against 0.831 / 1.222 / ~1.5 for the packed triangle in
// batches of 16384).  Above n = 352 it would need 16-step panels (132 registers, one workgroup
// per CU): 1.722 / 2.702 / 3.616 us at n = 384 / 460 / 512 against 1.867 / 2.932 / 3.737 in
// batches of 16384, but slower than the packed triangle in the low-rank hop ratios' batches of
// 512 Pfaffians of n ~ 460 (E_loc 363 against 349 ms per 32x32 sample), so the packed triangle
// keeps those sizes.
constexpr int kCmMaxWsN = 352;
static bool slogpf_column_form(int n) { return n >= kCmMinN && n <= kCmMaxWsN; }
static int slogpf_cm_panel(int n) { return n > 256 ? 12 : kPanelCm; }
static size_t slogpf_ws_stride(int n) {
  if (slogpf_column_form(n))
    return slogpf_cm_panel(n) == 12 ? pfl_matrix_floats<12>(n) : pfl_matrix_floats<kPanelCm>(n);
  return (((size_t)n * (n - 1)) / 2 + kWave + 63) & ~(size_t)63;
}

extern "C" size_t nh_slogpf_workspace_size(int32_t batch, int32_t n) {
  if (batch <= 0 || n <= 0 || n % 2 || n > 1024) return 0;
  if (n <= 64) return 256;
  return sizeof(float) * slogpf_ws_stride(n) * (size_t)batch + 256;
}

extern "C" int nh_slogpf_ws(void* stream, int32_t batch, int32_t n, const float* A, float* sign, double* logabs,
                            void* workspace, size_t workspace_bytes) {
  if (batch < 0 || n < 0 || (batch > 0 && (!A || !sign || !logabs))) return NH_ERR_ARG;
  if (n == 0 || n % 2 || n > 1024) return NH_ERR_SHAPE;
  if (batch == 0) return NH_OK;
  if (n <= 64) return nh_slogpf(stream, batch, n, A, sign, logabs);
  if (!workspace) return NH_ERR_ARG;
  const size_t gs = slogpf_ws_stride(n);
  float* gtri = (float*)(((size_t)workspace + 255) & ~(size_t)255);
  const size_t usable = workspace_bytes > 256 ? workspace_bytes - 256 : 0;
  const int64_t fit = (int64_t)(usable / (sizeof(float) * gs));
  if (fit < 1) return NH_ERR_WORKSPACE;
  const int chunk = fit < batch ? (int)fit : batch;
  hipStream_t st = (hipStream_t)stream;
  const bool cm = slogpf_column_form(n);
  const size_t lds = !cm       ? sizeof(float) * (size_t)pf2_scratch_floats<kPanel>(n)
                     : n <= 256 ? slogpf_cm_lds<256, kPanelCm>(n)
                                : slogpf_cm_lds<512, 12>(n);
  for (int b0 = 0; b0 < batch; b0 += chunk) {
    const int nb = batch - b0 < chunk ? batch - b0 : chunk;
    if (cm && n <= 256)
      hipLaunchKernelGGL((k_slogpf_cm<256, 8, kPanelCm>), dim3(nb), dim3(256), lds, st, n, b0, A, sign, logabs,
                         gtri, gs);
    else if (cm)  // 256 < n <= 352: 512 threads, 12-step panels
      hipLaunchKernelGGL((k_slogpf_cm<512, 10, 12>), dim3(nb), dim3(512), lds, st, n, b0, A, sign, logabs, gtri, gs);
    else if (n <= 256)
      hipLaunchKernelGGL(k_slogpf_ws<256>, dim3(nb), dim3(256), lds, st, n, b0, A, sign, logabs, gtri, gs);
    else
      hipLaunchKernelGGL(k_slogpf_ws<1024>, dim3(nb), dim3(1024), lds, st, n, b0, A, sign, logabs, gtri, gs);
    NH_LAUNCH_CHECK();
  }
  return NH_OK;
}

extern "C" int nh_slogdet(void* stream, int32_t batch, int32_t n, const float* A, float* sign,
                          double* logabs) {
  if (batch < 0 || n <= 0 || (batch > 0 && (!A || !sign || !logabs))) return NH_ERR_ARG;
  if (n > 192) return NH_ERR_SHAPE;
  if (batch == 0) return NH_OK;
  hipStream_t st = (hipStream_t)stream;
  const int small = 2 * n + 4, big = lu_scratch_words(n);
  const size_t lds = sizeof(float) * (size_t)(((n * (n + 1) + kWave + 3) & ~3) + (small > big ? small : big));
  if (n <= 64)
    hipLaunchKernelGGL(k_slogdet<64>, dim3(batch), dim3(64), lds, st, n, A, sign, logabs);
  else if (n <= 128)
    hipLaunchKernelGGL(k_slogdet<256>, dim3(batch), dim3(256), lds, st, n, A, sign, logabs);
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
  hipStream_t st = (hipStream_t)stream;
  const size_t lds = sizeof(float) * ((size_t)n * (d + 4) + n);
  const bool tiled = (d == 4 || d == 8 || d == 16 || d == 32 || d == 64) && lds <= kMaxLds && LX < (1 << 15) &&
                     LY < (1 << 15);
  if (tiled) {
    dim3 grid((n + kPairRows - 1) / kPairRows, batch);
    switch (d) {
      case 4: hipLaunchKernelGGL(k_pairing_occ_t<4>, grid, dim3(threads), lds, st, LX, LY, n, Y, idx, Wt, F); break;
      case 8: hipLaunchKernelGGL(k_pairing_occ_t<8>, grid, dim3(threads), lds, st, LX, LY, n, Y, idx, Wt, F); break;
      case 16: hipLaunchKernelGGL(k_pairing_occ_t<16>, grid, dim3(threads), lds, st, LX, LY, n, Y, idx, Wt, F); break;
      case 32: hipLaunchKernelGGL(k_pairing_occ_t<32>, grid, dim3(threads), lds, st, LX, LY, n, Y, idx, Wt, F); break;
      default: hipLaunchKernelGGL(k_pairing_occ_t<64>, grid, dim3(threads), lds, st, LX, LY, n, Y, idx, Wt, F);
    }
  } else {
    dim3 grid((n * n + threads - 1) / threads, batch);
    hipLaunchKernelGGL(k_pairing_occ, grid, dim3(threads), 0, st, LX, LY, d, n, Y, idx, Wt, F);
  }
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
