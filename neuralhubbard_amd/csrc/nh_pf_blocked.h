// Blocked Parlett-Reid Pfaffian of one packed skew triangle per workgroup (pf_blocked2) and its
// scratch size.  A fragment, not a standalone header: it is included inside namespace nh of
// nh_linalg.hip after kWave, toff, wave_max_u32 and nh_f32x4, and by the host emulation under
// tests/host, which supplies those and the wave-level builtins itself.
#pragma once

// Scratch floats of pf_blocked2: panel vectors row-interleaved UW (n + 16) x 2P and q-major
// UT, WT P x (n + 16) each, column buffer n + 4, pivot keys 16 words, pivots n / 2 + 4.
template <int P>
__host__ __device__ inline int pf2_scratch_floats(int n) {
  return (n + 16) * 4 * P + ((n + 7) & ~3) + 16 + ((n / 2 + 7) & ~3);
}

__device__ inline double wave_sum_f64(double x) {
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) x += __shfl_xor(x, o, kWave);
  return x;
}

// Blocked Parlett-Reid with a short per-step chain (64 < n <= NT, n <= 2^KB, n even), on the packed
// lower triangle `tri` (followed by kWave dummy floats).  Panels of up to P elimination steps:
//   (1) rows i > k: effective column c_i = A[i][k] + sum_{q<s} (U_q[i] W_q[k] - W_q[i] U_q[k])
//       (left-looking over the panel's earlier steps), one 32-bit key per row, |c_i| with its
//       low KB mantissa bits replaced by 2^KB - 1 - i, and a wave max of the keys;
//   (2) every thread reduces the wave keys to the pivot row b (largest |c| to 2^(KB-23) relative,
//       smallest index among those), reads the pivot, and rows i >= k + 2 form the new panel
//       column U_s[i] = A'[i][k] / A'[a][k], W_s[i] = A'[i][a] of A' = P A P (P the
//       transposition (a b)), and move row/column a's triangle entries to b;
//   trailing: A[i][j] += sum_q U_q[i] W_q[j] - W_q[i] U_q[j] for i > j >= k0 + 2P as 16 x 16
//       FP32 MFMA tiles [U | -W] [W | U]^T, each trailing element read and written once.
// Two barriers per step.  The panel vectors are kept twice: row-interleaved (UW[i][2q], UW[i][2q+1],
// one 16-byte load per two steps for the step reads) and q-major (UT[q][i], WT[q][i], conflict-free
// MFMA operands).  The pivot row takes row a's panel entries; others may still read row b's old
// ones in phase (2), so that store waits for phase (1) of the next step, whose readers of the row
// use row a's slot instead (row a is dead by then and also receives the new column entry).
// |pivots| go to `pivs` and their logs are summed in FP64 at the end; the sign is uniform.
template <int NT, int P, int KB = 8>
__device__ void pf_blocked2(float* tri, int n, float* scr, float& sign_out, double& log_out) {
  constexpr uint32_t kIdx = (1u << KB) - 1u;  // low key bits: kIdx - row (n <= 2^KB)
  constexpr int NW = NT / kWave;
  constexpr int LW = 2 * P;
  static_assert(NW <= 16 && NW % 4 == 0, "pivot keys: 4, 8 or 16 waves");
  static_assert(P % 2 == 0, "two steps per 16-byte load");
  const int ldt = n + 16;
  float* UW = scr;
  float* UT = UW + ldt * LW;
  float* WT = UT + P * ldt;
  float* cbuf = WT + P * ldt;
  uint32_t* keys = reinterpret_cast<uint32_t*>(cbuf + ((n + 7) & ~3));
  float* pivs = reinterpret_cast<float*>(keys + 16);
  const int t = threadIdx.x;
  const int lane = t & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(t / kWave);
  const int ntri = (n * (n - 1)) / 2;
  const int dummy = ntri + lane;
  const int row_dummy = n + (lane & 15);  // padding rows of UW / UT / WT, never read
  float sign = 1.f;
  bool singular = false;
  for (int k0 = 0; k0 < n; k0 += 2 * P) {
    const int steps = (n - k0) / 2 < P ? (n - k0) / 2 : P;
    float pu[P], pw[P];  // this thread's row of the panel vectors
#pragma unroll
    for (int q = 0; q < P; ++q) {
      pu[q] = 0.f;
      pw[q] = 0.f;
    }
    int bprev = -1, aprev = -1;  // uniform: exchange of the previous step of this panel
#pragma unroll
    for (int s = 0; s < P; ++s) {
      if (s >= steps) break;  // uniform
      const int k = k0 + 2 * s, a = k + 1;
      // (0) delayed store of the previous pivot row's adopted entries (q < s - 1)
      if (s > 1 && t == bprev && bprev != aprev) {
#pragma unroll
        for (int q = 0; q + 1 < s; ++q) {
          UW[t * LW + 2 * q] = pu[q];
          UW[t * LW + 2 * q + 1] = pw[q];
          UT[q * ldt + t] = pu[q];
          WT[q * ldt + t] = pw[q];
        }
      }
      // (1) effective column k and the wave's pivot key
      const bool row = t > k && t < n;
      float c = 0.f;
      const bool wrow = wave * kWave + kWave - 1 > k && wave * kWave < n;  // uniform
      uint32_t key = 0u;
      if (wrow) {
        c = tri[row ? toff(t, k) : dummy];
        const int kk = (k == bprev && bprev != aprev) ? aprev : k;
        const float4* uwk = reinterpret_cast<const float4*>(UW + kk * LW);
#pragma unroll
        for (int q2 = 0; 2 * q2 < s; ++q2) {
          const float4 x = uwk[q2];  // U_2q2, W_2q2, U_2q2+1, W_2q2+1 of row k
          c = fmaf(pu[2 * q2], x.y, fmaf(-pw[2 * q2], x.x, c));
          if (2 * q2 + 1 < s) c = fmaf(pu[2 * q2 + 1], x.w, fmaf(-pw[2 * q2 + 1], x.z, c));
        }
        cbuf[row ? t : n] = c;
        key = row ? ((__float_as_uint(fabsf(c)) & ~kIdx) | (kIdx - (uint32_t)t)) : 0u;
        key = wave_max_u32(key);
      }
      if (lane == 0) keys[wave] = key;
      __syncthreads();
      // (2) pivot row, new panel column, exchange of row/column a into b
      uint32_t km = 0u;
#pragma unroll
      for (int w = 0; w < NW; w += 4) {
        const uint4 kv = *reinterpret_cast<const uint4*>(keys + w);
        km = max(km, max(max(kv.x, kv.y), max(kv.z, kv.w)));
      }
      km = __builtin_amdgcn_readfirstlane(km);
      const int b = (int)(kIdx - (km & kIdx));
      const bool wlive = wave * kWave + kWave - 1 >= k + 2 && wave * kWave < n;  // uniform
      float piv;
      if (wlive) {
        // every load that depends on b is issued before the pivot is examined
        const bool live = t >= k + 2 && t < n;
        const bool isb = live && t == b && b != a;
        const int pi = isb ? a : t;  // P(i)
        const int hi = pi > b ? pi : b, lo = pi > b ? b : pi;
        piv = cbuf[b];                                        // A'[a][k] = A[b][k]
        const float raw = tri[live ? toff(hi, lo) : dummy];  // A[P i][b]
        const float ra = tri[live ? toff(t, a) : dummy];     // old A[i][a]
        const float ca = cbuf[a];
        const float4* uwb = reinterpret_cast<const float4*>(UW + b * LW);
        const float4* uwa = reinterpret_cast<const float4*>(UW + a * LW);
        float ub[P], wb[P];
#pragma unroll
        for (int q2 = 0; 2 * q2 < s; ++q2) {
          const float4 x = uwb[q2];
          ub[2 * q2] = x.x;
          wb[2 * q2] = x.y;
          ub[2 * q2 + 1] = x.z;
          wb[2 * q2 + 1] = x.w;
        }
        if (isb) {  // the pivot row takes row a's earlier entries
#pragma unroll
          for (int q2 = 0; 2 * q2 < s; ++q2) {
            const float4 x = uwa[q2];
            pu[2 * q2] = x.x;
            pw[2 * q2] = x.y;
            if (2 * q2 + 1 < s) {
              pu[2 * q2 + 1] = x.z;
              pw[2 * q2 + 1] = x.w;
            }
          }
        }
        float rp = __builtin_amdgcn_rcpf(piv);
        rp = fmaf(rp, fmaf(-piv, rp, 1.f), rp);  // one Newton step
        float v = pi > b ? raw : -raw;
#pragma unroll
        for (int q = 0; q < s; ++q) v = fmaf(pu[q], wb[q], fmaf(-pw[q], ub[q], v));
        const float tau = (isb ? ca : c) * rp;
        if (piv != 0.f && isfinite(piv)) {  // uniform
          const int r1 = live ? t : row_dummy;
          const int r2 = isb ? a : row_dummy;
          *reinterpret_cast<float2*>(UW + r1 * LW + 2 * s) = make_float2(tau, v);
          *reinterpret_cast<float2*>(UW + r2 * LW + 2 * s) = make_float2(tau, v);
          UT[s * ldt + r1] = tau;
          WT[s * ldt + r1] = v;
          pu[s] = live ? tau : 0.f;
          pw[s] = live ? v : 0.f;
          const bool ex = live && b != a && t != b;  // new A[i][b] = old A[i][a]
          tri[ex ? (t > b ? toff(t, b) : toff(b, t)) : dummy] = t > b ? ra : -ra;
        }
      } else {
        piv = cbuf[b];
      }
      if (!(piv != 0.f) || !isfinite(piv)) {
        singular = true;
        break;
      }
      if (b != a) sign = -sign;
      sign = (piv < 0.f) ? sign : -sign;  // Pf *= A[k][k+1] = -piv
      if (t == 0) pivs[k >> 1] = piv;
      bprev = b;
      aprev = a;
      __syncthreads();
    }
    if (singular) break;
    // the last step's delayed store
    if (t == bprev && bprev != aprev) {
#pragma unroll
      for (int q = 0; q + 1 < P; ++q)
        if (q + 1 < steps) {
          UW[t * LW + 2 * q] = pu[q];
          UW[t * LW + 2 * q + 1] = pw[q];
          UT[q * ldt + t] = pu[q];
          WT[q * ldt + t] = pw[q];
        }
    }
    __syncthreads();
    // trailing rank-2P update of rows/columns >= K
    const int K = k0 + 2 * steps;
    const int mtr = n - K;
    if (mtr <= 1) continue;
    const int nt = (mtr + 15) / 16;
    const int ntiles = nt * (nt + 1) / 2;
    for (int tile = wave; tile < ntiles; tile += NW) {
      // tile -> (ti, tj), tj <= ti, row-major over the lower triangle of tiles
      int ti = (int)((sqrtf(8.f * tile + 1.f) - 1.f) * 0.5f);
      while ((ti + 1) * (ti + 2) / 2 <= tile) ++ti;
      while (ti * (ti + 1) / 2 > tile) --ti;
      const int tj = tile - ti * (ti + 1) / 2;
      const int i0 = K + 16 * ti, j0 = K + 16 * tj;
      const int col = j0 + (lane & 15);
      int ad[4];
      nh_f32x4 cacc;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = i0 + 4 * (lane >> 4) + r;
        ad[r] = (i < n && col < i) ? toff(i, col) : dummy;
        cacc[r] = tri[ad[r]];
      }
      const int ia = i0 + (lane & 15);
      const bool iok = ia < n, jok = col < n;
#pragma unroll
      for (int kk = 0; kk < (2 * P) / 4; ++kk) {
        const int q = kk * 4 + (lane >> 4);  // 0 .. 2P-1
        // q < P: U_q[i] W_q[j];  q >= P: -W_q[i] U_q[j]
        const float* L = q < P ? UT + q * ldt : WT + (q - P) * ldt;
        const float* R = q < P ? WT + q * ldt : UT + (q - P) * ldt;
        const float av = iok ? (q < P ? L[ia] : -L[ia]) : 0.f;
        const float bv = jok ? R[col] : 0.f;
        cacc = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, cacc, 0, 0, 0);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) tri[ad[r]] = cacc[r];
    }
    __syncthreads();
  }
  if (singular) {
    sign_out = 0.f;
    log_out = -INFINITY;
    return;
  }
  sign_out = sign;
  log_out = 0.0;
  if (wave == 0) {
    double l = 0.0;
    for (int i = lane; i < n / 2; i += kWave) l += log((double)fabsf(pivs[i]));
    log_out = wave_sum_f64(l);
  }
}
