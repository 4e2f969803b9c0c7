// Blocked LU with partial pivoting (lu_blocked) for sign / log|det| of one dense matrix per
// workgroup, and its scratch size.  A fragment, not a standalone header: it is included inside
// namespace nh of nh_linalg.hip after kWave, wave_max_u32, wave_sum_f64 and nh_f32x4, and by the
// host emulation under tests/host, which supplies those and the wave-level builtins itself.
#pragma once

constexpr int kLuPanel = 16;  // columns per panel of the blocked LU

// Scratch (32-bit words) of lu_blocked: keys 2 x 16, pivot row per step n, step of each row n,
// |pivots| n, parity partials 16.
__host__ __device__ inline int lu_scratch_words(int n) { return 32 + 3 * ((n + 3) & ~3) + 16; }

// Blocked LU with partial pivoting for sign / log|det| (64 < n <= min(NT, 2^KB)), dense row-major
// M (leading dimension ld, followed by kWave dummy floats; LDS or global), one thread per row.  Rows are never
// moved: row p becomes "done" when it is chosen as pivot, and the permutation's parity is
// counted at the end from the list of pivot rows.  Panels of P columns:
//   step s (column j = k0 + s), one barrier: every live row holds its panel entries in
//   registers; the pivot row p* (largest |M[p][j]| to 2^(KB-23) relative, smallest index among
//   those; 32-bit keys with the row in the low KB bits, as in the Pfaffian) is read from LDS, live rows form
//   l = M[p][j] / M[p*][j], update their later panel entries M[p][j'] -= l M[p*][j'], store
//   them (so that any row can be the next pivot) and form the next column's key;
//   panel end: U12 rows of the panel's pivot rows by forward substitution (one thread per
//   column), then the trailing update M[i][c] -= sum_q L[i][q] U12[q][c] as 16 x 16 FP32 MFMA
//   tiles over all rows (rows done earlier hold dead values; the panel's own pivot rows,
//   read as U12, are not written).
template <int NT, int P, int KB = 8>
__device__ void lu_blocked(float* M, int n, int ld, uint32_t* scr, float& sign_out, double& log_out) {
  constexpr uint32_t kIdx = (1u << KB) - 1u;  // low key bits: kIdx - row (n <= 2^KB)
  constexpr int NW = NT / kWave;
  static_assert(NW <= 16 && NW % 4 == 0, "pivot keys: 4, 8 or 16 waves");
  static_assert(P % 4 == 0 && P <= 16, "panel of 4, 8, 12 or 16 columns");
  const int n4 = (n + 3) & ~3;
  uint32_t* keys0 = scr;
  int* pstar = reinterpret_cast<int*>(scr + 32);  // pivot row of each step
  int* stepof = pstar + n4;                       // step at which a row became pivot
  float* pivs = reinterpret_cast<float*>(stepof + n4);
  uint32_t* parts = reinterpret_cast<uint32_t*>(pivs + n4);
  const int t = threadIdx.x;
  const int lane = t & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(t / kWave);
  const int wlo = wave * kWave, whi = wlo + kWave - 1;
  const int dummy = n * ld + lane;
  if (t < n) stepof[t] = -1;
  bool live = t < n;  // this row has not been a pivot row
  float sign = 1.f;
  bool singular = false;
  for (int k0 = 0; k0 < n; k0 += P) {
    const int steps = n - k0 < P ? n - k0 : P;
    float rv[P];  // this row's panel entries
#pragma unroll
    for (int q = 0; q < P; ++q) rv[q] = M[(live && k0 + q < n) ? t * ld + k0 + q : dummy];
    {
      uint32_t key = live ? ((__float_as_uint(fabsf(rv[0])) & ~kIdx) | (kIdx - (uint32_t)t)) : 0u;
      if (wlo < n) key = wave_max_u32(key);
      if (lane == 0) keys0[wave] = wlo < n ? key : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < P; ++s) {
      if (s >= steps) break;  // uniform
      const int j = k0 + s;
      const uint32_t* kb = keys0 + (s & 1) * 16;
      uint32_t* kbn = keys0 + ((s + 1) & 1) * 16;
      uint32_t km = 0u;
#pragma unroll
      for (int w = 0; w < NW; w += 4) {
        const uint4 kv = *reinterpret_cast<const uint4*>(kb + w);
        km = max(km, max(max(kv.x, kv.y), max(kv.z, kv.w)));
      }
      km = __builtin_amdgcn_readfirstlane(km);
      const int ps = (int)(kIdx - (km & kIdx));
      const float* prow = M + ps * ld + k0;
      const float piv = prow[s];
      float urow[P];
#pragma unroll
      for (int q = s + 1; q < P; ++q) urow[q] = prow[q];  // within the row: k0 + q < ld
      if (!(piv != 0.f) || !isfinite(piv)) {
        singular = true;
        break;
      }
      if (piv < 0.f) sign = -sign;
      if (t == 0) {
        pstar[j] = ps;
        pivs[j] = piv;
      }
      if (t == ps) {
        live = false;
        stepof[t] = j;
      }
      uint32_t key = 0u;
      if (wlo < n && whi >= 0) {
        const bool act = live;  // rows other than the pivot row that are still live
        float rp = __builtin_amdgcn_rcpf(piv);
        rp = fmaf(rp, fmaf(-piv, rp, 1.f), rp);  // one Newton step
        const float l = rv[s] * rp;
        if (act) rv[s] = l;
#pragma unroll
        for (int q = s + 1; q < P; ++q)
          if (act) rv[q] = fmaf(-l, urow[q], rv[q]);
        M[act ? t * ld + j : dummy] = l;
#pragma unroll
        for (int q = s + 1; q < P; ++q) M[(act && k0 + q < n) ? t * ld + k0 + q : dummy] = rv[q];
        if (s + 1 < steps)
          key = act ? ((__float_as_uint(fabsf(rv[s + 1])) & ~kIdx) | (kIdx - (uint32_t)t)) : 0u;
        if (s + 1 < steps) key = wave_max_u32(key);
      }
      if (s + 1 < steps && lane == 0) kbn[wave] = wlo < n ? key : 0u;
      __syncthreads();
    }
    if (singular) break;
    const int K = k0 + steps;
    if (K >= n) continue;
    // U12: forward substitution down the panel's pivot rows, one thread per column c >= K
    {
      const int c = K + t;
      if (c < n) {
        float u[P];
#pragma unroll
        for (int s2 = 0; s2 < P; ++s2) {
          if (s2 < steps) {  // uniform
            const float* pr = M + pstar[k0 + s2] * ld;
            float x = pr[c];
#pragma unroll
            for (int q = 0; q < s2; ++q) x = fmaf(-pr[k0 + q], u[q], x);
            u[s2] = x;
            M[pstar[k0 + s2] * ld + c] = x;
          }
        }
      }
    }
    __syncthreads();
    // trailing update over all rows x columns >= K, skipping stores to this panel's pivot rows
    const int ntr = (n + 15) / 16, ntc = (n - K + 15) / 16;
    for (int tile = wave; tile < ntr * ntc; tile += NW) {
      const int ti = tile / ntc, tj = tile - ti * ntc;
      const int i0 = 16 * ti, j0 = K + 16 * tj;
      const int col = j0 + (lane & 15);
      int ad[4];
      nh_f32x4 cacc;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = i0 + 4 * (lane >> 4) + r;
        const bool ok = i < n && col < n && !(stepof[i < n ? i : 0] >= k0);
        ad[r] = ok ? i * ld + col : dummy;
        cacc[r] = M[ad[r]];
      }
      const int ia = i0 + (lane & 15);
      const bool iok = ia < n, jok = col < n;
#pragma unroll
      for (int kk = 0; kk < P / 4; ++kk) {
        const int q = kk * 4 + (lane >> 4);
        const bool qok = q < steps;
        const float av = (iok && qok) ? -M[ia * ld + k0 + q] : 0.f;
        const float bv = (jok && qok) ? M[pstar[k0 + (qok ? q : 0)] * ld + col] : 0.f;
        cacc = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, cacc, 0, 0, 0);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) M[ad[r]] = cacc[r];
    }
    __syncthreads();
  }
  if (singular) {
    sign_out = 0.f;
    log_out = -INFINITY;
    return;
  }
  // parity of the pivot-row permutation: inversions of pstar
  uint32_t inv = 0u;
  if (t < n) {
    const int me = pstar[t];
    for (int s2 = t + 1; s2 < n; ++s2) inv += pstar[s2] < me ? 1u : 0u;
  }
  inv &= 1u;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) inv ^= (uint32_t)__shfl_xor((int)inv, o, kWave);
  if (lane == 0) parts[wave] = inv;
  __syncthreads();
  uint32_t par = 0u;
  for (int w = 0; w < NW; ++w) par ^= parts[w];
  sign_out = (par & 1u) ? -sign : sign;
  log_out = 0.0;
  if (wave == 0) {
    double lsum = 0.0;
    for (int i = lane; i < n; i += kWave) lsum += log((double)fabsf(pivs[i]));
    log_out = wave_sum_f64(lsum);
  }
}
