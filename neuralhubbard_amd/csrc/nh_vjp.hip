// Reverse-mode pieces of the amplitude (gfx950): the VJPs of the occupied pairing block (Pfaffian
// states) and of the up x down pairing block (BCS states), and the Pfaffian and determinant VJPs.
//
// nh_pairing_occ_vjp replaces the backward of the reference's pairing_occ custom_vjp
// (code/backflow/pf_kernel.py:354-365): for the cotangent G of F = pairing_occ(Y, idx, Wt),
//   dY[p, a]      = sum_q (G[p, q] - G[q, p]) Wt[2 s_p + s_q][r(p, q)][a] Y[q, a]   (_dy_kernel)
//   dWt[b][r][a]  = sum over (p, q) with 2 s_p + s_q = b, r(p, q) = r of G[p, q] Y[p, a] Y[q, a]
//                                                                                 (_dw_kernel)
// per sample (the reference runs the custom_vjp per configuration under vmap).  The dY form uses
// the table identity Wt[2 s_q + s_p][-r] = -Wt[2 s_p + s_q][r] of build_wtilde.  dWt loops over the
// occupied p and finds the partner q of (p, r, s_q) through the inverse mode map, so every
// output element is one thread's plain sum (no atomics, fixed order).
//
// nh_pairing_ud_vjp replaces the backward of the reference's BCS pairing_ud custom_vjp
// (code/backflow/bcs_kernel.py:274-288): for the cotangent G (n_up x n_dn) of
// F[p, q] = sum_a Yu[p, a] Wt[r(iu_p - idn_q)][a] Yd[q, a],
//   dYu[p, a] = sum_q G[p, q] Wt[r(iu_p - idn_q)][a] Yd[q, a]      (k_vjp_ud_dy, side 0)
//   dYd[q, a] = sum_p G[p, q] Wt[r(iu_p - idn_q)][a] Yu[p, a]      (k_vjp_ud_dy, side 1)
//   dWt[r][a] = sum_p G[p, q(p, r)] Yu[p, a] Yd[q(p, r), a]          (k_vjp_ud_dw)
// with q(p, r) the occupied down site iu_p - r (at most one), found through the inverse site map.
//
// nh_slogpf_vjp is the Pfaffian VJP: d log|Pf F| / dF = (1/2) F^{-T} (d Pf = (1/2) Pf tr(F^{-1} dF)),
// scaled by the cotangent of log|Pf|.  It inverts the transposed copy of F in place in the output
// buffer by blocked Gauss-Jordan elimination with partial pivoting (k_inv_gj: 16-column panels,
// rank-16 MFMA updates), one block per matrix (256 threads up to n = 256, 1024 above); the row
// interchanges are undone at the end as one column permutation.

#include "nh_common.h"

namespace nh {

namespace {

constexpr int kVjpThreads = 256;
typedef float nh_vf32x4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(kVjpThreads) k_vjp_dy(int LX, int LY, int d, int n, const float* __restrict__ Y,
                                                        const int32_t* __restrict__ idx, const float* __restrict__ Wt,
                                                        const float* __restrict__ G, float* __restrict__ dY) {
  extern __shared__ int vj_lds[];
  int* coord = vj_lds;  // (spin << 30) | (x << 15) | y of each occupied mode
  const int b = blockIdx.y;
  const int N = LX * LY;
  const int32_t* ib = idx + (size_t)b * n;
  for (int q = threadIdx.x; q < n; q += kVjpThreads) {
    const int m = ib[q];
    const int sp = m >= N, site = m - (m >= N ? N : 0);
    coord[q] = (m < 0 || m >= 2 * N) ? -1 : ((sp << 30) | ((site / LY) << 15) | (site % LY));
  }
  __syncthreads();
  const int rows = kVjpThreads / d;
  const int pl = threadIdx.x / d, a = threadIdx.x - pl * d;
  const int p = blockIdx.x * rows + pl;
  if (p >= n) return;
  const float* Gb = G + (size_t)b * n * n;
  const float* Yb = Y + (size_t)b * n * d;
  const int cp = coord[p];
  if (cp < 0) {  // a mode outside [0, 2N)
    dY[((size_t)b * n + p) * d + a] = NAN;
    return;
  }
  const int sp = cp >> 30, xp = (cp >> 15) & 0x7fff, yp = cp & 0x7fff;
  float acc = 0.f;
  for (int q = 0; q < n; ++q) {
    const int cq = coord[q];
    if (cq < 0) continue;
    const int sq = cq >> 30;
    int dx = xp - ((cq >> 15) & 0x7fff), dy = yp - (cq & 0x7fff);
    dx += dx < 0 ? LX : 0;
    dy += dy < 0 ? LY : 0;
    const float g = Gb[(size_t)p * n + q] - Gb[(size_t)q * n + p];
    acc = fmaf(g * Wt[((size_t)(2 * sp + sq) * N + dx * LY + dy) * d + a], Yb[(size_t)q * d + a], acc);
  }
  dY[((size_t)b * n + p) * d + a] = acc;
}

__global__ void __launch_bounds__(kVjpThreads) k_vjp_dw(int LX, int LY, int d, int n, const float* __restrict__ Y,
                                                        const int32_t* __restrict__ idx, const float* __restrict__ G,
                                                        float* __restrict__ dWt) {
  extern __shared__ int vj_lds[];
  const int N = LX * LY;
  int* inv = vj_lds;         // [2N]: occupied position of each mode, -1 if empty
  int* coord = inv + 2 * N;  // [n]: (spin << 30) | site
  const int b = blockIdx.y;
  const int32_t* ib = idx + (size_t)b * n;
  for (int m = threadIdx.x; m < 2 * N; m += kVjpThreads) inv[m] = -1;
  __syncthreads();
  for (int q = threadIdx.x; q < n; q += kVjpThreads) {
    const int m = ib[q];
    const bool ok = m >= 0 && m < 2 * N;
    if (ok) inv[m] = q;
    coord[q] = ok ? (((m >= N) << 30) | (m - (m >= N ? N : 0))) : -1;
  }
  __syncthreads();
  const int rs = kVjpThreads / d;
  const int rl = threadIdx.x / d, a = threadIdx.x - rl * d;
  const int r = blockIdx.x * rs + rl;
  if (r >= N) return;
  const int dxr = r / LY, dyr = r - dxr * LY;
  const float* Gb = G + (size_t)b * n * n;
  const float* Yb = Y + (size_t)b * n * d;
  float acc[4] = {0.f, 0.f, 0.f, 0.f};
  for (int p = 0; p < n; ++p) {
    const int cp = coord[p];
    if (cp < 0) continue;
    const int sp = cp >> 30, site = cp & 0x3fffffff;
    int xq = site / LY - dxr, yq = site % LY - dyr;
    xq += xq < 0 ? LX : 0;
    yq += yq < 0 ? LY : 0;
    const int sq_site = xq * LY + yq;  // site_p - r
    const float ypa = Yb[(size_t)p * d + a];
#pragma unroll
    for (int sq = 0; sq < 2; ++sq) {
      const int q = inv[sq * N + sq_site];
      if (q >= 0) acc[2 * sp + sq] = fmaf(Gb[(size_t)p * n + q] * ypa, Yb[(size_t)q * d + a], acc[2 * sp + sq]);
    }
  }
  float* out = dWt + (size_t)b * 4 * N * d;
#pragma unroll
  for (int blk = 0; blk < 4; ++blk) out[((size_t)blk * N + r) * d + a] = acc[blk];
}

// One thread per (own row, a): side 0 rows are the up sites (partners down), side 1 the down
// sites (partners up); the displacement is always r(up site - down site) and G is [n_up][n_dn].
__global__ void __launch_bounds__(kVjpThreads)
    k_vjp_ud_dy(int LX, int LY, int d, int n_up, int n_dn, int side, const float* __restrict__ Yoth,
                const int32_t* __restrict__ iown, const int32_t* __restrict__ ioth, const float* __restrict__ Wt,
                const float* __restrict__ G, float* __restrict__ dYown) {
  extern __shared__ int vj_lds[];
  int* coord = vj_lds;  // (x << 15) | y of each partner site, -1 outside the lattice
  const int b = blockIdx.y;
  const int N = LX * LY;
  const int n_own = side ? n_dn : n_up, n_oth = side ? n_up : n_dn;
  const int32_t* ob = ioth + (size_t)b * n_oth;
  for (int q = threadIdx.x; q < n_oth; q += kVjpThreads) {
    const int site = ob[q];
    coord[q] = (site < 0 || site >= N) ? -1 : (((site / LY) << 15) | (site % LY));
  }
  __syncthreads();
  const int rows = kVjpThreads / d;
  const int pl = threadIdx.x / d, a = threadIdx.x - pl * d;
  const int p = blockIdx.x * rows + pl;
  if (p >= n_own) return;
  float* out = dYown + ((size_t)b * n_own + p) * d + a;
  const int site_p = iown[(size_t)b * n_own + p];
  if (site_p < 0 || site_p >= N) {
    *out = NAN;
    return;
  }
  const int xp = site_p / LY, yp = site_p % LY;
  const float* Gb = G + (size_t)b * n_up * n_dn;
  const float* Yb = Yoth + (size_t)b * n_oth * d;
  float acc = 0.f;
  for (int q = 0; q < n_oth; ++q) {
    const int cq = coord[q];
    if (cq < 0) {
      acc = NAN;
      continue;
    }
    int dx = xp - (cq >> 15), dy = yp - (cq & 0x7fff);
    if (side) {  // r(up - down) with the up site the partner
      dx = -dx;
      dy = -dy;
    }
    dx += dx < 0 ? LX : 0;
    dy += dy < 0 ? LY : 0;
    const float g = side ? Gb[(size_t)q * n_dn + p] : Gb[(size_t)p * n_dn + q];
    acc = fmaf(g * Wt[(size_t)(dx * LY + dy) * d + a], Yb[(size_t)q * d + a], acc);
  }
  *out = acc;
}

// One thread per (displacement r, a): the sum over the occupied up sites p of
// G[p, q] Yu[p, a] Yd[q, a] with q the occupied down site at site_p - r.
__global__ void __launch_bounds__(kVjpThreads)
    k_vjp_ud_dw(int LX, int LY, int d, int n_up, int n_dn, const float* __restrict__ Yu,
                const float* __restrict__ Yd, const int32_t* __restrict__ iu, const int32_t* __restrict__ idn,
                const float* __restrict__ G, float* __restrict__ dWt) {
  extern __shared__ int vj_lds[];
  const int N = LX * LY;
  int* inv = vj_lds;     // [N]: position of the down electron on each site, -1 if empty
  int* cu = inv + N;     // [n_up]: up site, -1 outside the lattice
  const int b = blockIdx.y;
  for (int m = threadIdx.x; m < N; m += kVjpThreads) inv[m] = -1;
  __syncthreads();
  for (int q = threadIdx.x; q < n_dn; q += kVjpThreads) {
    const int site = idn[(size_t)b * n_dn + q];
    if (site >= 0 && site < N) inv[site] = q;
  }
  for (int p = threadIdx.x; p < n_up; p += kVjpThreads) {
    const int site = iu[(size_t)b * n_up + p];
    cu[p] = (site < 0 || site >= N) ? -1 : site;
  }
  __syncthreads();
  const int rs = kVjpThreads / d;
  const int rl = threadIdx.x / d, a = threadIdx.x - rl * d;
  const int r = blockIdx.x * rs + rl;
  if (r >= N) return;
  const int dxr = r / LY, dyr = r - dxr * LY;
  const float* Gb = G + (size_t)b * n_up * n_dn;
  const float* Yub = Yu + (size_t)b * n_up * d;
  const float* Ydb = Yd + (size_t)b * n_dn * d;
  float acc = 0.f;
  for (int p = 0; p < n_up; ++p) {
    const int site = cu[p];
    if (site < 0) continue;
    int xq = site / LY - dxr, yq = site % LY - dyr;
    xq += xq < 0 ? LX : 0;
    yq += yq < 0 ? LY : 0;
    const int q = inv[xq * LY + yq];
    if (q >= 0) acc = fmaf(Gb[(size_t)p * n_dn + q] * Yub[(size_t)p * d + a], Ydb[(size_t)q * d + a], acc);
  }
  dWt[((size_t)b * N + r) * d + a] = acc;
}

// In-place inverse of the transposed copy of F (so X = F^{-T}), scaled by factor * cot (factor
// 1/2 for the Pfaffian, 1 for the determinant): blocked Gauss-Jordan with partial pivoting
// (Quintana-Orti, Quintana-Orti, Sun, van de Geijn, "A note on parallel matrix inversion", SIAM
// J. Sci. Comput. 22 (2001) 1762, variant GJE with the pivots of each panel's LU).  For the
// panel K of 16 columns at k0: the pivots of an LU with partial pivoting of its rows k0 ..
// n - 1 (in LDS); those row interchanges on the whole matrix; then with A11 the diagonal block,
// C the panel columns and R the panel rows outside it,
//   C <- -C A11^{-1},  A[i][j] += C[i] R[:, j] (i, j outside K; 16 x 16 FP32 MFMA tiles),
//   R <- A11^{-1} R,   A11 <- A11^{-1},
// and at the end the interchanges undone as one column permutation.  The matrix stays in the
// output buffer; each panel reads and writes it once (an unblocked sweep does so per column).
// LDS: the panel columns [nr][17], the panel rows [16][nr + 4], A11 [16][17], reduction slots,
// pivots and the final permutation (nr = n rounded up to 16).
constexpr int kGjB = 16;

__host__ __device__ inline size_t gj_lds_bytes(int n, int nt) {
  const size_t nr = (size_t)((n + 15) & ~15);
  return sizeof(float) * (nr * (kGjB + 1) + kGjB * (nr + 4) + kGjB * (kGjB + 1) + 2 * (nt / kWave)) +
         sizeof(int) * 2 * (size_t)n;
}

template <int NT>
__global__ void __launch_bounds__(NT) k_inv_gj(int n, const float* __restrict__ F, const float* __restrict__ cot,
                                               float factor, float* __restrict__ G) {
  static_assert(NT >= kGjB * kGjB, "one thread per entry of the diagonal block");
  constexpr int NW = NT / kWave;
  constexpr int LC = kGjB + 1;
  extern __shared__ float gj_lds[];
  const int nr = (n + 15) & ~15, ldr = nr + 4;
  float* Cp = gj_lds;                              // [nr][LC]
  float* Rp = Cp + (size_t)nr * LC;                // [16][ldr]
  float* Ai = Rp + (size_t)kGjB * ldr;             // [16][LC]
  float* redv = Ai + kGjB * LC;                    // [NW]
  int* redi = reinterpret_cast<int*>(redv + NW);   // [NW]
  int* piv = redi + NW;                            // [n]
  int* perm = piv + n;                             // [n]
  const int b = blockIdx.x, t = threadIdx.x, lane = t & (kWave - 1), wave = t / kWave;
  float* A = G + (size_t)b * n * n;
  const float* Fb = F + (size_t)b * n * n;
  for (int i = wave; i < n; i += NW)
    for (int j = lane; j < n; j += kWave) A[(size_t)i * n + j] = Fb[(size_t)j * n + i];
  __syncthreads();
  bool singular = false;
  for (int k0 = 0; k0 < n && !singular; k0 += kGjB) {
    const int kb = min(kGjB, n - k0), m = n - k0;
    // (a) the panel's pivots: LU with partial pivoting of its rows k0 .. n - 1
    for (int r = t; r < m; r += NT)
      for (int c = 0; c < kb; ++c) Cp[r * LC + c] = A[(size_t)(k0 + r) * n + k0 + c];
    __syncthreads();
    for (int j = 0; j < kb; ++j) {
      float best = -1.f;
      int bi = m;
      for (int r = j + t; r < m; r += NT) {
        const float v = fabsf(Cp[r * LC + j]);
        if (v > best) {
          best = v;
          bi = r;
        }
      }
#pragma unroll
      for (int o = kWave / 2; o > 0; o >>= 1) {
        const float ob = __shfl_xor(best, o, kWave);
        const int oi = __shfl_xor(bi, o, kWave);
        if (ob > best || (ob == best && oi < bi)) {
          best = ob;
          bi = oi;
        }
      }
      if (lane == 0) {
        redv[wave] = best;
        redi[wave] = bi;
      }
      __syncthreads();
      best = redv[0];
      bi = redi[0];
      for (int w = 1; w < NW; ++w)
        if (redv[w] > best || (redv[w] == best && redi[w] < bi)) {
          best = redv[w];
          bi = redi[w];
        }
      if (!(best > 0.f) || !isfinite(best)) {  // every thread reads the same slots: uniform
        singular = true;
        break;
      }
      if (t == 0) piv[k0 + j] = k0 + bi;
      if (bi != j)
        for (int c = t; c < kb; c += NT) {
          const float x = Cp[j * LC + c];
          Cp[j * LC + c] = Cp[bi * LC + c];
          Cp[bi * LC + c] = x;
        }
      __syncthreads();
      const float rinv = 1.f / Cp[j * LC + j];
      for (int r = j + 1 + t; r < m; r += NT) {
        const float l = Cp[r * LC + j] * rinv;
        for (int c = j + 1; c < kb; ++c) Cp[r * LC + c] = fmaf(-l, Cp[j * LC + c], Cp[r * LC + c]);
      }
      __syncthreads();
    }
    if (singular) break;
    // (b) the interchanges on the whole matrix, one column per thread
    for (int c = t; c < n; c += NT)
      for (int jj = 0; jj < kb; ++jj) {
        const int k = k0 + jj, p = piv[k];
        if (p != k) {
          const float x = A[(size_t)k * n + c];
          A[(size_t)k * n + c] = A[(size_t)p * n + c];
          A[(size_t)p * n + c] = x;
        }
      }
    __syncthreads();
    // (c) panel columns and rows after the interchanges (zero padded to 16), A11 padded with the identity
    for (int r = t; r < nr; r += NT)
      for (int c = 0; c < kGjB; ++c) Cp[r * LC + c] = (r < n && c < kb) ? A[(size_t)r * n + k0 + c] : 0.f;
    for (int r = 0; r < kGjB; ++r)
      for (int c = t; c < ldr; c += NT) Rp[r * ldr + c] = (r < kb && c < n) ? A[(size_t)(k0 + r) * n + c] : 0.f;
    __syncthreads();
    if (t < kGjB * kGjB) {
      const int i = t / kGjB, c = t % kGjB;
      Ai[i * LC + c] = (i < kb && c < kb) ? Cp[(k0 + i) * LC + c] : (i == c ? 1.f : 0.f);
    }
    __syncthreads();
    // A11^{-1} by Gauss-Jordan without interchanges (its rows are already pivoted)
    for (int s = 0; s < kGjB; ++s) {
      float v = 0.f;
      const int i = t / kGjB, c = t % kGjB;
      if (t < kGjB * kGjB) {
        const float d = 1.f / Ai[s * LC + s];
        const float ais = Ai[i * LC + s], asc = Ai[s * LC + c], aic = Ai[i * LC + c];
        if (i == s)
          v = c == s ? d : asc * d;
        else
          v = c == s ? -ais * d : fmaf(-ais * d, asc, aic);
      }
      __syncthreads();
      if (t < kGjB * kGjB) Ai[i * LC + c] = v;
      __syncthreads();
    }
    // (d) C <- -C A11^{-1} in place, one row per thread
    for (int r = t; r < nr; r += NT) {
      if (r >= k0 && r < k0 + kGjB) continue;
      float row[kGjB], out[kGjB];
#pragma unroll
      for (int c = 0; c < kGjB; ++c) row[c] = Cp[r * LC + c];
#pragma unroll
      for (int c = 0; c < kGjB; ++c) {
        float acc = 0.f;
#pragma unroll
        for (int q = 0; q < kGjB; ++q) acc = fmaf(row[q], Ai[q * LC + c], acc);
        out[c] = -acc;
      }
#pragma unroll
      for (int c = 0; c < kGjB; ++c) Cp[r * LC + c] = out[c];
    }
    __syncthreads();
    // (e) A[i][j] += C[i] R[:, j] outside the panel's rows and columns, 16 x 16 MFMA tiles
    const int nt = nr / kGjB, pb = k0 / kGjB;
    const int li = lane & 15, kk = lane >> 4;
    for (int tile = wave; tile < nt * nt; tile += NW) {
      const int I = tile / nt, J = tile - (tile / nt) * nt;
      if (I == pb || J == pb) continue;  // uniform over the wave
      const int col = J * kGjB + li;
      nh_vf32x4 acc;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = I * kGjB + 4 * kk + r;
        acc[r] = (row < n && col < n) ? A[(size_t)row * n + col] : 0.f;
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float av = Cp[(I * kGjB + li) * LC + 4 * q + kk];
        const float bv = Rp[(4 * q + kk) * ldr + col];
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc, 0, 0, 0);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = I * kGjB + 4 * kk + r;
        if (row < n && col < n) A[(size_t)row * n + col] = acc[r];
      }
    }
    // (d') C into the panel columns; (f) A11^{-1} R into the panel rows, A11^{-1} on the diagonal block
    for (int r = t; r < n; r += NT) {
      if (r >= k0 && r < k0 + kb) continue;
      for (int c = 0; c < kb; ++c) A[(size_t)r * n + k0 + c] = Cp[r * LC + c];
    }
    for (int rr = 0; rr < kb; ++rr)
      for (int c = t; c < n; c += NT) {
        float v;
        if (c >= k0 && c < k0 + kb) {
          v = Ai[rr * LC + (c - k0)];
        } else {
          v = 0.f;
#pragma unroll
          for (int q = 0; q < kGjB; ++q) v = fmaf(Ai[rr * LC + q], Rp[q * ldr + c], v);
        }
        A[(size_t)(k0 + rr) * n + c] = v;
      }
    __syncthreads();
  }
  if (singular) {
    for (int i = wave; i < n; i += NW)
      for (int j = lane; j < n; j += kWave) A[(size_t)i * n + j] = NAN;
    return;
  }
  // undo the row interchanges as one column permutation, then scale
  if (t == 0) {
    for (int j = 0; j < n; ++j) perm[j] = j;
    for (int k = n - 1; k >= 0; --k) {
      const int p = piv[k];
      if (p != k) {
        const int tmp = perm[k];
        perm[k] = perm[p];
        perm[p] = tmp;
      }
    }
  }
  const float scale = factor * (cot ? cot[b] : 1.0f);
  float* rowk = Cp;
  for (int i = 0; i < n; ++i) {
    __syncthreads();  // perm ready; the previous row's reads of rowk are done
    for (int j = t; j < n; j += NT) rowk[j] = A[(size_t)i * n + j];
    __syncthreads();
    for (int j = t; j < n; j += NT) A[(size_t)i * n + j] = scale * rowk[perm[j]];
  }
}

}  // namespace

}  // namespace nh

using namespace nh;

extern "C" int nh_pairing_occ_vjp(void* stream, int32_t LX, int32_t LY, int32_t d, int32_t batch, int32_t n,
                                  const float* Y, const int32_t* idx, const float* Wt, const float* G, float* dY,
                                  float* dWt) {
  if (batch < 0 || n < 0 || LX < 1 || LY < 1) return NH_ERR_ARG;
  if (batch == 0 || n == 0) return NH_OK;
  if (!Y || !idx || !Wt || !G || (!dY && !dWt)) return NH_ERR_ARG;
  const int N = LX * LY;
  if (N > 1024 || n > 2 * N || d < 1 || d > 64 || (kVjpThreads % d) != 0 || LX >= (1 << 15) || LY >= (1 << 15))
    return NH_ERR_SHAPE;
  if (batch > 65535) return NH_ERR_SHAPE;
  hipStream_t st = (hipStream_t)stream;
  if (dY) {
    const int rows = kVjpThreads / d;
    hipLaunchKernelGGL(k_vjp_dy, dim3((n + rows - 1) / rows, batch), dim3(kVjpThreads), sizeof(int) * n, st, LX, LY,
                       d, n, Y, idx, Wt, G, dY);
    NH_LAUNCH_CHECK();
  }
  if (dWt) {
    const int rs = kVjpThreads / d;
    hipLaunchKernelGGL(k_vjp_dw, dim3((N + rs - 1) / rs, batch), dim3(kVjpThreads), sizeof(int) * (2 * N + n), st,
                       LX, LY, d, n, Y, idx, G, dWt);
    NH_LAUNCH_CHECK();
  }
  return NH_OK;
}

static int inverse_transpose(void* stream, int32_t batch, int32_t n, const float* F, const float* cot,
                             float factor, float* G) {
  if (batch < 0 || n < 0) return NH_ERR_ARG;
  if (batch == 0 || n == 0) return NH_OK;
  if (!F || !G || F == G) return NH_ERR_ARG;
  if (n > 1024) return NH_ERR_SHAPE;
  if (n <= 256)
    hipLaunchKernelGGL(k_inv_gj<256>, dim3(batch), dim3(256), gj_lds_bytes(n, 256), (hipStream_t)stream, n, F, cot,
                       factor, G);
  else
    hipLaunchKernelGGL(k_inv_gj<1024>, dim3(batch), dim3(1024), gj_lds_bytes(n, 1024), (hipStream_t)stream, n, F,
                       cot, factor, G);
  NH_LAUNCH_CHECK();
  return NH_OK;
}

namespace nh {
int launch_inverse_t(hipStream_t st, int batch, int n, const float* F, float factor, float* G) {
  return inverse_transpose(st, batch, n, F, nullptr, factor, G);
}
}  // namespace nh

extern "C" int nh_slogpf_vjp(void* stream, int32_t batch, int32_t n, const float* F, const float* cot, float* G) {
  return inverse_transpose(stream, batch, n, F, cot, 0.5f, G);
}

extern "C" int nh_slogdet_vjp(void* stream, int32_t batch, int32_t n, const float* F, const float* cot, float* G) {
  return inverse_transpose(stream, batch, n, F, cot, 1.0f, G);
}

extern "C" int nh_pairing_ud_vjp(void* stream, int32_t LX, int32_t LY, int32_t d, int32_t batch, int32_t n_up,
                                 int32_t n_dn, const float* Yu, const float* Yd, const int32_t* iu,
                                 const int32_t* idn, const float* Wt, const float* G, float* dYu, float* dYd,
                                 float* dWt) {
  if (batch < 0 || n_up < 0 || n_dn < 0 || LX < 1 || LY < 1) return NH_ERR_ARG;
  if (batch == 0 || n_up == 0 || n_dn == 0) return NH_OK;
  if (!Yu || !Yd || !iu || !idn || !Wt || !G || (!dYu && !dYd && !dWt)) return NH_ERR_ARG;
  const int N = LX * LY;
  if (N > 1024 || n_up > N || n_dn > N || d < 1 || d > 64 || (kVjpThreads % d) != 0 || LX >= (1 << 15) ||
      LY >= (1 << 15) || batch > 65535)
    return NH_ERR_SHAPE;
  hipStream_t st = (hipStream_t)stream;
  const int rows = kVjpThreads / d;
  if (dYu) {
    hipLaunchKernelGGL(k_vjp_ud_dy, dim3((n_up + rows - 1) / rows, batch), dim3(kVjpThreads), sizeof(int) * n_dn, st,
                       LX, LY, d, n_up, n_dn, 0, Yd, iu, idn, Wt, G, dYu);
    NH_LAUNCH_CHECK();
  }
  if (dYd) {
    hipLaunchKernelGGL(k_vjp_ud_dy, dim3((n_dn + rows - 1) / rows, batch), dim3(kVjpThreads), sizeof(int) * n_up, st,
                       LX, LY, d, n_up, n_dn, 1, Yu, idn, iu, Wt, G, dYd);
    NH_LAUNCH_CHECK();
  }
  if (dWt) {
    hipLaunchKernelGGL(k_vjp_ud_dw, dim3((N + rows - 1) / rows, batch), dim3(kVjpThreads), sizeof(int) * (N + n_up),
                       st, LX, LY, d, n_up, n_dn, Yu, Yd, iu, idn, G, dWt);
    NH_LAUNCH_CHECK();
  }
  return NH_OK;
}
