// Occupied-mode gather and the Pfaffian image kernel's body for 64 < n <= 256 (image_pf_cm):
// the pairing block of one image built as a full skew square, column-major, in the global
// workspace, then pf_lazy on it.  A fragment, not a standalone header: it is included inside
// namespace nh of nh_linalg.hip after nh_pf_lazy.h, and by the host emulation under tests/host,
// which supplies the wave builtins and runs a whole workgroup of this body on the CPU.
#pragma once

// Ascending occupied modes of the item configuration, mapped by group element g, in the
// order of the item's own occupied list.  Wave 0 only; returns the count in lane 0's view.
__device__ inline int occupied_modes(const int8_t* sb, int32_t hop, int two_n, int N, int LX, int LY,
                                     int g, int* modes, int cap) {
  const int lane = threadIdx.x & (kWave - 1);
  constexpr int kPre = 32;  // occupations read ahead per lane: 2N <= 2048 takes one round
  int cnt = 0;
  for (int b0 = 0; b0 < two_n; b0 += kPre * kWave) {
    uint32_t occ = 0u;  // bit i: mode b0 + 64 i + lane occupied
#pragma unroll
    for (int i = 0; i < kPre; ++i) {
      const int m = b0 + i * kWave + lane;
      occ |= ((m < two_n) && item_occ(sb, hop, m, N) > 0) ? (1u << i) : 0u;
    }
#pragma unroll
    for (int i = 0; i < kPre; ++i) {
      const int m = b0 + i * kWave + lane;
      if (b0 + i * kWave >= two_n) break;  // uniform
      const bool o = (occ >> i) & 1u;
      const uint64_t mask = __ballot(o);
      const int pos = cnt + lanes_below(mask);
      if (o && pos < cap) modes[pos] = group_map(m, g, N, LX, LY);
      cnt += __popcll(mask);
    }
  }
  return cnt;
}

__host__ __device__ inline float parity_sign(long long k) { return (k & 1) ? -1.f : 1.f; }


// The build's LDS tables hold rows of D floats as D / 4 16-byte chunks, chunk a of row r at
// position a ^ swz<D>(r): rows of 16 to 64 B start on a few bank groups only, and the swizzle
// spreads the chunk a of 8 consecutive (or random) rows over all 8 groups of 4 banks, so the
// per-lane row gathers of the build run without bank conflicts.
template <int D>
__host__ __device__ inline int swz(int r) {
  constexpr int C = D / 4 < 8 ? D / 4 : 8;                     // chunks that take part
  constexpr int SH = D >= 32 ? 0 : (D == 16 ? 1 : (D == 8 ? 2 : 0));  // rows per 128 bytes: 2^SH
  return C <= 1 ? 0 : (r >> SH) & (C - 1);
}

// The pairing build stages the occupied features in LDS for D <= 16; for D >= 32 it reads them
// from the image's feature rows through the caches (a lane's own row as vector loads, the
// wave-uniform column rows through the scalar cache), which keeps the LDS region at pf_lazy's
// scratch, so four workgroups share a CU instead of two (released 16x16 state, d = 32: 3.13
// against 4.05 us per configuration of 8 images; at d = 16 the staged form is 1.4 % faster).
template <int D>
constexpr bool kPfYCached = D >= 32;

// Floats of the LDS region shared by the pairing build (occupied features [n][D] for D <= 16,
// one Wt spin block [N][D], both swizzled) and by pf_lazy's scratch.
template <int D, int NT, int P>
__host__ __device__ inline int pf_cm_region_floats(int n, int N) {
  const int build = (kPfYCached<D> ? 0 : n * D) + N * D;
  const int elim = pfl_scratch_floats<NT, P>(pfl_ld(n));
  return ((build > elim ? build : elim) + 3) & ~3;
}

// LDS bytes of image_pf_cm: the shared region, then the mode list, site coordinates and a flag.
template <int D, int NT, int P>
__host__ __device__ inline size_t pf_cm_lds_bytes(int n, int N) {
  return sizeof(float) * (size_t)pf_cm_region_floats<D, NT, P>(n, N) + sizeof(int) * (size_t)(2 * n + 4);
}

// F[p][q] = sum_a Y_pa Y_qa Wt[2 s_p + s_q][r(p - q)][a] (+ F0[m_p][m_q]) for every p, q into the
// column-major square A (A[p][q] at q * ld + p).  Three passes, each with one spin block of Wt in
// LDS: (0) rows and columns of the first n_up modes (spin s0 of the image), (1) rows of the other
// modes against columns of the first, (2) the other modes among themselves.  A wave takes 64 rows
// (one per lane) and 16 columns in groups of 4; a lane computes its row's 4 entries and writes
// them down the 4 columns (coalesced over the wave) and, negated, as one 16-byte piece of its own
// column (the mirrored entries).  In the 4 x 4 block on the diagonal a lane writes only its row,
// so every entry of the square is written once.  The formula itself is exactly antisymmetric
// (Wt[2 s_q + s_p][-r] = -Wt[2 s_p + s_q][r] bit for bit), so the square is.  n_up % 4 == 0.
template <int D, int NT>
__device__ void pf_build_cm(float* __restrict__ A, int ld, int n, int n_up, int N, int LX, int LY, int two_n,
                            int g, const float* yocc, const float* __restrict__ Yi, float* wtl,
                            const int* modes, const int* coord, const float* __restrict__ wt,
                            const float* __restrict__ F0) {
  constexpr int NW = NT / kWave;
  constexpr int C4 = D / 4;
  const int t = threadIdx.x, lane = t & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(t / kWave);
  const int s0 = (g >> 2) & 1, s1 = s0 ^ 1;
  for (int pass = 0; pass < 3; ++pass) {
    const int blk = pass == 0 ? 3 * s0 : (pass == 1 ? 2 * s1 + s0 : 3 * s1);
    const int pr0 = pass == 0 ? 0 : n_up, pr1 = pass == 0 ? n_up : n;
    const int qc0 = pass == 2 ? n_up : 0, qc1 = pass == 2 ? n : n_up;
    const bool diag = pass != 1;
    __syncthreads();  // the previous pass is done with wtl (and the features are staged)
    const float4* src = reinterpret_cast<const float4*>(wt + (size_t)blk * N * D);
    for (int e = t; e < N * C4; e += NT) {
      const int r = e / C4, a4 = e - r * C4;
      reinterpret_cast<float4*>(wtl)[r * C4 + (a4 ^ swz<D>(r))] = src[e];
    }
    __syncthreads();
    if (pr1 <= pr0 || qc1 <= qc0) continue;  // uniform
    const int nrc = (pr1 - pr0 + kWave - 1) / kWave, ncb = (qc1 - qc0 + 15) / 16;
    for (int it = wave; it < nrc * ncb; it += NW) {
      const int rc = it / ncb, cb = it - rc * ncb;
      const int p = pr0 + kWave * rc + lane;
      const int plast = min(pr0 + kWave * rc + kWave - 1, pr1 - 1);
      const int qb = qc0 + 16 * cb;
      if (diag && qb > plast) continue;  // uniform: the block lies above this chunk's rows
      const bool prow = p < pr1;
      const int mp = prow ? modes[p] : 0;
      float yp[D];
      {
        const float4* y4 = kPfYCached<D> ? reinterpret_cast<const float4*>(Yi + (size_t)mp * D)
                                         : reinterpret_cast<const float4*>(yocc + p * D);
        const int zp = kPfYCached<D> ? 0 : swz<D>(p);
#pragma unroll
        for (int a4 = 0; a4 < C4; ++a4) {
          const float4 y = prow ? y4[a4 ^ zp] : make_float4(0.f, 0.f, 0.f, 0.f);
          yp[4 * a4] = y.x;
          yp[4 * a4 + 1] = y.y;
          yp[4 * a4 + 2] = y.z;
          yp[4 * a4 + 3] = y.w;
        }
      }
      const int cp = prow ? coord[p] : 0;
      const int sxp = cp >> 16, syp = cp & 0xffff;
      for (int q0 = qb; q0 < qb + 16 && q0 < qc1; q0 += 4) {
        if (diag && q0 > plast) break;  // uniform
        // the four columns' table rows first, then four independent accumulations
        const float4* w4[4];
        const float4* y4[4];
        int zr[4], zq[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int q = q0 + j < qc1 ? q0 + j : q0;  // a column past the block repeats q0 (result unused)
          const int cq = coord[q];
          int dx = sxp - (cq >> 16), dy = syp - (cq & 0xffff);
          dx += dx < 0 ? LX : 0;
          dy += dy < 0 ? LY : 0;
          const int r = dx * LY + dy;
          w4[j] = reinterpret_cast<const float4*>(wtl + r * D);
          if constexpr (kPfYCached<D>) {  // one row for the whole wave, through the scalar cache
            const int mq = __builtin_amdgcn_readfirstlane(modes[q]);
            y4[j] = reinterpret_cast<const float4*>(Yi + (size_t)mq * D);
            zq[j] = 0;
          } else {
            y4[j] = reinterpret_cast<const float4*>(yocc + q * D);
            zq[j] = swz<D>(q);
          }
          zr[j] = swz<D>(r);
        }
        float acc[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int a4 = 0; a4 < C4; ++a4) {
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const float4 w = w4[j][a4 ^ zr[j]];
            const float4 yq = y4[j][a4 ^ zq[j]];
            acc[j] = fmaf(yp[4 * a4] * yq.x, w.x, acc[j]);
            acc[j] = fmaf(yp[4 * a4 + 1] * yq.y, w.y, acc[j]);
            acc[j] = fmaf(yp[4 * a4 + 2] * yq.z, w.z, acc[j]);
            acc[j] = fmaf(yp[4 * a4 + 3] * yq.w, w.w, acc[j]);
          }
        }
        if (F0 && prow) {
#pragma unroll
          for (int j = 0; j < 4; ++j)
            if (q0 + j < qc1) acc[j] += F0[(size_t)mp * two_n + modes[q0 + j]];
        }
        if (prow && p >= q0) {
#pragma unroll
          for (int j = 0; j < 4; ++j)
            if (q0 + j < qc1) A[(size_t)(q0 + j) * ld + p] = acc[j];
          if (!diag || p >= q0 + 4)  // rows q0 .. q0 + 3 of column p (rows >= n are padding)
            *reinterpret_cast<float4*>(A + (size_t)p * ld + q0) = make_float4(-acc[0], -acc[1], -acc[2], -acc[3]);
        }
      }
    }
  }
}

// One image of the Pfaffian amplitude (64 < n <= NT, n_up % 4 == 0): occupied modes of the item
// under group element g, their features, the pairing square in gmat (the image's slice of
// gstride >= pfl_matrix_floats<P>(n) floats), pf_lazy, and (sign, log|Pf|) with the reversal sign
// of the Ne creation operators (fermion_inverse_sign).  lds: pf_cm_lds_bytes<D, NT, P>(n, N).
template <int D, int NT, int P, int KB>
__device__ void image_pf_cm(int img, int N, int LX, int LY, int n, int n_up, const int8_t* __restrict__ s,
                            const int32_t* __restrict__ item_b, const int32_t* __restrict__ item_hop, int img0,
                            int ipi, const float* __restrict__ Y, const float* __restrict__ wt,
                            const float* __restrict__ F0, float* __restrict__ img_sign,
                            double* __restrict__ img_logabs, float* __restrict__ gmat, size_t gstride,
                            float* lds) {
  constexpr int C4 = D / 4;
  const int item = img / ipi;
  const int g = (ipi == kGroup) ? (img & 7) : 0;
  const int32_t hop = item_hop ? item_hop[item] : kHopNone;
  if (hop == kHopInvalid) return;
  const int b = item_b ? item_b[item] : item;
  const int two_n = 2 * N;
  float* A = gmat + (size_t)(img - img0) * gstride;
  float* yocc = lds;  // (D <= 16)
  float* wtl = kPfYCached<D> ? lds : yocc + n * D;
  int* modes = reinterpret_cast<int*>(lds + pf_cm_region_floats<D, NT, P>(n, N));
  int* coord = modes + n;  // (x << 16) | y of each mode's site
  int* flag = coord + n;

  const int t = threadIdx.x, lane = t & (kWave - 1), wave = t / kWave;
  if (wave == 0) {
    const int cnt = occupied_modes(s + (size_t)b * two_n, hop, two_n, N, LX, LY, g, modes, n);
    if (lane == 0) flag[0] = (cnt != n) ? 2 : 0;
  }
  __syncthreads();
  if (flag[0]) {  // a configuration with another electron count than the model's
    if (t == 0) {
      img_sign[img] = 0.f;
      img_logabs[img] = NAN;
    }
    return;
  }
  const float* Yi = Y + (size_t)(img - img0) * (two_n * D);
  for (int e = t; !kPfYCached<D> && e < n * C4; e += NT) {
    const int p = e / C4, a4 = e - p * C4;
    reinterpret_cast<float4*>(yocc)[p * C4 + (a4 ^ swz<D>(p))] =
        reinterpret_cast<const float4*>(Yi + (size_t)modes[p] * D)[a4];
  }
  for (int p = t; p < n; p += NT) {
    const int m = modes[p];
    const int site = m - (m >= N ? N : 0);
    coord[p] = ((site / LY) << 16) | (site % LY);
  }
  pf_build_cm<D, NT>(A, pfl_ld(n), n, n_up, N, LX, LY, two_n, g, yocc, Yi, wtl, modes, coord, wt, F0);
  __syncthreads();
  float sg;
  double la;
  pf_lazy<NT, P, KB>(A, n, lds, sg, la);
  if (t == 0) {
    sg *= parity_sign((long long)n * (n - 1) / 2);
    img_sign[img] = sg;
    img_logabs[img] = la;
  }
}
