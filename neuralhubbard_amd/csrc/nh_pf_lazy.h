// Blocked Parlett-Reid Pfaffian with lazy pivoting (pf_lazy) on a full skew square stored
// column-major, and its scratch size.  A fragment, not a standalone header: it is included
// inside namespace nh of nh_linalg.hip after kWave, wave_max_u32, row16_sum_f32, wave_sum_f64 and
// nh_f32x4, and by the host emulation under tests/host, which supplies those and the wave
// builtins itself.
//
// Why this form: every step of Parlett-Reid reads two columns of the remaining matrix (the
// pivot column and the column of the chosen pivot row).  With the matrix as a packed row-major
// triangle those reads, and the row/column exchange that follows, are one cache line per row;
// here thread t owns physical row t of a column-major square, nothing is ever exchanged (the
// elimination order is kept as a live set instead), so each column read is one coalesced load
// per wave and a step writes nothing to the matrix.
//
// Elimination order: step k takes kap = the smallest live row, chooses the pivot row beta =
// argmax |A[i][kap]| over the live rows i != kap (largest to 2^-(23-KB) relative, smallest row
// among those), and removes both.  With piv = A[beta][kap] (effective value),
//   Pf(A) = (-1)^(#live rows before beta, kap included, - 1) * (-piv) * Pf(S),
// S the Schur complement on the remaining live rows in their stored order, with
//   S[i][j] = A[i][j] + u_i w_j - w_i u_j,   u_i = A[i][kap] / piv,  w_i = A[i][beta]
// (Wimmer, ACM TOMS 38 (2012) 30, pfaffian_LTL, with the pivot swap replaced by the live-set
// order); the signs (-1)^(...) multiply to the sign of the elimination order as a permutation of
// the original rows, counted once at the end.  Steps are grouped in panels of P: inside a panel
// a column is corrected left-looking by the panel's earlier (u, w) pairs; after it, the Schur
// complement of the live rows is formed on 16 x 16 FP32 MFMA tiles of its lower block triangle
// and written, compacted to the live rows in their order, into the other of two buffers (each
// tile with its negated transpose), so the matrix shrinks by 2P rows per panel and no dead row
// is carried along.

#pragma once

// Leading dimension of a square of n rows: whole 16-row MFMA blocks.
__host__ __device__ inline int pfl_ld(int n) { return (n + 15) & ~15; }

// Operand row stride of the trailing update: rows of 16 floats padded so that the four rows of
// one MFMA K block (a wave's lanes 0-15, 16-31, ...) start 16 banks apart.
__host__ __device__ inline int pfl_ldq(int ld) { return ld % 32 == 0 ? ld + 16 : ld; }

// Floats of the two matrix buffers of pf_lazy for n rows: the n x n square (leading dimension
// pfl_ld(n)) and the first compacted Schur complement (n - 2P rows), each a multiple of 64.
template <int P>
__host__ __device__ inline size_t pfl_matrix_floats(int n) {
  const size_t l0 = (size_t)pfl_ld(n), l1 = n > 2 * P ? (size_t)pfl_ld(n - 2 * P) : 0;
  return ((l0 * l0 + 63) & ~(size_t)63) + ((l1 * l1 + 63) & ~(size_t)63);
}

// Scratch floats: the panel rows UW [ld][2P + 2] (u_q, w_q interleaved; after each panel the
// same space holds the q-major operands UT [P][ldq], WT [P][ldq]), the column buffer [NT], and
// for even and odd steps each: pivot keys (16 words), the waves' live-row masks (32 words) and
// two smallest live rows (32 words); pivots [NT / 2], the elimination order [NT], the original
// row of each current row (two arrays [NT]), the current row of each compacted row [NT], parity
// partials (16 words).
template <int NT, int P>
__host__ __device__ inline int pfl_scratch_floats(int ld) {
  const int uw = ld * (2 * P + 2), ops = 2 * P * pfl_ldq(ld);
  return ((uw > ops ? uw : ops) + NT + 2 * (16 + 32 + 32) + NT / 2 + NT + 3 * NT + 16 + 3) & ~3;
}

// Calls f(StepIndex<S>{}) for S = 0 .. N - 1 with S a compile-time constant, so that the
// per-step register arrays are indexed statically (a counted loop of P = 16 large steps is not
// unrolled by the compiler).
template <int S>
struct StepIndex {
  static constexpr int value = S;
};
template <int S, int N>
struct StaticSteps {
  template <class F>
  __device__ static void run(F& f) {
    f(StepIndex<S>{});
    StaticSteps<S + 1, N>::run(f);
  }
};
template <int N>
struct StaticSteps<N, N> {
  template <class F>
  __device__ static void run(F&) {}
};

constexpr int kNoRow = 1 << 20;  // "no such row" in the waves' smallest-row slots

__device__ inline float uniform_f32(float x) {
  return __uint_as_float((uint32_t)__builtin_amdgcn_readfirstlane((int)__float_as_uint(x)));
}

// A: n x n skew matrix, A[i][j] at j * pfl_ld(n) + i (both triangles; the diagonal is never
// read), followed by room for the compacted complements (pfl_matrix_floats<P>(n) floats in all,
// 16-byte aligned); it is overwritten.  n <= NT <= 2^KB, one workgroup of NT threads.  sign_out
// (+1/-1, 0 if singular) and log_out = log|Pf| summed in FP64 are valid in thread 0 (log_out
// -inf if singular).
template <int NT, int P, int KB>
__device__ void pf_lazy(float* __restrict__ A, int n, float* scr, float& sign_out, double& log_out) {
  constexpr int NW = NT / kWave;
  constexpr int LS = 2 * P + 2;  // UW row stride: 8-byte pairs, a wave's b64 writes 2-way at most
  constexpr uint32_t kIdx = (1u << KB) - 1u;
  static_assert(NT <= (1 << KB), "row index in the key bits");
  static_assert(NW >= 4 && NW <= 16 && NW % 4 == 0, "pivot keys: 4, 8 or 16 waves");
  static_assert(P % 4 == 0, "MFMA K blocks of 4 within the u or the w half");
  const int ld0 = pfl_ld(n);
  const int ldq = pfl_ldq(ld0);
  const int uw = ld0 * LS, ops = 2 * P * ldq;
  float* UW = scr;
  float* UT = scr;
  float* WT = scr + P * ldq;
  float* cbuf = scr + (uw > ops ? uw : ops);
  uint32_t* keys2 = reinterpret_cast<uint32_t*>(cbuf + NT);  // [2][16]: even and odd steps
  uint32_t* amask2 = keys2 + 32;  // [2][32], per wave: 64-bit mask of its live rows other than kap
  int* lows2 = reinterpret_cast<int*>(amask2 + 64);  // [2][32], per wave: its two smallest such rows
  float* pivs = reinterpret_cast<float*>(lows2 + 64);
  int* order = reinterpret_cast<int*>(pivs + NT / 2);  // original rows kap_0, beta_0, kap_1, ...
  int* orig = order + NT;                               // original row of each current row
  int* orig2 = orig + NT;
  int* lmap = orig2 + NT;  // current row of each compacted row
  uint32_t* parts = reinterpret_cast<uint32_t*>(lmap + NT);
  const int t = threadIdx.x;
  const int lane = t & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(t / kWave);

  float* Acur = A;  // the current matrix (m live rows, leading dimension ldc) and the other buffer
  float* Anext = A + (((size_t)ld0 * ld0 + 63) & ~(size_t)63);
  int ldc = ld0;
  int m = n;
  bool alive = t < n;
  if (t < n) orig[t] = t;
  float sign = 1.f;  // sign of the product of the pivots A[kap][beta]
  bool singular = false;
  int kap = 0, blast = 0;
  const int nsteps = n / 2;
  float cnext = alive ? A[t] : 0.f;  // column kap = 0
  for (int done = 0; done < nsteps; done += P) {
    const int steps = nsteps - done < P ? nsteps - done : P;
    float pu[P], pw[P];  // this thread's u_q, w_q of the panel
#pragma unroll
    for (int q = 0; q < P; ++q) {
      pu[q] = 0.f;
      pw[q] = 0.f;
    }
    float ku = 0.f, kw = 0.f;  // the newest pair (u, w) of row kap, not yet read back from UW
    auto step = [&](auto sidx) {
      constexpr int s = decltype(sidx)::value;
      if (s >= steps || singular) return;  // uniform
      uint32_t* keys = keys2 + 16 * (s & 1);
      uint32_t* amask = amask2 + 32 * (s & 1);
      int* lows = lows2 + 32 * (s & 1);
      // (1) column kap with the panel's earlier pairs, the wave's pivot key and live rows
      float c = cnext;
#pragma unroll
      for (int q = 0; q + 1 < s; ++q) {
        const float2 x = *reinterpret_cast<const float2*>(UW + kap * LS + 2 * q);  // u_q, w_q of row kap
        c = fmaf(pu[q], x.y, fmaf(-pw[q], x.x, c));
      }
      if (s > 0) c = fmaf(pu[s - 1], kw, fmaf(-pw[s - 1], ku, c));
      const bool cand = alive && t != kap;
      if (cand) cbuf[t] = c;
      uint32_t key = cand ? ((__float_as_uint(fabsf(c)) & ~kIdx) | (kIdx - (uint32_t)t)) : 0u;
      key = wave_max_u32(key);
      const uint64_t am = __ballot(cand);
      if (lane == 0) {  // the wave's live rows other than kap, and the two smallest of them
        const uint64_t am2 = am & (am - 1ull);
        keys[wave] = key;
        amask[2 * wave] = (uint32_t)am;
        amask[2 * wave + 1] = (uint32_t)(am >> 32);
        lows[2 * wave] = am ? kWave * wave + __builtin_ctzll(am) : kNoRow;
        lows[2 * wave + 1] = am2 ? kWave * wave + __builtin_ctzll(am2) : kNoRow;
      }
      // The step's only barrier.  The next step writes the other half of keys / amask / lows, and
      // cbuf only for its candidate rows, which are neither this step's beta (dead) nor its next
      // kap (the next step's kap): the two entries of cbuf read below.  UW rows are read below
      // at pairs q < s only, while the next step writes pair s + 1.
      __syncthreads();
      // (2) pivot row, column beta, the step's pair (u, w), the next kap and its pair
      uint32_t km = 0u;
#pragma unroll
      for (int w = 0; w < NW; w += 4) {
        const uint4 kv = *reinterpret_cast<const uint4*>(keys + w);
        km = max(km, max(max(kv.x, kv.y), max(kv.z, kv.w)));
      }
      km = (uint32_t)__builtin_amdgcn_readfirstlane((int)km);
      if (km == 0u) {  // no live row left besides kap: zero column
        singular = true;
        return;
      }
      const int beta = (int)(kIdx - (km & kIdx));
      // next kap: the smallest live row other than kap and beta.  Its column is untouched until
      // the panel's trailing update, so its load goes out together with column beta's: one memory
      // latency per step instead of two in a row
      int knext = kNoRow;
#pragma unroll
      for (int w = 0; w < NW; w += 2) {
        const int4 lw = *reinterpret_cast<const int4*>(lows + 2 * w);  // waves w, w + 1
        const int a = __builtin_amdgcn_readfirstlane(lw.x == beta ? lw.y : lw.x);
        const int b = __builtin_amdgcn_readfirstlane(lw.z == beta ? lw.w : lw.z);
        knext = min(knext, min(a, b));
      }
      knext = knext < m ? knext : 0;  // (a live row whenever the panel has another step)
      const bool alive_next = cand && t != beta;
      float v = alive ? Acur[beta * ldc + t] : 0.f;
      cnext = alive_next ? Acur[knext * ldc + t] : 0.f;
      // Every wave forms the next kap's pair itself, so that the next step has it in registers
      // and no barrier has to publish it: u = c[knext] / piv as row knext does, w = A[knext][beta]
      // + sum_{q<s} (u_q(knext) w_q(beta) - w_q(knext) u_q(beta)), one term per lane q of each
      // row of 16 lanes (the rounding of that sum may differ from row knext's own w; the pair
      // only corrects the column of kap, which is eliminated in the next step)
      const float ak = Acur[beta * ldc + knext];
      const float piv = uniform_f32(cbuf[beta]);
      const float ck = cbuf[knext];
      float dk = 0.f;
      if (s > 0) {
        const int q = lane & 15;
        if (q < s) {
          const float2 x = *reinterpret_cast<const float2*>(UW + beta * LS + 2 * q);
          const float2 y = *reinterpret_cast<const float2*>(UW + knext * LS + 2 * q);
          dk = fmaf(y.x, x.y, -y.y * x.x);
        }
        dk = row16_sum_f32(dk);
      }
#pragma unroll
      for (int q = 0; q < s; ++q) {
        const float2 x = *reinterpret_cast<const float2*>(UW + beta * LS + 2 * q);
        v = fmaf(pu[q], x.y, fmaf(-pw[q], x.x, v));
      }
      if (!(piv != 0.f) || !isfinite(piv)) {  // uniform
        singular = true;
        return;
      }
      float rp = __builtin_amdgcn_rcpf(piv);
      rp = fmaf(rp, fmaf(-piv, rp, 1.f), rp);  // one Newton step
      const float tau = c * rp;
      ku = ck * rp;
      kw = ak + dk;
      if (piv > 0.f) sign = -sign;  // Pf *= A[kap][beta] = -piv
      if (t == 0) {
        pivs[done + s] = piv;
        order[2 * (done + s)] = orig[kap];
        order[2 * (done + s) + 1] = orig[beta];
      }
      if (alive) {
        *reinterpret_cast<float2*>(UW + t * LS + 2 * s) = make_float2(tau, v);
        pu[s] = tau;
        pw[s] = v;
      }
      alive = alive_next;
      kap = knext;
      blast = beta;
    };
    StaticSteps<0, P>::run(step);
    __syncthreads();  // the steps' last reads of UW, cbuf and the masks are done
    if (singular) break;
    if (done + steps >= nsteps) break;
    // Schur complement of the live rows, compacted: a live row's new index is the number of live
    // rows before it (the last step's published masks without its beta)
    const int mn = m - 2 * steps;
    const int ldn = pfl_ld(mn);
    int pos = 0;
    const uint32_t* amask = amask2 + 32 * ((steps - 1) & 1);
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      uint64_t mw = (uint64_t)amask[2 * w] | ((uint64_t)amask[2 * w + 1] << 32);
      mw = (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)mw) |
           ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(mw >> 32)) << 32);
      if ((blast >> 6) == w) mw &= ~(1ull << (blast & 63));
      pos += w < wave ? __builtin_popcountll(mw) : (w == wave ? lanes_below(mw) : 0);
    }
    // q-major operands over UW (no step reads it any more), zero beyond the live rows
    if (alive) {
#pragma unroll
      for (int q = 0; q < P; ++q) {
        UT[q * ldq + pos] = pu[q];
        WT[q * ldq + pos] = pw[q];
      }
      lmap[pos] = t;
      orig2[pos] = orig[t];
    } else if (t >= mn && t < ldn) {
#pragma unroll
      for (int q = 0; q < P; ++q) {
        UT[q * ldq + t] = 0.f;
        WT[q * ldq + t] = 0.f;
      }
    }
    __syncthreads();
    const int nb = ldn / 16;
    const int ntiles = nb * (nb + 1) / 2;
    // each wave takes its tiles in batches of kTileBatch: the batch's loads go out together, so
    // one memory latency covers several tiles
    constexpr int kTileBatch = 4;
    for (int tile0 = wave; tile0 < ntiles; tile0 += NW * kTileBatch) {
      float cv[kTileBatch][4];
      int i0b[kTileBatch], j0b[kTileBatch];
#pragma unroll
      for (int b = 0; b < kTileBatch; ++b) {
        const int tile = tile0 + b * NW;
        i0b[b] = -1;
        j0b[b] = 0;
        if (tile < ntiles) {  // uniform
          // tile -> (ti, tj), tj <= ti, row-major over the lower triangle of blocks
          int ti = (int)((sqrtf(8.f * tile + 1.f) - 1.f) * 0.5f);
          while ((ti + 1) * (ti + 2) / 2 <= tile) ++ti;
          while (ti * (ti + 1) / 2 > tile) --ti;
          const int tj = tile - ti * (ti + 1) / 2;
          i0b[b] = 16 * ti;
          j0b[b] = 16 * tj;
          const int cn = j0b[b] + (lane & 15);
          const int src_c = cn < mn ? lmap[cn] : 0;
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int rn = i0b[b] + 4 * (lane >> 4) + r;
            cv[b][r] = (cn < mn && rn < mn) ? Acur[src_c * ldc + lmap[rn]] : 0.f;
          }
        }
      }
#pragma unroll
      for (int b = 0; b < kTileBatch; ++b) {
        if (i0b[b] < 0) break;  // uniform
        const int i0 = i0b[b], j0 = j0b[b];
        const int col = j0 + (lane & 15);
        const int r0 = i0 + 4 * (lane >> 4);
        nh_f32x4 acc;
        acc[0] = cv[b][0];
        acc[1] = cv[b][1];
        acc[2] = cv[b][2];
        acc[3] = cv[b][3];
        const int ia = i0 + (lane & 15);
#pragma unroll
        for (int kk = 0; kk < P / 2; ++kk) {
          const int q = 4 * kk + (lane >> 4);  // K index: u_q (q < P), then -w_(q-P)
          float av, bv;
          if (kk < P / 4) {
            av = UT[q * ldq + ia];
            bv = WT[q * ldq + col];
          } else {
            av = -WT[(q - P) * ldq + ia];
            bv = UT[(q - P) * ldq + col];
          }
          acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc, 0, 0, 0);
        }
        *reinterpret_cast<float4*>(Anext + col * ldn + r0) = make_float4(acc[0], acc[1], acc[2], acc[3]);
        if (i0 != j0) {  // the upper block: A[j][i] = -A[i][j]
#pragma unroll
          for (int r = 0; r < 4; ++r) Anext[(r0 + r) * ldn + col] = -acc[r];
        }
      }
    }
    __syncthreads();
    // the complement is the next panel's matrix: rows 0 .. mn - 1, in the live rows' order
    float* tmp = Acur;
    Acur = Anext;
    Anext = tmp;
    ldc = ldn;
    m = mn;
    int* to = orig;
    orig = orig2;
    orig2 = to;
    alive = t < m;
    kap = 0;
    cnext = alive ? Acur[t] : 0.f;
  }
  if (singular) {
    sign_out = 0.f;
    log_out = -INFINITY;
    return;
  }
  // Pf(A) = sgn(order) * prod A[kap][beta]: the sign of the elimination order by the parity of
  // its inversions (each thread counts the later entries below its own)
  uint32_t inv = 0u;
  if (t < n) {
    const int x = order[t];
    for (int j = t + 1; j < n; ++j) inv += order[j] < x ? 1u : 0u;
  }
  const uint64_t odd = __ballot((inv & 1u) != 0u);
  if (lane == 0) parts[wave] = (uint32_t)__popcll(odd) & 1u;
  __syncthreads();
  sign_out = sign;
  log_out = 0.0;
  if (wave == 0) {
    uint32_t par = 0u;
    for (int w = 0; w < NW; ++w) par ^= parts[w];
    if (par) sign_out = -sign;
    double l = 0.0;
    for (int i = lane; i < nsteps; i += kWave) l += log((double)fabsf(pivs[i]));
    log_out = wave_sum_f64(l);
  }
}
