// Projection combine, Hubbard hop enumeration and local-energy reduction on gfx950.

#include "nh_common.h"
#include "nh_internal.h"

namespace nh {

// Signed log-sum-exp of the ipi image amplitudes of every item, in FP64.
__global__ void k_combine(const int32_t* __restrict__ item_hop, int item0, int item_end, int ipi,
                          const float* __restrict__ img_sign, const double* __restrict__ img_logabs,
                          float* __restrict__ sign, double* __restrict__ logabs) {
  const int item = item0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (item >= item_end) return;
  if (item_hop && item_hop[item] == kHopInvalid) return;
  double mx = -INFINITY;
  bool nan = false;
  for (int g = 0; g < ipi; ++g) {
    const size_t i = (size_t)item * ipi + g;
    const double l = img_logabs[i];
    if (isnan(l)) nan = true;
    if (img_sign[i] != 0.f && l > mx) mx = l;
  }
  if (nan) {
    sign[item] = 0.f;
    logabs[item] = NAN;
    return;
  }
  double tot = 0.0;
  if (mx > -INFINITY)
    for (int g = 0; g < ipi; ++g) {
      const size_t i = (size_t)item * ipi + g;
      const float sg = img_sign[i];
      if (sg != 0.f) tot += (double)sg * exp(img_logabs[i] - mx);
    }
  if (tot == 0.0) {
    sign[item] = 0.f;
    logabs[item] = -INFINITY;
  } else {
    sign[item] = tot > 0.0 ? 1.f : -1.f;
    logabs[item] = mx + log(fabs(tot));
  }
}

int launch_combine(hipStream_t st, const int32_t* item_hop, int item0, int item_end, int ipi,
                   const float* img_sign, const double* img_logabs, float* sign, double* logabs) {
  const int n = item_end - item0;
  if (n <= 0) return NH_OK;
  hipLaunchKernelGGL(k_combine, dim3((n + 255) / 256), dim3(256), 0, st, item_hop, item0, item_end, ipi,
                     img_sign, img_logabs, sign, logabs);
  NH_LAUNCH_CHECK();
  return NH_OK;
}

// Neighbour offsets in table order: nearest (+x, -x, +y, -y) then diagonal.
__constant__ int c_ndx[8] = {1, -1, 0, 0, 1, 1, -1, -1};
__constant__ int c_ndy[8] = {0, 0, 1, -1, 1, -1, 1, -1};

// One wave per configuration: every hop of an occupied mode to an empty partner of the same
// spin, in the order (mode ascending, partner in table order).  <s'|H|s> = -t_ij (-1)^k with
// k the occupied modes strictly between the two modes.  Slots past the count get
// kHopInvalid.  `occ_before` is LDS scratch of 2N + 1 ints per wave.
__global__ void k_hops(int LX, int LY, float t, float tp, int z, int batch, int b0, int two_n,
                       const int8_t* __restrict__ s, int max_hops, int32_t* __restrict__ count,
                       int32_t* __restrict__ item_b, int32_t* __restrict__ item_hop,
                       float* __restrict__ coef, int32_t* __restrict__ out_from,
                       int32_t* __restrict__ out_to) {
  extern __shared__ int sh_cnt[];
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const int lb = blockIdx.x * (blockDim.x / kWave) + wave;  // local sample
  if (lb >= batch) return;
  const int N = LX * LY;
  const int8_t* sb = s + (size_t)(b0 + lb) * two_n;
  int* cntb = sh_cnt + wave * (two_n + 1);
  // occupied modes strictly below m, for all m
  int run = 0;
  for (int m0 = 0; m0 < two_n; m0 += kWave) {
    const int m = m0 + lane;
    const bool occ = m < two_n && sb[m] > 0;
    const uint64_t mask = __ballot(occ);
    if (m < two_n) cntb[m] = run + lanes_below(mask);
    run += __popcll(mask);
  }
  if (lane == 0) cntb[two_n] = run;
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();

  int base = 0;
  const size_t slot0 = (size_t)lb * max_hops;
  for (int m0 = 0; m0 < two_n; m0 += kWave) {
    const int m = m0 + lane;
    const bool occ = m < two_n && sb[m] > 0;
    int spin = 0, x = 0, y = 0;
    unsigned valid = 0;
    if (occ) {
      spin = m >= N;
      const int site = m - spin * N;
      x = site / LY;
      y = site - x * LY;
      for (int k = 0; k < z; ++k) {
        int xx = x + c_ndx[k], yy = y + c_ndy[k];
        xx += (xx < 0) ? LX : 0;
        xx -= (xx >= LX) ? LX : 0;
        yy += (yy < 0) ? LY : 0;
        yy -= (yy >= LY) ? LY : 0;
        const int mt = spin * N + xx * LY + yy;
        if (sb[mt] < 0) valid |= 1u << k;
      }
    }
    const int c = __popc(valid);
    // exclusive wave scan of c
    int incl = c;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      const int v = __shfl_up(incl, off);
      if (lane >= off) incl += v;
    }
    int pos = base + incl - c;
    for (int k = 0; k < z; ++k) {
      if (!(valid >> k & 1u)) continue;
      int xx = x + c_ndx[k], yy = y + c_ndy[k];
      xx += (xx < 0) ? LX : 0;
      xx -= (xx >= LX) ? LX : 0;
      yy += (yy < 0) ? LY : 0;
      yy -= (yy >= LY) ? LY : 0;
      const int mt = spin * N + xx * LY + yy;
      const int lo = m < mt ? m : mt, hi = m < mt ? mt : m;
      const int between = cntb[hi] - cntb[lo + 1];
      const float tt = (k < 4) ? t : tp;
      const float h = (between & 1) ? tt : -tt;
      if (pos < max_hops) {
        if (item_hop) item_hop[slot0 + pos] = pack_hop(m, mt);
        if (item_b) item_b[slot0 + pos] = b0 + lb;
        if (coef) coef[slot0 + pos] = h;
        if (out_from) out_from[slot0 + pos] = m;
        if (out_to) out_to[slot0 + pos] = mt;
      }
      ++pos;
    }
    base += __shfl(incl, kWave - 1);
  }
  for (int p = base + lane; p < max_hops; p += kWave) {
    if (item_hop) item_hop[slot0 + p] = kHopInvalid;
    if (item_b) item_b[slot0 + p] = b0 + lb;
    if (coef) coef[slot0 + p] = 0.f;
    if (out_from) out_from[slot0 + p] = -1;
    if (out_to) out_to[slot0 + p] = -1;
  }
  if (lane == 0 && count) count[lb] = base;
}

int launch_hops(hipStream_t st, int LX, int LY, float t, float tp, int z, int batch, int b0,
                const int8_t* s, int max_hops, int32_t* count, int32_t* item_b, int32_t* item_hop,
                float* coef, int32_t* out_from, int32_t* out_to) {
  if (batch <= 0) return NH_OK;
  const int two_n = 2 * LX * LY;
  const int wpb = 4;
  const size_t lds = sizeof(int) * (size_t)wpb * (two_n + 1);
  hipLaunchKernelGGL(k_hops, dim3((batch + wpb - 1) / wpb), dim3(wpb * kWave), lds, st, LX, LY, t, tp, z,
                     batch, b0, two_n, s, max_hops, count, item_b, item_hop, coef, out_from, out_to);
  NH_LAUNCH_CHECK();
  return NH_OK;
}

// E_loc[b] = U D(s_b) + sum_h coef_h psi(s_h) / psi(s_b); one wave per sample, FP64 sums.
__global__ void k_eloc_reduce(int N, double U, int batch, int b0, const int8_t* __restrict__ s,
                              const float* __restrict__ sign_s, const double* __restrict__ logabs_s,
                              int max_hops, const int32_t* __restrict__ count,
                              const float* __restrict__ coef, const float* __restrict__ hop_sign,
                              const double* __restrict__ hop_logabs, double* __restrict__ eloc) {
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const int lb = blockIdx.x * (blockDim.x / kWave) + wave;
  if (lb >= batch) return;
  const int b = b0 + lb;
  const int8_t* sb = s + (size_t)b * 2 * N;
  double acc = 0.0;
  for (int i = lane; i < N; i += kWave) acc += (sb[i] > 0 && sb[N + i] > 0) ? U : 0.0;
  const double l0 = logabs_s[b];
  const float s0 = sign_s[b];
  const int c = count[lb];
  if (c > max_hops) {  // more electrons than the model holds: only max_hops slots were written
    if (lane == 0) eloc[b] = NAN;
    return;
  }
  for (int h = lane; h < c; h += kWave) {
    const size_t i = (size_t)lb * max_hops + h;
    const float sg = hop_sign[i];
    if (sg != 0.f) acc += (double)coef[i] * (double)(sg * s0) * exp(hop_logabs[i] - l0);
    else if (isnan(hop_logabs[i])) acc += NAN;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
  if (lane == 0) eloc[b] = acc;
}

int launch_eloc_reduce(hipStream_t st, int N, double U, int batch, int b0, const int8_t* s,
                       const float* sign_s, const double* logabs_s, int max_hops, const int32_t* count,
                       const float* coef, const float* hop_sign, const double* hop_logabs, double* eloc) {
  if (batch <= 0) return NH_OK;
  const int wpb = 4;
  hipLaunchKernelGGL(k_eloc_reduce, dim3((batch + wpb - 1) / wpb), dim3(wpb * kWave), 0, st, N, U, batch, b0,
                     s, sign_s, logabs_s, max_hops, count, coef, hop_sign, hop_logabs, eloc);
  NH_LAUNCH_CHECK();
  return NH_OK;
}

}  // namespace nh
