// Residual CNN of the released backflow states on MFMA (gfx950), FP32 (v_mfma_f32_16x16x4_f32):
// an embedding of the four site occupations (empty, up, down, double) into C = 64 channels,
// n_blocks residual blocks h <- h + conv3x3(gelu(conv3x3(gelu(rmsnorm(h))))) with periodic 3x3
// convolutions and biases, RMSNorm over all channels and sites of one image without affine
// parameters, and the output h / sqrt(n_blocks + 1) split into y_up | y_dn of d = 32 features
// (network table of the reference: data/supplement/tableS1_network/network_settings.json:21-56;
// quantax ResConv as built at code/backflow/models.py:319-329; features at models.py:163-167).
//
// Lattices with LY == 16 sites per row and LX % 4 == 0, LX <= 16.  One block evaluates one
// image.  In k_resconv (16 LX threads) wave w owns lattice rows 4w .. 4w+3 for all 64 channels;
// the slab kernels k_resconv_fs / k_resconv_bf16 (32 LX threads) give those rows to two waves
// with two channel quarters each, so that two waves share every SIMD.  The residual stream h
// stays in registers in the MFMA result layout (lane l, register q -> site y = 4 (l / 16) + q,
// channel 16 quarter + l % 16).  A lattice row of 16 sites is the M = 16
// dimension of an MFMA tile, so the y-shift of a 3x3 tap is a DPP rotation inside rows of 16
// lanes (periodic for free) and the x-shift reads the neighbouring lattice row.  Each conv
// input sits in LDS as [site][64 + 4 pad], channel c at (c % 4) * 16 + c / 4, so a lane's 8
// K-values of consecutive 16x16x4 steps are two contiguous float4 reads.  Weights are read in
// their natural [tap][c_in][c_out] layout: each B fragment row is 16 consecutive floats.

#include "nh_common.h"
#include "nh_internal.h"

namespace nh {

namespace {

typedef float rc_f32x4 __attribute__((ext_vector_type(4)));

constexpr int kRcC = 64;              // channels
constexpr int kRcLY = 16;             // sites per lattice row
constexpr int kRcStride = kRcC + 4;   // LDS floats per site (padding spreads the banks)
constexpr int kRcClasses = 4;         // empty, up, down, double
constexpr float kRmsEps = 1e-5f;      // added to the mean square before the square root

__device__ inline int rc_pos(int c) { return (c & 3) * 16 + (c >> 2); }

// DPP rotation within rows of 16 lanes: row_ror:1 (0x121) gives lane y the value of lane y - 1,
// row_ror:15 (0x12F) that of lane y + 1.
template <int CTRL>
__device__ inline float rc_rot(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xf, 0xf, false));
}

__device__ inline float rc_wave_sum(float x) {
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) x += __shfl_xor(x, o, kWave);
  return x;
}

// Linear head of the BCS states: y[k][site][dd] = sum_c head[k][c][dd] x[site][c], x the scaled
// network output staged in LDS as [site][kRcStride] FP32 (natural channel order).
__device__ inline void rc_head(const float* xs, const float* __restrict__ head, int N, float* __restrict__ yo) {
  for (int o = threadIdx.x; o < 2 * N * kRcC; o += blockDim.x) {
    const int dd = o & (kRcC - 1);
    const int ks = o / kRcC;  // k * N + site
    const int site = ks % N, k = ks / N;
    const float* hk = head + (size_t)k * kRcC * kRcC + dd;
    const float* x = xs + site * kRcStride;
    float acc = 0.f;
#pragma unroll 8
    for (int c = 0; c < kRcC; ++c) acc = fmaf(hk[c * kRcC], x[c], acc);
    yo[o] = acc;
  }
}

// act[site][pos(c)] = gelu(scale * v) for the wave's rows x0 .. x0+3 and channel quarters
// q0 .. q0+QN-1
template <int QN>
__device__ inline void rc_store_q(float* act, const rc_f32x4 (&v)[4][QN], float scale, int x0, int q0, int lane) {
  const int j = lane & 15, g16 = lane >> 4;
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int qq = 0; qq < QN; ++qq) {
      const int pos = rc_pos(16 * (q0 + qq) + j);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int site = (x0 + r) * kRcLY + 4 * g16 + q;
        act[site * kRcStride + pos] = gelu_tanh(scale * v[r][qq][q]);
      }
    }
}

__device__ inline void rc_store(float* act, const rc_f32x4 (&v)[4][4], float scale, int x0, int lane) {
  rc_store_q<4>(act, v, scale, x0, 0, lane);
}

// acc = bias + periodic conv3x3 of act (weights [tap][c_in][c_out]) on rows x0 .. x0+3
template <int DYI>
__device__ inline void rc_taps(const float* act, const float* __restrict__ Wl, int LX, int x0, int lane, int kh,
                               rc_f32x4 (&acc)[4][4]) {
  const int j = lane & 15, g16 = lane >> 4;
  float fs[6][8];
#pragma unroll
  for (int u = 0; u < 6; ++u) {
    int xr = x0 - 1 + u;
    xr += xr < 0 ? LX : 0;
    xr -= xr >= LX ? LX : 0;
    const float4* p = reinterpret_cast<const float4*>(act + (xr * kRcLY + j) * kRcStride + 16 * g16 + 8 * kh);
    const float4 p0 = p[0], p1 = p[1];
    const float v[8] = {p0.x, p0.y, p0.z, p0.w, p1.x, p1.y, p1.z, p1.w};
#pragma unroll
    for (int kk = 0; kk < 8; ++kk) {
      if (DYI == 1) fs[u][kk] = v[kk];
      else if (DYI == 0) fs[u][kk] = rc_rot<0x121>(v[kk]);  // dy = -1: lane y <- y - 1
      else fs[u][kk] = rc_rot<0x12F>(v[kk]);                // dy = +1: lane y <- y + 1
    }
  }
#pragma unroll
  for (int dxi = 0; dxi < 3; ++dxi) {
    const int tap = dxi * 3 + DYI;
    // B fragment of K-step kk: lane l holds W[tap][c_in = 32 kh + 4 kk + l / 16][16 qq + l % 16]
    const float* wt = Wl + ((size_t)tap * kRcC + 32 * kh + g16) * kRcC + j;
#pragma unroll
    for (int qq = 0; qq < 4; ++qq) {
      float wv[8];
#pragma unroll
      for (int kk = 0; kk < 8; ++kk) wv[kk] = wt[4 * kk * kRcC + 16 * qq];
#pragma unroll
      for (int kk = 0; kk < 8; ++kk)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          acc[r][qq] = __builtin_amdgcn_mfma_f32_16x16x4f32(fs[r + dxi][kk], wv[kk], acc[r][qq], 0, 0, 0);
    }
  }
}

__device__ inline void rc_conv(const float* act, const float* __restrict__ Wl, const float* __restrict__ bl, int LX,
                               int x0, int lane, rc_f32x4 (&acc)[4][4]) {
  const int j = lane & 15;
#pragma unroll
  for (int qq = 0; qq < 4; ++qq) {
    const float bv = bl[16 * qq + j];
#pragma unroll
    for (int r = 0; r < 4; ++r) acc[r][qq] = rc_f32x4{bv, bv, bv, bv};
  }
#pragma unroll 1
  for (int kh = 0; kh < 2; ++kh) {
    rc_taps<0>(act, Wl, LX, x0, lane, kh, acc);
    rc_taps<1>(act, Wl, LX, x0, lane, kh, acc);
    rc_taps<2>(act, Wl, LX, x0, lane, kh, acc);
  }
}

// emb [4][64] (occupation class -> channels), Wc [2 n_blocks][9][64][64], Bc [2 n_blocks][64];
// Y[(img - img0)][spin][site][32]
__global__ void __launch_bounds__(256) k_resconv(int LX, int n_blocks, const float* __restrict__ emb,
                                                 const float* __restrict__ Wc, const float* __restrict__ Bc,
                                                 const float* __restrict__ head, float out_scale,
                                                 const int8_t* __restrict__ s,
                                                 const int32_t* __restrict__ item_b,
                                                 const int32_t* __restrict__ item_hop, int img0, int img_end,
                                                 int ipi, float* __restrict__ Y) {
  extern __shared__ float lds[];
  const int N = LX * kRcLY;
  float* act = lds;
  float* part = lds + N * kRcStride;  // one partial sum per wave
  const int img = img0 + blockIdx.x;
  if (img >= img_end) return;
  const int item = img / ipi;
  const int grp = (ipi == kGroup) ? (img & 7) : 0;
  const int32_t hop = item_hop ? item_hop[item] : kHopNone;
  if (hop == kHopInvalid) return;
  const int b = item_b ? item_b[item] : item;
  const int t = threadIdx.x, lane = t & (kWave - 1), wave = t / kWave, nw = blockDim.x / kWave;
  const int x0 = 4 * wave, j = lane & 15, g16 = lane >> 4;

  // embedding of the image's occupations: class = n_up + 2 n_dn
  rc_f32x4 h[4][4];
  {
    const int8_t* sb = s + (size_t)b * 2 * N;
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int site = (x0 + r) * kRcLY + 4 * g16 + q;
        const int up = item_occ(sb, hop, group_map(site, grp, N, LX, kRcLY), N) > 0;
        const int dn = item_occ(sb, hop, group_map(N + site, grp, N, LX, kRcLY), N) > 0;
        const float* e = emb + (up + 2 * dn) * kRcC + j;
#pragma unroll
        for (int qq = 0; qq < 4; ++qq) h[r][qq][q] = e[16 * qq];
      }
  }
  const float inv_count = 1.0f / (float)(kRcC * N);
  for (int blk = 0; blk < n_blocks; ++blk) {
    float ss = 0.f;
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int qq = 0; qq < 4; ++qq)
#pragma unroll
        for (int q = 0; q < 4; ++q) ss = fmaf(h[r][qq][q], h[r][qq][q], ss);
    ss = rc_wave_sum(ss);
    if (lane == 0) part[wave] = ss;
    __syncthreads();  // partial sums visible; every wave is done reading the previous conv input
    float tot = 0.f;
    for (int w = 0; w < nw; ++w) tot += part[w];
    const float rinv = 1.0f / sqrtf(fmaf(tot, inv_count, kRmsEps));
    rc_store(act, h, rinv, x0, lane);
    __syncthreads();
    const float* W1 = Wc + (size_t)(2 * blk) * 9 * kRcC * kRcC;
    rc_f32x4 acc[4][4];
    rc_conv(act, W1, Bc + (2 * blk) * kRcC, LX, x0, lane, acc);
    __syncthreads();  // every wave is done reading the first conv input
    rc_store(act, acc, 1.0f, x0, lane);
    __syncthreads();
    rc_conv(act, W1 + 9 * kRcC * kRcC, Bc + (2 * blk + 1) * kRcC, LX, x0, lane, acc);
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int qq = 0; qq < 4; ++qq) h[r][qq] += acc[r][qq];
  }
  if (head) {  // BCS: d = 64 features per spin from the linear head
    __syncthreads();  // every wave is done reading the last conv input
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int qq = 0; qq < 4; ++qq)
#pragma unroll
        for (int q = 0; q < 4; ++q)
          act[((x0 + r) * kRcLY + 4 * g16 + q) * kRcStride + 16 * qq + j] = out_scale * h[r][qq][q];
    __syncthreads();
    rc_head(act, head, N, Y + (size_t)(img - img0) * (2 * N * kRcC));
    return;
  }
  // features: channel c = 32 spin + a
  float* yo = Y + (size_t)(img - img0) * (2 * N * (kRcC / 2));
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int qq = 0; qq < 4; ++qq)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int site = (x0 + r) * kRcLY + 4 * g16 + q;
        yo[((size_t)(qq >> 1) * N + site) * (kRcC / 2) + 16 * (qq & 1) + j] = out_scale * h[r][qq][q];
      }
}

// ---------------------------------------------------------------- split-BF16 variant
//
// v_mfma_f32_16x16x32_bf16 with FP32 accumulation; every operand is carried as hi + lo BF16
// parts and the product accumulates lo*hi + hi*lo + hi*hi, which keeps the features within
// ~1e-5 of FP32.  The operand roles are swapped against the FP32 kernel: A = weights (M = 16
// output channels of a quarter), B = activations (N = 16 sites of a lattice row), so the
// result (lane l, register e -> channel 16 quarter + 4 (l / 16) + e, site y = l % 16) gives each
// lane four consecutive channels of one site, stored to LDS as one 8-byte hi and one 8-byte lo
// write.  Activations: two planes (hi, lo) of [site][64 + 8 pad] BF16; a lane's B fragment is
// the 8 channels 32 kh + 8 (l / 16) .. + 7 of its site, one 16-byte read per plane.
//
// Weights stream through LDS in slabs: a slab is one (conv, K-half kh, tap row dyi), i.e. the
// three taps dxi = 0..2 for all four quarters, hi then lo (24 KB).  The workgroup loads slab
// k + 1 from global memory into registers while its waves run the MFMAs of slab k from LDS,
// then writes it to the other LDS buffer; one barrier per slab.  Every weight byte is read from
// global memory once per image (not once per wave) and its latency hides behind the wave's 72
// MFMAs of the slab (two waves per SIMD: each wave owns four lattice rows and two of the four
// channel quarters).  Host packing (neuralhubbard_amd/mixed.py, pack_resconv_bf16): slab-major
// [conv][kh][dyi][plane][dxi][qq][lane][8 BF16] with W[tap = 3 dxi + dyi][32 kh + 8 (l / 16) + e]
// [16 qq + l % 16].

typedef short rb_bf16x8 __attribute__((ext_vector_type(8)));
typedef int rb_i32x4 __attribute__((ext_vector_type(4)));  // one 16-byte fragment

constexpr int kRbStride = kRcC + 8;            // BF16 per site per plane (144 bytes: conflict-free reads)
constexpr int kSlabFrag = 3 * 4 * kWave;       // 16-byte fragments per plane of one slab
constexpr int kSlabsPerConv = 6;               // 2 K-halves x 3 tap rows

__device__ inline uint32_t rb_bf16(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return u >> 16;
}

// x * sigmoid(2 u) = 0.5 x (1 + tanh u), u = sqrt(2/pi) (x + 0.044715 x^3), with the hardware
// exp2 and reciprocal
__device__ inline float rb_gelu(float x) {
  const float u2 = -2.0f * 0.7978845608028654f * 1.4426950408889634f * fmaf(0.044715f * x, x * x, x);
  return x * __builtin_amdgcn_rcpf(1.0f + exp2f(u2));
}

template <int CTRL>
__device__ inline int rb_rot(int v) { return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xf, 0xf, false); }

// act planes = gelu(scale * v), BF16 hi and lo, for the wave's rows x0 .. x0+3 and channel
// quarters q0 .. q0+QN-1
template <int QN>
__device__ inline void rb_store(uint16_t* act, const rc_f32x4 (&v)[4][QN], float scale, int x0, int q0, int lane,
                                int N) {
  const int j = lane & 15, g16 = lane >> 4;
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int qq = 0; qq < QN; ++qq) {
      uint32_t hi[4], lo[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const float g = rb_gelu(scale * v[r][qq][e]);
        hi[e] = rb_bf16(g);
        lo[e] = rb_bf16(g - __uint_as_float(hi[e] << 16));
      }
      uint16_t* dst = act + ((x0 + r) * kRcLY + j) * kRbStride + 16 * (q0 + qq) + 4 * g16;
      *reinterpret_cast<uint2*>(dst) = make_uint2(hi[0] | (hi[1] << 16), hi[2] | (hi[3] << 16));
      *reinterpret_cast<uint2*>(dst + N * kRbStride) = make_uint2(lo[0] | (lo[1] << 16), lo[2] | (lo[3] << 16));
    }
}

// fragments per thread when the workgroup copies one slab (hi + lo planes)
template <int NT>
constexpr int rb_per_thread() {
  static_assert(2 * kSlabFrag % NT == 0, "threads must divide the slab");
  return 2 * kSlabFrag / NT;
}

// MFMAs of one slab (tap row DYI, K-half kh) on rows x0 .. x0+3 and quarters q0 .. q0+QN-1;
// wb = the slab in LDS
template <int DYI, int QN>
__device__ __forceinline__ void rb_taps(const uint16_t* act, const rb_i32x4* wb, int LX, int N, int x0, int q0,
                                        int lane, int kh, rc_f32x4 (&acc)[4][QN]) {
  const int j = lane & 15, g16 = lane >> 4;
  rb_bf16x8 bh[6], bl[6];
#pragma unroll
  for (int u = 0; u < 6; ++u) {
    int xr = x0 - 1 + u;
    xr += xr < 0 ? LX : 0;
    xr -= xr >= LX ? LX : 0;
    const uint16_t* src = act + (xr * kRcLY + j) * kRbStride + 32 * kh + 8 * g16;
    int4 vh = *reinterpret_cast<const int4*>(src);
    int4 vl = *reinterpret_cast<const int4*>(src + N * kRbStride);
    if (DYI != 1) {
      constexpr int CTRL = DYI == 0 ? 0x121 : 0x12F;  // dy = -1: lane y <- y - 1; dy = +1: y <- y + 1
      vh.x = rb_rot<CTRL>(vh.x); vh.y = rb_rot<CTRL>(vh.y); vh.z = rb_rot<CTRL>(vh.z); vh.w = rb_rot<CTRL>(vh.w);
      vl.x = rb_rot<CTRL>(vl.x); vl.y = rb_rot<CTRL>(vl.y); vl.z = rb_rot<CTRL>(vl.z); vl.w = rb_rot<CTRL>(vl.w);
    }
    bh[u] = __builtin_bit_cast(rb_bf16x8, vh);
    bl[u] = __builtin_bit_cast(rb_bf16x8, vl);
  }
#pragma unroll
  for (int dxi = 0; dxi < 3; ++dxi) {
#pragma unroll
    for (int qq = 0; qq < QN; ++qq) {
      const int f = (dxi * 4 + q0 + qq) * kWave + lane;
      const rb_bf16x8 wh = __builtin_bit_cast(rb_bf16x8, wb[f]);
      const rb_bf16x8 wl = __builtin_bit_cast(rb_bf16x8, wb[kSlabFrag + f]);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        acc[r][qq] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wl, bh[r + dxi], acc[r][qq], 0, 0, 0);
        acc[r][qq] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wh, bl[r + dxi], acc[r][qq], 0, 0, 0);
        acc[r][qq] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wh, bh[r + dxi], acc[r][qq], 0, 0, 0);
      }
    }
  }
}

// One slab step: start loading slab k + 1 into registers, run slab k from LDS buffer k & 1,
// park slab k + 1 in the other buffer, barrier.  The barrier also ends every wave's reads of
// the conv input.
template <int NT, int DYI, int QN>
__device__ __forceinline__ void rb_step(const uint16_t* act, rb_i32x4* wbuf, const rb_i32x4* __restrict__ Wsl, int& k,
                                        int n_slabs, int LX, int N, int x0, int q0, int t, int kh,
                                        rc_f32x4 (&acc)[4][QN]) {
  constexpr int kPer = rb_per_thread<NT>();
  const bool more = k + 1 < n_slabs;
  rb_i32x4 v[kPer];
  if (more) {
    const rb_i32x4* src = Wsl + (size_t)(k + 1) * 2 * kSlabFrag;
#pragma unroll
    for (int i = 0; i < kPer; ++i) v[i] = src[t + i * NT];
  }
  rb_taps<DYI, QN>(act, wbuf + (k & 1) * 2 * kSlabFrag, LX, N, x0, q0, t & (kWave - 1), kh, acc);
  if (more) {
    rb_i32x4* dst = wbuf + ((k + 1) & 1) * 2 * kSlabFrag;
#pragma unroll
    for (int i = 0; i < kPer; ++i) dst[t + i * NT] = v[i];
  }
  __syncthreads();
  ++k;
}

template <int NT, int QN>
__device__ __forceinline__ void rb_conv(const uint16_t* act, rb_i32x4* wbuf, const rb_i32x4* __restrict__ Wsl, int& k,
                                        int n_slabs, const float* __restrict__ bl, int LX, int N, int x0, int q0,
                                        int t, rc_f32x4 (&acc)[4][QN]) {
  const int g16 = (t & (kWave - 1)) >> 4;
#pragma unroll
  for (int qq = 0; qq < QN; ++qq) {
    const float4 b4 = *reinterpret_cast<const float4*>(bl + 16 * (q0 + qq) + 4 * g16);
#pragma unroll
    for (int r = 0; r < 4; ++r) acc[r][qq] = rc_f32x4{b4.x, b4.y, b4.z, b4.w};
  }
#pragma unroll 1
  for (int kh = 0; kh < 2; ++kh) {
    rb_step<NT, 0, QN>(act, wbuf, Wsl, k, n_slabs, LX, N, x0, q0, t, kh, acc);
    rb_step<NT, 1, QN>(act, wbuf, Wsl, k, n_slabs, LX, N, x0, q0, t, kh, acc);
    rb_step<NT, 2, QN>(act, wbuf, Wsl, k, n_slabs, LX, N, x0, q0, t, kh, acc);
  }
}

__host__ __device__ constexpr size_t rb_lds_floats(int N) {
  return (size_t)N * kRbStride + 16 * kSlabFrag + 16;  // act planes, two slabs (2 x 2 planes of rb_i32x4), partials
}

// emb [4][64], Wsl = split-BF16 slabs of the 2 n_blocks convs, Bc [2 n_blocks][64].
// QN channel quarters per wave: QN = 4, one wave per 4 lattice rows (NT = 16 LX); QN = 2, two
// waves per 4 rows, each with two quarters (NT = 32 LX), so that two waves share each SIMD.
template <int NT, int QN>
__global__ void __launch_bounds__(NT) k_resconv_bf16(int LX, int n_blocks, const float* __restrict__ emb,
                                                     const rb_i32x4* __restrict__ Wsl, const float* __restrict__ Bc,
                                                     const float* __restrict__ head, float out_scale,
                                                     const int8_t* __restrict__ s,
                                                     const int32_t* __restrict__ item_b,
                                                     const int32_t* __restrict__ item_hop, int img0, int img_end,
                                                     int ipi, float* __restrict__ Y) {
  static_assert(QN == 4 || QN == 2, "four or two channel quarters per wave");
  extern __shared__ float lds[];
  const int N = LX * kRcLY;
  uint16_t* act = reinterpret_cast<uint16_t*>(lds);                    // hi plane [N][72], lo plane [N][72]
  rb_i32x4* wbuf = reinterpret_cast<rb_i32x4*>(lds + (size_t)N * kRbStride);  // 2 slabs
  float* part = lds + (size_t)N * kRbStride + 16 * kSlabFrag;
  const int img = img0 + blockIdx.x;
  if (img >= img_end) return;
  const int item = img / ipi;
  const int grp = (ipi == kGroup) ? (img & 7) : 0;
  const int32_t hop = item_hop ? item_hop[item] : kHopNone;
  if (hop == kHopInvalid) return;
  const int b = item_b ? item_b[item] : item;
  const int t = threadIdx.x, lane = t & (kWave - 1), wave = t / kWave;
  constexpr int nw = NT / kWave;
  constexpr int nrg = nw * QN / 4;  // row groups of 4 lattice rows (= LX / 4)
  const int x0 = 4 * (wave % nrg), q0 = QN * (wave / nrg), j = lane & 15, g16 = lane >> 4;
  const int n_slabs = 2 * n_blocks * kSlabsPerConv;

  int k = 0;
#pragma unroll
  for (int i = 0; i < rb_per_thread<NT>(); ++i) wbuf[t + i * NT] = Wsl[t + i * NT];  // slab 0, visible after
                                                                                     // the first barrier

  rc_f32x4 h[4][QN];
  {
    const int8_t* sb = s + (size_t)b * 2 * N;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int site = (x0 + r) * kRcLY + j;
      const int up = item_occ(sb, hop, group_map(site, grp, N, LX, kRcLY), N) > 0;
      const int dn = item_occ(sb, hop, group_map(N + site, grp, N, LX, kRcLY), N) > 0;
      const float* e = emb + (up + 2 * dn) * kRcC + 4 * g16;
#pragma unroll
      for (int qq = 0; qq < QN; ++qq) {
        const float4 v = *reinterpret_cast<const float4*>(e + 16 * (q0 + qq));
        h[r][qq] = rc_f32x4{v.x, v.y, v.z, v.w};
      }
    }
  }
  const float inv_count = 1.0f / (float)(kRcC * N);
  for (int blk = 0; blk < n_blocks; ++blk) {
    float ss = 0.f;
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int qq = 0; qq < QN; ++qq)
#pragma unroll
        for (int e = 0; e < 4; ++e) ss = fmaf(h[r][qq][e], h[r][qq][e], ss);
    ss = rc_wave_sum(ss);
    if (lane == 0) part[wave] = ss;
    __syncthreads();  // partial sums (and, first time, slab 0) visible
    float tot = 0.f;
#pragma unroll
    for (int w = 0; w < nw; ++w) tot += part[w];
    const float rinv = 1.0f / sqrtf(fmaf(tot, inv_count, kRmsEps));
    rb_store<QN>(act, h, rinv, x0, q0, lane, N);
    __syncthreads();
    rc_f32x4 acc[4][QN];
    rb_conv<NT, QN>(act, wbuf, Wsl, k, n_slabs, Bc + (2 * blk) * kRcC, LX, N, x0, q0, t, acc);
    rb_store<QN>(act, acc, 1.0f, x0, q0, lane, N);  // the last slab barrier ended every read of act
    __syncthreads();
    rb_conv<NT, QN>(act, wbuf, Wsl, k, n_slabs, Bc + (2 * blk + 1) * kRcC, LX, N, x0, q0, t, acc);
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int qq = 0; qq < QN; ++qq) h[r][qq] += acc[r][qq];
  }
  if (head) {  // BCS: the last slab barrier ended every read of act; stage FP32 [site][68] there
    float* xs = lds;
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int qq = 0; qq < QN; ++qq) {
        const rc_f32x4 v = out_scale * h[r][qq];
        *reinterpret_cast<float4*>(xs + ((x0 + r) * kRcLY + j) * kRcStride + 16 * (q0 + qq) + 4 * g16) =
            make_float4(v[0], v[1], v[2], v[3]);
      }
    __syncthreads();
    rc_head(xs, head, N, Y + (size_t)(img - img0) * (2 * N * kRcC));
    return;
  }
  float* yo = Y + (size_t)(img - img0) * (2 * N * (kRcC / 2));
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int site = (x0 + r) * kRcLY + j;
#pragma unroll
    for (int qq = 0; qq < QN; ++qq) {
      const int qg = q0 + qq;
      const rc_f32x4 v = out_scale * h[r][qq];
      *reinterpret_cast<float4*>(yo + ((size_t)(qg >> 1) * N + site) * (kRcC / 2) + 16 * (qg & 1) + 4 * g16) =
          make_float4(v[0], v[1], v[2], v[3]);
    }
  }
}

// ---------------------------------------------------------------- FP32 with weight slabs
//
// The FP32 kernel with the weight stream of the split-BF16 one: slabs of (conv, kh, dyi) in
// B-fragment order, [dxi][qq][lane][8 floats] with W[tap = 3 dxi + dyi][32 kh + 4 kk + l / 16]
// [16 qq + l % 16] at element kk (neuralhubbard_amd/mixed.py, pack_resconv_f32; 24 KB, the same
// size as a split-BF16 slab), double-buffered in LDS, the next slab in flight during the MFMAs.

template <int DYI, int QN>
__device__ __forceinline__ void rs_taps(const float* act, const float4* wb, int LX, int x0, int q0, int lane, int kh,
                                        rc_f32x4 (&acc)[4][QN]) {
  const int j = lane & 15, g16 = lane >> 4;
  float fs[6][8];
#pragma unroll
  for (int u = 0; u < 6; ++u) {
    int xr = x0 - 1 + u;
    xr += xr < 0 ? LX : 0;
    xr -= xr >= LX ? LX : 0;
    const float4* p = reinterpret_cast<const float4*>(act + (xr * kRcLY + j) * kRcStride + 16 * g16 + 8 * kh);
    const float4 p0 = p[0], p1 = p[1];
    const float v[8] = {p0.x, p0.y, p0.z, p0.w, p1.x, p1.y, p1.z, p1.w};
#pragma unroll
    for (int kk = 0; kk < 8; ++kk) {
      if (DYI == 1) fs[u][kk] = v[kk];
      else if (DYI == 0) fs[u][kk] = rc_rot<0x121>(v[kk]);
      else fs[u][kk] = rc_rot<0x12F>(v[kk]);
    }
  }
#pragma unroll
  for (int dxi = 0; dxi < 3; ++dxi) {
#pragma unroll
    for (int qq = 0; qq < QN; ++qq) {
      const float4* w = wb + ((dxi * 4 + q0 + qq) * kWave + lane) * 2;
      const float4 w0 = w[0], w1 = w[1];
      const float wv[8] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w};
#pragma unroll
      for (int kk = 0; kk < 8; ++kk)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          acc[r][qq] = __builtin_amdgcn_mfma_f32_16x16x4f32(fs[r + dxi][kk], wv[kk], acc[r][qq], 0, 0, 0);
    }
  }
}

template <int NT, int DYI, int QN>
__device__ __forceinline__ void rs_step(const float* act, rb_i32x4* wbuf, const rb_i32x4* __restrict__ Wsl, int& k,
                                        int n_slabs, int LX, int x0, int q0, int t, int kh,
                                        rc_f32x4 (&acc)[4][QN]) {
  constexpr int kPer = rb_per_thread<NT>();
  const bool more = k + 1 < n_slabs;
  rb_i32x4 v[kPer];
  if (more) {
    const rb_i32x4* src = Wsl + (size_t)(k + 1) * 2 * kSlabFrag;
#pragma unroll
    for (int i = 0; i < kPer; ++i) v[i] = src[t + i * NT];
  }
  rs_taps<DYI, QN>(act, reinterpret_cast<const float4*>(wbuf + (k & 1) * 2 * kSlabFrag), LX, x0, q0,
                   t & (kWave - 1), kh, acc);
  if (more) {
    rb_i32x4* dst = wbuf + ((k + 1) & 1) * 2 * kSlabFrag;
#pragma unroll
    for (int i = 0; i < kPer; ++i) dst[t + i * NT] = v[i];
  }
  __syncthreads();
  ++k;
}

template <int NT, int QN>
__device__ __forceinline__ void rs_conv(const float* act, rb_i32x4* wbuf, const rb_i32x4* __restrict__ Wsl, int& k,
                                        int n_slabs, const float* __restrict__ bl, int LX, int x0, int q0, int t,
                                        rc_f32x4 (&acc)[4][QN]) {
  const int j = t & 15;
#pragma unroll
  for (int qq = 0; qq < QN; ++qq) {
    const float bv = bl[16 * (q0 + qq) + j];
#pragma unroll
    for (int r = 0; r < 4; ++r) acc[r][qq] = rc_f32x4{bv, bv, bv, bv};
  }
#pragma unroll 1
  for (int kh = 0; kh < 2; ++kh) {
    rs_step<NT, 0, QN>(act, wbuf, Wsl, k, n_slabs, LX, x0, q0, t, kh, acc);
    rs_step<NT, 1, QN>(act, wbuf, Wsl, k, n_slabs, LX, x0, q0, t, kh, acc);
    rs_step<NT, 2, QN>(act, wbuf, Wsl, k, n_slabs, LX, x0, q0, t, kh, acc);
  }
}

__host__ __device__ constexpr size_t rs_lds_floats(int N) {
  return (size_t)N * kRcStride + 16 * kSlabFrag + 16;  // activations, two slabs, partials
}

// QN channel quarters per wave, as in k_resconv_bf16 (QN = 2: two waves per four lattice rows).
template <int NT, int QN>
__global__ void __launch_bounds__(NT) k_resconv_fs(int LX, int n_blocks, const float* __restrict__ emb,
                                                   const rb_i32x4* __restrict__ Wsl, const float* __restrict__ Bc,
                                                   const float* __restrict__ head, float out_scale,
                                                   const int8_t* __restrict__ s,
                                                   const int32_t* __restrict__ item_b,
                                                   const int32_t* __restrict__ item_hop, int img0, int img_end,
                                                   int ipi, float* __restrict__ Y) {
  static_assert(QN == 4 || QN == 2, "four or two channel quarters per wave");
  extern __shared__ float lds[];
  const int N = LX * kRcLY;
  float* act = lds;
  rb_i32x4* wbuf = reinterpret_cast<rb_i32x4*>(lds + (size_t)N * kRcStride);
  float* part = lds + (size_t)N * kRcStride + 16 * kSlabFrag;
  const int img = img0 + blockIdx.x;
  if (img >= img_end) return;
  const int item = img / ipi;
  const int grp = (ipi == kGroup) ? (img & 7) : 0;
  const int32_t hop = item_hop ? item_hop[item] : kHopNone;
  if (hop == kHopInvalid) return;
  const int b = item_b ? item_b[item] : item;
  const int t = threadIdx.x, lane = t & (kWave - 1), wave = t / kWave;
  constexpr int nw = NT / kWave;
  constexpr int nrg = nw * QN / 4;  // row groups of 4 lattice rows (= LX / 4)
  const int x0 = 4 * (wave % nrg), q0 = QN * (wave / nrg), j = lane & 15, g16 = lane >> 4;
  const int n_slabs = 2 * n_blocks * kSlabsPerConv;
  int k = 0;
#pragma unroll
  for (int i = 0; i < rb_per_thread<NT>(); ++i) wbuf[t + i * NT] = Wsl[t + i * NT];  // slab 0

  rc_f32x4 h[4][QN];
  {
    const int8_t* sb = s + (size_t)b * 2 * N;
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int site = (x0 + r) * kRcLY + 4 * g16 + q;
        const int up = item_occ(sb, hop, group_map(site, grp, N, LX, kRcLY), N) > 0;
        const int dn = item_occ(sb, hop, group_map(N + site, grp, N, LX, kRcLY), N) > 0;
        const float* e = emb + (up + 2 * dn) * kRcC + j;
#pragma unroll
        for (int qq = 0; qq < QN; ++qq) h[r][qq][q] = e[16 * (q0 + qq)];
      }
  }
  const float inv_count = 1.0f / (float)(kRcC * N);
  for (int blk = 0; blk < n_blocks; ++blk) {
    float ss = 0.f;
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int qq = 0; qq < QN; ++qq)
#pragma unroll
        for (int q = 0; q < 4; ++q) ss = fmaf(h[r][qq][q], h[r][qq][q], ss);
    ss = rc_wave_sum(ss);
    if (lane == 0) part[wave] = ss;
    __syncthreads();  // partial sums (and, first time, slab 0) visible
    float tot = 0.f;
#pragma unroll
    for (int w = 0; w < nw; ++w) tot += part[w];
    const float rinv = 1.0f / sqrtf(fmaf(tot, inv_count, kRmsEps));
    rc_store_q<QN>(act, h, rinv, x0, q0, lane);
    __syncthreads();
    rc_f32x4 acc[4][QN];
    rs_conv<NT, QN>(act, wbuf, Wsl, k, n_slabs, Bc + (2 * blk) * kRcC, LX, x0, q0, t, acc);
    rc_store_q<QN>(act, acc, 1.0f, x0, q0, lane);  // the last slab barrier ended every read of act
    __syncthreads();
    rs_conv<NT, QN>(act, wbuf, Wsl, k, n_slabs, Bc + (2 * blk + 1) * kRcC, LX, x0, q0, t, acc);
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int qq = 0; qq < QN; ++qq) h[r][qq] += acc[r][qq];
  }
  if (head) {  // BCS: the last slab barrier ended every read of act
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int qq = 0; qq < QN; ++qq)
#pragma unroll
        for (int q = 0; q < 4; ++q)
          act[((x0 + r) * kRcLY + 4 * g16 + q) * kRcStride + 16 * (q0 + qq) + j] = out_scale * h[r][qq][q];
    __syncthreads();
    rc_head(act, head, N, Y + (size_t)(img - img0) * (2 * N * kRcC));
    return;
  }
  float* yo = Y + (size_t)(img - img0) * (2 * N * (kRcC / 2));
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int qq = 0; qq < QN; ++qq)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int site = (x0 + r) * kRcLY + 4 * g16 + q;
        const int qg = q0 + qq;
        yo[((size_t)(qg >> 1) * N + site) * (kRcC / 2) + 16 * (qg & 1) + j] = out_scale * h[r][qq][q];
      }
}

// ---------------------------------------------------------------- other lattices
//
// Lattices outside the tile kernels (LY != 16, or more than 16 rows) run the network layer by
// layer over a chunk of images, FP32 (MFMA convolutions), with the activations in the workspace.  The
// residual stream H of image i is kept as [half][site][32] (channel c at half c / 32), the layout
// of the features y[spin][site][a], so that the scaled output lands in Y in place; A and B are
// two [site][channel] activation buffers of the same size and rinv one float per image.

__device__ inline int rg_at(int N, int site, int c) { return ((c >> 5) * N + site) * 32 + (c & 31); }

// H[i] = emb[occupation class] per site (zero for unused item slots)
__global__ void __launch_bounds__(256) k_rg_embed(int LX, int LY, const int8_t* __restrict__ s,
                                                  const int32_t* __restrict__ item_b,
                                                  const int32_t* __restrict__ item_hop, int img0, int ipi,
                                                  const float* __restrict__ emb, float* __restrict__ H) {
  const int N = LX * LY;
  const int i = blockIdx.y, img = img0 + i;
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= N * kRcC) return;
  const int site = e / kRcC, c = e % kRcC;
  const int item = img / ipi;
  const int grp = (ipi == kGroup) ? (img & 7) : 0;
  const int32_t hop = item_hop ? item_hop[item] : kHopNone;
  float v = 0.f;
  if (hop != kHopInvalid) {
    const int b = item_b ? item_b[item] : item;
    const int8_t* sb = s + (size_t)b * 2 * N;
    const int up = item_occ(sb, hop, group_map(site, grp, N, LX, LY), N) > 0;
    const int dn = item_occ(sb, hop, group_map(N + site, grp, N, LX, LY), N) > 0;
    v = emb[(up + 2 * dn) * kRcC + c];
  }
  H[(size_t)i * N * kRcC + rg_at(N, site, c)] = v;
}

// rinv[i] = 1 / sqrt(mean of H[i]^2 over channels and sites + eps), one block per image
__global__ void __launch_bounds__(256) k_rg_rinv(int N, const float* __restrict__ H, float* __restrict__ rinv) {
  __shared__ float part[4];
  const float* h = H + (size_t)blockIdx.x * N * kRcC;
  float ss = 0.f;
  for (int e = threadIdx.x; e < N * kRcC; e += 256) ss = fmaf(h[e], h[e], ss);
  ss = rc_wave_sum(ss);
  if ((threadIdx.x & (kWave - 1)) == 0) part[threadIdx.x / kWave] = ss;
  __syncthreads();
  if (threadIdx.x == 0) {
    const float tot = part[0] + part[1] + part[2] + part[3];
    rinv[blockIdx.x] = 1.0f / sqrtf(fmaf(tot, 1.0f / (float)(kRcC * N), kRmsEps));
  }
}

// A[i][site][c] = gelu(scale_i X[i]); X in the H layout (HL) or [site][c]
template <bool HL>
__global__ void __launch_bounds__(256) k_rg_act(int N, int nimg, const float* __restrict__ X,
                                                const float* __restrict__ rinv, float* __restrict__ A) {
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= (size_t)nimg * N * kRcC) return;
  const int i = (int)(e / ((size_t)N * kRcC));
  const int r = (int)(e - (size_t)i * N * kRcC);
  const int site = r / kRcC, c = r % kRcC;
  const float x = X[(size_t)i * N * kRcC + (HL ? rg_at(N, site, c) : r)];
  A[e] = gelu_tanh(rinv ? rinv[i] * x : x);
}

// Out = bias + periodic conv3x3 of A (W [tap][c_in][c_out], tap = (dx + 1) * 3 + dy + 1, input at
// (x + dx, y + dy)); RES: Out (H layout) += the result in place, otherwise Out is [site][c].
// FP32 MFMA (v_mfma_f32_16x16x4_f32): a block owns 32 sites (two M tiles of 16) x 64 output
// channels (four N tiles) of image blockIdx.y, one tap's weights and neighbour activations staged
// in LDS; wave w computes tiles (M = w / 2, N = 2 (w % 2) + {0,1})
// over K = 9 taps x 64 input channels.  A fragment: lane l holds site 16 M + l % 16, channel
// 4 kk + l / 16 of the tap's neighbour rows; B fragment: lane l holds W[tap][4 kk + l / 16]
// [16 N + l % 16]; result lane l, register r -> site 16 M + 4 (l / 16) + r, channel 16 N + l % 16.
// LDS rows are padded by 4 floats so that both fragment reads are conflict-free.
constexpr int kRgPad = kRcC + 4;

template <bool RES>
__global__ void __launch_bounds__(256) k_rg_conv_mfma(int LX, int LY, const float* __restrict__ A,
                                                      const float* __restrict__ W, const float* __restrict__ bias,
                                                      float* __restrict__ Out) {
  __shared__ float wl[kRcC * kRgPad];
  __shared__ float al[32 * kRgPad];
  const int N = LX * LY;
  const int i = blockIdx.y;
  const int s0 = blockIdx.x * 32;
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const int mt = wave >> 1, nt0 = 2 * (wave & 1);
  const int m = lane & 15, k = lane >> 4;
  const float* Ai = A + (size_t)i * N * kRcC;
  rc_f32x4 acc[2];
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const float bv = bias[16 * (nt0 + q) + m];
    acc[q] = rc_f32x4{bv, bv, bv, bv};
  }
  for (int tap = 0; tap < 9; ++tap) {
    const int dx = tap / 3 - 1, dy = tap % 3 - 1;
    __syncthreads();  // the previous tap's reads of wl / al are done
    for (int e = threadIdx.x; e < kRcC * kRcC; e += 256) wl[(e / kRcC) * kRgPad + e % kRcC] = W[tap * kRcC * kRcC + e];
    for (int e = threadIdx.x; e < 32 * kRcC; e += 256) {
      const int ls = e / kRcC, c = e % kRcC, site = s0 + ls;
      float v = 0.f;
      if (site < N) {
        const int x = site / LY, y = site % LY;
        const int xn = (x + dx + LX) % LX, yn = (y + dy + LY) % LY;
        v = Ai[(xn * LY + yn) * kRcC + c];
      }
      al[ls * kRgPad + c] = v;
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < kRcC / 4; ++kk) {
      const float a = al[(16 * mt + m) * kRgPad + 4 * kk + k];
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const float b = wl[(4 * kk + k) * kRgPad + 16 * (nt0 + q) + m];
        acc[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[q], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int co = 16 * (nt0 + q) + m;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int site = s0 + 16 * mt + 4 * k + r;
      if (site >= N) continue;
      if (RES) {
        float* o = Out + (size_t)i * N * kRcC + rg_at(N, site, co);
        *o += acc[q][r];
      } else {
        Out[(size_t)i * N * kRcC + site * kRcC + co] = acc[q][r];
      }
    }
  }
}

// Split-BF16 form of the same conv (the mixed setting): activations as BF16 hi and lo planes
// [image][plane][site][64], weights the split-BF16 A fragments of the tile kernel's slabs
// (conv_w_lp, [conv][kh][dyi][plane][dxi][qq][lane][8]).  A block owns 32 sites (two N tiles of
// 16) x 64 output channels; wave w computes quarter w (M tile) for both site tiles with
// v_mfma_f32_16x16x32_bf16, lo*hi + hi*lo + hi*hi per product.  B fragment: lane l holds the 8
// channels 32 kh + 8 (l / 16) .. + 7 of site 16 nt + l % 16 (one 16-byte LDS read per plane);
// result lane l, register r -> channel 16 w + 4 (l / 16) + r, site 16 nt + l % 16.

// A16[i] = split BF16 of gelu(scale_i X[i]); X in the H layout (HL) or [site][c]
template <bool HL>
__global__ void __launch_bounds__(256) k_rg_act_bf16(int N, int nimg, const float* __restrict__ X,
                                                     const float* __restrict__ rinv, uint16_t* __restrict__ A16) {
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= (size_t)nimg * N * kRcC) return;
  const int i = (int)(e / ((size_t)N * kRcC));
  const int r = (int)(e - (size_t)i * N * kRcC);
  const int site = r / kRcC, c = r % kRcC;
  const float x = X[(size_t)i * N * kRcC + (HL ? rg_at(N, site, c) : r)];
  const float g = rb_gelu(rinv ? rinv[i] * x : x);
  const uint32_t hi = rb_bf16(g);
  const uint32_t lo = rb_bf16(g - __uint_as_float(hi << 16));
  uint16_t* a = A16 + (size_t)i * 2 * N * kRcC;
  a[r] = (uint16_t)hi;
  a[(size_t)N * kRcC + r] = (uint16_t)lo;
}

template <bool RES>
__global__ void __launch_bounds__(256) k_rg_conv_bf16(int LX, int LY, const uint16_t* __restrict__ A16,
                                                      const rb_i32x4* __restrict__ Wl, const float* __restrict__ bias,
                                                      float* __restrict__ Out) {
  __shared__ uint16_t ah[32 * kRbStride];
  __shared__ uint16_t alo[32 * kRbStride];
  const int N = LX * LY;
  const int i = blockIdx.y;
  const int s0 = blockIdx.x * 32;
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  const int j16 = lane & 15, g16 = lane >> 4;
  const uint16_t* Ai = A16 + (size_t)i * 2 * N * kRcC;
  rc_f32x4 acc[2];
  {
    const float4 b4 = *reinterpret_cast<const float4*>(bias + 16 * w + 4 * g16);
    acc[0] = acc[1] = rc_f32x4{b4.x, b4.y, b4.z, b4.w};
  }
  for (int tap = 0; tap < 9; ++tap) {
    const int dxi = tap / 3, dyi = tap % 3;
    const int dx = dxi - 1, dy = dyi - 1;
    __syncthreads();  // the previous tap's reads of ah / alo are done
    // 32 sites x 8 chunks of 8 channels x 2 planes, 16 bytes each: two per thread
    for (int e = threadIdx.x; e < 32 * 8 * 2; e += 256) {
      const int pl = e >> 8, ls = (e >> 3) & 31, ch = e & 7, site = s0 + ls;
      int4 v = make_int4(0, 0, 0, 0);
      if (site < N) {
        const int x = site / LY, y = site % LY;
        const int xn = (x + dx + LX) % LX, yn = (y + dy + LY) % LY;
        v = *reinterpret_cast<const int4*>(Ai + (size_t)pl * N * kRcC + (xn * LY + yn) * kRcC + 8 * ch);
      }
      *reinterpret_cast<int4*>((pl ? alo : ah) + ls * kRbStride + 8 * ch) = v;
    }
    __syncthreads();
#pragma unroll
    for (int kh = 0; kh < 2; ++kh) {
      // A fragments of (kh, dyi, dxi, quarter w), hi and lo planes
      const rb_i32x4* base = Wl + ((size_t)((kh * 3 + dyi) * 2) * 3 * 4) * kWave;
      const rb_bf16x8 wh = __builtin_bit_cast(rb_bf16x8, base[((0 * 3 + dxi) * 4 + w) * kWave + lane]);
      const rb_bf16x8 wlo = __builtin_bit_cast(rb_bf16x8, base[((1 * 3 + dxi) * 4 + w) * kWave + lane]);
#pragma unroll
      for (int nt = 0; nt < 2; ++nt) {
        const int off = (16 * nt + j16) * kRbStride + 32 * kh + 8 * g16;
        const rb_bf16x8 bh = __builtin_bit_cast(rb_bf16x8, *reinterpret_cast<const int4*>(ah + off));
        const rb_bf16x8 bl = __builtin_bit_cast(rb_bf16x8, *reinterpret_cast<const int4*>(alo + off));
        acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wlo, bh, acc[nt], 0, 0, 0);
        acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wh, bl, acc[nt], 0, 0, 0);
        acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wh, bh, acc[nt], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int nt = 0; nt < 2; ++nt) {
    const int site = s0 + 16 * nt + j16;
    if (site >= N) continue;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int co = 16 * w + 4 * g16 + r;
      if (RES) {
        float* o = Out + (size_t)i * N * kRcC + rg_at(N, site, co);
        *o += acc[nt][r];
      } else {
        Out[(size_t)i * N * kRcC + site * kRcC + co] = acc[nt][r];
      }
    }
  }
}

// BCS linear head: Y[i][k][site][dd] = sum_c head[k][c][dd] scale H[i][c at site] (head [k][c][dd])
__global__ void __launch_bounds__(256) k_rg_head(int N, float scale, const float* __restrict__ H,
                                                 const float* __restrict__ head, float* __restrict__ Y) {
  __shared__ float xs[4][kRcC];
  const int i = blockIdx.y;
  const int ls = threadIdx.x >> 6, dd = threadIdx.x & (kRcC - 1);
  const int site = blockIdx.x * 4 + ls;
  xs[ls][dd] = site < N ? scale * H[(size_t)i * N * kRcC + rg_at(N, site, dd)] : 0.f;
  __syncthreads();
  if (site >= N) return;
  for (int k = 0; k < 2; ++k) {
    const float* hk = head + (size_t)k * kRcC * kRcC + dd;
    float acc = 0.f;
#pragma unroll 8
    for (int c = 0; c < kRcC; ++c) acc = fmaf(hk[c * kRcC], xs[ls][c], acc);
    Y[(size_t)i * 2 * N * kRcC + ((size_t)k * N + site) * kRcC + dd] = acc;
  }
}

__global__ void __launch_bounds__(256) k_rg_scale(size_t n, float scale, float* __restrict__ H) {
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (e < n) H[e] *= scale;
}

bool resconv_tiled(const nh_model* m) {
  return m->LY == kRcLY && m->LX >= 4 && m->LX <= 16 && m->LX % 4 == 0;
}

}  // namespace

bool resconv_supported(const nh_model* m) {
  return m->channels == kRcC && m->n_layers >= 1;
}

size_t resconv_scratch_floats(const nh_model* m, int64_t images) {
  if ((m->network != NH_NET_RESCONV && m->network != NH_NET_RESCONV_HEAD) || resconv_tiled(m)) return 0;
  // two activation buffers (and, with the BCS head, the residual stream: Y then holds 2 x 64
  // features per site) and one float per image
  const size_t bufs = m->network == NH_NET_RESCONV_HEAD ? 3 : 2;
  return (size_t)images * (bufs * (size_t)m->LX * m->LY * kRcC + 1);
}

// the layer-by-layer network of lattices outside the tile kernels
static int launch_resconv_layers(hipStream_t st, const nh_model* m, const int8_t* s, const int32_t* item_b,
                                 const int32_t* item_hop, int img0, int img_end, int ipi, float* Y,
                                 float* scratch) {
  if (!scratch) return NH_ERR_UNSUPPORTED;  // nh_features carries no workspace
  const int nimg = img_end - img0;
  const int N = m->LX * m->LY;
  const size_t per = (size_t)N * kRcC;
  const bool head = m->network == NH_NET_RESCONV_HEAD;
  float* A = scratch;
  float* B = A + (size_t)nimg * per;
  float* H = head ? B + (size_t)nimg * per : Y;  // without the head: [image][half][site][32] = Y's layout
  float* rinv = (head ? H : B) + (size_t)nimg * per;
  const float* emb = m->conv_w;
  const float* Wc = m->conv_w + kRcClasses * kRcC;
  const unsigned ge = (unsigned)((per + 255) / 256);
  const unsigned ga = (unsigned)(((size_t)nimg * per + 255) / 256);
  // 32 sites per block on FP32 MFMA: 37.9 us per 24x24 image (VALU FMA: 48 us with 32 sites per
  // block, 52.8 with 16, 62.2 with 64); split BF16 12.5 us
  const dim3 gc((unsigned)((N + 31) / 32), (unsigned)nimg);
  hipLaunchKernelGGL(k_rg_embed, dim3(ge, nimg), dim3(256), 0, st, m->LX, m->LY, s, item_b, item_hop, img0, ipi,
                     emb, H);
  // the mixed setting: split-BF16 convolutions on the BF16 slabs of the tile kernel
  const bool bf16 = m->precision == NH_PREC_MIXED && m->conv_w_lp;
  const rb_i32x4* Wlp = reinterpret_cast<const rb_i32x4*>(m->conv_w_lp);
  uint16_t* A16 = reinterpret_cast<uint16_t*>(A);  // hi and lo planes fill the FP32 buffer's bytes
  for (int blk = 0; blk < m->n_layers; ++blk) {
    const float* W1 = Wc + (size_t)(2 * blk) * 9 * kRcC * kRcC;
    hipLaunchKernelGGL(k_rg_rinv, dim3(nimg), dim3(256), 0, st, N, H, rinv);
    if (bf16) {
      const rb_i32x4* L1 = Wlp + (size_t)(2 * blk) * kSlabsPerConv * 2 * kSlabFrag;
      hipLaunchKernelGGL(k_rg_act_bf16<true>, dim3(ga), dim3(256), 0, st, N, nimg, H, rinv, A16);
      hipLaunchKernelGGL((k_rg_conv_bf16<false>), gc, dim3(256), 0, st, m->LX, m->LY, A16, L1,
                         m->conv_b + (2 * blk) * kRcC, B);
      hipLaunchKernelGGL(k_rg_act_bf16<false>, dim3(ga), dim3(256), 0, st, N, nimg, B, nullptr, A16);
      hipLaunchKernelGGL((k_rg_conv_bf16<true>), gc, dim3(256), 0, st, m->LX, m->LY, A16,
                         L1 + kSlabsPerConv * 2 * kSlabFrag, m->conv_b + (2 * blk + 1) * kRcC, H);
      continue;
    }
    hipLaunchKernelGGL(k_rg_act<true>, dim3(ga), dim3(256), 0, st, N, nimg, H, rinv, A);
    hipLaunchKernelGGL((k_rg_conv_mfma<false>), gc, dim3(256), 0, st, m->LX, m->LY, A, W1,
                       m->conv_b + (2 * blk) * kRcC, B);
    hipLaunchKernelGGL(k_rg_act<false>, dim3(ga), dim3(256), 0, st, N, nimg, B, nullptr, A);
    hipLaunchKernelGGL((k_rg_conv_mfma<true>), gc, dim3(256), 0, st, m->LX, m->LY, A, W1 + 9 * kRcC * kRcC,
                       m->conv_b + (2 * blk + 1) * kRcC, H);
  }
  const float scale = (float)(1.0 / sqrt((double)m->n_layers + 1.0));
  if (head) {
    const float* hw = m->conv_w + kRcClasses * kRcC + (size_t)2 * m->n_layers * 9 * kRcC * kRcC;
    hipLaunchKernelGGL(k_rg_head, dim3((unsigned)((N + 3) / 4), (unsigned)nimg), dim3(256), 0, st, N, scale, H, hw, Y);
  } else {
    hipLaunchKernelGGL(k_rg_scale, dim3(ga), dim3(256), 0, st, (size_t)nimg * per, scale, H);
  }
  NH_LAUNCH_CHECK();
  return NH_OK;
}

int launch_resconv(hipStream_t st, const nh_model* m, const int8_t* s, const int32_t* item_b,
                   const int32_t* item_hop, int img0, int img_end, int ipi, float* Y, float* scratch) {
  if (!resconv_supported(m)) return NH_ERR_UNSUPPORTED;
  const int nimg = img_end - img0;
  if (nimg <= 0) return NH_OK;
  if (!resconv_tiled(m)) return launch_resconv_layers(st, m, s, item_b, item_hop, img0, img_end, ipi, Y, scratch);
  const int N = m->LX * kRcLY;
  const size_t lds = sizeof(float) * ((size_t)N * kRcStride + 16);
  const float scale = (float)(1.0 / sqrt((double)m->n_layers + 1.0));
  const float* head = m->network == NH_NET_RESCONV_HEAD
                          ? m->conv_w + kRcClasses * kRcC + (size_t)2 * m->n_layers * 9 * kRcC * kRcC
                          : nullptr;
  if (m->precision == NH_PREC_MIXED) {
    if (!m->conv_w_lp) return NH_ERR_UNSUPPORTED;
    const size_t lds_bf16 = sizeof(float) * rb_lds_floats(N);
    const rb_i32x4* wsl = reinterpret_cast<const rb_i32x4*>(m->conv_w_lp);
    // two waves per four lattice rows, two channel quarters each: 1.73 us per 16x16 image against
    // 1.89 us with one wave per four rows (one wave per SIMD)
#define NH_RB_LAUNCH(NT)                                                                                      \
  hipLaunchKernelGGL((k_resconv_bf16<NT, 2>), dim3(nimg), dim3(NT), lds_bf16, st, m->LX, m->n_layers, m->conv_w, \
                     wsl, m->conv_b, head, scale, s, item_b, item_hop, img0, img_end, ipi, Y)
    switch (m->LX) {
      case 4: NH_RB_LAUNCH(128); break;
      case 8: NH_RB_LAUNCH(256); break;
      case 12: NH_RB_LAUNCH(384); break;
      default: NH_RB_LAUNCH(512);
    }
#undef NH_RB_LAUNCH
    NH_LAUNCH_CHECK();
    return NH_OK;
  }
  if (m->conv_w_mfma) {  // FP32 weight slabs
    const size_t lds_fs = sizeof(float) * rs_lds_floats(N);
    const rb_i32x4* wsl = reinterpret_cast<const rb_i32x4*>(m->conv_w_mfma);
    // two waves per four lattice rows, as in the split-BF16 kernel: 5.05 us per 16x16 image
    // against 5.43 us with one wave per four rows
#define NH_RS_LAUNCH(NT)                                                                                     \
  hipLaunchKernelGGL((k_resconv_fs<NT, 2>), dim3(nimg), dim3(NT), lds_fs, st, m->LX, m->n_layers, m->conv_w, \
                     wsl, m->conv_b, head, scale, s, item_b, item_hop, img0, img_end, ipi, Y)
    switch (m->LX) {
      case 4: NH_RS_LAUNCH(128); break;
      case 8: NH_RS_LAUNCH(256); break;
      case 12: NH_RS_LAUNCH(384); break;
      default: NH_RS_LAUNCH(512);
    }
#undef NH_RS_LAUNCH
    NH_LAUNCH_CHECK();
    return NH_OK;
  }
  hipLaunchKernelGGL(k_resconv, dim3(nimg), dim3(16 * m->LX), lds, st, m->LX, m->n_layers, m->conv_w,
                     m->conv_w + kRcClasses * kRcC, m->conv_b, head, scale, s, item_b, item_hop, img0, img_end, ipi,
                     Y);
  NH_LAUNCH_CHECK();
  return NH_OK;
}

}  // namespace nh
