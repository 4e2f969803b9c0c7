// Metropolis sweep with optional delayed acceptance (gfx950).
//
// Step t of chain c draws Philox4x32-10 (Salmon et al., SC'11) at counter (c, t_lo, t_hi, 0),
// key (seed_lo, seed_hi); r2 / r3 -> uniforms u1 / u2 in [0, 1) (24 bits).  Proposals:
//   exchange:      r0 -> site i = r0 mod N, r1 & 3 -> nearest-neighbour direction, r1 >> 2 & 1
//                  -> spin; the occupations of modes (spin, i) and (spin, j) trade places, j the
//                  neighbour (a spin-preserving hop, or nothing when both modes agree);
//   particle hop:  r0 -> particle k = r0 mod (n_up + n_dn), the k-th occupied mode in ascending
//                  order, r1 & (z - 1) -> one of its z = 4 nearest (8: and next-nearest)
//                  neighbours; the particle moves there unless that mode is occupied;
//   site exchange: r0 -> site i = r0 mod N, r1 & (z - 1) -> neighbour j; the occupations of
//                  sites i and j trade places in both spin blocks (an electron hops, or a doublon
//                  and a hole, or an up and a down electron swap), nothing when they agree;
//   mix:           r1 >> 4 & 1 picks particle hop (0) or site exchange (1) per chain and step.
// Each is symmetric (the pair (s, s') is drawn with the same probability from either end), so
//   plain:   accept with min(1, |psi(s')/psi(s)|^2);
//   delayed: stage 1 accepts with min(1, exp(2 (f(s') - f(s)))), f the surrogate log|psi|;
//            only stage-1 survivors get the exact psi(s'), and stage 2 accepts with
//            min(1, |psi(s')/psi(s)|^2 exp(-2 (f(s') - f(s)))) (Christen & Fox 2005), which
//            leaves |psi|^2 stationary exactly.

#include "nh_common.h"
#include "nh_internal.h"

namespace nh {

int launch_sur_forward(hipStream_t st, const nh_surrogate* q, const int8_t* s, const int32_t* item_b,
                       const int32_t* item_hop, int n_items, float* out);
int check_sur(const nh_surrogate* q);

__device__ inline void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
    const uint32_t n1 = (uint32_t)p1;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
    const uint32_t n3 = (uint32_t)p0;
    c[0] = n0;
    c[1] = n1;
    c[2] = n2;
    c[3] = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

__device__ inline float u24(uint32_t r) { return (float)(r >> 8) * (1.0f / 16777216.0f); }

// neighbour directions: the four nearest, then the four next-nearest
__constant__ int c_sdx[8] = {1, -1, 0, 0, 1, 1, -1, -1};
__constant__ int c_sdy[8] = {0, 0, 1, -1, 1, -1, 1, -1};

__device__ inline int neighbour_site(int i, int dir, int LX, int LY) {
  const int x = i / LY, y = i - (i / LY) * LY;
  int xx = x + c_sdx[dir], yy = y + c_sdy[dir];
  xx += (xx < 0) ? LX : 0;
  xx -= (xx >= LX) ? LX : 0;
  yy += (yy < 0) ? LY : 0;
  yy -= (yy >= LY) ? LY : 0;
  return xx * LY + yy;
}

constexpr int kProposeWaves = 4;  // chains per 256-thread block, one wave each

// Proposal of every chain: item_hop[c] (kHopInvalid for no change), uniforms u1/u2.  One wave
// per chain: every lane draws the same Philox block (the values are wave-uniform) and the
// particle hop finds its k-th occupied mode by ballots over the configuration.
__global__ void __launch_bounds__(kProposeWaves * kWave)
    k_propose(int LX, int LY, int n_chains, const int8_t* __restrict__ s, uint64_t seed, int64_t step, int kind,
              int n_occ, int ph_z, int se_z, int32_t* __restrict__ item_hop, float* __restrict__ u1,
              float* __restrict__ u2) {
  const int lane = threadIdx.x & (kWave - 1);
  const int c = blockIdx.x * kProposeWaves + (int)(threadIdx.x / kWave);
  if (c >= n_chains) return;  // whole waves leave together
  const int N = LX * LY, two_n = 2 * N;
  uint32_t r[4] = {(uint32_t)c, (uint32_t)step, (uint32_t)((uint64_t)step >> 32), 0u};
  philox4x32_10(r, (uint32_t)seed, (uint32_t)(seed >> 32));
  const int8_t* sc = s + (size_t)c * two_n;
  if (kind == NH_MOVE_MIX) kind = ((r[1] >> 4) & 1) ? NH_MOVE_SITE_EXCHANGE : NH_MOVE_PARTICLE_HOP;
  int32_t h = kHopInvalid;
  if (kind == NH_MOVE_EXCHANGE) {
    const int i = (int)(r[0] % (uint32_t)N);
    const int spin = (r[1] >> 2) & 1;
    const int mi = spin * N + i, mj = spin * N + neighbour_site(i, r[1] & 3, LX, LY);
    const int8_t oi = sc[mi], oj = sc[mj];
    h = (oi == oj) ? kHopInvalid : (oi > 0 ? pack_hop(mi, mj) : pack_hop(mj, mi));
  } else if (kind == NH_MOVE_SITE_EXCHANGE) {
    const int i = (int)(r[0] % (uint32_t)N);
    const int j = neighbour_site(i, r[1] & (se_z - 1), LX, LY);
    const bool same = sc[i] == sc[j] && sc[N + i] == sc[N + j];
    h = same ? kHopInvalid : pack_exchange(i, j);
  } else {  // particle hop
    const int k = (int)(r[0] % (uint32_t)n_occ);
    int from = -1, base = 0;
    for (int m0 = 0; m0 < two_n; m0 += kWave) {
      const int m = m0 + lane;
      const bool occ = m < two_n && sc[m] > 0;
      const uint64_t mask = __ballot(occ);
      const int cnt = __popcll(mask);
      if (k < base + cnt) {  // wave-uniform
        const uint64_t hit = __ballot(occ && base + lanes_below(mask) == k);
        from = m0 + __ffsll((long long)hit) - 1;
        break;
      }
      base += cnt;
    }
    if (from >= 0) {  // fewer than n_occ particles (a malformed configuration): no move
      const int off = from >= N ? N : 0;
      const int to = off + neighbour_site(from - off, r[1] & (ph_z - 1), LX, LY);
      h = sc[to] > 0 ? kHopInvalid : pack_hop(from, to);
    }
  }
  if (lane == 0) {
    item_hop[c] = h;
    u1[c] = u24(r[2]);
    u2[c] = u24(r[3]);
  }
}

// Stage 1 of delayed acceptance: survivors keep their hop, the rest become kHopInvalid.
// Per-chain counters cnt[k * n_chains + c]: k = 0 proposals that change the configuration,
// 1 stage-1 survivors, 2 accepted, 3 exact evaluations.
__global__ void k_stage1(int n_chains, int32_t* __restrict__ item_hop, const float* __restrict__ f_new,
                         const float* __restrict__ f_cur, const float* __restrict__ u1, int32_t* __restrict__ cnt) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_chains) return;
  const int32_t h = item_hop[c];
  if (h == kHopInvalid) return;
  const float d = 2.0f * (f_new[c] - f_cur[c]);
  const bool pass = d >= 0.f || u1[c] < expf(d);
  if (!pass) item_hop[c] = kHopInvalid;
  else cnt[1 * n_chains + c] += 1;
}

// Final accept: plain (ratio) or stage 2 (ratio corrected by the surrogate).
__global__ void k_accept(int N, int n_chains, int8_t* __restrict__ s, float* __restrict__ sign,
                         double* __restrict__ logabs, float* __restrict__ f_cur, const int32_t* __restrict__ item_hop,
                         const float* __restrict__ new_sign, const double* __restrict__ new_logabs,
                         const float* __restrict__ f_new, const float* __restrict__ u, int delayed,
                         int32_t* __restrict__ cnt) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_chains) return;
  const int32_t h = item_hop[c];
  if (h == kHopInvalid) return;
  cnt[3 * n_chains + c] += 1;  // exact evaluations
  double d = 2.0 * (new_logabs[c] - logabs[c]);
  if (delayed) d -= 2.0 * ((double)f_new[c] - (double)f_cur[c]);
  const bool ok = new_sign[c] != 0.f && !isnan(new_logabs[c]) && (d >= 0.0 || (double)u[c] < exp(d));
  if (!ok) return;
  int8_t* sc = s + (size_t)c * 2 * N;
  if (is_exchange(h)) {
    const int i = exchange_i(h), j = exchange_j(h);
    for (int off = 0; off <= N; off += N) {
      const int8_t t = sc[off + i];
      sc[off + i] = sc[off + j];
      sc[off + j] = t;
    }
  } else {
    sc[hop_from(h)] = -1;
    sc[hop_to(h)] = 1;
  }
  sign[c] = new_sign[c];
  logabs[c] = new_logabs[c];
  if (delayed) f_cur[c] = f_new[c];
  cnt[2 * n_chains + c] += 1;
}

__global__ void k_count_valid(int n_chains, const int32_t* __restrict__ item_hop, int32_t* __restrict__ cnt) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n_chains && item_hop[c] != kHopInvalid) cnt[c] += 1;
}

}  // namespace nh

using namespace nh;

extern "C" size_t nh_metropolis_workspace_size(const nh_model* model, int32_t n_chains) {
  if (check_model(model) != NH_OK || n_chains <= 0) return 0;
  const size_t per = 4 * sizeof(float) + sizeof(int32_t) + sizeof(float) + sizeof(double);
  return (size_t)n_chains * per + 8 * 256 + 32 + nh_log_psi_workspace_size(model, n_chains);
}

int nh_log_psi_items(hipStream_t st, const nh_model* m, const int8_t* s, const int32_t* item_b,
                     const int32_t* item_hop, int n_items, float* sign, double* logabs, void* ws, size_t wsb);

extern "C" int nh_metropolis_sweep(void* stream, const nh_model* model, const nh_surrogate* sur,
                                   const nh_proposal* proposal, int32_t n_chains, int8_t* s, float* sign,
                                   double* logabs, float* f_cur, uint64_t seed, int64_t step0, int32_t n_steps,
                                   int32_t* counters, void* workspace, size_t workspace_bytes) {
  int rc = check_model(model);
  if (rc != NH_OK) return rc;
  const nh_proposal prop = proposal ? *proposal : nh_proposal{NH_MOVE_EXCHANGE, 4, 4};
  if (prop.kind < NH_MOVE_EXCHANGE || prop.kind > NH_MOVE_MIX) return NH_ERR_ARG;
  const bool ph = prop.kind == NH_MOVE_PARTICLE_HOP || prop.kind == NH_MOVE_MIX;
  const bool se = prop.kind == NH_MOVE_SITE_EXCHANGE || prop.kind == NH_MOVE_MIX;
  if (ph && prop.ph_neighbors != 4 && prop.ph_neighbors != 8) return NH_ERR_ARG;
  if (se && prop.se_neighbors != 4 && prop.se_neighbors != 8) return NH_ERR_ARG;
  const int n_occ = model->n_up + model->n_dn;
  if (ph && n_occ <= 0) return NH_ERR_SHAPE;
  if (sur) {
    rc = check_sur(sur);
    if (rc != NH_OK) return rc;
    if (!f_cur || sur->LX != model->LX || sur->LY != model->LY) return NH_ERR_ARG;
  }
  if (n_chains <= 0 || n_steps < 0 || !s || !sign || !logabs || !counters || !workspace) return NH_ERR_ARG;
  if (workspace_bytes < nh_metropolis_workspace_size(model, n_chains)) return NH_ERR_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  char* p = (char*)(((size_t)workspace + 255) & ~(size_t)255);
  auto take = [&](size_t bytes) {
    char* q = p;
    p += (bytes + 255) & ~(size_t)255;
    return (void*)q;
  };
  int32_t* hop = (int32_t*)take(sizeof(int32_t) * n_chains);
  float* u1 = (float*)take(sizeof(float) * n_chains);
  float* u2 = (float*)take(sizeof(float) * n_chains);
  float* f_new = (float*)take(sizeof(float) * n_chains);
  float* nsign = (float*)take(sizeof(float) * n_chains);
  double* nlog = (double*)take(sizeof(double) * n_chains);
  void* eng = (void*)p;
  const size_t eng_bytes = workspace_bytes - (size_t)(p - (char*)workspace);
  const int N = model->LX * model->LY;
  const int tb = 256, nb = (n_chains + tb - 1) / tb;
  const int nbp = (n_chains + kProposeWaves - 1) / kProposeWaves;
  for (int t = 0; t < n_steps; ++t) {
    const int64_t step = step0 + t;
    hipLaunchKernelGGL(k_propose, dim3(nbp), dim3(kProposeWaves * kWave), 0, st, model->LX, model->LY, n_chains, s,
                       seed, step, prop.kind, n_occ, ph ? prop.ph_neighbors : 4, se ? prop.se_neighbors : 4, hop,
                       u1, u2);
    NH_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_count_valid, dim3(nb), dim3(tb), 0, st, n_chains, hop, counters);
    NH_LAUNCH_CHECK();
    if (sur) {
      rc = launch_sur_forward(st, sur, s, nullptr, hop, n_chains, f_new);
      if (rc != NH_OK) return rc;
      hipLaunchKernelGGL(k_stage1, dim3(nb), dim3(tb), 0, st, n_chains, hop, f_new, f_cur, u1, counters);
      NH_LAUNCH_CHECK();
    }
    rc = nh_log_psi_items(st, model, s, nullptr, hop, n_chains, nsign, nlog, eng, eng_bytes);
    if (rc != NH_OK) return rc;
    hipLaunchKernelGGL(k_accept, dim3(nb), dim3(tb), 0, st, N, n_chains, s, sign, logabs, f_cur, hop, nsign, nlog, f_new,
                       sur ? u2 : u1, sur ? 1 : 0, counters);
    NH_LAUNCH_CHECK();
  }
  return NH_OK;
}
