// Local energy of a Pfaffian state with every hop's amplitude ratio as a low-rank update of the
// sample's image matrices instead of a full recompute (the hop ratios of Hubbard.Oloc,
// code/quick_start.py:33, for the synthetic network whose receptive field is small against the
// lattice: the 32x32 workload).
//
// For a sample s, image g and hop a -> b, let F = F_g(s) (n x n, modes in the ascending order of
// s mapped by g) and F' the image matrix of s' with b in a's position.  F' differs from F in the
// rows and columns K of the positions whose features changed (the hop's cone: sites within the
// network's receptive radius of either end) and of a's position.  With D the K columns of
// dF = F' - F (rows in K halved), B = [D, P_K] and J = [[0, I], [-I, 0]] (2k x 2k),
// dF = B J B^T and, from the Pfaffian Schur complements of [[F, B], [-B^T, J^{-1}]] (its
// Pfaffian is Pf(J^{-1}) Pf(F + B J B^T) and Pf(F) Pf(J^{-1} + B^T F^{-1} B)) with J^{-1} = -J
// and Pf(-J) = (-1)^k Pf(J) = (-1)^(k(k+1)/2),
//   Pf(F') / Pf(F) = (-1)^(k(k+1)/2) Pf(-J + B^T F^{-1} B),
//   B^T G B = [[D^T X, -X_K^T], [X_K, G_KK]],  X = G D, G = F^{-1},
// of which the Pfaffian routine reads the strict lower triangle.  Per sample and image: F (the
// pairing block of the sample's features), G by the blocked Gauss-Jordan inverse refined by one
// Newton step in FP64 (R = I - F G, G <- G + G R, rocBLAS DGEMMs; the update's error is G's
// error amplified, so G is made accurate to FP32 rounding), psi_g(s) through the product path.  Per hop: the item's features, K by comparing them with the
// sample's, D, then X = G D for all hops of the sample at once (one GEMM: G is shared) and D^T X
// per hop (a strided batched GEMM) on rocBLAS, the 2k x 2k Pfaffians by nh_slogpf_ws, and
// psi_g(s') = psi_g(s) Pf(F') / Pf(F) times the sign of moving b to its sorted position (the
// product's order of s').  The image sum and the E_loc reduction are the full path's.

#include <rocblas/rocblas.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "nh_common.h"
#include "nh_internal.h"

namespace nh {
namespace {

constexpr int kLT = 256;
constexpr int kLrItems = 512;  // hop items per pass
constexpr int kLrMaxDev = 16;

// An upper bound of |K| for the synthetic network (n_layers + 1 convolutions of 3 x 3, so the
// features of a site depend on the occupations within n_layers + 1 sites in x and y): both spins
// of every site in the union of the two squares around a nearest- or next-nearest-neighbour
// hop's ends.
int lr_kbound(const nh_model* m) {
  const long R = m->n_layers + 1;
  const long side = 2 * R + 1;
  const long sites = std::min<long>((long)m->LX * m->LY, std::max(side * (side + 1), 2 * side * side - 4 * R * R));
  return (int)std::min<long>(m->n_up + m->n_dn, 2 * sites);
}

// Sample image modes (ascending occupied modes of s mapped by g) and, per sample, the number of
// occupied modes below each mode.  One block per image; missing entries (wrong particle number)
// are 0.
__global__ void __launch_bounds__(kLT) k_lr_prep(int N, int LX, int LY, int n, const int8_t* __restrict__ s,
                                                 int32_t* __restrict__ modes, int32_t* __restrict__ cnt) {
  __shared__ int wsum[kLT / kWave];
  const int img = blockIdx.x, lb = img / kGroup, g = img % kGroup;
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const int8_t* sb = s + (size_t)lb * 2 * N;
  int32_t* out = modes + (size_t)img * n;
  int base = 0;
  for (int c0 = 0; c0 < 2 * N; c0 += kLT) {
    const int i = c0 + threadIdx.x;
    const bool occ = i < 2 * N && sb[i] > 0;
    const uint64_t mask = __ballot(occ);
    if (lane == 0) wsum[wave] = __popcll(mask);
    __syncthreads();
    int off = base, tot = 0;
    for (int w = 0; w < kLT / kWave; ++w) {
      off += w < wave ? wsum[w] : 0;
      tot += wsum[w];
    }
    const int pos = off + lanes_below(mask);
    if (g == 0 && i < 2 * N) cnt[(size_t)lb * (2 * N + 1) + i] = pos;
    if (occ && pos < n) out[pos] = group_map(i, g, N, LX, LY);
    __syncthreads();
    base += tot;
  }
  if (g == 0 && threadIdx.x == 0) cnt[(size_t)lb * (2 * N + 1) + 2 * N] = base;
  for (int p = base + threadIdx.x; p < n; p += kLT) out[p] = 0;
}

// Yo[img][p][a] = Y[img][modes[p]][a]
__global__ void k_lr_gather(int n, int d, int two_n, const float* __restrict__ Y, const int32_t* __restrict__ modes,
                            float* __restrict__ Yo) {
  const int img = blockIdx.y, e = blockIdx.x * kLT + threadIdx.x;
  if (e >= n * d) return;
  const int p = e / d, a = e % d;
  Yo[(size_t)img * n * d + e] = Y[((size_t)img * two_n + modes[(size_t)img * n + p]) * d + a];
}

// F[img][p][q] += F0[modes[p]][modes[q]]
__global__ void k_lr_base(int n, int two_n, const float* __restrict__ F0, const int32_t* __restrict__ modes,
                          float* __restrict__ F) {
  const int img = blockIdx.y, e = blockIdx.x * kLT + threadIdx.x;
  if (e >= n * n) return;
  const int32_t* md = modes + (size_t)img * n;
  F[(size_t)img * n * n + e] += F0[(size_t)md[e / n] * two_n + md[e % n]];
}

// Per hop item c of the pass (slot lb * max_hops + c0 + c) and image g: the positions whose row
// of F changes, ascending, then the others (list[c][0 .. n)), each position's index in that list
// (lpos), their count k (kcnt) and the pass maximum (kmax, by atomicMax).
__global__ void __launch_bounds__(kLT) k_lr_cone(int N, int n, int d, int g, int lb, int c0, int max_hops,
                                                 const float* __restrict__ Ys, const float* __restrict__ Yi,
                                                 const int32_t* __restrict__ item_hop,
                                                 const int32_t* __restrict__ modes, const int32_t* __restrict__ cnt,
                                                 int32_t* __restrict__ list, int32_t* __restrict__ lpos,
                                                 int32_t* __restrict__ kcnt, int32_t* __restrict__ kmax) {
  __shared__ int wsum[kLT / kWave];
  const int c = blockIdx.x;
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const int32_t hop = item_hop[(size_t)lb * max_hops + c0 + c];
  const int a = hop_from(hop);
  const int pos_a = cnt[(size_t)lb * (2 * N + 1) + a];
  const int32_t* md = modes + (size_t)(lb * kGroup + g) * n;
  const float* ys = Ys + (size_t)(lb * kGroup + g) * 2 * N * d;
  const float* yi = Yi + (size_t)(c * kGroup + g) * 2 * N * d;
  int32_t* lst = list + (size_t)c * n;
  int32_t* lp = lpos + (size_t)c * n;
  int nchg = 0;
  for (int pass = 0; pass < 2; ++pass) {  // changed positions first, then the others
    int base = pass ? nchg : 0;
    for (int p0 = 0; p0 < n; p0 += kLT) {
      const int p = p0 + threadIdx.x;
      bool chg = false;
      if (p < n) {
        chg = p == pos_a;
        const int m = md[p];
        for (int k = 0; k < d && !chg; ++k) chg = yi[(size_t)m * d + k] != ys[(size_t)m * d + k];
      }
      const bool take = p < n && (pass ? !chg : chg);
      const uint64_t mask = __ballot(take);
      if (lane == 0) wsum[wave] = __popcll(mask);
      __syncthreads();
      int off = base, tot = 0;
      for (int w = 0; w < kLT / kWave; ++w) {
        off += w < wave ? wsum[w] : 0;
        tot += wsum[w];
      }
      if (take) {
        const int j = off + lanes_below(mask);
        lst[j] = p;
        lp[p] = j;
      }
      __syncthreads();
      base += tot;
    }
    if (!pass) nchg = base;
  }
  if (threadIdx.x == 0) {
    kcnt[c] = nchg;
    atomicMax(kmax, nchg);
  }
}

// D[c][j][i] (column-major n x kmax per hop): dF[i][q_j] = F'[i][q_j] - F[i][q_j] with rows of
// the padded set (list index < kmax) halved, exactly 0 where neither row nor column changed.
// F' from the item's features (position pos_a holds the hop's target), F[i][q] = -F[q][i] from
// the sample's matrix.  A thread keeps its row's features (d / 4 float4) and walks kLrCols
// columns, whose data are the same for the whole block; the table rows are float4 gathers.
constexpr int kLrCols = 8;

template <int D>
__global__ void __launch_bounds__(kLT) k_lr_dtilde(int N, int LX, int LY, int n, int g, int lb, int c0, int max_hops,
                                                   int kmax, const float* __restrict__ Yi,
                                                   const float* __restrict__ Fs, const float* __restrict__ wt,
                                                   const float* __restrict__ F0, const int32_t* __restrict__ item_hop,
                                                   const int32_t* __restrict__ modes, const int32_t* __restrict__ cnt,
                                                   const int32_t* __restrict__ list, const int32_t* __restrict__ lpos,
                                                   const int32_t* __restrict__ kcnt, float* __restrict__ Dt) {
  constexpr int D4 = D / 4;
  const int c = blockIdx.z, j0 = blockIdx.y * kLrCols, i = blockIdx.x * kLT + threadIdx.x;
  const int32_t hop = item_hop[(size_t)lb * max_hops + c0 + c];
  const int pos_a = cnt[(size_t)lb * (2 * N + 1) + hop_from(hop)];
  const int mb = group_map(hop_to(hop), g, N, LX, LY);
  const int kc = kcnt[c];
  const int32_t* md = modes + (size_t)(lb * kGroup + g) * n;
  const float4* yi4 = reinterpret_cast<const float4*>(Yi + (size_t)(c * kGroup + g) * 2 * N * D);
  const float* fs = Fs + (size_t)(lb * kGroup + g) * n * n;
  const bool row = i < n;
  const int li = row ? lpos[(size_t)c * n + i] : kmax;
  const int mi = !row ? 0 : (i == pos_a ? mb : md[i]);
  const int si = mi >= N, sitei = mi - si * N;
  const int xi = sitei / LY, yi_ = sitei - xi * LY;
  float4 ya[D4];
#pragma unroll
  for (int a = 0; a < D4; ++a) ya[a] = row ? yi4[(size_t)mi * D4 + a] : make_float4(0.f, 0.f, 0.f, 0.f);
  for (int jj = 0; jj < kLrCols; ++jj) {
    const int j = j0 + jj;
    if (j >= kmax) break;  // uniform
    const int q = list[(size_t)c * n + j];
    const int mq = q == pos_a ? mb : md[q];
    const int sq = mq >= N, siteq = mq - sq * N;
    const int xq = siteq / LY, yq = siteq - xq * LY;
    if (!row) continue;
    float v = 0.f;
    if (li < kc || j < kc) {
      int dx = xi - xq, dy = yi_ - yq;
      dx += dx < 0 ? LX : 0;
      dy += dy < 0 ? LY : 0;
      const float4* w4 = reinterpret_cast<const float4*>(wt + ((size_t)(2 * si + sq) * N + dx * LY + dy) * D);
      const float4* yb4 = yi4 + (size_t)mq * D4;
      float f = 0.f;
#pragma unroll
      for (int a = 0; a < D4; ++a) {
        const float4 w = w4[a], yb = yb4[a];
        f = fmaf(ya[a].x * yb.x, w.x, f);
        f = fmaf(ya[a].y * yb.y, w.y, f);
        f = fmaf(ya[a].z * yb.z, w.z, f);
        f = fmaf(ya[a].w * yb.w, w.w, f);
      }
      if (F0) f += F0[(size_t)mi * 2 * N + mq];
      v = f + fs[(size_t)q * n + i];
      if (li < kmax) v *= 0.5f;
    }
    Dt[((size_t)c * kmax + j) * n + i] = v;
  }
}

// M[c] (2k x 2k row-major, strict lower triangle): [[D^T X, .], [X_K + I, G_KK]], with E = D^T X
// column-major k x k, X column-major n x k, G the sample image's inverse (row-major).
__global__ void k_lr_assemble(int n, int kmax, int img_s, const int32_t* __restrict__ list,
                              const float* __restrict__ X, const float* __restrict__ E, const float* __restrict__ Gs,
                              float* __restrict__ M) {
  const int c = blockIdx.y;
  const int two_k = 2 * kmax;
  const int e = blockIdx.x * kLT + threadIdx.x;
  if (e >= two_k * two_k) return;
  const int i = e / two_k, j = e % two_k;
  float v = 0.f;
  if (i > j) {
    const int32_t* lst = list + (size_t)c * n;
    if (i < kmax) {
      v = E[(size_t)c * kmax * kmax + (size_t)j * kmax + i];
    } else if (j < kmax) {
      v = X[((size_t)c * kmax + j) * n + lst[i - kmax]] + (i - kmax == j ? 1.f : 0.f);
    } else {
      v = Gs[(size_t)img_s * n * n + (size_t)lst[i - kmax] * n + lst[j - kmax]];
    }
  }
  M[(size_t)c * two_k * two_k + e] = v;
}

// psi_g(s') of each hop of the pass into the item images' slots.
__global__ void k_lr_write(int N, int g, int lb, int c0, int ic, int max_hops, int kmax,
                           const int32_t* __restrict__ item_hop, const int32_t* __restrict__ cnt,
                           const float* __restrict__ sg_s, const double* __restrict__ lg_s,
                           const float* __restrict__ psg, const double* __restrict__ plg, float* __restrict__ isg,
                           double* __restrict__ ilg) {
  const int c = blockIdx.x * kLT + threadIdx.x;
  if (c >= ic) return;
  const size_t slot = (size_t)lb * max_hops + c0 + c;
  const int32_t hop = item_hop[slot];
  const int a = hop_from(hop), b = hop_to(hop);
  const int lo = a < b ? a : b, hi = a < b ? b : a;
  const int32_t* cb = cnt + (size_t)lb * (2 * N + 1);
  const int between = cb[hi] - cb[lo + 1];  // occupied modes strictly between a and b
  float sg = sg_s[lb * kGroup + g] * psg[c];
  if (((kmax * (kmax + 1) / 2) & 1) ^ (between & 1)) sg = -sg;  // 1 / Pf(-J), and b moved to its sorted place
  isg[slot * kGroup + g] = sg;
  ilg[slot * kGroup + g] = lg_s[lb * kGroup + g] + plg[c];
}

__global__ void k_lr_zero(int32_t* p) { *p = 0; }

// FP64 copies of F and G, R = I (the Newton step's accumulator)
__global__ void k_lr_to_f64(size_t total, int n, const float* __restrict__ F, const float* __restrict__ G,
                            double* __restrict__ Fd, double* __restrict__ Gd, double* __restrict__ R) {
  const size_t e = (size_t)blockIdx.x * kLT + threadIdx.x;
  if (e >= total) return;
  const size_t w = e % ((size_t)n * n);
  Fd[e] = F[e];
  Gd[e] = G[e];
  R[e] = (w / n == w % n) ? 1.0 : 0.0;
}

__global__ void k_lr_to_f32(size_t total, const double* __restrict__ Gd, float* __restrict__ G) {
  const size_t e = (size_t)blockIdx.x * kLT + threadIdx.x;
  if (e < total) G[e] = (float)Gd[e];
}

// One rocBLAS handle per device, created on first use; a call holds its device's lock for its
// whole duration (the handle carries the caller's stream), so calls on one device serialise.
std::mutex g_rb_mu[kLrMaxDev];
rocblas_handle g_rb[kLrMaxDev] = {};

int rb_device(int* dev) {
  NH_HIP(hipGetDevice(dev));
  return (*dev < 0 || *dev >= kLrMaxDev) ? NH_ERR_HIP : NH_OK;
}

int rb_handle(int dev, hipStream_t st, rocblas_handle* h) {
  if (!g_rb[dev] && rocblas_create_handle(&g_rb[dev]) != rocblas_status_success) return NH_ERR_HIP;
  if (rocblas_set_stream(g_rb[dev], st) != rocblas_status_success) return NH_ERR_HIP;
  *h = g_rb[dev];
  return NH_OK;
}

struct LrCarver {
  char* base;
  size_t off = 0;
  template <class T>
  T* take(size_t k) {
    T* p = base ? (T*)(base + off) : nullptr;
    off += (k * sizeof(T) + 255) & ~(size_t)255;
    return p;
  }
};

struct LrBuffers {
  int32_t *count, *item_b, *item_hop;
  float *coef, *hsign, *isg;
  double *hlog, *ilg;
  float *Ys, *gtri, *net, *sg_s, *Yo, *Fs, *Gs, *Yi, *Dt, *X, *E, *M, *psg;
  double *lg_s, *plg, *Fd, *Gd, *Rd;
  int32_t *modes, *cnt, *list, *lpos, *kcnt, *kmax;
  void* pws;
  size_t pws_bytes;
};

int lr_max_hops(const nh_model* m, double tprime) { return (m->n_up + m->n_dn) * (tprime != 0.0 ? 8 : 4); }

LrBuffers lr_carve(LrCarver& c, const nh_model* m, double tprime, int cs) {
  const size_t N = (size_t)m->LX * m->LY, n = (size_t)(m->n_up + m->n_dn), d = (size_t)feature_dim(m);
  const size_t mh = (size_t)lr_max_hops(m, tprime), kb = (size_t)lr_kbound(m);
  const size_t imgs = (size_t)cs * kGroup, items = (size_t)cs * mh, ic = kLrItems;
  LrBuffers b;
  b.count = c.take<int32_t>(cs);
  b.item_b = c.take<int32_t>(items);
  b.item_hop = c.take<int32_t>(items);
  b.coef = c.take<float>(items);
  b.hsign = c.take<float>(items);
  b.hlog = c.take<double>(items);
  b.isg = c.take<float>(items * kGroup);
  b.ilg = c.take<double>(items * kGroup);
  b.Ys = c.take<float>(imgs * 2 * N * d);
  const size_t tri = image_global_floats(m);
  b.gtri = tri ? c.take<float>(imgs * tri) : nullptr;
  const size_t net = resconv_scratch_floats(m, (int64_t)std::max(imgs, ic * kGroup));
  b.net = net ? c.take<float>(net) : nullptr;
  b.sg_s = c.take<float>(imgs);
  b.lg_s = c.take<double>(imgs);
  b.modes = c.take<int32_t>(imgs * n);
  b.cnt = c.take<int32_t>((size_t)cs * (2 * N + 1));
  b.Yo = c.take<float>(imgs * n * d);
  b.Fs = c.take<float>(imgs * n * n);
  b.Gs = c.take<float>(imgs * n * n);
  b.Fd = c.take<double>(imgs * n * n);
  b.Gd = c.take<double>(imgs * n * n);
  b.Rd = c.take<double>(imgs * n * n);
  b.Yi = c.take<float>(ic * kGroup * 2 * N * d);
  b.list = c.take<int32_t>(ic * n);
  b.lpos = c.take<int32_t>(ic * n);
  b.kcnt = c.take<int32_t>(ic);
  b.kmax = c.take<int32_t>(1);
  b.Dt = c.take<float>(ic * n * kb);
  b.X = c.take<float>(ic * n * kb);
  b.E = c.take<float>(ic * kb * kb);
  b.M = c.take<float>(ic * 4 * kb * kb);
  b.psg = c.take<float>(ic);
  b.plg = c.take<double>(ic);
  // a pass's 2k x 2k Pfaffians have 2k <= 2 kb; the workspace form's column form takes more bytes
  // per matrix than the packed triangle of a larger n, so the size is the largest over every 2k
  b.pws_bytes = 0;
  for (int32_t two_k = 2; two_k <= (int32_t)(2 * kb); two_k += 2) {
    const size_t need = nh_slogpf_workspace_size((int32_t)ic, two_k);
    if (need > b.pws_bytes) b.pws_bytes = need;
  }
  b.pws = c.take<char>(b.pws_bytes);
  return b;
}

int lr_check(const nh_model* m, double tprime, int32_t chunk) {
  int rc = check_model(m);
  if (rc != NH_OK) return rc;
  if (m->kind != NH_KIND_PFAFFIAN || m->network != NH_NET_CONV || !m->project) return NH_ERR_UNSUPPORTED;
  if (feature_dim(m) != 4 && feature_dim(m) != 8 && feature_dim(m) != 16) return NH_ERR_UNSUPPORTED;
  if (2 * lr_kbound(m) > 1024) return NH_ERR_UNSUPPORTED;  // the 2k x 2k Pfaffians
  if (chunk <= 0 || (int64_t)chunk * kGroup > image_chunk(m)) return NH_ERR_SHAPE;
  if ((int64_t)chunk * lr_max_hops(m, tprime) * kGroup > INT32_MAX / 2) return NH_ERR_SHAPE;
  return NH_OK;
}

}  // namespace
}  // namespace nh

using namespace nh;

extern "C" size_t nh_local_energy_lowrank_workspace_size(const nh_model* model, double tprime, int32_t chunk) {
  if (lr_check(model, tprime, chunk) != NH_OK) return 0;
  LrCarver c{nullptr};
  lr_carve(c, model, tprime, chunk);
  return c.off + 256;
}

extern "C" int nh_local_energy_lowrank(void* stream, const nh_model* model, double t, double tprime, double U,
                                       int32_t batch, const int8_t* s, const float* sign_s, const double* logabs_s,
                                       double* eloc, int32_t chunk, void* workspace, size_t workspace_bytes) {
  int rc = lr_check(model, tprime, chunk);
  if (rc != NH_OK) return rc;
  if (batch < 0) return NH_ERR_ARG;
  if (batch == 0) return NH_OK;
  if (!s || !sign_s || !logabs_s || !eloc || !workspace) return NH_ERR_ARG;
  if (workspace_bytes < nh_local_energy_lowrank_workspace_size(model, tprime, chunk)) return NH_ERR_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  int dev = 0;
  rc = rb_device(&dev);
  if (rc != NH_OK) return rc;
  std::lock_guard<std::mutex> lk(g_rb_mu[dev]);
  rocblas_handle rb;
  rc = rb_handle(dev, st, &rb);
  if (rc != NH_OK) return rc;
  LrCarver cv{(char*)(((size_t)workspace + 255) & ~(size_t)255)};
  const LrBuffers bf = lr_carve(cv, model, tprime, chunk);
  const int N = model->LX * model->LY, n = model->n_up + model->n_dn, d = feature_dim(model);
  const int mh = lr_max_hops(model, tprime), kb = lr_kbound(model);
  const int z = tprime != 0.0 ? 8 : 4;
  std::vector<int32_t> count(chunk);
  for (int b0 = 0; b0 < batch; b0 += chunk) {
    const int cs = std::min(chunk, batch - b0);
    const int8_t* sc = s + (size_t)b0 * 2 * N;
    rc = launch_hops(st, model->LX, model->LY, (float)t, (float)tprime, z, cs, b0, s, mh, bf.count, bf.item_b,
                     bf.item_hop, bf.coef, nullptr, nullptr);
    if (rc != NH_OK) return rc;
    // the samples' images: features, amplitudes, matrices and inverses
    const int imgs = cs * kGroup;
    rc = launch_cnn(st, model, sc, nullptr, nullptr, 0, imgs, kGroup, bf.Ys, bf.net);
    if (rc != NH_OK) return rc;
    rc = launch_image_amplitude(st, model, sc, nullptr, nullptr, 0, imgs, kGroup, bf.Ys, bf.sg_s, bf.lg_s, bf.gtri);
    if (rc != NH_OK) return rc;
    hipLaunchKernelGGL(k_lr_prep, dim3(imgs), dim3(kLT), 0, st, N, model->LX, model->LY, n, sc, bf.modes, bf.cnt);
    hipLaunchKernelGGL(k_lr_gather, dim3((n * d + kLT - 1) / kLT, imgs), dim3(kLT), 0, st, n, d, 2 * N, bf.Ys,
                       bf.modes, bf.Yo);
    NH_LAUNCH_CHECK();
    rc = nh_pairing_occ(stream, model->LX, model->LY, d, imgs, n, bf.Yo, bf.modes, model->wt, bf.Fs);
    if (rc != NH_OK) return rc;
    if (model->base) {
      hipLaunchKernelGGL(k_lr_base, dim3((n * n + kLT - 1) / kLT, imgs), dim3(kLT), 0, st, n, 2 * N, model->base,
                         bf.modes, bf.Fs);
      NH_LAUNCH_CHECK();
    }
    rc = launch_inverse_t(st, imgs, n, bf.Fs, -1.f, bf.Gs);  // F^{-1} = -F^{-T} for skew F
    if (rc != NH_OK) return rc;
    {  // one Newton step in FP64: R = I - F G, G <- G + G R (row-major; column-major views transposed)
      const size_t tot = (size_t)imgs * n * n;
      const rocblas_stride sn = (rocblas_stride)n * n;
      hipLaunchKernelGGL(k_lr_to_f64, dim3((unsigned)((tot + kLT - 1) / kLT)), dim3(kLT), 0, st, tot, n, bf.Fs, bf.Gs,
                         bf.Fd, bf.Gd, bf.Rd);
      NH_LAUNCH_CHECK();
      const double m1 = -1.0, p1 = 1.0;
      if (rocblas_dgemm_strided_batched(rb, rocblas_operation_none, rocblas_operation_none, n, n, n, &m1, bf.Gd, n, sn,
                                        bf.Fd, n, sn, &p1, bf.Rd, n, sn, imgs) != rocblas_status_success)
        return NH_ERR_HIP;
      NH_HIP(hipMemcpyAsync(bf.Fd, bf.Gd, sizeof(double) * tot, hipMemcpyDeviceToDevice, st));
      if (rocblas_dgemm_strided_batched(rb, rocblas_operation_none, rocblas_operation_none, n, n, n, &p1, bf.Rd, n, sn,
                                        bf.Gd, n, sn, &p1, bf.Fd, n, sn, imgs) != rocblas_status_success)
        return NH_ERR_HIP;
      hipLaunchKernelGGL(k_lr_to_f32, dim3((unsigned)((tot + kLT - 1) / kLT)), dim3(kLT), 0, st, tot, bf.Fd, bf.Gs);
      NH_LAUNCH_CHECK();
    }
    NH_HIP(hipMemcpyAsync(count.data(), bf.count, sizeof(int32_t) * cs, hipMemcpyDeviceToHost, st));
    NH_HIP(hipStreamSynchronize(st));
    for (int lb = 0; lb < cs; ++lb) {
      if (count[lb] > mh) return NH_ERR_SHAPE;  // more electrons than the model holds
      for (int c0 = 0; c0 < count[lb]; c0 += kLrItems) {
        const int ic = std::min(kLrItems, count[lb] - c0);
        const int slot0 = lb * mh + c0;
        rc = launch_cnn(st, model, s, bf.item_b, bf.item_hop, slot0 * kGroup, (slot0 + ic) * kGroup, kGroup, bf.Yi,
                        bf.net);
        if (rc != NH_OK) return rc;
        for (int g = 0; g < kGroup; ++g) {
          hipLaunchKernelGGL(k_lr_zero, dim3(1), dim3(1), 0, st, bf.kmax);
          hipLaunchKernelGGL(k_lr_cone, dim3(ic), dim3(kLT), 0, st, N, n, d, g, lb, c0, mh, bf.Ys, bf.Yi,
                             bf.item_hop, bf.modes, bf.cnt, bf.list, bf.lpos, bf.kcnt, bf.kmax);
          NH_LAUNCH_CHECK();
          int32_t kmax = 0;
          NH_HIP(hipMemcpyAsync(&kmax, bf.kmax, sizeof(int32_t), hipMemcpyDeviceToHost, st));
          NH_HIP(hipStreamSynchronize(st));
          if (kmax < 1 || kmax > kb) return NH_ERR_SHAPE;  // the receptive-field bound does not hold
          const int imgs_g = lb * kGroup + g;
          {
            const dim3 gd((n + kLT - 1) / kLT, (kmax + kLrCols - 1) / kLrCols, ic);
#define NH_LR_DT(DD)                                                                                             \
  hipLaunchKernelGGL(k_lr_dtilde<DD>, gd, dim3(kLT), 0, st, N, model->LX, model->LY, n, g, lb, c0, mh, kmax, bf.Yi, \
                     bf.Fs, model->wt, model->base, bf.item_hop, bf.modes, bf.cnt, bf.list, bf.lpos, bf.kcnt, bf.Dt)
            switch (d) {
              case 4: NH_LR_DT(4); break;
              case 8: NH_LR_DT(8); break;
              default: NH_LR_DT(16);
            }
#undef NH_LR_DT
            NH_LAUNCH_CHECK();
          }
          const float one = 1.f, zero = 0.f;
          // X = G D for all hops of the pass (G row-major: its column-major view is G^T)
          if (rocblas_sgemm(rb, rocblas_operation_transpose, rocblas_operation_none, n, ic * kmax, n, &one,
                            bf.Gs + (size_t)imgs_g * n * n, n, bf.Dt, n, &zero, bf.X, n) != rocblas_status_success)
            return NH_ERR_HIP;
          // E = D^T X per hop
          if (rocblas_sgemm_strided_batched(rb, rocblas_operation_transpose, rocblas_operation_none, kmax, kmax, n,
                                            &one, bf.Dt, n, (rocblas_stride)n * kmax, bf.X, n,
                                            (rocblas_stride)n * kmax, &zero, bf.E, kmax,
                                            (rocblas_stride)kmax * kmax, ic) != rocblas_status_success)
            return NH_ERR_HIP;
          const int two_k = 2 * kmax;
          hipLaunchKernelGGL(k_lr_assemble, dim3((two_k * two_k + kLT - 1) / kLT, ic), dim3(kLT), 0, st, n, kmax,
                             imgs_g, bf.list, bf.X, bf.E, bf.Gs, bf.M);
          NH_LAUNCH_CHECK();
          rc = nh_slogpf_ws(stream, ic, two_k, bf.M, bf.psg, bf.plg, bf.pws, bf.pws_bytes);
          if (rc != NH_OK) return rc;
          hipLaunchKernelGGL(k_lr_write, dim3((ic + kLT - 1) / kLT), dim3(kLT), 0, st, N, g, lb, c0, ic, mh, kmax,
                             bf.item_hop, bf.cnt, bf.sg_s, bf.lg_s, bf.psg, bf.plg, bf.isg, bf.ilg);
          NH_LAUNCH_CHECK();
        }
      }
    }
    rc = launch_combine(st, bf.item_hop, 0, cs * mh, kGroup, bf.isg, bf.ilg, bf.hsign, bf.hlog);
    if (rc != NH_OK) return rc;
    rc = launch_eloc_reduce(st, N, U, cs, b0, s, sign_s, logabs_s, mh, bf.count, bf.coef, bf.hsign, bf.hlog, eloc);
    if (rc != NH_OK) return rc;
  }
  return NH_OK;
}
