/* nh.h - C ABI of the MI355X (gfx950) Pfaffian/BCS backflow amplitude and local-energy path.
 *
 * Every pointer named "device" is caller-owned device memory; no entry point allocates,
 * frees or throws.  Each call is asynchronous on `stream` (a hipStream_t passed as void*;
 * NULL is the default stream) and returns an NH_* status.  Work that needs scratch memory
 * takes a caller-provided workspace whose size comes from the matching *_workspace_size
 * query.  Calls on different streams with different workspaces are independent.
 *
 * Conventions (bit-exact with the reference, ChenAo-Phys/NeuralHubbard):
 *   configuration s: int8 [2N], +1 occupied / -1 empty, spin-up block then spin-down block,
 *   site (x, y) -> x * LY + y (README.md:85-86; code/backflow/models.py:96-97).
 *   mode m = spin * N + site; occupied modes in ascending order (quantax fermion_idx).
 *   projection: sum over Flip(axis=0) @ Flip(axis=1) @ SpinInverse (code/states.py:274-279),
 *   group element g = fx | fy << 1 | si << 2.
 *   log amplitude: (sign, logabs) pairs, the quantax LogArray (code/backflow/models.py:211).
 */
#ifndef NH_H
#define NH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  NH_OK = 0,
  NH_ERR_ARG = 1,         /* null pointer or invalid scalar argument */
  NH_ERR_SHAPE = 2,       /* sizes outside what the kernels support */
  NH_ERR_WORKSPACE = 3,   /* workspace smaller than the query returned */
  NH_ERR_HIP = 4,         /* a HIP runtime call failed */
  NH_ERR_UNSUPPORTED = 5  /* option combination not implemented */
};

enum { NH_KIND_PFAFFIAN = 0, NH_KIND_BCS = 1 };
enum { NH_PREC_FP32 = 0, NH_PREC_MIXED = 1 };
enum { NH_NET_CONV = 0, NH_NET_RESCONV = 1, NH_NET_RESCONV_HEAD = 2 };

/* One backflow network and its pairing parameters (all device pointers).
 * Features y[spin][site][a] = h[spin * d + a][site], d = C / 2 (code/backflow/models.py:163-167).
 * Conv weights are packed per conv as [tap = (dx+1)*3 + (dy+1)][c_in][c_out].
 * network NH_NET_CONV: embedding 3x3 periodic conv 2 -> C of the +-1 occupancy, then n_layers
 *   residual layers h <- h + conv3x3(gelu_tanh(h)); conv_w = the n_layers + 1 convs (the first
 *   with c_in = 2), conv_b [(n_layers + 1)][C].  For LY == 16 and C == 32 the network runs on
 *   MFMA when the fragment-ordered copies are given (layout in neuralhubbard_amd/mixed.py):
 *   conv_w_mfma for NH_PREC_FP32, conv_w_lp (BF16 operands, FP32 accumulation) for
 *   NH_PREC_MIXED.
 * network NH_NET_RESCONV: the residual CNN of the released states (quantax ResConv as built at
 *   code/backflow/models.py:319-329; data/supplement/tableS1_network/network_settings.json:21-56):
 *   h = emb[occupation class of each site] (class = n_up + 2 n_dn: empty, up, down, double),
 *   n_layers blocks h <- h + conv3x3(gelu_tanh(conv3x3(gelu_tanh(rmsnorm(h))))) with
 *   rmsnorm(h) = h / sqrt(mean over all channels and sites of h^2 + 1e-5), then
 *   h / sqrt(n_layers + 1).  conv_w = emb [4][C] followed by the 2 n_layers convs,
 *   conv_b [2 n_layers][C].  C == 64; on LY == 16, LX in {4, 8, 12, 16} NH_PREC_FP32 runs FP32
 *   MFMA; NH_PREC_MIXED runs split-BF16 MFMA (FP32 accumulation) on conv_w_lp, the BF16
 *   A-fragments of the convs (layout in neuralhubbard_amd/mixed.py, pack_resconv_bf16).
 *   conv_w_mfma, when given, holds the FP32 convs as B-fragment slabs (pack_resconv_f32) that
 *   the FP32 kernel streams through LDS; NULL reads conv_w directly.  Other lattices
 *   (N <= 1024) run the network (and the BCS head) layer by layer with the activations in the
 *   workspace: FP32 MFMA convolutions, or split-BF16 MFMA on conv_w_lp under NH_PREC_MIXED.
 * network NH_NET_RESCONV_HEAD: the same network followed by the BCS states' linear head
 *   (code/backflow/models.py:261-270): y[k][site][a] = sum_c head[k][a][c] x[c][site], x the
 *   scaled output, d = C features per spin; conv_w ends with the head as [k][c][a] (2 x C x C);
 *   a missing bias (the trunk's last conv) is a zero row of conv_b.
 * pfaffian: wt = [4][N][d] spin-block table [uu, ud, du, dd] with the distance envelope
 *           (code/backflow/pf_kernel.py:107-124); base = F0 [2N][2N] antisymmetric or NULL.
 * bcs:      wt = [N][d] (code/backflow/bcs_kernel.py:63-70); base = phi0 [N] or NULL,
 *           added as phi0[r(up_p - dn_q)].
 * Sizes: N = LX * LY <= 1024; pfaffian n_up + n_dn even and <= 1024; bcs n_up == n_dn <= 512
 * (the 32x32 lattice at 7/8 filling has n_up = 448); matrices above n = 64 (pfaffian) or
 * n = 192 (bcs, and from n = 65 where the features take most of the LDS, d = 64) live in the
 * workspace. */
typedef struct nh_model {
  int32_t kind;       /* NH_KIND_* */
  int32_t LX, LY;     /* periodic lattice, both >= 3 */
  int32_t n_up, n_dn; /* electrons per spin */
  int32_t channels;   /* C: 8, 16 or 32 (NH_NET_CONV), 64 (NH_NET_RESCONV[_HEAD]) */
  int32_t n_layers;   /* residual layers after the embedding conv, or residual blocks */
  int32_t precision;  /* NH_PREC_FP32 or NH_PREC_MIXED */
  int32_t project;    /* 1: 8-element projection, 0: the bare network */
  int32_t network;    /* NH_NET_* */
  const float* conv_w;
  const float* conv_b;
  const float* wt;
  const float* base;
  const void* conv_w_lp;     /* NH_PREC_MIXED: BF16 conv weights in MFMA fragment order, else NULL */
  const float* conv_w_mfma;  /* FP32 conv weights in MFMA fragment order, or NULL (VALU kernel) */
} nh_model;

/* ---- a6: occupied pairing block  (replaces pf_kernel.pairing_occ, code/backflow/pf_kernel.py:335-351)
 * F[b,p,q] = sum_a Y[b,p,a] Y[b,q,a] Wt[2 s_p + s_q][r(p,q)][a], r = site_p - site_q
 * periodic, s_p = idx[b,p] / N.  Exactly antisymmetric.  Y [batch][n][d], idx [batch][n],
 * Wt [4][N][d], F [batch][n][n] (device).  d % 4 == 0, n <= 1024. */
int nh_pairing_occ(void* stream, int32_t LX, int32_t LY, int32_t d, int32_t batch, int32_t n,
                   const float* Y, const int32_t* idx, const float* Wt, float* F);

/* ---- a9: up x down pairing block  (replaces bcs_kernel.pairing_ud, code/backflow/bcs_kernel.py:257-271)
 * F[b,p,q] = sum_a Yu[b,p,a] Wt[r(iu_p - idn_q)][a] Yd[b,q,a]. */
int nh_pairing_ud(void* stream, int32_t LX, int32_t LY, int32_t d, int32_t batch, int32_t n_up,
                  int32_t n_dn, const float* Yu, const float* Yd, const int32_t* iu,
                  const int32_t* idn, const float* Wt, float* F);

/* ---- a8: batched Pfaffian  (replaces fermix.slogpf, called at code/backflow/models.py:210)
 * A [batch][n][n] antisymmetric (only the strict lower triangle is read), n even, n <= 256.
 * Outputs sign (+1/-1, 0 if singular) and logabs (-inf if singular) per matrix.
 * Parlett-Reid elimination with partial pivoting, FP32, pivot logs summed in FP64. */
int nh_slogpf(void* stream, int32_t batch, int32_t n, const float* A, float* sign, double* logabs);
/* The same for n <= 1024 (fermix.slogpf at the 32x32 sizes, n = 896) with a workspace: each
 * matrix is eliminated in its workspace slice by a blocked Parlett-Reid routine of the image
 * kernels, for 100 <= n <= 352 as a column-major skew square with lazy pivoting (the two
 * triangles, 348 KB per matrix at n = 224), for the other n as the packed strict lower
 * triangle (panel vectors in LDS).  Only the strict lower triangle of A is read.  A workspace of
 * nh_slogpf_workspace_size(batch, n) bytes takes the batch in one launch; a smaller one (at
 * least one matrix: nh_slogpf_workspace_size(1, n)) runs the batch in chunks that fit.
 * n <= 64 needs no workspace (the result of nh_slogpf). */
size_t nh_slogpf_workspace_size(int32_t batch, int32_t n);
int nh_slogpf_ws(void* stream, int32_t batch, int32_t n, const float* A, float* sign, double* logabs,
                 void* workspace, size_t workspace_bytes);

/* ---- a9: batched determinant  (replaces fermix.slogdet, called at code/backflow/models.py:297)
 * A [batch][n][n], n <= 192.  LU with partial pivoting, FP32, logs summed in FP64. */
int nh_slogdet(void* stream, int32_t batch, int32_t n, const float* A, float* sign, double* logabs);

/* ---- f2: reverse mode of the occupied pairing block
 *      (replaces the backward of pf_kernel.pairing_occ's custom_vjp, code/backflow/pf_kernel.py:354-365,
 *      kernels _dy_kernel :193-213 and _dw_kernel :239-290)
 * For the cotangent G [batch][n][n] of F = nh_pairing_occ(Y, idx, Wt), per sample:
 *   dY[b,p,a]      = sum_q (G[p,q] - G[q,p]) Wt[2 s_p + s_q][r(p,q)][a] Y[b,q,a]
 *   dWt[b][k][r][a] = sum over (p, q) with 2 s_p + s_q = k, r(p,q) = r of G[p,q] Y[p,a] Y[q,a].
 * dY [batch][n][d], dWt [batch][4][N][d] (device; either may be NULL to skip it).  Wt must have
 * the antisymmetry of build_wtilde (Wt[2 s_q + s_p][-r] = -Wt[2 s_p + s_q][r]).  d divides 256
 * and d <= 64, N <= 1024, batch <= 65535; a mode outside [0, 2N) gives NaN rows of dY. */
int nh_pairing_occ_vjp(void* stream, int32_t LX, int32_t LY, int32_t d, int32_t batch, int32_t n,
                       const float* Y, const int32_t* idx, const float* Wt, const float* G, float* dY,
                       float* dWt);

/* ---- f2: Pfaffian VJP (the reverse mode of fermix.slogpf, code/backflow/models.py:210)
 * G[b] = cot[b] * d log|Pf F[b]| / dF[b] = cot[b] * (1/2) F[b]^{-T}, every entry of F taken as
 * independent; cot NULL means 1.  F, G [batch][n][n] (device, distinct buffers), n <= 1024.
 * Blocked Gauss-Jordan with partial pivoting (16-column panels, rank-16 MFMA updates) in FP32 on
 * G itself; a singular F gives NaN. */
int nh_slogpf_vjp(void* stream, int32_t batch, int32_t n, const float* F, const float* cot, float* G);

/* ---- f2: reverse mode of the BCS up x down pairing block
 *      (replaces the backward of bcs_kernel.pairing_ud's custom_vjp, code/backflow/bcs_kernel.py:274-288,
 *      kernels _dy_kernel :123-141 (both sides) and _dw_kernel :169-206)
 * For the cotangent G [batch][n_up][n_dn] of F = nh_pairing_ud(Yu, Yd, iu, idn, Wt), per sample:
 *   dYu[b,p,a] = sum_q G[p,q] Wt[r(iu_p - idn_q)][a] Yd[b,q,a]
 *   dYd[b,q,a] = sum_p G[p,q] Wt[r(iu_p - idn_q)][a] Yu[b,p,a]
 *   dWt[b][r][a] = sum over (p, q) with r(iu_p - idn_q) = r of G[p,q] Yu[p,a] Yd[q,a].
 * dYu [batch][n_up][d], dYd [batch][n_dn][d], dWt [batch][N][d] (device; any may be NULL to skip
 * it).  The down sites of a sample are distinct; d divides 256 and d <= 64, N <= 1024,
 * batch <= 65535; a site outside [0, N) gives NaN rows. */
int nh_pairing_ud_vjp(void* stream, int32_t LX, int32_t LY, int32_t d, int32_t batch, int32_t n_up,
                      int32_t n_dn, const float* Yu, const float* Yd, const int32_t* iu, const int32_t* idn,
                      const float* Wt, const float* G, float* dYu, float* dYd, float* dWt);

/* ---- f2: determinant VJP (the reverse mode of fermix.slogdet, code/backflow/models.py:297)
 * G[b] = cot[b] * d log|det F[b]| / dF[b] = cot[b] * F[b]^{-T}; cot NULL means 1.  F, G
 * [batch][n][n] (device, distinct buffers), n <= 1024; the Gauss-Jordan kernel of nh_slogpf_vjp;
 * a singular F gives NaN. */
int nh_slogdet_vjp(void* stream, int32_t batch, int32_t n, const float* F, const float* cot, float* G);

/* ---- a4: backflow features  (replaces PfaffianBackflow.features, code/backflow/models.py:163-167)
 * s int8 [batch][2N] -> y float [batch][2][N][d]  (no projection; the identity image). */
int nh_features(void* stream, const nh_model* model, int32_t batch, const int8_t* s, float* y);
/* The same with a workspace: the residual CNN on lattices outside its tile kernels (LY != 16 or
 * LX > 16, e.g. the released 24x24 state) runs layer by layer with its activations there, and
 * nh_features returns NH_ERR_UNSUPPORTED for it.  The query returns 256 bytes where none is needed. */
size_t nh_features_workspace_size(const nh_model* model, int32_t batch);
int nh_features_ws(void* stream, const nh_model* model, int32_t batch, const int8_t* s, float* y, void* workspace,
                   size_t workspace_bytes);

/* ---- a12 + model __call__: projected amplitude  (replaces quantax Variational.__call__ over
 * PfaffianBackflow/BCSBackflow.__call__, code/backflow/models.py:208-211, 294-302,
 * code/states.py:318-320).  s int8 [batch][2N] -> sign float [batch], logabs double [batch]. */
size_t nh_log_psi_workspace_size(const nh_model* model, int32_t batch);
int nh_log_psi(void* stream, const nh_model* model, int32_t batch, const int8_t* s, float* sign,
               double* logabs, void* workspace, size_t workspace_bytes);

/* ---- f2: per-sample Jacobian of log|psi_P|  (replaces the per-sample gradient quantax's AdamSR
 * takes through the model, code/train_pfaffian_16x16_tpm02.py:100-101, with the pairing block's
 * custom_vjp backward, code/backflow/pf_kernel.py:354-365)
 * J[b][k] = d log|psi_P(s_b)| / d theta_k (the bare network's amplitude when model->project is
 * 0), columns in the order
 *   the packed conv weights model->conv_w (synthetic network: the first convolution [9][2][C],
 *     then [9][C][C] per layer; residual CNN: the embedding [4][C], [9][C][C] per convolution,
 *     the BCS head [2][C][d]) -- [tap][c_in][c_out], tap = (dx + 1) * 3 + (dy + 1);
 *   the conv biases model->conv_b;
 *   W of build_wtilde ([3][N][d]: W_ud, V_uu, V_dd) or of build_wt_ud ([N][d]);
 *   theta (the envelope parameter, lambda = softplus(theta); model->wt must be built from it);
 *   with model->base: F0 [2N][2N] (every entry independent) or phi0 [N].
 * nh_log_psi_jacobian_size gives the column count (0 for an unsupported model).  sign / logabs
 * receive psi_P as nh_log_psi computes it.  FP32 throughout (conv_w, not the mixed form's
 * low-precision weights); the image matrices' inverses by the Gauss-Jordan kernel of
 * nh_slogpf_vjp, so an image with a singular matrix gives NaN rows.  batch * 8 <= 65535 (with
 * the projection); the workspace grows with every layer's activations of every image
 * (nh_log_psi_jacobian_workspace_size). */
size_t nh_log_psi_jacobian_size(const nh_model* model);
size_t nh_log_psi_jacobian_workspace_size(const nh_model* model, int32_t batch);
int nh_log_psi_jacobian(void* stream, const nh_model* model, double theta, int32_t batch, const int8_t* s,
                        float* sign, double* logabs, float* J, void* workspace, size_t workspace_bytes);

/* ---- allocation-free projected amplitude (no counterpart in the reference: for callers that
 * cannot allocate device memory -- code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  The same result as nh_log_psi, with the workspace taken from a static
 * device array of NH_NOALLOC_SCRATCH_BYTES that each GPU the library runs on holds until the
 * process ends; no hipMalloc, hipFree or allocating library call.  Any batch is processed in
 * passes that fit the array; a model of which one configuration needs more workspace than that
 * returns NH_ERR_WORKSPACE.  Every call on a GPU shares that GPU's array: calls that overlap on
 * one GPU (other streams, other host threads) race and must be serialised by the caller. */
#define NH_NOALLOC_SCRATCH_BYTES ((size_t)256 << 20)
int nh_log_psi_noalloc(void* stream, const nh_model* model, int32_t batch, const int8_t* s, float* sign,
                       double* logabs);

/* ---- a13: Hubbard local energy  (replaces quantax Hubbard.Oloc, code/quick_start.py:33,
 * built at code/states.py:244-247).  E_loc(s) = U sum_i n_iu n_id
 *   - sum_{<ij>,spin} t_ij JW ψ(s^{i->j}) / ψ(s), t_ij = t (nearest) or tprime (diagonal,
 * used when tprime != 0).  sign_s / logabs_s are ψ(s) from nh_log_psi.  eloc double [batch].
 * Connected configurations are evaluated in full (8 images each), in chunks that fit the
 * workspace; the query returns the size for `chunk` samples per chunk. */
size_t nh_local_energy_workspace_size(const nh_model* model, double tprime, int32_t chunk);
int nh_local_energy(void* stream, const nh_model* model, double t, double tprime, double U,
                    int32_t batch, const int8_t* s, const float* sign_s, const double* logabs_s,
                    double* eloc, int32_t chunk, void* workspace, size_t workspace_bytes);

/* ---- a13 with low-rank hop ratios (the north star's low-rank Pfaffian-ratio updates for the
 * nearest- and next-nearest-neighbour hops of the local energy; replaces the ratios inside
 * quantax Hubbard.Oloc, code/quick_start.py:33, for the synthetic network)
 * The same E_loc as nh_local_energy, with ψ(s') / ψ(s) of every hop as a rank-2k update of the
 * sample's eight image matrices (k = the occupied modes whose features the hop changes: its
 * two ends' receptive-field squares), Pf(F') / Pf(F) = (-1)^(k(k+1)/2) Pf(-J + B^T F^{-1} B),
 * F^{-1} once per sample image by the blocked Gauss-Jordan kernel refined by one FP64 Newton
 * step, the GEMMs on rocBLAS.  Pfaffian
 * states with the projection and network NH_NET_CONV, where 2k <= 1024 holds by the receptive
 * field (the 32x32 workload: k <= 508); other models return NH_ERR_UNSUPPORTED.  chunk samples
 * per pass (chunk * 8 <= the image chunk: 128 at n = 896); the call synchronises the stream, and
 * calls on one device run one at a time (they share that device's rocBLAS handle). */
size_t nh_local_energy_lowrank_workspace_size(const nh_model* model, double tprime, int32_t chunk);
int nh_local_energy_lowrank(void* stream, const nh_model* model, double t, double tprime, double U,
                            int32_t batch, const int8_t* s, const float* sign_s, const double* logabs_s,
                            double* eloc, int32_t chunk, void* workspace, size_t workspace_bytes);

/* Hop enumeration used by nh_local_energy, exposed for checking:
 * for each sample, hops in order (occupied mode ascending, partner in the order
 * (+x, -x, +y, -y, +x+y, +x-y, -x+y, -x-y)); out_from/out_to/out_coef [batch][max_hops],
 * count [batch].  max_hops = (n_up + n_dn) * z, z = 4 or 8. */
int nh_hops(void* stream, int32_t LX, int32_t LY, double t, double tprime, int32_t batch,
            int32_t n_occ, const int8_t* s, int32_t max_hops, int32_t* count, int32_t* out_from,
            int32_t* out_to, float* out_coef);

/* ---- surrogate log-amplitude network (no reference counterpart: the learned stage one of
 * delayed-acceptance Metropolis).  z1 = conv3x3(x0) + b1, z2 = conv3x3(gelu(z1)) + b2,
 * z3 = conv3x3(gelu(z2)) + b3 (1 channel), f(s) = sum over sites of z3; x0 = the 2-channel +-1
 * occupancy, periodic 3x3 convolutions, channels 16 or 32 for N <= 256, 16 for N <= 1024.
 * params (device, FP32):
 * w1 [9][2][C], b1 [C], w2 [9][C][C], b2 [C], w3 [9][C], b3 [1]; nh_sur_nparams gives the count. */
typedef struct nh_surrogate {
  int32_t LX, LY, channels, reserved;
  float* params;
} nh_surrogate;
int32_t nh_sur_nparams(int32_t channels);
/* f(s) for s int8 [batch][2N] -> out float [batch]. */
int nh_sur_forward(void* stream, const nh_surrogate* sur, int32_t batch, const int8_t* s, float* out);
/* Gradient of sum_b 0.5 * weight * (f(s_b) - target_b)^2 w.r.t. params -> grad [nparams];
 * loss [1] (may be NULL).  Per-configuration gradients are summed in a fixed order. */
size_t nh_sur_grad_workspace_size(const nh_surrogate* sur, int32_t batch);
int nh_sur_grad(void* stream, const nh_surrogate* sur, int32_t batch, const int8_t* s, const float* target,
                float weight, float* grad, float* loss, void* workspace, size_t workspace_bytes);
/* Adam step (Kingma & Ba) on n parameters with bias correction at step >= 1. */
int nh_adam(void* stream, int32_t n, float* w, const float* g, float* m, float* v, float lr, float beta1,
            float beta2, float eps, int32_t step);

/* ---- a14 (+ a16): Metropolis sweep  (replaces quantax sampler.sweep of the MixSampler of
 * ParticleHop and SiteExchange built at code/states.py:355-408 and
 * code/train_pfaffian_16x16_tpm02.py:93-99, with n_neighbor 1 or (1, 2) per state,
 * code/states.py:95-96,105-119; the delayed-acceptance stage is new).
 * n_steps proposals per chain; step t of chain c uses Philox4x32-10 at counter
 * (c, step0 + t, 0) and key seed: r0, r1 draw the move, r2, r3 two 24-bit uniforms.
 * Proposal kinds (nh_proposal, NULL = NH_MOVE_EXCHANGE):
 *   NH_MOVE_EXCHANGE       site i = r0 mod N, direction r1 & 3 of (+x, -x, +y, -y), spin
 *                          r1 >> 2 & 1: the occupations of modes (spin, i) and (spin, j) trade
 *                          places (a spin-preserving nearest-neighbour hop);
 *   NH_MOVE_PARTICLE_HOP   particle r0 mod (n_up + n_dn) (occupied modes in ascending order)
 *                          to neighbour r1 & (z - 1) of (+x, -x, +y, -y, +x+y, +x-y, -x+y,
 *                          -x-y), z = ph_neighbors (4: nearest, 8: and next-nearest); no move
 *                          when that mode is occupied;
 *   NH_MOVE_SITE_EXCHANGE  site i = r0 mod N and neighbour r1 & (z - 1), z = se_neighbors: the
 *                          two sites' occupations trade places in both spin blocks;
 *   NH_MOVE_MIX            particle hop or site exchange by r1 >> 4 & 1 (probability 1/2
 *                          each, per chain and step).
 * Every kind is a symmetric proposal.  sur == NULL: accept with min(1, |psi'/psi|^2).
 * sur != NULL: delayed acceptance (stage 1 with the surrogate, stage 2 exact), f_cur
 * [n_chains] the surrogate value of the current configurations.
 * s / sign / logabs (/ f_cur) are updated in place; counters int32 [4][n_chains] are
 * incremented: changing proposals, stage-1 survivors, accepted, exact evaluations. */
enum { NH_MOVE_EXCHANGE = 0, NH_MOVE_PARTICLE_HOP = 1, NH_MOVE_SITE_EXCHANGE = 2, NH_MOVE_MIX = 3 };
typedef struct nh_proposal {
  int32_t kind;         /* NH_MOVE_* */
  int32_t ph_neighbors; /* partners of a particle hop: 4 or 8 (unused unless hops are drawn) */
  int32_t se_neighbors; /* partners of a site exchange: 4 or 8 (unused unless exchanges are drawn) */
} nh_proposal;
size_t nh_metropolis_workspace_size(const nh_model* model, int32_t n_chains);
int nh_metropolis_sweep(void* stream, const nh_model* model, const nh_surrogate* sur, const nh_proposal* proposal,
                        int32_t n_chains, int8_t* s, float* sign, double* logabs, float* f_cur, uint64_t seed,
                        int64_t step0, int32_t n_steps, int32_t* counters, void* workspace, size_t workspace_bytes);

/* ---- measurement hook (no reference counterpart; used by bench.py): while enabled, every
 * launch of a kernel class is bracketed by HIP events recorded on the launch stream.
 * nh_profile_query waits for the recorded events and returns the launch count and the summed
 * kernel time of one class; nh_profile_reset drops the records. */
enum { NH_PROF_FEATURES = 0, NH_PROF_PAIRING_LINALG = 1, NH_PROF_OTHER = 2, NH_PROF_CLASSES = 3 };
int nh_profile_enable(int32_t on);
int nh_profile_query(int32_t kernel_class, int64_t* launches, double* total_ms);
int nh_profile_reset(void);

#ifdef __cplusplus
}
#endif
#endif /* NH_H */
