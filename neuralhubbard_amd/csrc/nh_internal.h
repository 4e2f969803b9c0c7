// Internal launchers shared by the translation units of libnh.
#pragma once

#include "nh_common.h"

namespace nh {

int check_model(const nh_model* m);

// features per spin: C / 2 (the two halves of the channels), or C from the linear head
inline int feature_dim(const nh_model* m) { return m->network == NH_NET_RESCONV_HEAD ? m->channels : m->channels / 2; }

// slots (threads) of one feature-network block for an N-site lattice
int cnn_slots(int N);

// Features of images [img0, img_end) of an item list (ipi images per item) into
// Y[(img - img0)][2N][d].
int launch_cnn(hipStream_t st, const nh_model* m, const int8_t* s, const int32_t* item_b,
               const int32_t* item_hop, int img0, int img_end, int ipi, float* Y,
               float* scratch = nullptr);
int launch_cnn_mfma(hipStream_t st, const nh_model* m, bool bf16, const int8_t* s, const int32_t* item_b,
                    const int32_t* item_hop, int img0, int img_end, int ipi, float* Y);
bool cnn_mfma_supported(const nh_model* m);
// the residual CNN of the released states (network == NH_NET_RESCONV)
int launch_resconv(hipStream_t st, const nh_model* m, const int8_t* s, const int32_t* item_b,
                   const int32_t* item_hop, int img0, int img_end, int ipi, float* Y, float* scratch);
bool resconv_supported(const nh_model* m);
// floats of workspace the residual CNN needs for `images` images (lattices outside the tile
// kernels run layer by layer with two activation buffers; 0 otherwise)
size_t resconv_scratch_floats(const nh_model* m, int64_t images);

// Pairing block + Pfaffian (or determinant) of images [img0, img_end), writing the signed
// log of eta_g(s) psi(g.s) per image into img_sign / img_logabs (indexed by global image).
// Pfaffians above n = 64 keep their triangle, BCS determinants above n = 192 (or with d = 64
// features) their matrix, in gtri, image_global_floats(m) floats per image of the chunk (NULL
// otherwise).
int launch_image_amplitude(hipStream_t st, const nh_model* m, const int8_t* s,
                           const int32_t* item_b, const int32_t* item_hop, int img0, int img_end,
                           int ipi, const float* Y, float* img_sign, double* img_logabs, float* gtri);
size_t image_global_floats(const nh_model* m);
// images whose features (and global matrices) are resident at once
int64_t image_chunk(const nh_model* m);

// The projected amplitudes of `batch` configurations with every image's signed log kept: the
// nh_log_psi pass on a 256-byte aligned buffer of image_amplitudes_bytes(m, batch) bytes; the
// image values (indexed img = b * ipi + g) point into it.
size_t image_amplitudes_bytes(const nh_model* m, int64_t batch);
int image_amplitudes(hipStream_t st, const nh_model* m, const int8_t* s, int64_t batch, void* ws, float** img_sign,
                     double** img_logabs, float* sign, double* logabs);

// G[b] = factor * F[b]^{-T} for batch n x n matrices (blocked Gauss-Jordan; NaN if singular);
// factor -1 gives F^{-1} for skew F.
int launch_inverse_t(hipStream_t st, int batch, int n, const float* F, float factor, float* G);

// The hop list of samples [b0, b0 + batch): max_hops slots per sample, valid hops first
// (count[lb] of them), coef = -t_ij (-1)^(occupied modes strictly between the two modes).
int launch_hops(hipStream_t st, int LX, int LY, float t, float tp, int z, int batch, int b0, const int8_t* s,
                int max_hops, int32_t* count, int32_t* item_b, int32_t* item_hop, float* coef, int32_t* out_from,
                int32_t* out_to);
// E_loc from the hop items' amplitudes (hop_sign / hop_logabs per slot).
int launch_eloc_reduce(hipStream_t st, int N, double U, int batch, int b0, const int8_t* s, const float* sign_s,
                       const double* logabs_s, int max_hops, const int32_t* count, const float* coef,
                       const float* hop_sign, const double* hop_logabs, double* eloc);

// Signed log-sum over the ipi images of each item in [item0, item_end).
int launch_combine(hipStream_t st, const int32_t* item_hop, int item0, int item_end, int ipi,
                   const float* img_sign, const double* img_logabs, float* sign, double* logabs);

}  // namespace nh
