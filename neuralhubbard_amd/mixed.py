"""Fragment-ordered conv weights for the MFMA feature networks (C == 32 for the conv stack at
LY == 16 or 32, C == 64 for the residual CNN, `pack_resconv_bf16`).  The layout does not depend
on the lattice.

FP32 (v_mfma_f32_16x16x4_f32): for layer l, tap, output half hf, K-step kk and lane:
    W[tap][ci = 4 kk + lane // 16][co = 16 hf + lane % 16]
FP16 (v_mfma_f32_16x16x32_f16): for layer l, tap, half hf, lane, j < 8:
    LP_SCALE * W[tap][ci = 8 (lane // 16) + j][co = 16 hf + lane % 16]
  stored as split FP16: the hi parts of all layers, then the lo parts, LP_SCALE * W - hi
  rounded.  The power-of-two pre-scale keeps the lo parts of the small conv weights FP16
  normals; the kernel divides it out of each layer's sum exactly.
Layer 0 has 2 input channels; the missing ones are zero.  Tap = (dx + 1) * 3 + (dy + 1).
"""

from __future__ import annotations

import numpy as np

C = 32
LP_SCALE = 64.0  # kLpScale of nh_cnn_mfma.hip


def _tap_major(params):
    """[layer][tap][ci (padded to C)][co] float32."""
    out = []
    for w in params.conv_w:
        w = np.asarray(w, np.float32)  # (co, ci, 3, 3)
        co, ci = w.shape[:2]
        t = np.zeros((9, C, C), np.float32)
        t[:, :ci, :co] = np.transpose(w, (2, 3, 1, 0)).reshape(9, ci, co)
        out.append(t)
    return np.stack(out)


def pack_mfma_f32(params) -> np.ndarray:
    W = _tap_major(params)  # (L, 9, ci, co)
    lane = np.arange(64)
    kk = np.arange(8)
    ci = 4 * kk[:, None] + lane[None, :] // 16  # (8, 64)
    out = np.empty((W.shape[0], 9, 2, 8, 64), np.float32)
    for hf in range(2):
        co = 16 * hf + lane % 16  # (64,)
        out[:, :, hf] = W[:, :, ci, co[None, :]]
    return out.reshape(-1)


def bf16_bits(x: np.ndarray) -> np.ndarray:
    """float32 -> bf16 bit patterns, round to nearest even."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def pack_f16(params) -> np.ndarray:
    """Split-FP16 fragments of LP_SCALE * W: hi parts of every layer, then lo parts, as uint16
    bit patterns (LP_SCALE * W = hi + lo + O(2^-22 |W|))."""
    W = _tap_major(params)
    lane = np.arange(64)
    j = np.arange(8)
    ci = 8 * (lane[:, None] // 16) + j[None, :]  # (64, 8)
    out = np.empty((W.shape[0], 9, 2, 64, 8), np.float32)
    for hf in range(2):
        co = (16 * hf + lane % 16)[:, None]  # (64, 1)
        out[:, :, hf] = W[:, :, ci, co]
    x = out.reshape(-1) * np.float32(LP_SCALE)
    hi = x.astype(np.float16)
    lo = (x - hi.astype(np.float32)).astype(np.float16)
    return np.concatenate([hi.view(np.uint16), lo.view(np.uint16)])


def pack_resconv_bf16(params) -> np.ndarray:
    """Split-BF16 weights of the residual CNN's convolutions (C = 64) as A fragments of
    v_mfma_f32_16x16x32_bf16, in slabs: for conv, K-half kh, tap row dyi, plane (hi, lo),
    tap column dxi, quarter qq, lane, j < 8:
        W[tap = 3 dxi + dyi][ci = 32 kh + 8 (lane // 16) + j][co = 16 qq + lane % 16]
    (plane lo = W - hi rounded).  One slab (kh, dyi) is 2 x 3 x 4 x 64 fragments of 16 bytes."""
    W = np.stack([np.asarray(w, np.float32) for w in params.conv_w])  # (n, co, ci, 3, 3)
    n, Cc = W.shape[0], W.shape[1]
    T = np.transpose(W, (0, 3, 4, 2, 1)).reshape(n, 9, Cc, Cc)  # [conv][tap][ci][co]
    kh = np.arange(2).reshape(2, 1, 1, 1, 1, 1)
    dyi = np.arange(3).reshape(1, 3, 1, 1, 1, 1)
    dxi = np.arange(3).reshape(1, 1, 3, 1, 1, 1)
    qq = np.arange(4).reshape(1, 1, 1, 4, 1, 1)
    lane = np.arange(64).reshape(1, 1, 1, 1, 64, 1)
    j = np.arange(8).reshape(1, 1, 1, 1, 1, 8)
    shape = (2, 3, 3, 4, 64, 8)
    tap = np.broadcast_to(3 * dxi + dyi, shape)
    ci = np.broadcast_to(32 * kh + 8 * (lane // 16) + j, shape)
    co = np.broadcast_to(16 * qq + lane % 16, shape)
    frag = T[:, tap, ci, co]  # (n, kh, dyi, dxi, qq, lane, j)
    hi = bf16_bits(frag.reshape(-1)).reshape(frag.shape)
    lo = bf16_bits((frag.reshape(-1) - (hi.reshape(-1).astype(np.uint32) << 16).view(np.float32))).reshape(frag.shape)
    return np.stack([hi, lo], axis=3).reshape(-1)  # (n, kh, dyi, plane, dxi, qq, lane, j)


def pack_resconv_f32(params) -> np.ndarray:
    """FP32 weights of the residual CNN's convolutions as B fragments of v_mfma_f32_16x16x4_f32, in
    the same slabs as `pack_resconv_bf16`: for conv, K-half kh, tap row dyi, tap column dxi,
    quarter qq, lane and K-step kk < 8:
        W[tap = 3 dxi + dyi][ci = 32 kh + 4 kk + lane // 16][co = 16 qq + lane % 16]."""
    W = np.stack([np.asarray(w, np.float32) for w in params.conv_w])  # (n, co, ci, 3, 3)
    n, Cc = W.shape[0], W.shape[1]
    T = np.transpose(W, (0, 3, 4, 2, 1)).reshape(n, 9, Cc, Cc)  # [conv][tap][ci][co]
    kh = np.arange(2).reshape(2, 1, 1, 1, 1, 1)
    dyi = np.arange(3).reshape(1, 3, 1, 1, 1, 1)
    dxi = np.arange(3).reshape(1, 1, 3, 1, 1, 1)
    qq = np.arange(4).reshape(1, 1, 1, 4, 1, 1)
    lane = np.arange(64).reshape(1, 1, 1, 1, 64, 1)
    kk = np.arange(8).reshape(1, 1, 1, 1, 1, 8)
    shape = (2, 3, 3, 4, 64, 8)
    tap = np.broadcast_to(3 * dxi + dyi, shape)
    ci = np.broadcast_to(32 * kh + 4 * kk + lane // 16, shape)
    co = np.broadcast_to(16 * qq + lane % 16, shape)
    return np.ascontiguousarray(T[:, tap, ci, co]).reshape(-1)  # (n, kh, dyi, dxi, qq, lane, kk)


def pack_lowp(model):
    import torch

    if model.workload.network == "resconv":
        return torch.as_tensor(pack_resconv_bf16(model.params).view(np.int16), device="cuda")
    return torch.as_tensor(pack_f16(model.params).view(np.int16), device="cuda")
