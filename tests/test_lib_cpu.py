"""CPU-side checks of the C ABI library: it loads, and it exports every entry point that
include/nh.h declares (no compute call is made without a GPU)."""

import ctypes
import os
import re

import numpy as np
import pytest

from neuralhubbard_amd import _lib
from neuralhubbard_amd import backflow as bf
from neuralhubbard_amd.configs import WORKLOADS, make_params, small_workload
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "nh.h")).read()
    return sorted(set(re.findall(r"\b(nh_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libnh.so is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 10
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(_lib.SIGNATURES), "ctypes signatures out of sync with include/nh.h"


def test_workspace_queries_need_no_gpu():
    lib = _lib.load()
    # a model description with dummy (non-null) pointers: size queries never dereference them
    m = _lib.NhModel(0, 16, 16, 112, 112, 32, 6, 0, 1, 0, 16, 16, 16, 0, 0)
    ws = lib.nh_log_psi_workspace_size(ctypes.byref(m), 8192)
    assert ws > 8192 * 8 * 12
    assert lib.nh_local_energy_workspace_size(ctypes.byref(m), -0.2, 4) > 0
    bad = _lib.NhModel(0, 2, 16, 112, 112, 32, 6, 0, 1, 0, 16, 16, 16, 0, 0)
    assert lib.nh_log_psi_workspace_size(ctypes.byref(bad), 8) == 0


def test_residual_cnn_model_checks():
    lib = _lib.load()
    R = _lib.NH_NET_RESCONV
    ok = _lib.NhModel(0, 16, 16, 112, 112, 64, 16, 0, 1, R, 16, 16, 16, 0, 0)
    assert lib.nh_log_psi_workspace_size(ctypes.byref(ok), 64) > 64 * 8 * 2 * 256 * 32 * 4
    # other lattices run layer by layer: the workspace then also holds two activation buffers
    for other in (_lib.NhModel(0, 24, 24, 252, 252, 64, 16, 0, 1, R, 16, 16, 16, 0, 0),    # LY != 16
                  _lib.NhModel(0, 20, 16, 140, 140, 64, 16, 0, 1, R, 16, 16, 16, 0, 0)):   # LX > 16
        N = other.LX * other.LY
        assert lib.nh_log_psi_workspace_size(ctypes.byref(other), 64) > 64 * 8 * 3 * N * 64 * 4
        assert lib.nh_features_workspace_size(ctypes.byref(other), 64) >= 64 * (2 * N * 64 + 1) * 4
    assert lib.nh_features_workspace_size(ctypes.byref(ok), 64) == 256  # tile kernels: none needed
    # the BCS head off the tiles keeps the residual stream in the workspace too
    bcs32 = _lib.NhModel(1, 32, 32, 448, 448, 64, 16, 0, 1, _lib.NH_NET_RESCONV_HEAD, 16, 16, 16, 0, 0)
    assert lib.nh_features_workspace_size(ctypes.byref(bcs32), 8) >= 8 * (3 * 1024 * 64 + 1) * 4
    for bad in (_lib.NhModel(0, 16, 16, 112, 112, 32, 16, 0, 1, R, 16, 16, 16, 0, 0),   # C != 64
                _lib.NhModel(0, 40, 40, 700, 700, 64, 16, 0, 1, R, 16, 16, 16, 0, 0),   # N > 1024
                _lib.NhModel(0, 16, 16, 112, 112, 64, 0, 0, 1, R, 16, 16, 16, 0, 0),    # no blocks
                _lib.NhModel(0, 16, 16, 112, 112, 64, 16, 1, 1, R, 16, 16, 16, 0, 0),    # mixed, no BF16 copy
                _lib.NhModel(0, 16, 16, 112, 112, 64, 16, 0, 1, 7, 16, 16, 16, 0, 0)):   # network
        assert lib.nh_log_psi_workspace_size(ctypes.byref(bad), 64) == 0
    mixed = _lib.NhModel(0, 16, 16, 112, 112, 64, 16, 1, 1, R, 16, 16, 16, 0, 16)
    assert lib.nh_log_psi_workspace_size(ctypes.byref(mixed), 64) > 0


def test_pack_resconv_layout():
    wk = small_workload("pfaffian", LX=4, LY=16, n_up=12, n_dn=12, channels=64, n_layers=2, network="resconv")
    p = make_params(wk)
    w, b = bf.pack_conv(p)
    C = 64
    assert w.size == 4 * C + 4 * 9 * C * C and b.size == 4 * C
    assert w[2 * C + 7] == p.emb[7, 2]  # [class][c]
    # conv 3 (second conv of block 1), tap (dx, dy) = (0, +1), c_in 9, c_out 40
    assert w[4 * C + 3 * 9 * C * C + ((1 * 3 + 2) * C + 9) * C + 40] == p.conv_w[3][40, 9, 1, 2]
    assert b[3 * C + 40] == p.conv_b[3][40]


def test_pack_f16_layout():
    """Split-FP16 fragments of the conv stack: hi + lo = 2^6 W to FP16's 2^-22 relative (plus half
    the FP16 subnormal spacing 2^-25 of the lo part, 2^-31 in W), in the
    v_mfma_f32_16x16x32_f16 A layout [layer][tap][half][lane][j] -> W[tap][8 (lane // 16) + j][16 half + lane % 16]."""
    from neuralhubbard_amd.mixed import LP_SCALE, pack_f16

    wk = small_workload("pfaffian", LX=4, LY=16, n_up=12, n_dn=12, channels=32, n_layers=2)
    p = make_params(wk)
    packed = pack_f16(p)
    L = wk.n_layers + 1
    assert packed.size == 2 * L * 9 * 2 * 64 * 8
    f = packed.view(np.float16).astype(np.float64).reshape(2, L, 9, 2, 64, 8)
    rec = (f[0] + f[1]) / LP_SCALE
    lane = np.arange(64)
    for l in range(L):
        w = np.asarray(p.conv_w[l], np.float64)  # (co, ci, 3, 3)
        for tap in (0, 4, 8):
            for hf in range(2):
                ci = 8 * (lane // 16)[:, None] + np.arange(8)[None, :]
                co = (16 * hf + lane % 16)[:, None]
                want = np.zeros((64, 8))
                ok = ci < w.shape[1]
                want[ok] = w[np.broadcast_to(co, ci.shape)[ok], ci[ok], tap // 3, tap % 3]
                assert np.all(np.abs(rec[l, tap, hf] - want) <= 2.0**-21 * np.abs(want) + 2.0**-31)


def test_pack_resconv_bf16_layout():
    from neuralhubbard_amd.mixed import pack_resconv_bf16

    wk = small_workload("pfaffian", LX=4, LY=16, n_up=12, n_dn=12, channels=64, n_layers=1, network="resconv")
    p = make_params(wk)
    packed = pack_resconv_bf16(p)
    assert packed.size == 2 * 6 * 2 * 3 * 4 * 64 * 8  # two convs of six slabs, hi and lo planes
    f = (packed.astype(np.uint32) << 16).view(np.float32).reshape(2, 2, 3, 2, 3, 4, 64, 8)
    rec = f[:, :, :, 0] + f[:, :, :, 1]  # hi + lo
    # conv 1, kh 1, tap row dyi 0 (dy = -1), tap column dxi 2 (dx = +1), quarter 2, lane 37, j 5:
    # ci = 32 + 8 * 2 + 5 = 53, co = 32 + 5 = 37
    w = p.conv_w[1][37, 53, 2, 0]
    assert abs(rec[1, 1, 0, 2, 2, 37, 5] - w) <= 2.0**-16 * abs(w)
    full = np.stack([np.transpose(c, (2, 3, 1, 0)).reshape(9, 64, 64) for c in p.conv_w])  # [conv][tap][ci][co]
    lane = np.arange(64)
    for kh, dyi, dxi, qq in ((0, 1, 1, 0), (1, 2, 0, 3)):
        want = full[:, 3 * dxi + dyi][:, 32 * kh + 8 * (lane // 16)[:, None] + np.arange(8)[None, :],
                                      (16 * qq + lane % 16)[:, None]]
        assert np.all(np.abs(rec[:, kh, dyi, dxi, qq] - want) <= 2.0**-16 * np.abs(want))


def test_pack_resconv_f32_layout():
    from neuralhubbard_amd.mixed import pack_resconv_f32

    wk = small_workload("pfaffian", LX=4, LY=16, n_up=12, n_dn=12, channels=64, n_layers=1, network="resconv")
    p = make_params(wk)
    f = pack_resconv_f32(p).reshape(2, 2, 3, 3, 4, 64, 8)
    # conv 1, kh 1, dyi 2 (dy = +1), dxi 0 (dx = -1), quarter 3, lane 50, kk 6:
    # ci = 32 + 24 + 50 // 16 = 59, co = 48 + 2 = 50
    assert f[1, 1, 2, 0, 3, 50, 6] == p.conv_w[1][50, 59, 0, 2]
    full = np.stack([np.transpose(c, (2, 3, 1, 0)).reshape(9, 64, 64) for c in p.conv_w])
    lane = np.arange(64)
    for kh, dyi, dxi, qq in ((0, 0, 0, 0), (1, 1, 2, 1)):
        want = full[:, 3 * dxi + dyi][:, 32 * kh + 4 * np.arange(8)[None, :] + (lane // 16)[:, None],
                                      (16 * qq + lane % 16)[:, None]]
        np.testing.assert_array_equal(f[:, kh, dyi, dxi, qq], want)


def test_argument_errors_without_gpu():
    lib = _lib.load()
    m = _lib.NhModel(0, 16, 16, 112, 112, 32, 6, 0, 1, 0, 0, 0, 0, 0, 0)  # null weights
    assert lib.nh_log_psi(None, ctypes.byref(m), 4, None, None, None, None, 0) == _lib.NH_ERR_ARG
    assert lib.nh_slogpf(None, 4, 7, None, None, None) == _lib.NH_ERR_ARG
    assert lib.nh_slogpf(None, 1, 300, ctypes.c_void_p(16), ctypes.c_void_p(16), ctypes.c_void_p(16)) == _lib.NH_ERR_SHAPE
    assert lib.nh_slogdet(None, 1, 500, ctypes.c_void_p(16), ctypes.c_void_p(16), ctypes.c_void_p(16)) == _lib.NH_ERR_SHAPE
    v = ctypes.c_void_p(16)
    assert lib.nh_slogpf_ws(None, 1, 1026, v, v, v, v, 1 << 30) == _lib.NH_ERR_SHAPE
    assert lib.nh_slogpf_ws(None, 1, 895, v, v, v, v, 1 << 30) == _lib.NH_ERR_SHAPE
    assert lib.nh_slogpf_ws(None, 2, 896, v, v, v, v, lib.nh_slogpf_workspace_size(1, 896) - 1024) == _lib.NH_ERR_WORKSPACE
    assert lib.nh_slogpf_workspace_size(1, 896) == 4 * ((896 * 895 // 2 + 64 + 63) // 64 * 64) + 256
    # 100 <= n <= 256: the column-major square (leading dimension n rounded up to 16) and the
    # first compacted complement (n - 32 rows); n = 98: the packed triangle
    assert lib.nh_slogpf_workspace_size(3, 224) == 4 * 3 * (224 * 224 + 192 * 192) + 256
    assert lib.nh_slogpf_workspace_size(1, 146) == 4 * (160 * 160 + 128 * 128) + 256
    assert lib.nh_slogpf_workspace_size(1, 100) == 4 * (112 * 112 + 80 * 80) + 256
    assert lib.nh_slogpf_workspace_size(1, 98) == 4 * ((98 * 97 // 2 + 64 + 63) // 64 * 64) + 256
    # 256 < n <= 352: 512 threads, 12-step panels (first complement n - 24 rows); above: the
    # packed triangle
    assert lib.nh_slogpf_workspace_size(1, 258) == 4 * (272 * 272 + 240 * 240) + 256
    assert lib.nh_slogpf_workspace_size(1, 352) == 4 * (352 * 352 + 336 * 336) + 256
    assert lib.nh_slogpf_workspace_size(1, 354) == 4 * ((354 * 353 // 2 + 64 + 63) // 64 * 64) + 256
    # the layer-by-layer network needs its workspace: too small a one is refused before any launch
    R = _lib.NH_NET_RESCONV
    big = _lib.NhModel(0, 24, 24, 252, 252, 64, 16, 0, 1, R, 16, 16, 16, 0, 0)
    p = ctypes.c_void_p(256)
    need = lib.nh_features_workspace_size(ctypes.byref(big), 2)
    assert need >= 2 * (2 * 576 * 64 + 1) * 4
    assert lib.nh_features_ws(None, ctypes.byref(big), 2, p, p, p, need // 2) == _lib.NH_ERR_WORKSPACE
    assert lib.nh_features_ws(None, ctypes.byref(big), 2, p, p, None, need) == _lib.NH_ERR_WORKSPACE


def test_host_tables_match_oracle():
    rng = np.random.default_rng(0)
    for LX, LY in [(8, 8), (16, 16), (5, 3)]:
        N, d = LX * LY, 4
        W = rng.standard_normal((3, N, d))
        np.testing.assert_allclose(bf.build_wtilde(W, 0.55, LX, LY), O.build_wtilde(W, 0.55, LX, LY), rtol=2e-7,
                                   atol=1e-12)
        W2 = rng.standard_normal((N, d))
        np.testing.assert_allclose(bf.build_wt_ud(W2, 0.55, LX, LY), O.build_wt_ud(W2, 0.55, LX, LY), rtol=2e-7,
                                   atol=1e-12)


def test_pack_conv_layout():
    wk = WORKLOADS["pf8"]
    p = make_params(wk)
    w, b = bf.pack_conv(p)
    C = wk.channels
    assert w.size == 9 * 2 * C + wk.n_layers * 9 * C * C and b.size == (wk.n_layers + 1) * C
    # layer 1, tap (dx, dy) = (+1, -1), c_in 3, c_out 5
    off = 9 * 2 * C + ((2 * 3 + 0) * C + 3) * C + 5
    assert w[off] == p.conv_w[1][5, 3, 2, 0]


def test_gpu_entry_points_refuse_without_gpu():
    import torch

    if torch.cuda.is_available():
        pytest.skip("GPU present")
    with pytest.raises(_lib.NhError):
        bf.slogpf(torch.zeros((1, 4, 4)))


def _jac_model(wk, p):
    kind = _lib.NH_KIND_PFAFFIAN if wk.kind == "pfaffian" else _lib.NH_KIND_BCS
    if wk.network == "conv":
        net = _lib.NH_NET_CONV
    else:
        net = _lib.NH_NET_RESCONV_HEAD if p.head is not None else _lib.NH_NET_RESCONV
    base = 0 if (p.F0 is None and p.phi0 is None) else 16
    return _lib.NhModel(kind, wk.LX, wk.LY, wk.n_up, wk.n_dn, wk.channels, wk.n_layers, 0, 1, net, 16, 16, 16, base,
                        0, 0)


@pytest.mark.parametrize("name", ["pf_base", "bcs", "pf_res", "bcs_head", "pf16", "tpm02_L16"])
def test_jacobian_columns_match_param_vector(name):
    """nh_log_psi_jacobian_size counts exactly the entries of `param_vector` (the column order the
    Jacobian documents) for every network and kind."""
    wk = {
        "pf_base": small_workload(kind="pfaffian", n_up=3, n_dn=3, channels=8, n_layers=2, base_f0=True),
        "bcs": small_workload(kind="bcs", n_up=3, n_dn=3, channels=8, n_layers=2),
        "pf_res": small_workload(kind="pfaffian", n_up=3, n_dn=3, channels=64, n_layers=2, network="resconv"),
        "bcs_head": small_workload(kind="bcs", n_up=3, n_dn=3, channels=64, n_layers=2, network="resconv",
                                   features=64),
    }.get(name) or WORKLOADS[name]
    p = make_params(wk)
    lib = _lib.load()
    m = _jac_model(wk, p)
    assert lib.nh_log_psi_jacobian_size(ctypes.byref(m)) == bf.param_vector(p).size
    assert lib.nh_log_psi_jacobian_workspace_size(ctypes.byref(m), 2) > lib.nh_log_psi_jacobian_workspace_size(
        ctypes.byref(m), 1)


def test_jacobian_argument_errors_without_gpu():
    """Refusals happen before any launch: bad batch, null buffers, too small a workspace, more
    than 65535 images."""
    lib = _lib.load()
    wk = small_workload(kind="pfaffian", n_up=3, n_dn=3, channels=8, n_layers=2)
    m = _jac_model(wk, make_params(wk))
    v = ctypes.c_void_p(256)
    need = lib.nh_log_psi_jacobian_workspace_size(ctypes.byref(m), 4)
    assert lib.nh_log_psi_jacobian(None, ctypes.byref(m), 0.5, -1, v, v, v, v, v, need) == _lib.NH_ERR_ARG
    assert lib.nh_log_psi_jacobian(None, ctypes.byref(m), 0.5, 4, v, v, v, None, v, need) == _lib.NH_ERR_ARG
    assert lib.nh_log_psi_jacobian(None, ctypes.byref(m), float("nan"), 4, v, v, v, v, v, need) == _lib.NH_ERR_ARG
    assert lib.nh_log_psi_jacobian(None, ctypes.byref(m), 0.5, 4, v, v, v, v, v, need - 1) == _lib.NH_ERR_WORKSPACE
    assert lib.nh_log_psi_jacobian(None, ctypes.byref(m), 0.5, 8192, v, v, v, v, v, 1 << 40) == _lib.NH_ERR_SHAPE
    m.channels = 12  # not a network this library builds
    assert lib.nh_log_psi_jacobian_size(ctypes.byref(m)) == 0
