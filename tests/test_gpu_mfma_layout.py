"""Lane layouts of the gfx950 MFMA builtins and the DPP row rotation used by the MFMA feature
network, checked against numpy products on the CPU (test-only probe library)."""

import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
HERE = os.path.dirname(os.path.abspath(__file__))


def _run(which, A, B, nout):
    lib = ctypes.CDLL(os.path.join(HERE, "probe", "libnhprobe.so"))
    lib.probe_run.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    a = torch.tensor(A, dtype=torch.float32, device="cuda")
    b = torch.tensor(B, dtype=torch.float32, device="cuda")
    d = torch.zeros(nout, dtype=torch.float32, device="cuda")
    assert lib.probe_run(which, a.data_ptr(), b.data_ptr(), d.data_ptr()) == 0
    return d.cpu().numpy()


def test_mfma_f32_16x16x4_layout():
    rng = np.random.default_rng(0)
    A = rng.standard_normal((16, 4)).astype(np.float32)
    B = rng.standard_normal((4, 16)).astype(np.float32)
    D = _run(0, A, B, 256).reshape(16, 16)
    np.testing.assert_allclose(D, A.astype(np.float64) @ B, rtol=1e-5, atol=1e-5)


def test_mfma_bf16_16x16x32_layout():
    rng = np.random.default_rng(1)
    A = rng.integers(-4, 5, (16, 32)).astype(np.float32)  # exact in bf16
    B = rng.integers(-4, 5, (32, 16)).astype(np.float32)
    D = _run(1, A, B, 256).reshape(16, 16)
    np.testing.assert_array_equal(D, A @ B)


def test_dpp_row_rotation_direction():
    x = np.arange(64, dtype=np.float32)
    out = _run(2, x, x, 128)
    lane = np.arange(64)
    # row_ror:1 -> lane l receives lane (l - 1) mod 16 of its row; row_ror:15 -> (l + 1) mod 16
    np.testing.assert_array_equal(out[:64], (lane // 16) * 16 + (lane - 1) % 16)
    np.testing.assert_array_equal(out[64:], (lane // 16) * 16 + (lane + 1) % 16)
