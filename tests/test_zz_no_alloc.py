"""nh_log_psi_noalloc (the allocation-free projected amplitude, workspace in a static device array)
against nh_log_psi with a caller workspace, through the C ABI (marked gpu): two calls in a row on
the 16x16 Pfaffian workload (224 x 224 triangles in the scratch), the first larger than one pass
of the scratch so the entry point loops, the second smaller; results bit-identical (the same
kernels evaluate every image independently of how the batch is split)."""

import ctypes

import numpy as np
import pytest

from neuralhubbard_amd.configs import WORKLOADS, make_configurations, make_params

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible MI355X")


@pytest.mark.parametrize("precision", ["fp32", "mixed"])
def test_log_psi_noalloc_matches_workspace_form(precision):
    from neuralhubbard_amd import _lib
    from neuralhubbard_amd.backflow import Variational, build_backflow

    lib = _lib.load()
    wk = WORKLOADS["pf16"]
    st = Variational(build_backflow(wk, make_params(wk), precision))
    m = st.model.struct(True)
    per_pass = None
    for batch, seed in ((600, 31), (45, 32)):
        s = torch.tensor(make_configurations(wk, batch, seed=seed), device="cuda")
        ref = st(s)
        sign = torch.full((batch,), 7.0, dtype=torch.float32, device="cuda")
        logabs = torch.full((batch,), 7.0, dtype=torch.float64, device="cuda")
        rc = lib.nh_log_psi_noalloc(_lib.stream_handle(), ctypes.byref(m), batch, _lib.ptr(s), _lib.ptr(sign),
                                    _lib.ptr(logabs))
        _lib.check(rc, "nh_log_psi_noalloc")
        torch.cuda.synchronize()
        if per_pass is None:  # the first call must take more than one pass of the scratch
            cap = 256 << 20
            per_pass = max(k for k in range(1, batch + 1) if lib.nh_log_psi_workspace_size(ctypes.byref(m), k) <= cap)
            print(f"{precision}: {per_pass} configurations per pass of the scratch", flush=True)
            assert per_pass < batch
        assert torch.equal(sign, ref.sign), (batch, precision)
        assert torch.equal(logabs, ref.logabs), (batch, precision)
