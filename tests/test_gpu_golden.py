"""The HIP path against the golden fixtures of tests/golden (marked gpu).

Every stage is checked on its own fixture inputs: the feature network, the pairing block (from
the fixture's features), the Pfaffian / determinant (of the fixture's matrix), then the bare and
projected amplitudes and E_loc end to end.  Tolerances as in test_gpu_parity: log|psi| 1e-4
(scaled by the projection's cancellation e^c), signs exact, E_loc 1e-5 relative + 1e-4."""

import importlib.util
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests._util import cancellation, log_tol

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

HERE = os.path.dirname(os.path.abspath(__file__))
spec = importlib.util.spec_from_file_location("make_golden", os.path.join(HERE, "golden", "make_golden.py"))
G = importlib.util.module_from_spec(spec)
spec.loader.exec_module(G)


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible MI355X")


def load(name):
    return dict(np.load(os.path.join(HERE, "golden", f"{name}.npz"), allow_pickle=False))


@pytest.mark.parametrize("name", list(G.CASES))
def test_stages_match_golden(name):
    from neuralhubbard_amd.backflow import build_backflow, pairing_occ, pairing_ud, slogdet, slogpf
    from neuralhubbard_amd.configs import make_params

    wk = G.workload(name)
    f = load(name)
    p = make_params(wk)
    dev = build_backflow(wk, p)
    S = torch.tensor(f["s"], device="cuda")
    B, N = len(f["s"]), wk.N
    y = dev.features(S).cpu().numpy()
    scale = np.abs(f["features"]).max()
    np.testing.assert_allclose(y, f["features"], rtol=1e-4, atol=1e-5 * scale)
    Y = f["features"].reshape(B, 2 * N, -1)
    if wk.kind == "pfaffian":
        modes = f["modes"]
        Yo = np.stack([Y[b][modes[b]] for b in range(B)]).astype(np.float32)
        F = pairing_occ(torch.tensor(Yo, device="cuda"), torch.tensor(modes, device="cuda"), dev.Wt, wk.LX,
                        wk.LY).cpu().numpy().astype(np.float64)
        if p.F0 is not None:
            F = F + np.stack([p.F0[np.ix_(m, m)] for m in modes])
        sg, la = slogpf(torch.tensor(f["F"].astype(np.float32), device="cuda"))
        sg, la = sg.cpu().numpy(), la.cpu().numpy()
        ne = wk.n_up + wk.n_dn
        ref_sign = f["image_sign"] * (-1) ** (ne * (ne - 1) // 2)  # the image amplitude's extra sign
    else:
        modes = f["modes"]
        up = np.stack([m[m < N] for m in modes])
        dn = np.stack([m[m >= N] - N for m in modes])
        Yu = np.stack([Y[b][up[b]] for b in range(B)]).astype(np.float32)
        Yd = np.stack([Y[b][N + dn[b]] for b in range(B)]).astype(np.float32)
        F = pairing_ud(torch.tensor(Yu, device="cuda"), torch.tensor(Yd, device="cuda"),
                       torch.tensor(up, device="cuda"), torch.tensor(dn, device="cuda"), dev.Wt, wk.LX,
                       wk.LY).cpu().numpy().astype(np.float64)
        disp, _, _ = O.lattice_tables(wk.LX, wk.LY)
        F = F + np.stack([p.phi0[disp[np.ix_(up[b], dn[b])]] for b in range(B)])
        sg, la = slogdet(torch.tensor(f["F"].astype(np.float32), device="cuda"))
        sg, la = sg.cpu().numpy(), la.cpu().numpy()
        n, ne = wk.n_up, wk.n_up + wk.n_dn
        ref_sign = f["image_sign"] * (-1) ** (n * (n - 1) // 2) * (-1) ** (ne * (ne - 1) // 2)
    fs = np.abs(f["F"]).max()
    np.testing.assert_allclose(F, f["F"], rtol=1e-4, atol=1e-5 * fs)
    assert np.array_equal(sg, ref_sign)
    np.testing.assert_allclose(la, f["image_logabs"], rtol=0, atol=1e-4 * max(1.0, wk.N / 16))


@pytest.mark.parametrize("name", list(G.CASES))
def test_amplitudes_and_eloc_match_golden(name):
    from neuralhubbard_amd.backflow import Hubbard, Samples, Variational, build_backflow
    from neuralhubbard_amd.configs import make_params

    wk = G.workload(name)
    f = load(name)
    p = make_params(wk)
    dev = build_backflow(wk, p)
    ref = O.Model(wk, p)
    S = torch.tensor(f["s"], device="cuda")
    sg0, la0 = Variational(dev, symm=False)(S).numpy()
    assert np.array_equal(sg0, f["image_sign"])
    np.testing.assert_allclose(la0, f["image_logabs"], rtol=0, atol=1e-4 * max(1.0, wk.N / 16))
    state = Variational(dev)
    psi = state(S)
    sg, la = psi.numpy()
    el = Hubbard(wk.U, 1.0, wk.tprime).Oloc(state, Samples(S, psi)).cpu().numpy()
    for b in range(len(f["s"])):
        c = cancellation(ref, f["s"][b])
        if not np.isfinite(c) or c > 4:
            continue  # near-total cancellation of the projection: FP32 cannot resolve psi_P
        assert sg[b] == f["psi_sign"][b]
        assert abs(la[b] - f["psi_logabs"][b]) < log_tol(1e-4 * max(1.0, wk.N / 16), c)
        assert abs(el[b] - f["eloc"][b]) < 1e-5 * abs(f["eloc"][b]) + 1e-4 * max(1.0, np.exp(c))
