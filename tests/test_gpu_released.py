"""The residual CNN of the released states on the GPU (marked gpu).

* synthetic residual CNNs on 4x16 and 8x16 lattices against the FP64 oracle: features,
  bare and projected log|psi| and E_loc;
* the released tpm02_L16_pfaffian state (fixture tests/golden/tpm02_L16_pfaffian.npz, written by
  make_released.py from the reference's weight and chain files): features, log|psi| and E_loc
  against the FP64 oracle values stored in the fixture, and the energy per site over the
  released chains against the paper's E/N = -0.74533(1), sigma^2/N = 0.0272(5)
  (code/states.py:114-115).

Tolerances: features 1e-4 relative to their largest magnitude (FP32 MFMA through 2 x blocks
convolutions), log|psi| 1e-4 on the small lattices and 1e-3 at 16x16 (the bar for the 16x16
Pfaffian), signs exact, E_loc 1e-4 per site.  The E/N check allows four standard errors of the
chain mean, sigma^2/N a factor 1.5 either way."""

import os

import numpy as np
import pytest

from oracle import oracle as O
from oracle.cpu_port import CpuPath

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "tpm02_L16_pfaffian.npz")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible MI355X")


def _small(LX, n_el, blocks, seed):
    from neuralhubbard_amd.configs import make_configurations, make_params, small_workload

    wk = small_workload("pfaffian", LX=LX, LY=16, n_up=n_el, n_dn=n_el, tprime=-0.2, channels=64,
                        n_layers=blocks, seed=seed, network="resconv")
    return wk, make_params(wk), make_configurations(wk, 4)


@pytest.mark.parametrize("LX,n_el,blocks", [(4, 12, 2), (8, 28, 3)])
def test_small_resconv_matches_oracle(LX, n_el, blocks):
    from neuralhubbard_amd.backflow import Hubbard, Samples, Variational, build_backflow

    wk, p, S = _small(LX, n_el, blocks, seed=11 + LX)
    print(f"resconv {LX}x16 n_el={n_el} blocks={blocks}", flush=True)
    ref = O.Model(wk, p)
    dev = build_backflow(wk, p)
    y = dev.features(torch.tensor(S, device="cuda")).cpu().numpy()
    yr = ref.features_batch(S)
    np.testing.assert_allclose(y, yr, rtol=0, atol=1e-4 * np.abs(yr).max())
    for symm in (False, True):
        psi = Variational(dev, symm=symm)(torch.tensor(S, device="cuda"))
        sg, la = psi.numpy()
        rs, rl = ref.psi_batch(S, project=symm)
        np.testing.assert_array_equal(sg, rs)
        np.testing.assert_allclose(la, rl, rtol=0, atol=1e-4)
    state = Variational(dev, symm=True)
    psi = state(torch.tensor(S, device="cuda"))
    el = Hubbard(U=wk.U, tprime=wk.tprime, chunk=2).Oloc(state, Samples(torch.tensor(S, device="cuda"), psi))
    el = el.cpu().numpy()
    rs, rl = ref.psi_batch(S)
    er = CpuPath(wk, p, bits=64).local_energy(S, rs, rl)
    np.testing.assert_allclose(el / wk.N, er / wk.N, rtol=0, atol=1e-4)


def _released():
    from neuralhubbard_amd.configs import Params
    from neuralhubbard_amd.released import STATES, unpack_spins

    f = dict(np.load(FIXTURE, allow_pickle=False))
    info = STATES["tpm02_L16_pfaffian"]
    p = Params(list(f["conv_w"]), list(f["conv_b"]), f["W"], float(f["theta"]), None, None, f["emb"])
    return info, info.workload(), p, unpack_spins(f["bits"], f["shape"]), f


def test_released_state_matches_oracle():
    from neuralhubbard_amd.backflow import Hubbard, Samples, Variational, build_backflow

    info, wk, p, S, f = _released()
    dev = build_backflow(wk, p)
    nf, npsi, nel = len(f["features"]), len(f["psi_sign"]), len(f["eloc"])
    print(f"released {info.name}: features of {nf}, psi of {npsi}, E_loc of {nel} chains", flush=True)
    y = dev.features(torch.tensor(S[:nf], device="cuda")).cpu().numpy()
    np.testing.assert_allclose(y, f["features"], rtol=0, atol=1e-4 * np.abs(f["features"]).max())
    for symm, key in ((False, "bare"), (True, "psi")):
        sg, la = Variational(dev, symm=symm)(torch.tensor(S[:npsi], device="cuda")).numpy()
        np.testing.assert_array_equal(sg, f[f"{key}_sign"])
        np.testing.assert_allclose(la, f[f"{key}_logabs"], rtol=0, atol=1e-3)
    state = Variational(dev, symm=True)
    s = torch.tensor(S[:nel], device="cuda")
    el = Hubbard(U=wk.U, tprime=wk.tprime, chunk=nel).Oloc(state, Samples(s, state(s))).cpu().numpy()
    print("E_loc / N", el / wk.N, "oracle", f["eloc"] / wk.N, flush=True)
    np.testing.assert_allclose(el / wk.N, f["eloc"] / wk.N, rtol=0, atol=1e-4)


def test_released_state_energy_per_site():
    from neuralhubbard_amd.backflow import Hubbard, Samples, Variational, build_backflow

    info, wk, p, S, _ = _released()
    K = 512
    dev = build_backflow(wk, p)
    state = Variational(dev, symm=True)
    H = Hubbard(U=wk.U, tprime=wk.tprime, chunk=64)
    el = []
    for b0 in range(0, K, 128):
        s = torch.tensor(S[b0:b0 + 128], device="cuda")
        el.append(H.Oloc(state, Samples(s, state(s))).cpu().numpy())
        print(f"E_loc of chains {b0}..{b0 + len(s)}", flush=True)
    e = np.concatenate(el) / wk.N
    assert np.all(np.isfinite(e))
    mean, err = e.mean(), e.std(ddof=1) / np.sqrt(len(e))
    var_per_site = e.var(ddof=1) * wk.N
    print(f"E/N = {mean:.5f} +- {err:.5f} (paper {info.E_per_site}), sigma^2/N = {var_per_site:.4f} "
          f"(paper {info.var_per_site})", flush=True)
    assert abs(mean - info.E_per_site) < 4 * err + info.E_err
    assert info.var_per_site / 1.5 < var_per_site < 1.5 * info.var_per_site
