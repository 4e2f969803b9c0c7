"""The residual CNN of the released states on the GPU (marked gpu).

* synthetic residual CNNs on 4x16 and 8x16 lattices against the FP64 oracle: features,
  bare and projected log|psi| and E_loc;
* the released tpm02_L16_pfaffian, tp0_L16_pfaffian and tp0_L16_bcs states (fixtures
  tests/golden/<name>.npz, written by make_released.py from the reference's weight and chain
  files): features, log|psi| and E_loc against the FP64 oracle values stored in the fixtures, and
  the energy per site over the released chains against the paper: E/N = -0.74533(1),
  sigma^2/N = 0.0272(5) (code/states.py:114-115), E/N = -0.75813(1), sigma^2/N = 0.0259(2)
  (:106-107) and E/N = -0.75878(2), sigma^2/N = 0.0255(20) (:104-105).

Both precisions of the network: FP32 MFMA (with the LDS weight slabs, and reading the weights
directly: "fp32-direct") and split-BF16 MFMA ("mixed").  Tolerances: features
1e-4 (FP32) or 1e-3 (split BF16) relative to their largest magnitude, log|psi| as described in
the small-lattice test and 1e-3 at 16x16 (the bar for the 16x16 Pfaffian), signs exact, E_loc
1e-4 per site.  The E/N check (2048 fixture chains split-BF16, 512 FP32) allows four standard
errors of the chain mean, sigma^2/N (2048 chains) a factor 1.5 either way."""

import os

import numpy as np
import pytest

from oracle import oracle as O
from oracle.cpu_port import CpuPath

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible MI355X")


def _small(LX, n_el, blocks, seed, kind="pfaffian"):
    from neuralhubbard_amd.configs import make_configurations, make_params, small_workload

    wk = small_workload(kind, LX=LX, LY=16, n_up=n_el, n_dn=n_el, tprime=-0.2 if kind == "pfaffian" else 0.0,
                        channels=64, n_layers=blocks, seed=seed, network="resconv",
                        features=64 if kind == "bcs" else 0)
    return wk, make_params(wk), make_configurations(wk, 4)


class _Perturbed(O.Model):
    """The oracle with every feature multiplied by (1 + rel * N(0, 1)): the FP32 rounding
    scale of the features, to measure how strongly a configuration's Pfaffian amplifies it."""

    def __init__(self, wk, p, rel, seed):
        super().__init__(wk, p)
        self.rel, self.rng = rel, np.random.default_rng(seed)

    def features(self, s):
        y = super().features(s)
        return y * (1.0 + self.rel * self.rng.standard_normal(y.shape))


def _sensitivity(wk, p, S, project, rel=1e-7, trials=3):
    base = O.Model(wk, p).psi_batch(S, project)[1]
    return np.max([np.abs(_Perturbed(wk, p, rel, k).psi_batch(S, project)[1] - base) for k in range(trials)], 0)


@pytest.mark.parametrize("prec", ["fp32", "fp32-direct", "mixed"])
@pytest.mark.parametrize("kind,LX,n_el,blocks", [("pfaffian", 4, 12, 2), ("pfaffian", 8, 28, 3), ("bcs", 4, 12, 2)])
def test_small_resconv_matches_oracle(kind, LX, n_el, blocks, prec):
    """Random weights give some configurations nearly singular pairing matrices (condition
    numbers ~1e4 at 8x16), whose log|Pf| moves by ~1e-4 under feature changes of one FP32 unit;
    the log|psi| tolerance is 1e-4 plus 30 times that measured sensitivity (at relative feature
    changes of 1e-7 for FP32, 1e-5 for split BF16).  E_loc takes the oracle's psi(s) on both
    sides and is held to 1e-4 per site on 4x16, 1e-3 on 8x16 (ten times that in split BF16)."""
    from neuralhubbard_amd.backflow import Hubbard, LogArray, Samples, Variational, build_backflow

    wk, p, S = _small(LX, n_el, blocks, seed=11 + LX, kind=kind)
    print(f"resconv {kind} {LX}x16 n_el={n_el} blocks={blocks} {prec}", flush=True)
    mixed = prec == "mixed"
    ref = O.Model(wk, p)
    dev = build_backflow(wk, p, "mixed" if mixed else "fp32")
    if prec == "fp32-direct":  # the FP32 kernel reading conv_w without the weight slabs
        dev.conv_w_mfma = None
    y = dev.features(torch.tensor(S, device="cuda")).cpu().numpy()
    yr = ref.features_batch(S)
    np.testing.assert_allclose(y, yr, rtol=0, atol=(1e-3 if mixed else 1e-4) * np.abs(yr).max())
    for symm in (False, True):
        psi = Variational(dev, symm=symm)(torch.tensor(S, device="cuda"))
        sg, la = psi.numpy()
        rs, rl = ref.psi_batch(S, project=symm)
        tol = 1e-4 + 30 * _sensitivity(wk, p, S, symm, rel=1e-5 if mixed else 1e-7)
        print(f"project={symm}: |d log psi| {np.abs(la - rl)}, tolerance {tol}", flush=True)
        np.testing.assert_array_equal(sg, rs)
        assert np.all(np.abs(la - rl) <= tol)
    rs, rl = ref.psi_batch(S)
    er = CpuPath(wk, p, bits=64).local_energy(S, rs, rl)
    state = Variational(dev, symm=True)
    psi = LogArray(torch.tensor(rs, dtype=torch.float32, device="cuda"), torch.tensor(rl, device="cuda"))
    el = Hubbard(U=wk.U, tprime=wk.tprime, chunk=2).Oloc(state, Samples(torch.tensor(S, device="cuda"), psi))
    el = el.cpu().numpy()
    print(f"E_loc / N {el / wk.N}, oracle {er / wk.N}", flush=True)
    np.testing.assert_allclose(el / wk.N, er / wk.N, rtol=0, atol=(1e-4 if LX == 4 else 1e-3) * (10 if mixed else 1))


@pytest.mark.parametrize("prec", ["fp32", "mixed"])
@pytest.mark.parametrize("LX,LY,n_el,blocks", [(4, 6, 8, 2), (6, 6, 14, 2), (20, 4, 30, 1)])
def test_other_lattice_resconv_matches_oracle(LX, LY, n_el, blocks, prec):
    """The residual CNN on lattices outside the tile kernels (layer by layer: FP32 MFMA, or split
    BF16 MFMA in the mixed setting): features through nh_features_ws, bare and projected log|psi|
    and E_loc against the FP64 oracle, tolerances as in the 16-row test."""
    from neuralhubbard_amd.backflow import Hubbard, LogArray, Samples, Variational, build_backflow
    from neuralhubbard_amd.configs import make_configurations, make_params, small_workload

    wk = small_workload("pfaffian", LX=LX, LY=LY, n_up=n_el, n_dn=n_el, tprime=-0.2, channels=64,
                        n_layers=blocks, seed=LX * LY, network="resconv")
    p = make_params(wk)
    S = make_configurations(wk, 4)
    print(f"resconv {LX}x{LY} n_el={n_el} blocks={blocks} {prec}", flush=True)
    mixed = prec == "mixed"
    ref = O.Model(wk, p)
    dev = build_backflow(wk, p, prec)
    y = dev.features(torch.tensor(S, device="cuda")).cpu().numpy()
    yr = ref.features_batch(S)
    print("max |d y| / max |y|", np.abs(y - yr).max() / np.abs(yr).max(), flush=True)
    np.testing.assert_allclose(y, yr, rtol=0, atol=(1e-3 if mixed else 1e-4) * np.abs(yr).max())
    for symm in (False, True):
        sg, la = Variational(dev, symm=symm)(torch.tensor(S, device="cuda")).numpy()
        rs, rl = ref.psi_batch(S, project=symm)
        tol = 1e-4 + 30 * _sensitivity(wk, p, S, symm, rel=1e-5 if mixed else 1e-7)
        print(f"project={symm}: |d log psi| {np.abs(la - rl)}, tolerance {tol}", flush=True)
        np.testing.assert_array_equal(sg, rs)
        assert np.all(np.abs(la - rl) <= tol)
    rs, rl = ref.psi_batch(S)
    er = CpuPath(wk, p, bits=64).local_energy(S, rs, rl)
    psi = LogArray(torch.tensor(rs, dtype=torch.float32, device="cuda"), torch.tensor(rl, device="cuda"))
    el = Hubbard(U=wk.U, tprime=wk.tprime, chunk=2).Oloc(Variational(dev), Samples(torch.tensor(S, device="cuda"), psi))
    el = el.cpu().numpy()
    print(f"E_loc / N {el / wk.N}, oracle {er / wk.N}", flush=True)
    np.testing.assert_allclose(el / wk.N, er / wk.N, rtol=0, atol=1e-4 * (10 if mixed else 1))


NAMES = ("tpm02_L16_pfaffian", "tp0_L16_pfaffian", "tp0_L16_bcs")


def _released(name="tpm02_L16_pfaffian"):
    from neuralhubbard_amd.configs import Params
    from neuralhubbard_amd.released import STATES, unpack_spins

    f = dict(np.load(os.path.join(HERE, "golden", f"{name}.npz"), allow_pickle=False))
    info = STATES[name]
    p = Params(list(f["conv_w"]), list(f["conv_b"]), f["W"], float(f["theta"]), None, f.get("phi0"), f["emb"],
               f.get("head"))
    return info, info.workload(), p, unpack_spins(f["bits"], f["shape"]), f


@pytest.mark.parametrize("prec", ["fp32", "mixed"])
@pytest.mark.parametrize("name", NAMES + ("tpm02_L24_pfaffian", "tp0_L32_bcs"))
def test_released_state_matches_oracle(name, prec):
    from neuralhubbard_amd.backflow import Hubbard, Samples, Variational, build_backflow

    info, wk, p, S, f = _released(name)
    dev = build_backflow(wk, p, prec)
    nf, npsi, nel = len(f["features"]), len(f["psi_sign"]), len(f["eloc"])
    print(f"released {info.name} {prec}: features of {nf}, psi of {npsi}, E_loc of {nel} chains", flush=True)
    y = dev.features(torch.tensor(S[:nf], device="cuda")).cpu().numpy()
    print("max |d y| / max |y|", np.abs(y - f["features"]).max() / np.abs(f["features"]).max(), flush=True)
    np.testing.assert_allclose(y, f["features"], rtol=0,
                               atol=(1e-3 if prec == "mixed" else 1e-4) * np.abs(f["features"]).max())
    for symm, key in ((False, "bare"), (True, "psi")):
        sg, la = Variational(dev, symm=symm)(torch.tensor(S[:npsi], device="cuda")).numpy()
        print(f"project={symm}: max |d log psi| {np.abs(la - f[f'{key}_logabs']).max():.2e}", flush=True)
        np.testing.assert_array_equal(sg, f[f"{key}_sign"])
        np.testing.assert_allclose(la, f[f"{key}_logabs"], rtol=0, atol=1e-3)
    if nel:  # (the 24x24 fixture carries no oracle E_loc)
        state = Variational(dev, symm=True)
        s = torch.tensor(S[:nel], device="cuda")
        el = Hubbard(U=wk.U, tprime=wk.tprime, chunk=nel).Oloc(state, Samples(s, state(s))).cpu().numpy()
        print("E_loc / N", el / wk.N, "oracle", f["eloc"] / wk.N, flush=True)
        np.testing.assert_allclose(el / wk.N, f["eloc"] / wk.N, rtol=0, atol=1e-4)


@pytest.mark.parametrize("prec", ["fp32", "mixed"])
@pytest.mark.parametrize("name", NAMES)
def test_released_state_energy_per_site(name, prec):
    from neuralhubbard_amd.backflow import Hubbard, Samples, Variational, build_backflow

    info, wk, p, S, _ = _released(name)
    # E_loc / N has heavy tails (tp0_L16_pfaffian chain 353 sits at -0.951, the FP64 oracle
    # agrees), so sigma^2/N is checked over all 2048 fixture chains, in the faster split-BF16
    # form; the FP32 form checks E/N over the first 512
    K = len(S) if prec == "mixed" else 512
    dev = build_backflow(wk, p, prec)
    state = Variational(dev, symm=True)
    H = Hubbard(U=wk.U, tprime=wk.tprime, chunk=64)
    el = []
    for b0 in range(0, K, 256):
        s = torch.tensor(S[b0:b0 + 256], device="cuda")
        el.append(H.Oloc(state, Samples(s, state(s))).cpu().numpy())
        print(f"{prec}: E_loc of chains {b0}..{b0 + len(s)}", flush=True)
    e = np.concatenate(el) / wk.N
    assert np.all(np.isfinite(e))
    mean, err = e.mean(), e.std(ddof=1) / np.sqrt(len(e))
    var_per_site = e.var(ddof=1) * wk.N
    print(f"E/N = {mean:.5f} +- {err:.5f} (paper {info.E_per_site}), sigma^2/N = {var_per_site:.4f} "
          f"(paper {info.var_per_site})", flush=True)
    assert abs(mean - info.E_per_site) < 4 * err + info.E_err
    if K == len(S):
        assert info.var_per_site / 1.5 < var_per_site < 1.5 * info.var_per_site


def test_released_chains_under_the_papers_sampler():
    """The paper's sampler (`released.make_sampler`: particle hops and site exchanges with the
    state's nearest + next-nearest ranges, code/states.py:355-408) moves 1024 released
    tpm02_L16_pfaffian chains by N steps each: both move kinds are accepted, the particle numbers
    hold, and E/N of the moved chains stays at the paper's -0.74533(1) within four standard
    errors (the released chains are equilibrated under the reference's sampler, so a sampler
    that left |psi_P|^2 would drift them)."""
    from neuralhubbard_amd.backflow import Hubbard, Samples, Variational, build_backflow
    from neuralhubbard_amd.released import make_sampler

    info, wk, p, S, _ = _released("tpm02_L16_pfaffian")
    state = Variational(build_backflow(wk, p, "mixed"), symm=True)
    K, N = 1024, wk.N
    sm = make_sampler(state, info, K, S[:K], sweep_steps=N, thermal_steps=0, seed=5)
    assert sm.proposal.kind == 3 and sm.proposal.ph_neighbors == 8 and sm.proposal.se_neighbors == 8
    smp = sm.sweep()
    st = sm.stats()
    x = smp.spins.cpu().numpy()
    moved = (x != S[:K]).any(1).mean()
    print(f"mix sampler: {st}; chains moved {moved:.3f}", flush=True)
    assert ((x[:, :N] > 0).sum(1) == info.n_up).all() and ((x[:, N:] > 0).sum(1) == info.n_up).all()
    assert st["accepted"] > 0 and moved > 0.9
    # moves that change the doublon count (particle hops onto or off doubly occupied sites) and
    # moves that keep it (site exchanges of an up and a down electron among them) both occur
    d0 = ((S[:K, :N] > 0) & (S[:K, N:] > 0)).sum(1)
    d1 = ((x[:, :N] > 0) & (x[:, N:] > 0)).sum(1)
    assert (d0 != d1).any()
    H = Hubbard(U=wk.U, tprime=wk.tprime, chunk=64)
    el = []
    for b0 in range(0, K, 256):
        s = smp.spins[b0:b0 + 256]
        el.append(H.Oloc(state, Samples(s, state(s))).cpu().numpy())
    e = np.concatenate(el) / N
    mean, err = e.mean(), e.std(ddof=1) / np.sqrt(len(e))
    print(f"after N mixed steps: E/N = {mean:.5f} +- {err:.5f} (paper {info.E_per_site})", flush=True)
    assert abs(mean - info.E_per_site) < 4 * err + info.E_err


def test_released_24x24_energy_per_site():
    """The released tpm02_L24_pfaffian state (24x24, 252 + 252 electrons, 504 x 504 Pfaffians, the
    residual CNN layer by layer): E/N over the first 32 fixture chains against the paper's
    -0.74585(1) (code/states.py:118-119), four standard errors of the chain mean."""
    from neuralhubbard_amd.backflow import Hubbard, Samples, Variational, build_backflow

    info, wk, p, S, _ = _released("tpm02_L24_pfaffian")
    state = Variational(build_backflow(wk, p, "fp32"))
    H = Hubbard(U=wk.U, tprime=wk.tprime, chunk=4)
    el = []
    for b0 in range(0, 32, 8):
        s = torch.tensor(S[b0:b0 + 8], device="cuda")
        el.append(H.Oloc(state, Samples(s, state(s))).cpu().numpy())
        print(f"24x24: E_loc of chains {b0}..{b0 + 8}", flush=True)
    e = np.concatenate(el) / wk.N
    assert np.all(np.isfinite(e))
    mean, err = e.mean(), e.std(ddof=1) / np.sqrt(len(e))
    print(f"E/N = {mean:.5f} +- {err:.5f} (paper {info.E_per_site})", flush=True)
    assert abs(mean - info.E_per_site) < 4 * err + info.E_err


def test_released_32x32_bcs_energy_per_site():
    """The released tp0_L32_bcs state (32x32, 448 + 448 electrons, 448 x 448 determinants in the
    workspace, the residual CNN and its linear head layer by layer): E/N over the first 16
    fixture chains against the paper's -0.75617(1) (code/states.py:110-111), four standard
    errors of the chain mean."""
    from neuralhubbard_amd.backflow import Hubbard, Samples, Variational, build_backflow

    info, wk, p, S, _ = _released("tp0_L32_bcs")
    state = Variational(build_backflow(wk, p, "fp32"))
    H = Hubbard(U=wk.U, tprime=wk.tprime, chunk=4)
    el = []
    for b0 in range(0, 16, 4):
        s = torch.tensor(S[b0:b0 + 4], device="cuda")
        el.append(H.Oloc(state, Samples(s, state(s))).cpu().numpy())
        print(f"32x32 BCS: E_loc of chains {b0}..{b0 + 4}", flush=True)
    e = np.concatenate(el) / wk.N
    assert np.all(np.isfinite(e))
    mean, err = e.mean(), e.std(ddof=1) / np.sqrt(len(e))
    print(f"E/N = {mean:.5f} +- {err:.5f} (paper {info.E_per_site})", flush=True)
    assert abs(mean - info.E_per_site) < 4 * err + info.E_err
