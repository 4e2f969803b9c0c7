"""Per-sample Jacobian of log|psi_P| (nh_log_psi_jacobian) against central differences of the FP64
oracle's amplitude: along random directions in each parameter group (packed conv weights,
biases, W, theta, base table), and column by column on single parameters.

Directional check: J v against (L(p + eps dp) - L(p - eps dp)) / (2 eps), L = oracle log|psi_P|
in FP64, v = the change of `param_vector` along dp (exact: `param_vector` is linear).  The
tolerance is 2e-3 relative plus 2e-4 of sum_k |J_k v_k| (the scale of the terms summed in
FP32), times e^c for a projection that cancels by c e-folds (the FP32 image weights psi_g /
psi_P carry that factor)."""

import copy
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests._util import cancellation
from neuralhubbard_amd.configs import WORKLOADS, make_params, make_configurations, small_workload

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible MI355X")
    from neuralhubbard_amd import _lib

    _lib.load()


GROUPS = ("conv_w", "conv_b", "emb", "head", "W", "theta", "F0", "phi0")


def _perturbed(p, dp, eps):
    """Params p + eps * dp (dp a dict of arrays by group name)."""
    q = copy.deepcopy(p)
    for k, v in dp.items():
        if k in ("conv_w", "conv_b"):
            setattr(q, k, [np.asarray(a, np.float64) + eps * d for a, d in zip(getattr(p, k), v)])
        elif k == "theta":
            q.theta = p.theta + eps * v
        else:
            setattr(q, k, np.asarray(getattr(p, k), np.float64) + eps * v)
    return q


def _direction(p, group, rng, one_hot=False):
    """A random direction in one parameter group, scaled like the group's entries (or a unit
    step in one random entry)."""
    if group == "theta":
        return {"theta": 1.0}
    if group in ("conv_w", "conv_b"):
        arrs = getattr(p, group)
        if one_hot:
            layer = int(rng.integers(len(arrs)))
            out = [np.zeros_like(np.asarray(a, np.float64)) for a in arrs]
            out[layer].flat[int(rng.integers(out[layer].size))] = 1.0
            return {group: out}
        return {group: [rng.standard_normal(np.shape(a)) * (np.std(a) + 1e-3) for a in arrs]}
    a = np.asarray(getattr(p, group), np.float64)
    if group == "F0":  # an antisymmetric change keeps F0 a skew matrix
        d = rng.standard_normal(a.shape) * np.std(a)
        d = d - d.T
        if one_hot:
            d = np.zeros_like(a)
            i, j = rng.choice(a.shape[0], 2, replace=False)
            d[i, j], d[j, i] = 1.0, -1.0
        return {"F0": d}
    if one_hot:
        d = np.zeros_like(a)
        d.flat[int(rng.integers(a.size))] = 1.0
        return {group: d}
    return {group: rng.standard_normal(a.shape) * (np.std(a) + 1e-3)}


def _groups(p):
    out = ["conv_w", "conv_b", "W", "theta"]
    if p.emb is not None:
        out.append("emb")
    if p.head is not None:
        out.append("head")
    if p.F0 is not None:
        out.append("F0")
    if p.phi0 is not None:
        out.append("phi0")
    return out


def _check(wk, n_conf, seed, symm=True, one_hot=0, eps=1e-4):
    from neuralhubbard_amd.backflow import Variational, build_backflow, param_vector

    p = make_params(wk)
    dev = build_backflow(wk, p, "fp32")
    st = Variational(dev, symm=symm)
    # configurations whose projection does not cancel by more than 3 e-folds (small lattices
    # hold configurations whose projected amplitude vanishes by symmetry)
    ref = O.Model(wk, p)
    cand = make_configurations(wk, 8 * n_conf, seed=seed)
    keep = [s for s in cand if not symm or cancellation(ref, s) < 3.0][:n_conf]
    assert len(keep) == n_conf
    S = np.stack(keep)
    psi, J = st.jacobian(torch.tensor(S, device="cuda"))
    sg, la = psi.numpy()
    J = J.cpu().numpy().astype(np.float64)
    assert J.shape == (n_conf, param_vector(p).size)
    assert np.isfinite(J).all()
    # psi_P itself: the nh_log_psi values
    sg2, la2 = st(torch.tensor(S, device="cuda")).numpy()
    np.testing.assert_array_equal(sg, sg2)
    np.testing.assert_array_equal(la, la2)
    rng = np.random.default_rng(seed)
    base = param_vector(p)
    worst = 0.0
    cases = [(g, False) for g in _groups(p)] + [(g, True) for g in rng.choice(_groups(p), one_hot)]
    for group, oh in cases:
        dp = _direction(p, group, rng, one_hot=oh)
        pp, pm = _perturbed(p, dp, eps), _perturbed(p, dp, -eps)
        v = (param_vector(pp) - param_vector(pm)) / (2 * eps)
        mp, mm = O.Model(wk, pp), O.Model(wk, pm)
        for b in range(n_conf):
            fd = (mp.psi(S[b], symm)[1] - mm.psi(S[b], symm)[1]) / (2 * eps)
            jv = float(J[b] @ v)
            scale = float(np.abs(J[b] * v).sum())
            c = cancellation(ref, S[b]) if symm else 0.0
            tol = (2e-3 * abs(fd) + 2e-4 * scale + 1e-6) * max(1.0, math.exp(c))
            err = abs(jv - fd)
            worst = max(worst, err / tol)
            print(f"{wk.name} {group}{' one-hot' if oh else ''} b={b}: J.v={jv:+.6e} fd={fd:+.6e} "
                  f"|d|={err:.2e} tol={tol:.2e}", flush=True)
            assert err < tol, (group, oh, b, jv, fd)
    print(f"{wk.name}: worst error / tolerance {worst:.3f}", flush=True)
    assert np.array_equal(base, param_vector(p))


@pytest.mark.parametrize("symm", [True, False])
def test_jacobian_pfaffian_conv_with_base(symm):
    """4x4 Pfaffian, synthetic network (8 channels, 2 layers), base F0; projected and bare."""
    wk = small_workload(kind="pfaffian", n_up=3, n_dn=3, channels=8, n_layers=2, base_f0=True)
    _check(wk, 3, seed=11, symm=symm, one_hot=6)


def test_jacobian_bcs_conv():
    """4x4 BCS, synthetic network, base pair function phi0."""
    wk = small_workload(kind="bcs", n_up=3, n_dn=3, channels=8, n_layers=2)
    _check(wk, 3, seed=12, one_hot=6)


def test_jacobian_pfaffian_resconv():
    """The released states' residual CNN (embedding, RMS norm, two convolutions per block,
    output / sqrt(blocks + 1)) at 4x4 with 64 channels and 2 blocks."""
    wk = small_workload(kind="pfaffian", n_up=3, n_dn=3, channels=64, n_layers=2, network="resconv")
    _check(wk, 2, seed=13, one_hot=4)


def test_jacobian_bcs_resconv_head():
    """Residual CNN with the BCS states' linear head (d = 64 features per spin); 4 + 4 electrons
    (with an odd number per spin the spin-inverted images cancel the projection almost fully)."""
    wk = small_workload(kind="bcs", n_up=4, n_dn=4, channels=64, n_layers=2, network="resconv", features=64)
    _check(wk, 2, seed=14, one_hot=4)


def test_jacobian_pf16():
    """The 16x16 Pfaffian workload (112 + 112 electrons, 32 channels, 6 layers, n = 224)."""
    _check(WORKLOADS["pf16"], 2, seed=15, eps=1e-4)


def test_jacobian_mixed_model_uses_fp32_weights():
    """A mixed-precision model's Jacobian is computed with its FP32 weights: bit-identical to
    the FP32 model's."""
    from neuralhubbard_amd.backflow import Variational, build_backflow

    wk = WORKLOADS["pf16"]
    p = make_params(wk)
    S = torch.tensor(make_configurations(wk, 3, seed=5), device="cuda")
    _, J32 = Variational(build_backflow(wk, p, "fp32")).jacobian(S)
    _, Jm = Variational(build_backflow(wk, p, "mixed")).jacobian(S)
    assert torch.equal(J32, Jm)


def test_jacobian_chunked_batch():
    """A batch run in chunks (small workspace cap) gives the same rows as one launch."""
    from neuralhubbard_amd.backflow import Variational, build_backflow

    wk = small_workload(kind="pfaffian", n_up=3, n_dn=3, channels=8, n_layers=2, base_f0=True)
    st = Variational(build_backflow(wk, make_params(wk), "fp32"))
    S = torch.tensor(make_configurations(wk, 7, seed=3), device="cuda")
    _, J1 = st.jacobian(S)
    _, J2 = st.jacobian(S, workspace_cap=1)
    assert torch.equal(J1, J2)
