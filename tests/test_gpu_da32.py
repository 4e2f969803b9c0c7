"""Delayed-acceptance Metropolis at 32x32 (marked gpu): the 32x32 Pfaffian workload (t' = -0.2,
448 + 448 electrons, 896 x 896 Pfaffians) in its mixed form (split-FP16 features on 32-site rows),
screened by a 16-channel conv surrogate (the 1024-site forward / gradient kernels) trained by the
library's own kernels on (configuration, exact log|psi|) pairs recorded from the path's chains.
Checked: every decision of the delayed-acceptance sweep at this size replays bit for bit (the
oracle's chain replay driven by the GPU amplitude and surrogate, so proposals, stage one and stage
two are the rule itself at 1024 sites); the surrogate learned (held-out error below the variance
of log|psi|); stage one passes a fraction of the changing proposals, stage two rejects some, and
the sweeps need fewer exact amplitudes than plain Metropolis from the same chains.  Stationarity of
the rule is pinned on an enumerated lattice (test_gpu_sampler.py), the amplitude's error against
FP64 at this size by test_log_psi_pf32_matches_oracle."""

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible MI355X")


@pytest.fixture(scope="module")
def da32():
    from neuralhubbard_amd.backflow import Variational, build_backflow
    from neuralhubbard_amd.configs import WORKLOADS, make_configurations, make_params
    from neuralhubbard_amd.sampler import MetropolisSampler, Surrogate

    wk = WORKLOADS["pf32"]
    st = Variational(build_backflow(wk, make_params(wk), "mixed"))
    s0 = torch.as_tensor(make_configurations(wk, 256, seed=5), device="cuda")
    rec = MetropolisSampler(st, s0, 4, seed=11)
    X, Y = [], []
    for _ in range(64):
        smp = rec.sweep()
        X.append(smp.spins)
        Y.append(smp.psi.logabs)
    X = torch.cat(X).cpu().numpy()
    Y = torch.cat(Y).cpu().numpy()
    # the chains start from random configurations and still relax over the recording, so the
    # held-out pairs are a random eighth of all records rather than the last ones
    ho = np.random.default_rng(3).permutation(len(X))
    tr, te = ho[len(X) // 8:], ho[:len(X) // 8]
    sur = Surrogate(wk.LX, wk.LY, 16, seed=2)
    hist = sur.train(X[tr], Y[tr], epochs=20, batch=256, lr=2e-3)
    held = float(np.mean((sur(X[te]).cpu().numpy() - Y[te]) ** 2))
    print(f"32x32 surrogate: {len(tr)} pairs, train mse {hist[-1]:.2f}, held-out mse {held:.2f}, "
          f"var(log|psi|) {Y[te].var():.2f}", flush=True)
    return wk, st, sur, rec.s.clone(), held, float(Y[te].var())


def test_surrogate_learns_the_32x32_amplitude(da32):
    _, _, _, _, held, var = da32
    assert held < 0.9 * var


def test_delayed_acceptance_pf32_saves_exact_evaluations(da32):
    from neuralhubbard_amd.sampler import MetropolisSampler

    wk, st, sur, start, _, _ = da32
    out = {}
    for name, s_ in (("plain", None), ("delayed", sur)):
        sm = MetropolisSampler(st, start, 64, seed=23, surrogate=s_)
        sm.sweep()
        out[name] = sm.stats()
        print(name, out[name], flush=True)
    da = out["delayed"]
    assert 0.05 < da["stage1_acceptance"] < 0.95
    assert da["stage2_acceptance"] < 0.999
    assert da["exact_evaluations"] < 0.95 * out["plain"]["exact_evaluations"]


def test_delayed_acceptance_pf32_replays(da32):
    from neuralhubbard_amd.sampler import MetropolisSampler

    wk, st, sur, start, _, _ = da32
    s0 = start[:4].clone()
    sm = MetropolisSampler(st, s0, 24, seed=77, surrogate=sur)
    sm.sweep()
    got = sm.s.cpu().numpy()

    def psi_gpu(x):
        sg, la = st(torch.as_tensor(np.asarray(x)[None], device="cuda")).numpy()
        return float(sg[0]), float(la[0])

    def sur_gpu(x):
        return float(sur(np.asarray(x)[None]).cpu().numpy()[0])

    init = s0.cpu().numpy()
    for c in range(4):
        sf, acc, _ = O.metropolis_chain(psi_gpu, init[c], c, 77, 0, 24, wk.LX, wk.LY, sur=sur_gpu)
        assert np.array_equal(got[c], sf), c
