"""Delayed-acceptance Metropolis on the 16x16 Pfaffian workload (marked gpu): the surrogate (3-layer
3x3 periodic conv, 32 channels) is trained by the library's own forward / backward / Adam kernels
on 2^20 (configuration, exact log|psi|) pairs recorded from 8192 equilibrated chains
(tests/golden/pf16_chains_8192.npz), then delayed acceptance and plain Metropolis run from the same
ensemble.  Checked: the surrogate learned (held-out error below the variance of log|psi|), stage
one passes a fraction of the changing proposals and stage two rejects some of them, the
ensemble's double occupancy and mean log|psi| stay where plain Metropolis keeps them (a sampler
that left |psi_P|^2 would drift them away from the equilibrated values), and E/N of the two
samplers agrees within its error bars.  Exact stationarity of delayed acceptance is pinned
separately by the chi-square test on an enumerated lattice (test_gpu_sampler.py)."""

import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

HERE = os.path.dirname(os.path.abspath(__file__))
CHAINS = os.path.join(HERE, "golden", "pf16_chains_8192.npz")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible MI355X")


def test_delayed_acceptance_pf16_matches_plain_metropolis():
    from neuralhubbard_amd.backflow import Hubbard, LogArray, Samples, Variational, build_backflow
    from neuralhubbard_amd.configs import WORKLOADS, make_params
    from neuralhubbard_amd.measure import energy, within_error_bars
    from neuralhubbard_amd.released import unpack_spins
    from neuralhubbard_amd.sampler import MetropolisSampler, Surrogate

    wk = WORKLOADS["pf16"]
    N = wk.N
    st = Variational(build_backflow(wk, make_params(wk), "mixed"))
    with np.load(CHAINS, allow_pickle=False) as f:
        ens = torch.as_tensor(unpack_spins(f["bits"], f["shape"]), device="cuda")
    C = ens.shape[0]
    # 2^20 pairs: 128 recorded states per chain, 4 exact proposals apart
    rec = MetropolisSampler(st, ens, 4, seed=11)
    X, Y = [], []
    for _ in range((1 << 20) // C):
        smp = rec.sweep()
        X.append(smp.spins)
        Y.append(smp.psi.logabs)
    X = torch.cat(X).cpu().numpy()
    Y = torch.cat(Y).cpu().numpy()
    assert len(X) == 1 << 20
    ntr = len(X) - 32768
    sur = Surrogate(wk.LX, wk.LY, 32, seed=2)
    hist = sur.train(X[:ntr], Y[:ntr], epochs=2, batch=1024, lr=2e-3)
    held = float(np.mean((sur(X[ntr:]).cpu().numpy() - Y[ntr:]) ** 2))
    print(f"surrogate: train mse {hist[-1]:.2f}, held-out mse {held:.2f}, var(log|psi|) {Y.var():.2f}", flush=True)
    assert held < 0.9 * Y.var()
    del X, Y
    start = rec.s.clone()

    H = Hubbard(wk.U, 1.0, wk.tprime, chunk=16)
    out = {}
    for name, s_ in (("plain", None), ("delayed", sur)):
        sm = MetropolisSampler(st, start, 512, seed=23, surrogate=s_)
        sm.sweep()
        sm.sweep()
        stats = sm.stats()
        s = sm.s
        D = ((s[:, :N] > 0) & (s[:, N:] > 0)).sum(1).double() / N
        la = sm.logabs.double()
        meas = slice(0, 2048)
        e = H.Oloc(st, Samples(s[meas].clone(), LogArray(sm.sign[meas].clone(), sm.logabs[meas].clone())))
        e = e.double().cpu().numpy()
        # 16 groups of 128 chains stand for 16 sweeps of the jackknife
        g = e.reshape(16, -1)
        out[name] = {"stats": stats, "D": D.cpu().numpy(), "logabs": la.cpu().numpy(),
                     "energy": energy(g.mean(1), (g * g).mean(1), N)}
        print(f"{name}: {stats}; D/N {D.mean().item():.5f}, <log|psi|> {la.mean().item():.3f}, "
              f"E/N {out[name]['energy']['E_per_site']:.4f} +- {out[name]['energy']['E_err']:.4f}", flush=True)
    st_da = out["delayed"]["stats"]
    assert 0.05 < st_da["stage1_acceptance"] < 0.95
    assert st_da["stage2_acceptance"] < 0.999
    assert st_da["exact_evaluations"] < 0.95 * out["plain"]["stats"]["exact_evaluations"]
    # the equilibrated ensemble stays equilibrated: paired differences over the 8192 chains
    for key in ("D", "logabs"):
        d = out["delayed"][key] - out["plain"][key]
        se = d.std(ddof=1) / np.sqrt(len(d))
        print(f"{key}: delayed - plain = {d.mean():.5f} +- {se:.5f}", flush=True)
        assert abs(d.mean()) < 4 * se + (1e-4 if key == "D" else 1e-2)
    assert within_error_bars(out["delayed"]["energy"], out["plain"]["energy"])
