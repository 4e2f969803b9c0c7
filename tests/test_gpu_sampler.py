"""Sampler and surrogate on the GPU against the oracle (marked gpu).

* surrogate forward vs the FP64 oracle; gradients vs FP64 central differences of the oracle;
* GPU chains replayed step by step by the oracle (same Philox draws), plain and delayed
  acceptance, for every move kind (same-spin exchange, particle hop, site exchange and their
  mix, 4 or 8 neighbours): configurations bit-exact;
* stationarity: the chain histogram on a 3x3 lattice (1296 states) against the exact
  |psi_P|^2 / Z (chi-square), for plain Metropolis and for delayed acceptance, with the
  exchange and with the mix of particle hops and site exchanges; the chain energy of a 4x4
  lattice against exact enumeration with both move sets;
* the quantax-style constructors (ParticleHop, SiteExchange, MixSampler) run the mix;
* delayed acceptance with a trained surrogate needs fewer exact evaluations.
"""

import itertools

import numpy as np
import pytest
from scipy import stats as sps

from oracle import oracle as O
from neuralhubbard_amd import lattice as L
from neuralhubbard_amd.configs import make_params, small_workload

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible MI355X")


def _sur(LX, LY, C, seed=3):
    from neuralhubbard_amd.sampler import Surrogate

    return Surrogate(LX, LY, C, seed=seed)


@pytest.mark.parametrize("LX,LY,C", [(4, 4, 16), (16, 16, 32), (3, 5, 16), (20, 20, 16), (32, 32, 16)])
def test_surrogate_forward_matches_oracle(LX, LY, C):
    sur = _sur(LX, LY, C)
    # larger last layer so the output is not tiny
    p = sur.params.cpu().numpy()
    p[-(9 * C + 1):] *= 10.0
    sur.params.copy_(torch.tensor(p))
    geom = L.Geometry(LX, LY, (LX * LY) // 3, (LX * LY) // 4)
    s = L.random_configurations(geom, 5, 11)
    f = sur(torch.tensor(s, device="cuda")).cpu().numpy()
    for b in range(5):
        ref = O.sur_forward(p, s[b], LX, LY, C)
        assert abs(f[b] - ref) < 1e-4 * max(1.0, abs(ref)), (b, f[b], ref)


@pytest.mark.parametrize("LX,LY,n_up,n_dn", [(4, 4, 5, 4), (32, 32, 448, 448)])
def test_surrogate_gradient_matches_finite_differences(LX, LY, n_up, n_dn):
    """The gradient kernel (one 64-thread block per configuration at 4x4, the 1024-thread form of
    lattices above 256 sites at 32x32) against central differences of the FP64 oracle's loss."""
    C = 16
    sur = _sur(LX, LY, C, seed=5)
    p = sur.params.cpu().numpy().astype(np.float64)
    p[-(9 * C + 1):] *= 10.0
    sur.params.copy_(torch.tensor(p.astype(np.float32)))
    p = sur.params.cpu().numpy().astype(np.float64)
    geom = L.Geometry(LX, LY, n_up, n_dn)
    S = L.random_configurations(geom, 3, 12)
    T = np.array([0.3, -0.7, 1.1])
    g, loss = sur.grad(torch.tensor(S, device="cuda"), torch.tensor(T, dtype=torch.float32, device="cuda"), 0.5)
    g = g.cpu().numpy()
    assert abs(loss.item() - O.sur_loss(p, S, T, LX, LY, C, 0.5)) < 1e-4 * max(1, loss.item())
    rng = np.random.default_rng(0)
    npar = p.size
    sections = [0, 9 * 2 * C, 9 * 2 * C + C, 9 * 2 * C + C + 9 * C * C, npar - 9 * C - 1, npar - 9 * C - 1 + 9 * C]
    idx = sorted(set(int(i) for a, b in zip(sections, sections[1:] + [npar]) for i in rng.integers(a, b, size=6)))
    fd = O.sur_grad_fd(p, S, T, LX, LY, C, 0.5, idx)
    np.testing.assert_allclose(g[idx], fd, rtol=2e-3, atol=2e-4 * np.abs(fd).max())


def test_adam_step_matches_formula():
    from neuralhubbard_amd import _lib

    n = 1000
    rng = np.random.default_rng(1)
    w0, g0 = rng.standard_normal(n).astype(np.float32), rng.standard_normal(n).astype(np.float32)
    w = torch.tensor(w0, device="cuda")
    g = torch.tensor(g0, device="cuda")
    m = torch.zeros(n, device="cuda")
    v = torch.zeros(n, device="cuda")
    for step in (1, 2):
        _lib.check(_lib.load().nh_adam(None, n, w.data_ptr(), g.data_ptr(), m.data_ptr(), v.data_ptr(), 0.01, 0.9,
                                       0.999, 1e-8, step), "nh_adam")
    mm, vv, ww = np.zeros(n), np.zeros(n), w0.astype(np.float64)
    for step in (1, 2):
        mm = 0.9 * mm + 0.1 * g0
        vv = 0.999 * vv + 0.001 * g0.astype(np.float64) ** 2
        ww -= 0.01 * (mm / (1 - 0.9**step)) / (np.sqrt(vv / (1 - 0.999**step)) + 1e-8)
    np.testing.assert_allclose(w.cpu().numpy(), ww, rtol=1e-5, atol=1e-6)


def _small_state(LX=4, LY=4, n_up=3, n_dn=3, channels=8, project=True):
    from neuralhubbard_amd.backflow import Variational, build_backflow

    wk = small_workload(LX=LX, LY=LY, n_up=n_up, n_dn=n_dn, channels=channels, n_layers=2, base_f0=True,
                        tprime=-0.2)
    p = make_params(wk)
    return wk, p, Variational(build_backflow(wk, p), symm=project), O.Model(wk, p)


MOVES = {  # name: (kind, ph_z, se_z)
    "exchange": (O.MOVE_EXCHANGE, 4, 4),
    "hop4": (O.MOVE_PARTICLE_HOP, 4, 4),
    "swap8": (O.MOVE_SITE_EXCHANGE, 4, 8),
    "mix8": (O.MOVE_MIX, 8, 8),
    "mix48": (O.MOVE_MIX, 4, 8),
}


def _proposal(name):
    from neuralhubbard_amd.sampler import Proposal

    return Proposal(*MOVES[name])


@pytest.mark.parametrize("delayed,moves", [(False, "exchange"), (True, "exchange"), (False, "hop4"),
                                           (False, "swap8"), (False, "mix8"), (True, "mix48")])
def test_metropolis_replays_oracle(delayed, moves):
    from neuralhubbard_amd.sampler import MetropolisSampler

    kind, ph_z, se_z = MOVES[moves]
    wk, p, state, ref = _small_state()
    s0 = L.random_configurations(wk.geom, 6, 21)
    sur = None
    if delayed:
        sur = _sur(4, 4, 16, seed=9)
        pp = sur.params.cpu().numpy()
        pp[-(9 * 16 + 1):] *= 20.0
        sur.params.copy_(torch.tensor(pp))
    sm = MetropolisSampler(state, s0, sweep_steps=25, seed=1234, surrogate=sur, proposal=_proposal(moves))
    sm.sweep()
    got = sm.s.cpu().numpy()
    sur_fn = None
    if delayed:
        pp = sur.params.cpu().numpy().astype(np.float64)
        sur_fn = lambda x: O.sur_forward(pp, x, 4, 4, 16)
    checked = 0
    for c in range(6):
        sf, acc, tie = O.metropolis_chain(ref.psi, s0[c], c, 1234, 0, 25, 4, 4, sur=sur_fn, margin=1e-4,
                                          kind=kind, ph_z=ph_z, se_z=se_z)
        if tie:
            continue
        checked += 1
        assert np.array_equal(got[c], sf), c
    assert checked >= 4
    st = sm.stats()
    print(st)
    assert st["accepted"] > 0


def _exact_distribution(ref, wk):
    N = wk.N
    states = []
    for u in itertools.combinations(range(N), wk.n_up):
        for d in itertools.combinations(range(N), wk.n_dn):
            s = -np.ones(2 * N, np.int8)
            s[list(u)] = 1
            s[[N + x for x in d]] = 1
            states.append(s)
    states = np.stack(states)
    la = np.array([ref.psi(s)[1] for s in states])
    w = np.exp(2 * (la - la[np.isfinite(la)].max()))
    w[~np.isfinite(w)] = 0.0
    return states, w / w.sum()


@pytest.mark.parametrize("delayed,moves", [(False, "exchange"), (True, "exchange"), (False, "mix8"), (True, "mix8")])
def test_metropolis_stationary_distribution(delayed, moves):
    """Chain histogram on a 3x3 lattice with 2 + 2 electrons vs |psi_P|^2 / Z: 32768
    independent chains, 1000 proposals of burn-in, then the state of every chain after each of
    two sweeps of 200 proposals; each sweep's 32768 states are independent draws (the two
    sweeps are not independent of each other, so each gets its own chi-square)."""
    from neuralhubbard_amd.sampler import MetropolisSampler

    wk, p, state, ref = _small_state(3, 3, 2, 2, channels=8)
    states, prob = _exact_distribution(ref, wk)
    key = {s.tobytes(): k for k, s in enumerate(states)}
    n_chains = 32768
    s0 = L.random_configurations(wk.geom, n_chains, 5)
    sur = _sur(3, 3, 16, seed=4) if delayed else None
    sm = MetropolisSampler(state, s0, sweep_steps=200, seed=99, surrogate=sur, thermal_steps=1000,
                           proposal=_proposal(moves))
    for sweep in range(2):
        x = sm.sweep().spins.cpu().numpy()
        counts = np.bincount(np.array([key[s.tobytes()] for s in x]), minlength=len(states)).astype(float)
        exp = prob * counts.sum()
        mask = exp >= 5
        chi2 = ((counts[mask] - exp[mask]) ** 2 / exp[mask]).sum()
        rest = exp[~mask].sum()
        dof = int(mask.sum()) - 1
        if rest > 0:
            chi2 += (counts[~mask].sum() - rest) ** 2 / rest
            dof += 1
        pval = sps.chi2.sf(chi2, dof)
        print(f"delayed={delayed} {moves} sweep {sweep}: chi2 {chi2:.1f} dof {dof} p {pval:.3g}")
        assert pval > 1e-4
    print(sm.stats())


def test_trained_surrogate_reduces_exact_evaluations():
    """Train the surrogate on exact (s, log|psi|) pairs from plain sampling of an 8x8 lattice
    with 28 + 28 electrons; delayed acceptance then accepts more hops per exact evaluation
    than plain Metropolis (measured on MI355X: about 1.4x)."""
    from neuralhubbard_amd.backflow import Variational, build_backflow
    from neuralhubbard_amd.sampler import MetropolisSampler

    wk = small_workload(LX=8, LY=8, n_up=28, n_dn=28, channels=16, n_layers=4, base_f0=False, tprime=0.0)
    state = Variational(build_backflow(wk, make_params(wk)), symm=True)
    s0 = L.random_configurations(wk.geom, 2048, 7)
    plain = MetropolisSampler(state, s0, sweep_steps=32, seed=1, thermal_steps=256)
    data_s, data_l = [], []
    for _ in range(32):
        smp = plain.sweep()
        data_s.append(smp.spins.cpu().numpy())
        data_l.append(smp.psi.logabs.cpu().numpy())
    X, y = np.concatenate(data_s), np.concatenate(data_l)
    sur = _sur(8, 8, 16, seed=2)
    hist = sur.train(X, y, epochs=20, batch=512, lr=2e-3)
    print("surrogate mse per epoch:", np.round(hist[::4], 3), "target variance", y.var())
    assert hist[-1] < 0.5 * y.var()
    start = plain.s.cpu().numpy()
    da = MetropolisSampler(state, start, sweep_steps=128, seed=7, surrogate=sur)
    pl = MetropolisSampler(state, start, sweep_steps=128, seed=7)
    da.sweep()
    pl.sweep()
    a, b = da.stats(), pl.stats()
    gain = (a["accepted"] / a["exact_evaluations"]) / (b["accepted"] / b["exact_evaluations"])
    print("delayed:", a, "\nplain:", b, "\naccepted per exact evaluation, delayed / plain:", gain)
    assert a["exact_evaluations"] < 0.7 * b["exact_evaluations"]
    assert gain > 1.2


@pytest.mark.parametrize("delayed,moves", [(False, "exchange"), (True, "exchange"), (False, "mix8")])
def test_chain_energy_matches_enumeration(delayed, moves):
    """E/N from 2048 chains x 32 sweeps (block jackknife over sweeps) against the exact
    variational energy sum_s |psi_P(s)|^2 E_loc(s) / sum_s |psi_P(s)|^2 over all 14400
    configurations of a 4x4 lattice with 2 + 2 electrons."""
    from neuralhubbard_amd.backflow import Hubbard, Samples
    from neuralhubbard_amd.sampler import MetropolisSampler

    wk, p, state, ref = _small_state(4, 4, 2, 2, channels=8)
    H = Hubbard(U=wk.U, t=1.0, tprime=wk.tprime)
    states, prob = _exact_distribution(ref, wk)
    S = torch.tensor(states, device="cuda")
    el = H.Oloc(state, Samples(S)).cpu().numpy()
    live = prob > 0  # E_loc is NaN where psi_P vanishes; those states carry no weight
    assert np.isfinite(el[live]).all()
    E_exact = float((prob[live] * el[live]).sum()) / wk.N
    sur = _sur(4, 4, 16, seed=3) if delayed else None
    sm = MetropolisSampler(state, L.random_configurations(wk.geom, 2048, 9), sweep_steps=32, seed=11, surrogate=sur,
                           thermal_steps=512, proposal=_proposal(moves))
    per_sweep = []
    for _ in range(32):
        smp = sm.sweep()
        per_sweep.append(H.Oloc(state, smp).mean().item() / wk.N)
    E, err, _ = O.jackknife(lambda m: m, np.array(per_sweep))
    print(f"delayed={delayed} {moves}: E/N {E:.6f} +- {err:.6f}, exact {E_exact:.6f}, stats {sm.stats()}")
    assert abs(E - E_exact) < 4 * err + 1e-6


def test_mix_sampler_constructors_run_the_mix():
    """MixSampler([ParticleHop(n/2, n_neighbor=(1, 2)), SiteExchange(n/2, (1, 2))]) as the reference
    builds it: the chains run the mixed proposal (replayed by the oracle), keep their particle
    numbers, and site exchanges change the doublon count only together with particle hops."""
    from neuralhubbard_amd.sampler import MixSampler, ParticleHop, SiteExchange

    wk, p, state, ref = _small_state()
    s0 = L.random_configurations(wk.geom, 8, 33)
    sm = MixSampler([ParticleHop(state, 4, n_neighbor=(1, 2)), SiteExchange(state, 4, n_neighbor=(1, 2))],
                    sweep_steps=20, initial_spins=s0, seed=77)
    assert sm.n_chains == 8 and sm.proposal.kind == O.MOVE_MIX and sm.proposal.ph_neighbors == 8
    smp = sm.sweep()
    got = smp.spins.cpu().numpy()
    N = wk.N
    assert ((got[:, :N] > 0).sum(1) == wk.n_up).all() and ((got[:, N:] > 0).sum(1) == wk.n_dn).all()
    for c in range(8):
        sf, _, tie = O.metropolis_chain(ref.psi, s0[c], c, 77, 0, 20, 4, 4, margin=1e-4, kind=O.MOVE_MIX, ph_z=8,
                                        se_z=8)
        if not tie:
            assert np.array_equal(got[c], sf), c
    # amplitudes carried by the chains equal a fresh evaluation
    psi = state(smp.spins)
    np.testing.assert_allclose(smp.psi.logabs.cpu().numpy(), psi.logabs.cpu().numpy(), rtol=0, atol=1e-9)
    # a single-move sampler started lazily
    hop = ParticleHop(state, 4, n_neighbor=1, initial_spins=s0[:4], sweep_steps=5, seed=3)
    hop.sweep()
    assert hop.stats()["proposals"] == 20

