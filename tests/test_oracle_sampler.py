"""Pins of the oracle's surrogate network and Metropolis chain (CPU only).

* sur_forward: equals the same network written with torch's conv2d (circular padding) in
  float64, whose tap layout follows the documented parameter order; f is invariant under
  lattice translations; zero weights leave f = N b3.
* sur_grad_fd: equals torch autograd of that network.
* metropolis_chain: plain and delayed acceptance (with an arbitrary screening function) both
  leave |psi|^2 stationary, checked by chi-square over an exactly enumerated 3x3 space; a
  chain that accepted with |psi'/psi| instead of its square fails the same test.
"""

import math

import numpy as np
import pytest

from oracle import oracle as O

torch = pytest.importorskip("torch")


def torch_surrogate(p, S, LX, LY, C):
    """Same network in torch (float64): params [w1 (3,3,2,C) | b1 | w2 (3,3,C,C) | b2 |
    w3 (3,3,C,1) | b3], tap (kx, ky) <-> displacement (kx-1, ky-1), periodic."""
    p = torch.as_tensor(p, dtype=torch.float64)
    sizes = [9 * 2 * C, C, 9 * C * C, C, 9 * C, 1]
    w1, b1, w2, b2, w3, b3 = torch.split(p, sizes)
    x = torch.as_tensor(np.asarray(S, np.float64)).reshape(-1, 2, LX, LY)

    def conv(h, w, cin, cout, b):
        k = w.reshape(3, 3, cin, cout).permute(3, 2, 0, 1)
        return torch.nn.functional.conv2d(torch.nn.functional.pad(h, (1, 1, 1, 1), mode="circular"), k, b)

    def gelu(z):
        return 0.5 * z * (1 + torch.tanh(math.sqrt(2 / math.pi) * (z + 0.044715 * z**3)))

    z1 = conv(x, w1, 2, C, b1)
    z2 = conv(gelu(z1), w2, C, C, b2)
    z3 = conv(gelu(z2), w3, C, 1, b3)
    return z3.sum(dim=(1, 2, 3))


def _params(C, seed):
    rng = np.random.default_rng(seed)
    n = 9 * 2 * C + C + 9 * C * C + C + 9 * C + 1
    return rng.standard_normal(n) * 0.3


def _configs(LX, LY, n, seed):
    rng = np.random.default_rng(seed)
    return np.where(rng.random((n, 2 * LX * LY)) < 0.4, 1, -1).astype(np.int8)


@pytest.mark.parametrize("LX,LY", [(4, 4), (3, 5)])
def test_oracle_surrogate_forward(LX, LY):
    C = 4
    p = _params(C, 1)
    S = _configs(LX, LY, 3, 2)
    ref = torch_surrogate(p, S, LX, LY, C).numpy()
    got = np.array([O.sur_forward(p, s, LX, LY, C) for s in S])
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12)
    # translation invariance of the site sum
    s = S[0].reshape(2, LX, LY)
    st = np.roll(s, shift=(1, 2), axis=(1, 2)).reshape(-1)
    assert abs(O.sur_forward(p, st, LX, LY, C) - got[0]) < 1e-10
    # zero weights: f = N b3
    q = np.zeros_like(p)
    q[-1] = 0.7
    assert abs(O.sur_forward(q, S[1], LX, LY, C) - 0.7 * LX * LY) < 1e-12


def test_oracle_surrogate_gradient():
    LX, LY, C = 3, 4, 3
    p = _params(C, 4)
    S = _configs(LX, LY, 2, 5)
    T = np.array([0.4, -1.2])
    w = 0.5
    pt = torch.tensor(p, requires_grad=True)
    f = torch_surrogate(pt, S, LX, LY, C)
    loss = (0.5 * w * (f - torch.as_tensor(T)) ** 2).sum()
    loss.backward()
    assert abs(loss.item() - O.sur_loss(p, S, T, LX, LY, C, w)) < 1e-10
    idx = list(range(0, p.size, 7)) + [p.size - 1]
    fd = O.sur_grad_fd(p, S, T, LX, LY, C, w, idx)
    np.testing.assert_allclose(fd, pt.grad.numpy()[idx], rtol=1e-6, atol=1e-8)


def _space_3x3():
    """All configurations of a 3x3 lattice with one up and one down electron (81 states)
    and a fixed, deliberately rough amplitude table."""
    N = 9
    states = []
    for u in range(N):
        for d in range(N):
            s = -np.ones(2 * N, np.int8)
            s[u] = 1
            s[N + d] = 1
            states.append(s)
    rng = np.random.default_rng(11)
    logabs = rng.normal(0, 0.6, len(states))
    sign = np.where(rng.random(len(states)) < 0.5, -1.0, 1.0)
    return np.stack(states), sign, logabs


def _chi2_pvalue(counts, prob):
    from scipy import stats as sps

    exp = prob * counts.sum()
    chi2 = ((counts - exp) ** 2 / exp).sum()
    return sps.chi2.sf(chi2, len(counts) - 1)


def _run_chains(psi, key, states, sur, n_chains=600, n_steps=600, burn=200, every=100):
    """Counts of 4 states per chain, 100 proposals apart (a move of one electron needs
    about 9 proposals here, so these are close to independent draws)."""
    counts = np.zeros(len(states))
    rng = np.random.default_rng(3)
    for c in range(n_chains):
        s = states[rng.integers(len(states))]
        s, _, _ = O.metropolis_chain(psi, s, c, 77, 0, burn, 3, 3, sur=sur)
        for t in range(burn, n_steps, every):
            s, _, _ = O.metropolis_chain(psi, s, c, 77, t, every, 3, 3, sur=sur)
            counts[key[s.tobytes()]] += 1
    return counts


@pytest.mark.parametrize("delayed", [False, True])
def test_oracle_metropolis_stationary(delayed):
    states, sign, logabs = _space_3x3()
    key = {s.tobytes(): k for k, s in enumerate(states)}
    psi = lambda s: (sign[key[s.tobytes()]], logabs[key[s.tobytes()]])
    # the screening function only has to be a function of s: a noisy imitation of log|psi|
    noise = np.random.default_rng(5).normal(0, 0.5, len(states))
    sur = (lambda s: logabs[key[s.tobytes()]] + noise[key[s.tobytes()]]) if delayed else None
    prob = np.exp(2 * logabs)
    prob /= prob.sum()
    counts = _run_chains(psi, key, states, sur)
    assert _chi2_pvalue(counts, prob) > 1e-4
    # a wrong target (|psi| instead of |psi|^2) is detected by the same statistic
    half = lambda s: (sign[key[s.tobytes()]], 0.5 * logabs[key[s.tobytes()]])
    if not delayed:
        wrong = _run_chains(half, key, states, None)
        assert _chi2_pvalue(wrong, prob) < 1e-6


def test_oracle_metropolis_zero_amplitude_is_never_entered():
    states, sign, logabs = _space_3x3()
    key = {s.tobytes(): k for k, s in enumerate(states)}
    sign = sign.copy()
    sign[::3] = 0.0
    logabs = logabs.copy()
    logabs[::3] = -np.inf
    psi = lambda s: (sign[key[s.tobytes()]], logabs[key[s.tobytes()]])
    s, acc, _ = O.metropolis_chain(psi, states[1], 0, 5, 0, 400, 3, 3)
    assert sign[key[s.tobytes()]] != 0 and acc > 0
