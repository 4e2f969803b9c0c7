"""Pins of the oracle's surrogate network and Metropolis chain (CPU only).

* sur_forward: equals the same network written with torch's conv2d (circular padding) in
  float64, whose tap layout follows the documented parameter order; f is invariant under
  lattice translations; zero weights leave f = N b3.
* sur_grad_fd: equals torch autograd of that network.
* metropolis_chain: plain and delayed acceptance (with an arbitrary screening function) both
  leave |psi|^2 stationary, checked by chi-square over an exactly enumerated 3x3 space; a
  chain that accepted with |psi'/psi| instead of its square fails the same test.
* propose: over every configuration of a 3x3 lattice (2 + 1 electrons), the exact
  proposal matrix of each move kind (all Philox words reduced to the residues the moves use)
  equals the move graph built by brute force from its definition (each particle to each
  neighbour / each site with each neighbour, equal weights) and is symmetric, which is what
  plain Metropolis needs; the same for sampled rows of the mix on a 4x3 (3 + 2) lattice; a
  particle hop that draws only among empty neighbours (a plausible shortcut) is not.  Site exchange alone keeps the doublon count (so it cannot sample alone);
  the mix of particle hops and site exchanges connects the whole space, and its chain leaves
  |psi|^2 stationary by chi-square.
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


def _run_chains(psi, key, states, sur, n_chains=600, n_steps=600, burn=200, every=100, **moves):
    """Counts of 4 states per chain, 100 proposals apart (a move of one electron needs
    about 9 proposals here, so these are close to independent draws)."""
    counts = np.zeros(len(states))
    rng = np.random.default_rng(3)
    for c in range(n_chains):
        s = states[rng.integers(len(states))]
        s, _, _ = O.metropolis_chain(psi, s, c, 77, 0, burn, 3, 3, sur=sur, **moves)
        for t in range(burn, n_steps, every):
            s, _, _ = O.metropolis_chain(psi, s, c, 77, t, every, 3, 3, sur=sur, **moves)
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


# ------------------------------------------------------------------ proposal kinds

KINDS = [  # (kind, ph_z, se_z)
    (O.MOVE_EXCHANGE, 4, 4),
    (O.MOVE_PARTICLE_HOP, 4, 4),
    (O.MOVE_PARTICLE_HOP, 8, 4),
    (O.MOVE_SITE_EXCHANGE, 4, 4),
    (O.MOVE_SITE_EXCHANGE, 4, 8),
    (O.MOVE_MIX, 4, 4),
    (O.MOVE_MIX, 8, 8),
]


def _space(LX, LY, n_up, n_dn):
    import itertools

    N = LX * LY
    out = []
    for u in itertools.combinations(range(N), n_up):
        for d in itertools.combinations(range(N), n_dn):
            s = -np.ones(2 * N, np.int8)
            s[list(u)] = 1
            s[[N + x for x in d]] = 1
            out.append(s)
    return np.stack(out)


def _draws(LX, LY, n_occ, kind, ph_z, se_z):
    """Every (r0, r1) residue class the move reads, each of equal probability: r0 over
    lcm(N, n_occ) values (uniform in r0 mod N and in r0 mod n_occ), r1 over its direction bits
    (and the spin bit of the exchange, the kind bit of the mix)."""
    R0 = range(math.lcm(LX * LY, n_occ))
    if kind == O.MOVE_EXCHANGE:
        R1 = range(8)
    elif kind == O.MOVE_MIX:
        R1 = [d | (k << 4) for k in (0, 1) for d in range(max(ph_z, se_z))]
    else:
        R1 = range(ph_z if kind == O.MOVE_PARTICLE_HOP else se_z)
    return [(r0, r1) for r0 in R0 for r1 in R1]


def _proposal_rows(rows, states, LX, LY, kind, ph_z, se_z):
    """Q[a, b] = P(propose states[b] | rows[a]) over all draws (dense rows)."""
    n_occ = int((rows[0] > 0).sum())
    key = {s.tobytes(): k for k, s in enumerate(states)}
    draws = _draws(LX, LY, n_occ, kind, ph_z, se_z)
    Q = np.zeros((len(rows), len(states)))
    for a, s in enumerate(rows):
        for r0, r1 in draws:
            sp = O.propose(s, r0, r1, LX, LY, kind, ph_z, se_z)
            if sp is not None:
                Q[a, key[sp.tobytes()]] += 1.0 / len(draws)
    return Q


def _neighbours(i, z, LX, LY):
    x, y = divmod(i, LY)
    steps = [(1, 0), (-1, 0), (0, 1), (0, -1), (1, 1), (1, -1), (-1, 1), (-1, -1)][:z]
    return [((x + dx) % LX) * LY + (y + dy) % LY for dx, dy in steps]


def _brute_moves(states, LX, LY, kind, ph_z, se_z, only_empty=False):
    """The move graph from the definitions: a particle hop picks one of the n_occ particles and
    one of the z neighbouring sites of its spin (no move onto an occupied mode); a site exchange
    picks one of the N sites and one of its z neighbours and swaps the two sites' occupations;
    the mix is either with probability 1/2; the exchange picks a site, a spin and one of 4
    neighbours and swaps the two modes.  only_empty: the particle draws among its empty
    neighbours only (a different, asymmetric proposal)."""
    from collections import defaultdict

    N = LX * LY
    key = {s.tobytes(): k for k, s in enumerate(states)}
    Q = defaultdict(float)

    def hop(a, s, w, z):
        occ = np.flatnonzero(s > 0)
        for m in occ:
            off = N if m >= N else 0
            targets = [off + j for j in _neighbours(m - off, z, LX, LY)]
            if only_empty:
                targets = [t for t in targets if s[t] < 0]
                if not targets:
                    continue
                wt = w / (len(occ) * len(targets))
            else:
                wt = w / (len(occ) * z)
            for t in targets:
                if s[t] > 0:
                    continue
                sp = s.copy()
                sp[m], sp[t] = -1, 1
                Q[a, key[sp.tobytes()]] += wt

    def swap_sites(a, s, w, z):
        for i in range(N):
            for j in _neighbours(i, z, LX, LY):
                sp = s.copy()
                sp[[i, j]] = s[[j, i]]
                sp[[N + i, N + j]] = s[[N + j, N + i]]
                if not np.array_equal(sp, s):
                    Q[a, key[sp.tobytes()]] += w / (N * z)

    def exchange(a, s, w):
        for i in range(N):
            for spin in (0, 1):
                for j in _neighbours(i, 4, LX, LY):
                    mi, mj = spin * N + i, spin * N + j
                    if s[mi] != s[mj]:
                        sp = s.copy()
                        sp[[mi, mj]] = s[[mj, mi]]
                        Q[a, key[sp.tobytes()]] += w / (N * 2 * 4)

    for a, s in enumerate(states):
        if kind == O.MOVE_EXCHANGE:
            exchange(a, s, 1.0)
        if kind in (O.MOVE_PARTICLE_HOP, O.MOVE_MIX):
            hop(a, s, 1.0 if kind == O.MOVE_PARTICLE_HOP else 0.5, ph_z)
        if kind in (O.MOVE_SITE_EXCHANGE, O.MOVE_MIX):
            swap_sites(a, s, 1.0 if kind == O.MOVE_SITE_EXCHANGE else 0.5, se_z)
    from scipy import sparse

    ij = np.array(list(Q.keys()), dtype=np.int64).reshape(-1, 2)
    return sparse.csr_matrix((np.array(list(Q.values())), (ij[:, 0], ij[:, 1])), shape=(len(states), len(states)))


def _connected(Q):
    from scipy.sparse.csgraph import connected_components

    return connected_components(Q, directed=True, connection="strong")[0] == 1


def _asym(B):
    return abs(B - B.T).max() if B.nnz else 0.0


@pytest.mark.parametrize("kind,ph_z,se_z", KINDS)
def test_oracle_proposals_match_brute_force_and_are_symmetric(kind, ph_z, se_z):
    LX, LY = 3, 3
    N = LX * LY
    states = _space(LX, LY, 2, 1)
    Q = _proposal_rows(states, states, LX, LY, kind, ph_z, se_z)
    B = _brute_moves(states, LX, LY, kind, ph_z, se_z)
    np.testing.assert_allclose(Q, B.toarray(), atol=1e-15)
    np.testing.assert_allclose(Q, Q.T, atol=1e-15)
    assert (Q.sum(1) <= 1 + 1e-12).all()
    dbl = ((states[:, :N] > 0) & (states[:, N:] > 0)).sum(1)
    a, b = np.nonzero(Q)
    if kind == O.MOVE_SITE_EXCHANGE:
        assert (dbl[a] == dbl[b]).all() and not _connected(B)
    if kind == O.MOVE_MIX:
        assert (dbl[a] != dbl[b]).any() and _connected(B)
        # one step of the mix swaps an up and a down electron of two neighbouring sites
        assert any((states[x, :N] != states[y, :N]).sum() == 2 and (states[x, N:] != states[y, N:]).sum() == 2
                   for x, y in zip(a, b))


@pytest.mark.parametrize("kind,ph_z,se_z", [(O.MOVE_MIX, 8, 8), (O.MOVE_MIX, 4, 8)])
def test_oracle_proposals_4x3_rows_match_brute_force(kind, ph_z, se_z):
    """A larger space (4x3, 3 + 2 electrons, 14 520 states): the brute-force matrix is symmetric
    and connected, and every 97th row of the drawn proposal matrix equals it."""
    LX, LY = 4, 3
    states = _space(LX, LY, 3, 2)
    B = _brute_moves(states, LX, LY, kind, ph_z, se_z)
    assert _asym(B) < 1e-15 and _connected(B)
    rows = np.arange(0, len(states), 97)
    Q = _proposal_rows(states[rows], states, LX, LY, kind, ph_z, se_z)
    np.testing.assert_allclose(Q, B[rows].toarray(), atol=1e-15)


def test_oracle_proposal_drawing_only_empty_neighbours_is_not_symmetric():
    states = _space(3, 3, 2, 1)
    B = _brute_moves(states, 3, 3, O.MOVE_PARTICLE_HOP, 4, 4, only_empty=True)
    assert _asym(B) > 1e-3


@pytest.mark.parametrize("kind,ph_z,se_z", [(O.MOVE_MIX, 8, 8), (O.MOVE_MIX, 4, 4)])
def test_oracle_proposals_mixed_chain_stationary(kind, ph_z, se_z):
    states, sign, logabs = _space_3x3()
    key = {s.tobytes(): k for k, s in enumerate(states)}
    psi = lambda s: (sign[key[s.tobytes()]], logabs[key[s.tobytes()]])
    prob = np.exp(2 * logabs)
    prob /= prob.sum()
    counts = _run_chains(psi, key, states, None, kind=kind, ph_z=ph_z, se_z=se_z)
    assert _chi2_pvalue(counts, prob) > 1e-4

