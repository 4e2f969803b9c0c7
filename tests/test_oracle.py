"""Pins of the FP64 oracle: closed forms, brute force and independent constructions.

The reference ships no tests or golden vectors for this path (its ``.gitignore:6-7`` drops
``code/tests``), so every oracle function is checked against mathematics that fixes it, chosen
so that a dropped term, a wrong sign or index, or a transposed operand fails.
"""

import itertools
import math
from fractions import Fraction

import numpy as np
import pytest
import scipy.ndimage

from oracle import oracle as O
from neuralhubbard_amd import lattice as L
from neuralhubbard_amd.configs import make_params, small_workload


def rand_skew(n, rng, scale=1.0):
    A = rng.standard_normal((n, n)) * scale
    return np.triu(A, 1) - np.triu(A, 1).T


# ----------------------------------------------------------------- tables


def test_oracle_tables():
    for LX, LY in [(4, 4), (3, 5), (8, 8), (16, 16)]:
        disp, neg, dist = O.lattice_tables(LX, LY)
        N = LX * LY
        for i in range(0, N, max(1, N // 13)):
            for j in range(0, N, max(1, N // 11)):
                xi, yi = divmod(i, LY)
                xj, yj = divmod(j, LY)
                r = ((xi - xj) % LX) * LY + (yi - yj) % LY
                assert disp[i, j] == r
        for r in range(N):
            x, y = divmod(r, LY)
            assert neg[r] == ((-x) % LX) * LY + (-y) % LY
            assert disp[r, 0] == r and disp[0, r] == neg[r]
            best = min(math.hypot(x + a * LX, y + b * LY) for a in (-1, 0, 1) for b in (-1, 0, 1))
            assert abs(dist[r] - best) < 1e-12
        # product table agrees with the package table used by the host
        d2, n2, di2 = L.displacement_tables(LX, LY)
        assert np.array_equal(d2, disp) and np.array_equal(n2, neg) and np.allclose(di2, dist)


def test_group_perms_involutions_and_group_law():
    LX, LY = 4, 6
    P = [O.group_perm(g, LX, LY) for g in range(8)]
    ident = np.arange(2 * LX * LY)
    for g in range(8):
        assert np.array_equal(P[g][P[g]], ident)
        for h in range(8):
            assert np.array_equal(P[g][P[h]], P[g ^ h])
    assert np.array_equal(np.asarray(L.group_mode_perms(LX, LY)), np.stack(P))


# ----------------------------------------------------------------- network


def test_oracle_conv():
    rng = np.random.default_rng(0)
    h = rng.standard_normal((3, 5, 7))
    w = rng.standard_normal((4, 3, 3, 3))
    b = rng.standard_normal(4)
    out = O.conv3x3(h, w, b)
    for o in range(4):
        ref = b[o] + sum(
            scipy.ndimage.correlate(h[c], w[o, c], mode="wrap") for c in range(3)
        )
        np.testing.assert_allclose(out[o], ref, rtol=1e-12, atol=1e-12)
    # a delta kernel at (dx, dy) is a periodic shift
    w = np.zeros((1, 1, 3, 3))
    w[0, 0, 2, 0] = 1.0  # dx = +1, dy = -1
    out = O.conv3x3(h[:1], w, [0.0])
    np.testing.assert_array_equal(out[0], np.roll(h[0], (-1, 1), axis=(0, 1)))


def test_features_translation_equivariant():
    wk = small_workload(LX=4, LY=4, channels=4, n_layers=2)
    p = make_params(wk)
    s = L.random_configurations(wk.geom, 1, 3)[0]
    y = O.features(s, p.conv_w, p.conv_b, 4, 4)
    st = np.roll(s.reshape(2, 4, 4), (1, 2), axis=(1, 2)).reshape(-1)
    yt = O.features(st, p.conv_w, p.conv_b, 4, 4)
    np.testing.assert_allclose(
        np.roll(y.reshape(2, 4, 4, -1), (1, 2), axis=(1, 2)).reshape(y.shape), yt, atol=1e-12
    )


def test_gelu_values():
    assert O.gelu(0.0) == 0.0
    np.testing.assert_allclose(O.gelu(np.array([1.0, -1.0, 3.0])),
                               [0.8411919906082768, -0.15880800939172324, 2.99636259], rtol=1e-7)


# ----------------------------------------------------------------- W tilde and pairing


def test_oracle_wtilde():
    rng = np.random.default_rng(1)
    LX, LY, d = 4, 6, 3
    N = LX * LY
    W = rng.standard_normal((3, N, d))
    Wt = O.build_wtilde(W, 0.55, LX, LY)
    _, neg, dist = O.lattice_tables(LX, LY)
    for sp in range(2):
        for sq in range(2):
            # Wt[2 sq + sp, -r] = -Wt[2 sp + sq, r] exactly (pf_kernel.py:111-113)
            np.testing.assert_array_equal(Wt[2 * sq + sp][neg], -Wt[2 * sp + sq])
    # envelope normalisation sum_ij exp(-lam D_ij) = N sum_r exp(-lam D_r)
    lam = math.log1p(math.exp(0.55))
    disp, _, _ = O.lattice_tables(LX, LY)
    tot = np.exp(-lam * dist[disp]).sum()
    np.testing.assert_allclose(Wt[1] / W[0], np.broadcast_to(np.exp(-lam * dist)[:, None] / tot, (N, d)),
                               rtol=1e-12)


@pytest.mark.parametrize("LX,LY", [(4, 4), (3, 5)])
def test_oracle_pairing(LX, LY):
    """Gather path == the reference's full-matrix einsum path, sliced (pf_kernel.py:388)."""
    rng = np.random.default_rng(2)
    N, d = LX * LY, 3
    W = rng.standard_normal((3, N, d))
    y = rng.standard_normal((2, N, d))
    Ffull = O.pairing_full(y, W, 0.3, LX, LY)
    Wt = O.build_wtilde(W, 0.3, LX, LY)
    idx = np.sort(rng.choice(2 * N, 9, replace=False))
    F = O.pairing_occ(y.reshape(2 * N, d), idx, Wt, LX, LY)
    np.testing.assert_allclose(F, Ffull[np.ix_(idx, idx)], rtol=1e-12, atol=1e-15)
    np.testing.assert_array_equal(F, -F.T)
    # explicit triple loop
    disp, _, _ = O.lattice_tables(LX, LY)
    p, q = idx[1], idx[7]
    val = sum(y.reshape(2 * N, d)[p, a] * y.reshape(2 * N, d)[q, a]
              * Wt[2 * (p // N) + q // N, disp[p % N, q % N], a] for a in range(d))
    assert abs(F[1, 7] - val) < 1e-14


# ----------------------------------------------------------------- Pfaffian / determinant


def pf_exact(A):
    """Exact Pfaffian by expansion along the first row, in rationals."""
    n = len(A)
    if n == 0:
        return Fraction(1)
    if n % 2:
        return Fraction(0)
    tot = Fraction(0)
    for j in range(1, n):
        if A[0][j] == 0:
            continue
        rest = [k for k in range(1, n) if k != j]
        sub = [[A[a][b] for b in rest] for a in rest]
        tot += (-1) ** (j + 1) * A[0][j] * pf_exact(sub)
    return tot


def test_oracle_pfaffian_exact_rational():
    rng = np.random.default_rng(3)
    for n in (2, 4, 6, 8, 10):
        M = rng.integers(-5, 6, size=(n, n))
        M = np.triu(M, 1) - np.triu(M, 1).T
        exact = pf_exact([[Fraction(int(v)) for v in row] for row in M])
        sg, la = O.slogpf(M.astype(float))
        if exact == 0:
            assert sg == 0 or la < -20
            continue
        assert sg == (1 if exact > 0 else -1)
        assert abs(la - math.log(abs(float(exact)))) < 1e-10


def test_oracle_pfaffian_identities():
    rng = np.random.default_rng(4)
    for n in (2, 8, 56, 120):
        A = rand_skew(n, rng)
        sg, la = O.slogpf(A)
        ds, dl = np.linalg.slogdet(A)
        assert ds > 0 and abs(2 * la - dl) < 1e-8 * max(1, abs(dl))
        B = rng.standard_normal((n, n))
        s2, l2 = O.slogpf(B @ A @ B.T)
        bs, bl = np.linalg.slogdet(B)
        assert s2 == sg * bs and abs(l2 - (la + bl)) < 1e-8 * max(1, abs(l2))
    a = rng.standard_normal(5)
    A = np.zeros((10, 10))
    for k in range(5):
        A[2 * k, 2 * k + 1], A[2 * k + 1, 2 * k] = a[k], -a[k]
    sg, la = O.slogpf(A)
    assert sg == np.sign(np.prod(a)) and abs(la - np.log(np.abs(a)).sum()) < 1e-12
    # swapping two rows/columns flips the sign
    P = np.eye(10)[[1, 0] + list(range(2, 10))]
    s3, l3 = O.slogpf(P @ A @ P.T)
    assert s3 == -sg and abs(l3 - la) < 1e-12
    assert O.slogpf(np.zeros((4, 4)))[0] == 0


def test_oracle_det():
    rng = np.random.default_rng(5)
    for n in (1, 3, 28, 112):
        A = rng.standard_normal((n, n))
        s, l = O.slogdet(A)
        s2, l2 = np.linalg.slogdet(A)
        assert s == s2 and abs(l - l2) < 1e-9 * max(1, abs(l2))


# ----------------------------------------------------------------- fermion algebra on bit strings


def apply_cdag_c(occ, j, i):
    """c+_j c_i |occ> for a sorted tuple of occupied modes; returns (sign, new) or None."""
    if i not in occ:
        return None
    k = occ.index(i)
    sign = (-1) ** k
    rest = occ[:k] + occ[k + 1 :]
    if j in rest:
        return None
    pos = sum(1 for m in rest if m < j)
    sign *= (-1) ** pos
    return sign, tuple(sorted(rest + (j,)))


def permute_state(occ, perm):
    """U_g c+_{m1} ... c+_{mn} |0> = c+_{pi m1} ... c+_{pi mn} |0>, sorted with its sign."""
    mapped = [int(perm[m]) for m in occ]
    inv = sum(1 for a in range(len(mapped)) for b in range(a + 1, len(mapped)) if mapped[a] > mapped[b])
    return (-1) ** inv, tuple(sorted(mapped))


def config_of(occ, N2):
    s = -np.ones(N2, np.int8)
    s[list(occ)] = 1
    return s


def hamiltonian_row(occ, LX, LY, t, tp, U):
    """<s|H|s'> for all s' by operator algebra: H = -sum t_ij (c+_i c_j + c+_j c_i) + U sum n n."""
    N = LX * LY
    bonds = set()
    for i in range(N):
        x, y = divmod(i, LY)
        for (dx, dy), tt in [((1, 0), t), ((0, 1), t), ((1, 1), tp), ((1, -1), tp)]:
            if tt == 0:
                continue
            j = ((x + dx) % LX) * LY + (y + dy) % LY
            if j != i:
                bonds.add((min(i, j), max(i, j), tt))
    row = {}
    for i, j, tt in bonds:
        for sp in range(2):
            for a, b in ((i, j), (j, i)):
                r = apply_cdag_c(occ, sp * N + b, sp * N + a)
                if r is not None:
                    row[r[1]] = row.get(r[1], 0.0) - tt * r[0]
    diag = U * sum(1 for m in occ if m < N and (m + N) in occ)
    return diag, row


@pytest.mark.parametrize("kind", ["pfaffian", "bcs"])
def test_oracle_eloc_matches_operator_algebra(kind):
    wk = small_workload(kind=kind, LX=4, LY=4, n_up=3, n_dn=1 if kind == "pfaffian" else 3,
                        tprime=-0.3, channels=4, n_layers=1, base_f0=(kind == "pfaffian"))
    p = make_params(wk)
    m = O.Model(wk, p)
    rng = np.random.default_rng(6)
    for trial in range(2):
        s = L.random_configurations(wk.geom, 1, 100 + trial)[0]
        occ = tuple(np.flatnonzero(s > 0))
        diag, row = hamiltonian_row(occ, 4, 4, 1.0, -0.3, 8.0)
        ps = m.psi(s)
        ref = diag
        for occ2, h in row.items():
            sg, la = m.psi(config_of(occ2, 32))
            ref += h * sg * ps[0] * math.exp(la - ps[1])
        e = O.eloc(m, s, 8.0, 1.0, -0.3)
        assert abs(e - ref) < 1e-9 * max(1, abs(ref)), (e, ref)
        # same set of connected states, with the same matrix elements
        hl = O.hops(s, 4, 4, 1.0, -0.3)
        assert len(hl) == len(row)
        for a, b, h in hl:
            occ2 = tuple(sorted(set(occ) - {a} | {b}))
            assert abs(row[occ2] - h) < 1e-15


def test_oracle_projection_invariance():
    """U_g |psi_P> = |psi_P>: psi_P(s) = eps * psi_P(g.s) with eps from operator algebra, and
    psi_P = sum_g eps_g(s) psi(g.s) computed with explicitly sorted modes."""
    wk = small_workload(LX=4, LY=4, n_up=2, n_dn=2, channels=4, n_layers=1, base_f0=True)
    p = make_params(wk)
    m = O.Model(wk, p)
    s = L.random_configurations(wk.geom, 1, 11)[0]
    occ = tuple(np.flatnonzero(s > 0))
    ps = m.psi(s)
    terms = []
    for g in range(8):
        perm = O.group_perm(g, 4, 4)
        eps, occ_g = permute_state(occ, perm)
        pg = m.psi(config_of(occ_g, 32))
        assert pg[0] * eps == ps[0] and abs(pg[1] - ps[1]) < 1e-9
        # explicit: psi(g.s) from the sorted modes of g.s
        sgs = config_of(occ_g, 32)
        F = O.pairing_occ(m.features(sgs).reshape(32, -1), np.array(occ_g), m.Wt, 4, 4)
        F = F + p.F0.astype(np.float64)[np.ix_(occ_g, occ_g)]
        sg, la = O.slogpf(F)
        terms.append((eps * sg, la))
    sg, la = O.signed_logsumexp([t[0] for t in terms], [t[1] for t in terms])
    assert sg == ps[0] and abs(la - ps[1]) < 1e-10


def test_oracle_bcs_is_pfaffian_of_block_matrix():
    """The BCS image amplitude equals Pf of the 2N x 2N [[0, Phi], [-Phi^T, 0]] over the
    image's modes in the order of s (spin-down-first images included)."""
    wk = small_workload(kind="bcs", LX=4, LY=4, n_up=3, n_dn=3, tprime=0.0, channels=4, n_layers=1)
    p = make_params(wk)
    m = O.Model(wk, p)
    s = L.random_configurations(wk.geom, 1, 12)[0]
    N = 16
    disp, _, _ = O.lattice_tables(4, 4)
    for g in range(8):
        perm = O.group_perm(g, 4, 4)
        gs = np.empty_like(s)
        gs[perm] = s
        y = m.features(gs)
        Phi = np.einsum("pa,pqa,qa->pq", y[0], m.Wt[disp], y[1]) + p.phi0.astype(np.float64)[disp]
        A = np.zeros((2 * N, 2 * N))
        A[:N, N:] = Phi
        A[N:, :N] = -Phi.T
        modes = perm[np.flatnonzero(s > 0)]
        sg, la = O.slogpf(A[np.ix_(modes, modes)])
        ne = 6
        sg *= (-1) ** (ne * (ne - 1) // 2)
        s2, l2 = m.image_amplitude(s, g)
        assert s2 == sg and abs(l2 - la) < 1e-10


def test_oracle_variational_identity_3x3():
    """sum_s |psi|^2 E_loc / sum |psi|^2 = <psi|H|psi> / <psi|psi> with H from operator algebra
    over the whole Hilbert space of a 3x3 lattice with 2 + 2 electrons."""
    wk = small_workload(LX=3, LY=3, n_up=2, n_dn=2, tprime=-0.2, channels=2, n_layers=1)
    p = make_params(wk)
    m = O.Model(wk, p)
    N = 9
    basis = [tuple(sorted(u + tuple(N + x for x in d)))
             for u in itertools.combinations(range(N), 2) for d in itertools.combinations(range(N), 2)]
    index = {b: k for k, b in enumerate(basis)}
    psi = np.array([(lambda r: r[0] * math.exp(r[1]))(m.psi(config_of(b, 2 * N), project=False))
                    for b in basis])
    H = np.zeros((len(basis), len(basis)))
    for k, b in enumerate(basis):
        diag, row = hamiltonian_row(b, 3, 3, 1.0, -0.2, 8.0)
        H[k, k] += diag
        for b2, h in row.items():
            H[k, index[b2]] += h
    assert np.allclose(H, H.T)
    E_exact = psi @ H @ psi / (psi @ psi)
    el = np.array([O.eloc(m, config_of(b, 2 * N), 8.0, 1.0, -0.2, project=False) for b in basis])
    E_mc = np.sum(psi**2 * el) / np.sum(psi**2)
    assert abs(E_mc - E_exact) < 1e-10


# ----------------------------------------------------------------- RNG and statistics


def test_oracle_philox():
    # Random123 known-answer vectors for philox4x32-10
    assert O.philox4x32([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert O.philox4x32([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert O.philox4x32([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == [
        0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_oracle_jackknife():
    rng = np.random.default_rng(7)
    x = rng.standard_normal(64)
    v, e, errs = O.jackknife(lambda m: m, x)
    assert abs(v - x.mean()) < 1e-15
    assert abs(errs[1] - x.std(ddof=1) / math.sqrt(64)) < 1e-12
    xb = x.reshape(8, 8).mean(1)
    assert abs(errs[8] - xb.std(ddof=1) / math.sqrt(8)) < 1e-12
    assert e == max(errs.values())


def test_product_energy_statistics_match_oracle():
    """neuralhubbard_amd.measure (host statistics) against the oracle's jackknife, and the
    closed form for i.i.d. sweeps: error with blocks of 1 = std(ddof=1)/sqrt(n)."""
    from neuralhubbard_amd import measure as M

    rng = np.random.default_rng(3)
    Em = rng.normal(-0.7 * 256, 2.0, 64)
    Eq = Em**2 + rng.uniform(50, 60, 64)
    r = M.energy(Em, Eq, 256)
    v, e, errs = O.jackknife(lambda m: m / 256, Em)
    assert abs(r["E_per_site"] - v) < 1e-12 and abs(r["E_err"] - e) < 1e-12
    assert set(r["E_err_blocks"]) == {1, 2, 4, 8}
    assert abs(r["E_err_blocks"][1] - Em.std(ddof=1) / np.sqrt(64) / 256) < 1e-12
    vv, ve, _ = O.jackknife(lambda m, q: (q - m * m) / 256, Em, Eq)
    assert abs(r["var_per_site"] - vv) < 1e-9 and abs(r["var_err"] - ve) < 1e-9
    few = M.energy(Em[:7], Eq[:7], 256)
    assert np.isnan(few["E_err"])


# ------------------------------------------------------------------------ VJPs (f2)


def _valid_wt(LX, LY, d, seed):
    rng = np.random.default_rng(seed)
    return O.build_wtilde(rng.standard_normal((3, LX * LY, d)), 0.3, LX, LY)


def test_oracle_vjp_pairing_finite_differences():
    """<G, F(Y + h dY, Wt + h dW)> - <G, F(Y - h dY, ...)> over 2h against <dY, VJP_Y> + <dW, VJP_W>,
    with random directions, for both the table argument and the features."""
    rng = np.random.default_rng(3)
    LX, LY, d = 4, 3, 4
    N = LX * LY
    Wt = _valid_wt(LX, LY, d, 4)
    modes = np.sort(rng.choice(2 * N, 10, replace=False))
    Y = rng.standard_normal((10, d))
    G = rng.standard_normal((10, 10))
    dY, dWt = O.pairing_occ_vjp(Y, modes, Wt, G, LX, LY)
    for _ in range(3):
        vY = rng.standard_normal(Y.shape)
        vW = rng.standard_normal(Wt.shape)
        h = 1e-6

        def f(t):
            Yfull = np.zeros((2 * N, d))  # pairing_occ takes features by mode
            Yfull[modes] = Y + t * vY
            return np.sum(G * O.pairing_occ(Yfull, modes, Wt + t * vW, LX, LY))

        fd = (f(h) - f(-h)) / (2 * h)
        assert fd == pytest.approx(np.sum(dY * vY) + np.sum(dWt * vW), rel=1e-7, abs=1e-9)
    # the reference's dY form G - G^T holds for tables with the antisymmetry identity
    disp, _, _ = O.lattice_tables(LX, LY)
    sp, site = np.divmod(modes, N)
    Wpq = Wt[2 * sp[:, None] + sp[None, :], disp[site[:, None], site[None, :]]]
    np.testing.assert_allclose(np.einsum("pq,pqa,qa->pa", G - G.T, Wpq, Y), dY, rtol=1e-12, atol=1e-12)


def test_oracle_vjp_bcs_pairing_finite_differences():
    """<G, F_ud(Yu + h vu, Yd + h vd, Wt + h vW)> by central differences against
    <vu, dYu> + <vd, dYd> + <vW, dWt>; dWt also against a direct loop over the pairs."""
    rng = np.random.default_rng(8)
    LX, LY, d = 4, 3, 3
    N = LX * LY
    iu = np.sort(rng.choice(N, 5, replace=False))
    idn = np.sort(rng.choice(N, 4, replace=False))
    Yu, Yd = rng.standard_normal((5, d)), rng.standard_normal((4, d))
    Wt = rng.standard_normal((N, d))
    G = rng.standard_normal((5, 4))
    dYu, dYd, dWt = O.pairing_ud_vjp(Yu, Yd, iu, idn, Wt, G, LX, LY)
    for _ in range(3):
        vu, vd, vW = rng.standard_normal(Yu.shape), rng.standard_normal(Yd.shape), rng.standard_normal(Wt.shape)
        h = 1e-6

        def f(t):
            return np.sum(G * O.pairing_ud(Yu + t * vu, Yd + t * vd, iu, idn, Wt + t * vW, LX, LY))

        fd = (f(h) - f(-h)) / (2 * h)
        assert fd == pytest.approx(np.sum(dYu * vu) + np.sum(dYd * vd) + np.sum(dWt * vW), rel=1e-7, abs=1e-9)
    ref = np.zeros((N, d))
    for p in range(5):
        for q in range(4):
            (xp, yp), (xq, yq) = divmod(int(iu[p]), LY), divmod(int(idn[q]), LY)
            ref[((xp - xq) % LX) * LY + (yp - yq) % LY] += G[p, q] * Yu[p] * Yd[q]
    np.testing.assert_allclose(dWt, ref, rtol=1e-12, atol=1e-12)


def test_oracle_vjp_determinant_finite_differences():
    rng = np.random.default_rng(9)
    n = 9
    A = rng.standard_normal((n, n))
    G = O.slogdet_vjp(A)
    for _ in range(3):
        V = rng.standard_normal((n, n))
        h = 1e-6
        fd = (O.slogdet(A + h * V)[1] - O.slogdet(A - h * V)[1]) / (2 * h)
        assert fd == pytest.approx(np.sum(G * V), rel=1e-6, abs=1e-8)


def test_oracle_vjp_pfaffian_finite_differences():
    rng = np.random.default_rng(5)
    n = 12
    A = rng.standard_normal((n, n))
    A = A - A.T
    G = O.slogpf_vjp(A)
    np.testing.assert_allclose(G, -G.T, atol=1e-12)
    for _ in range(3):
        V = rng.standard_normal((n, n))
        V = V - V.T
        h = 1e-6
        fd = (O.slogpf(A + h * V)[1] - O.slogpf(A - h * V)[1]) / (2 * h)
        assert fd == pytest.approx(np.sum(G * V), rel=1e-6, abs=1e-8)


# ------------------------------------------------------------------------ spin correlations (f4)


def _fock_apply(v, m, create, nmodes):
    """c_m or c+_m on a dense Fock-space vector, Jordan-Wigner by bit arithmetic: basis index
    sum_k n_k 2^(nmodes-1-k) (mode 0 the most significant bit), sign (-1)^(occupied modes < m)."""
    x = np.arange(v.size, dtype=np.int64)
    bit = np.int64(1) << (nmodes - 1 - m)
    above = (np.int64(1) << nmodes) - (bit << 1)  # the bits of the modes below m
    sign = 1 - 2 * (np.bitwise_count(x & above).astype(np.int64) & 1)
    has = (x & bit) != 0
    out = np.zeros_like(v)
    src = has if not create else ~has
    dst = x[src] ^ bit
    out[dst] = sign[src] * v[src]
    return out


def test_oracle_transverse_matches_jordan_wigner():
    """The exact |psi|^2 average of the oracle's local values equals <psi|O|psi> / <psi|psi> with
    O = S+_a S-_j applied to the dense Fock vector by an independent Jordan-Wigner routine, for
    every j and both terms, on a 3x3 lattice with 2 + 2 electrons; and the two terms agree
    (<S+_a S-_j> = <S-_a S+_j> for real psi)."""
    import itertools

    LX = LY = 3
    N = 9
    wk = small_workload(LX=LX, LY=LY, n_up=2, n_dn=2, tprime=0.0, channels=4, n_layers=1, base_f0=True, seed=4)
    p = make_params(wk)
    m = O.Model(wk, p)
    basis = [u + tuple(N + d for d in dd) for u in itertools.combinations(range(N), 2)
             for dd in itertools.combinations(range(N), 2)]
    table = {}
    for occ in basis:
        table[occ] = m.psi(config_of(occ, 2 * N))

    def psi_fn(s):
        return table[tuple(int(x) for x in np.flatnonzero(s > 0))]

    nm = 2 * N
    v = np.zeros(1 << nm)
    for occ, (sg, la) in table.items():
        v[sum(1 << (nm - 1 - k) for k in occ)] = sg * math.exp(la)
    norm = v @ v
    w = {occ: math.exp(2 * (la - max(t[1] for t in table.values()))) for occ, (sg, la) in table.items()}
    wsum = sum(w.values())
    for a in (0, 4):
        # psi_P vanishes on configurations that a group element maps to minus themselves: weight 0
        est = sum(w[occ] * O.transverse_local(config_of(occ, nm), a, LX, LY, psi_fn) for occ in basis
                  if table[occ][0] != 0) / wsum
        ax, ay = divmod(a, LY)
        for j in range(N):
            r = ((j // LY - ax) % LX) * LY + (j % LY - ay) % LY
            for chan, (i, k) in ((0, (a, j)), (1, (j, a))):  # S+_i S-_k: row 0 S+_a S-_j, row 1 S-_a S+_j
                u = v
                for mode, create in reversed(O.spin_flip(i, k, N)):
                    u = _fock_apply(u, mode, create, nm)
                exact = (v @ u) / norm
                assert abs(est[chan, r] - exact) < 1e-10, (a, j, chan, est[chan, r], exact)
        np.testing.assert_allclose(est[0], est[1], atol=1e-10)
    # S^z S^z: direct sums against the FFT form
    S = np.stack([config_of(occ, nm) for occ in basis[:50]])
    from neuralhubbard_amd.observables import autocorr, sz_grid
    import torch

    t = torch.tensor(S)
    np.testing.assert_allclose(autocorr(sz_grid(t, N), LX, LY).numpy(), O.sz_autocorr(S, LX, LY), atol=1e-12)


# ------------------------------------------------------------------------ d-wave pair terms (f4)


def _dwave_exact_setup(kind="pfaffian"):
    LX = LY = 3
    N = 9
    # no base orbital matrix F0: the state is then translation invariant (periodic convolutions,
    # displacement-indexed pairing), which the translation fill of the assembly assumes
    wk = small_workload(kind=kind, LX=LX, LY=LY, n_up=2, n_dn=2, tprime=-0.2, channels=4, n_layers=1,
                        base_f0=False, seed=6)
    m = O.Model(wk, make_params(wk))
    basis = [u + tuple(N + d for d in dd) for u in itertools.combinations(range(N), 2)
             for dd in itertools.combinations(range(N), 2)]
    table = {occ: m.psi(config_of(occ, 2 * N)) for occ in basis}
    nm = 2 * N
    v = np.zeros(1 << nm)
    for occ, (sg, la) in table.items():
        v[sum(1 << (nm - 1 - k) for k in occ)] = sg * math.exp(la)
    return LX, LY, N, basis, table, v


def _fock_expect(v, ops, nm):
    """<v|O|v> / <v|v> for O written left to right as (mode, create) operators."""
    u = v
    for mode, create in reversed(ops):
        u = _fock_apply(u, mode, create, nm)
    return float(v @ u) / float(v @ v)


def _pair_ops(a, b, c, d, N):
    """A+ A' = c+_{a up} c+_{b dn} c_{d dn} c_{c up} (A+ = c+_{a up} c+_{b dn}, A' its (c, d) pair)."""
    return [(a, True), (N + b, True), (N + d, False), (c, False)]


def test_oracle_dwave_terms_match_jordan_wigner():
    """The exact |psi|^2 averages of the oracle's local F and G values equal <psi|A+(bl, oL)
    A(br, oR)|psi> and <psi|c+_{i s} c_{j s}|psi> built on the dense Fock space by the independent
    Jordan-Wigner routine, for both left bonds and orientations, every right bond and orientation
    and every hopping term, on a 3x3 lattice with 2 + 2 electrons."""
    LX, LY, N, basis, table, v = _dwave_exact_setup()
    nm = 2 * N

    def psi_fn(s):
        return table[tuple(int(x) for x in np.flatnonzero(s > 0))]

    top = max(t[1] for t in table.values())
    w = {occ: math.exp(2 * (la - top)) for occ, (sg, la) in table.items()}
    keep = [occ for occ in basis if table[occ][0] != 0]
    wsum = sum(w[o] for o in basis)
    F = np.zeros((4, 4 * N))
    G = np.zeros((3, N, 2))
    for occ in keep:
        f, g = O.dwave_local(config_of(occ, nm), LX, LY, psi_fn)
        F += w[occ] * f
        G += w[occ] * g
    F /= wsum
    G /= wsum
    for l, ax in enumerate((0, 1)):
        lo, hi = (0, LY) if ax == 0 else (0, 1)
        for oL in (0, 1):
            a, b = (lo, hi) if oL == 0 else (hi, lo)
            for site in range(N):
                x, y = divmod(site, LY)
                for bax in (0, 1):
                    clo = site
                    chi = ((x + 1) % LX) * LY + y if bax == 0 else x * LY + (y + 1) % LY
                    for oR in (0, 1):
                        c, d = (clo, chi) if oR == 0 else (chi, clo)
                        exact = _fock_expect(v, _pair_ops(a, b, c, d, N), nm)
                        got = F[2 * l + oL, 2 * (2 * site + bax) + oR]
                        assert abs(got - exact) < 1e-10, (l, oL, site, bax, oR, got, exact)
    for li, i in enumerate((0, 1, LY)):
        for sp in (0, 1):
            for j in range(N):
                exact = _fock_expect(v, [(sp * N + i, True), (sp * N + j, False)], nm)
                assert abs(G[li, j, sp] - exact) < 1e-10, (i, j, sp, G[li, j, sp], exact)
    # keep the exact means for the assembly test
    test_oracle_dwave_terms_match_jordan_wigner.means = (F, G)


def test_oracle_dwave_pair_correlation_matches_direct_fock():
    """C_p(r) assembled by `observables.DwaveIndex` from the measured terms at R = (0, 0) (the other
    left bonds and sites by translation) equals (1/16) sum h h' <B+(bl) B(br)> built directly on
    the Fock space for every pair of bonds, minus the Wick products of the exact <c+ c> of every
    pair of sites (``code/measure/assemble.py:103-175``); without the subtraction it is
    <Delta+_0 Delta_r> of main-text Eq. (2)."""
    from neuralhubbard_amd.observables import DIRS, FORM_FACTOR, DwaveIndex, _bond_sites, _canon_bond

    LX, LY, N, basis, table, v = _dwave_exact_setup()
    nm = 2 * N
    means = getattr(test_oracle_dwave_terms_match_jordan_wigner, "means", None)
    if means is None:
        test_oracle_dwave_terms_match_jordan_wigner()
        means = test_oracle_dwave_terms_match_jordan_wigner.means
    F, G = means
    idx = DwaveIndex(LX, LY)
    full = idx.pair_correlation(F, G)
    bare = idx.pair_correlation(F, G, connected=False)

    def ix(p):
        return p[0] * LY + p[1]

    cache = {}

    def pair(a, b, c, d):
        if (a, b, c, d) not in cache:
            cache[a, b, c, d] = _fock_expect(v, _pair_ops(a, b, c, d, N), nm)
        return cache[a, b, c, d]

    def hop(i, j, sp):
        if (i, j, sp) not in cache:
            cache[i, j, sp] = _fock_expect(v, [(sp * N + i, True), (sp * N + j, False)], nm)
        return cache[i, j, sp]

    for n in range(N):
        r2 = divmod(n, LY)
        direct = wick = 0.0
        for dl in DIRS:
            a, b = map(ix, _bond_sites(_canon_bond((0, 0), dl, LX, LY), LX, LY))
            for d2 in DIRS:
                c, d = map(ix, _bond_sites(_canon_bond(r2, d2, LX, LY), LX, LY))
                h = FORM_FACTOR[dl] * FORM_FACTOR[d2]
                # B+ = (A+(o=0) + A+(o=1)) / sqrt 2: (a, b) and (b, a) on the left, (c, d), (d, c) right
                direct += h * 0.5 * sum(pair(p, q, r, t) for p, q in ((a, b), (b, a)) for r, t in ((c, d), (d, c)))
                wick += h * 0.5 * (hop(a, c, 0) * hop(b, d, 1) + hop(a, d, 0) * hop(b, c, 1)
                                   + hop(b, c, 0) * hop(a, d, 1) + hop(b, d, 0) * hop(a, c, 1))
        assert abs(bare[n] - direct / 16) < 1e-10, (n, bare[n], direct / 16)
        assert abs(full[n] - (direct - wick) / 16) < 1e-10, (n, full[n], (direct - wick) / 16)
    assert np.abs(full).max() > 1e-3  # the test is not trivially zero
