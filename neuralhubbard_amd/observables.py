"""Spin, charge and d-wave pair correlations of a state.

Mirrors the spin channel of the reference's measurement (``code/measure/observables.py:179-206``,
``code/measure/transverse_spin.py``, ``code/measure/assemble.py:230-238``):

* ``C_zz(r)`` = batch mean of (1/N) sum_i S^z_i S^z_{i+r}, S^z = (n_up - n_dn) / 2 (`sz_grid`,
  `autocorr`);
* ``T[0, r]`` = <S^+_a S^-_{a+r}>, ``T[1, r]`` = <S^-_a S^+_{a+r}> for a reference site a, from
  local values sum_{s'} <s|O|s'> psi(s') / psi(s) over the configurations s' that the terms connect
  (`transverse_local`; r = 0 holds n_up (1 - n_dn) and n_dn (1 - n_up) of site a);
* ``C_s(r) = C_zz(r) + (T[0, r] + T[1, r]) / 2`` (`spin_correlation`), its structure factor and
  the stripe order ``M^2_str`` (`structure_factor`, `m2_str`, ``code/measure/assemble.py:241-276``),
  and the charge correlation ``C_c(r)`` of the densities (`charge_correlation`).

A term S^+_a S^-_j connects s to s' with the up electron of a moved to j and the down electron of
j moved to a (S^-_a S^+_j: the up electron of j to a, the down electron of a to j), the order in
which ``transverse_spin.py`` identifies terms.  The amplitudes psi(s') come from the projected
state (``nh_log_psi``) on the GPU; the matrix elements carry the fermionic signs of the operators
applied in mode order (site x * LY + y, up block then down block), the convention of the hopping
terms of E_loc.

The d-wave channel (``code/measure/dwave_packed.py``, ``code/measure/assemble.py:103-230``):
`dwave_local` gives per-sample local values of the pair terms F and the hopping terms G at
R = (0, 0) in the reference's key order (`dwave_keys`), `DwaveIndex` assembles the pair
correlation C_p(r) of main-text Eq. (2) from their means, and `dwave` adds Delta^2_SC over the far
displacements with block jackknife errors.
"""

from __future__ import annotations

import numpy as np


def _torch():
    import torch

    return torch


def reference_sites(LX: int, LY: int, nref: int):
    """nref reference sites (x, x mod LY) on the diagonal, x every max(LX // nref, 1) rows from 0
    (the reference's choice of sites for the transverse terms)."""
    step = max(LX // nref, 1)
    return [(int(x), int(x % LY)) for x in np.arange(0, LX, step)[:nref]]


def sz_grid(spins, N: int):
    """S^z_i = (n_up - n_dn) / 2 per sample, float64 (B, N)."""
    s = spins.double()
    return (s[:, :N] - s[:, N:]) / 4


def dn_grid(spins, N: int, nbar: float):
    """n_i - nbar per sample, float64 (B, N) (``observables.dn_grid``)."""
    s = spins.double()
    return (s[:, :N] + s[:, N:]) / 2 + 1 - nbar


def charge_correlation(spins, LX: int, LY: int, nbar: float):
    """C_c(r) = <(n_0 - nbar)(n_r - nbar)>, the batch mean of the translation average (N,)."""
    return autocorr(dn_grid(spins, LX * LY, nbar), LX, LY)


def autocorr(g, LX: int, LY: int):
    """Batch mean of (1/N) sum_i g_i g_{i+r}, shape (N,), r = rx * LY + ry (by FFT)."""
    torch = _torch()
    G = g.reshape(-1, LX, LY)
    Fk = torch.fft.fft2(G)
    A = torch.fft.ifft2(Fk.abs() ** 2).real
    return A.mean(0).reshape(-1) / (LX * LY)


# (mode offset of the site: 0 up, 1 down, times N; create?) of O^dagger, right-to-left order, for
# O = S^+_a S^-_j (channel 0): O^dagger = S^+_j S^-_a = c+_{j up} c_{j dn} c+_{a dn} c_{a up};
# O = S^-_a S^+_j (channel 1): O^dagger = S^+_a S^-_j = c+_{a up} c_{a dn} c+_{j dn} c_{j up}.
_OPS = {0: (("a", 0, False), ("a", 1, True), ("j", 1, False), ("j", 0, True)),
        1: (("j", 0, False), ("j", 1, True), ("a", 1, False), ("a", 0, True))}


def _matrix_elements(occ, a, j, N, chan):
    """<s|O|s'> of the connected s' for items with occupations occ (I, 2N) bool of s: O^dagger
    applied to |s> in the Jordan-Wigner order of the modes, (-1)^(occupied modes below)."""
    torch = _torch()
    occ = occ.clone()
    sign = torch.ones(occ.shape[0], dtype=torch.float64, device=occ.device)
    rows = torch.arange(occ.shape[0], device=occ.device)
    for who, spin, create in _OPS[chan]:
        m = (a if who == "a" else j) + spin * N
        below = occ.cumsum(1, dtype=torch.int32).gather(1, m[:, None])[:, 0] - occ[rows, m].int()
        sign = sign * (1 - 2 * (below & 1)).double()
        occ[rows, m] = create
    return sign


def transverse_local(state, spins, psi, ref: int, LX: int, LY: int):
    """(B, 2, N) float64 local values of S^+_a S^-_{a+r} (channel 0) and S^-_a S^+_{a+r}
    (channel 1) for reference site `ref` (index x * LY + y): <s|O|s'> psi(s') / psi(s) summed over
    the connected s' of each term (``transverse_spin._transverse_sum`` and ``_finalize``).
    `spins` (B, 2N) int8 on the GPU, `psi` their LogArray under `state`."""
    torch = _torch()
    N = LX * LY
    s = spins
    B = s.shape[0]
    dev = s.device
    up, dn = s[:, :N] > 0, s[:, N:] > 0
    su, sd = up & ~dn, dn & ~up  # singly occupied by up / by down
    ax, ay = divmod(ref, LY)
    j_all = torch.arange(N, device=dev)
    rj = ((j_all // LY - ax) % LX) * LY + (j_all % LY - ay) % LY  # r of j relative to a
    T = torch.zeros((B, 2, N), dtype=torch.float64, device=dev)
    T[:, 0, 0] = su[:, ref].double()
    T[:, 1, 0] = sd[:, ref].double()
    for chan in (0, 1):
        cond = (su[:, ref:ref + 1] & sd) if chan == 0 else (sd[:, ref:ref + 1] & su)
        cond[:, ref] = False
        b, j = torch.nonzero(cond, as_tuple=True)
        if b.numel() == 0:
            continue
        a = torch.full_like(j, ref)
        sp = s[b].clone()
        rows = torch.arange(b.numel(), device=dev)
        src, dst = (a, j) if chan == 0 else (j, a)  # the up electron moves src -> dst
        sp[rows, src] = -1
        sp[rows, dst] = 1
        sp[rows, N + dst] = -1
        sp[rows, N + src] = 1
        me = _matrix_elements(s[b] > 0, a, j, N, chan)
        p2 = state(sp)
        val = me * p2.sign.double() * psi.sign[b].double() * torch.exp(p2.logabs - psi.logabs[b])
        T.index_put_((b, torch.full_like(b, chan), rj[j]), val, accumulate=True)
    return T


def spin_correlation_samples(state, spins, psi, LX: int, LY: int, refs=None):
    """(B, N) per-sample local values of C_s(r): (1/N) sum_i S^z_i S^z_{i+r} of the sample plus half
    the two transverse terms of its reference site, sample k taking refs[k % len(refs)] (site
    indices; default every site), so that the sample mean is the translation average."""
    torch = _torch()
    from .backflow import LogArray

    N = LX * LY
    refs = list(range(N)) if refs is None else list(refs)
    B = spins.shape[0]
    out = torch.zeros((B, N), dtype=torch.float64, device=spins.device)
    for k, a in enumerate(refs):
        sel = torch.arange(k, B, len(refs), device=spins.device)
        if sel.numel() == 0:
            continue
        T = transverse_local(state, spins[sel], LogArray(psi.sign[sel], psi.logabs[sel]), a, LX, LY)
        out[sel] = 0.5 * (T[:, 0] + T[:, 1])
    g = sz_grid(spins, N).reshape(-1, LX, LY)
    Fk = torch.fft.fft2(g)
    out += torch.fft.ifft2(Fk.abs() ** 2).real.reshape(B, N) / N
    return out


def spin_correlation(state, spins, psi, LX: int, LY: int, refs=None):
    """(C_s (N,), C_zz (N,)) sample means (``code/measure/assemble.py:230-238``)."""
    cs = spin_correlation_samples(state, spins, psi, LX, LY, refs)
    return cs.mean(0), autocorr(sz_grid(spins, LX * LY), LX, LY)


FAMILIES = {"pi8": (1, 8), "pi4": (1, 4)}  # stripe wavevector q / pi of each family


def _grid_q(LX: int, LY: int, family: str):
    """q of the family as integer momenta (qx, qy) on the LX x LY grid (q L / 2 pi)."""
    from fractions import Fraction

    q = Fraction(*FAMILIES[family]) / 2  # q / 2 pi
    mx, my = q * LX, q * LY
    if mx.denominator != 1 or my.denominator != 1:
        raise ValueError(f"q = {FAMILIES[family][0]}/{FAMILIES[family][1]} pi is not on the momentum grid of {LX}x{LY}")
    return int(mx), int(my)


def stripe_k(LX: int, LY: int, family: str):
    """The four stripe momenta k_alpha = (pi - q, pi), (pi + q, pi), (pi, pi - q), (pi, pi + q)
    as integer grid momenta."""
    qx, qy = _grid_q(LX, LY, family)
    px, py = LX // 2, LY // 2
    return ((px - qx, py), (px + qx, py), (px, py - qy), (px, py + qy))


def charge_q(LX: int, LY: int, family: str):
    """The four charge momenta (2q, 0), (-2q, 0), (0, 2q), (0, -2q) as integer grid momenta in
    [0, L)."""
    qx, qy = _grid_q(LX, LY, family)
    return ((2 * qx, 0), (LX - 2 * qx, 0), (0, 2 * qy), (0, LY - 2 * qy))


def _phases(L: int):
    k = np.arange(L)
    return np.exp(2j * np.pi * np.outer(k, k) / L)  # [k, x] -> e^{i 2 pi k x / L}


def structure_factor(C, LX: int, LY: int):
    """S(k) = Re sum_r e^{i k.r} C(r) for every grid momentum, C (..., N) with r = rx LY + ry ->
    (..., LX, LY)."""
    C = np.asarray(C, np.float64).reshape(*np.shape(C)[:-1], LX, LY)
    return np.einsum("ax,...xy,by->...ab", _phases(LX), C, _phases(LY)).real


def stripe_weights(LX: int, LY: int, momenta):
    """w (N,) with w . C = (1/4N) sum_alpha S(k_alpha): the mean of cos(k_alpha . r) / N over the
    momenta (S(k) is the real part, so only the cosines of k . r enter)."""
    rx, ry = np.divmod(np.arange(LX * LY), LY)
    w = np.zeros(LX * LY)
    for kx, ky in momenta:
        w += np.cos(2 * np.pi * (kx * rx / LX + ky * ry / LY))
    return w / (len(momenta) * LX * LY)


def m2_str(C, LX: int, LY: int, family: str):
    """Stripe order M^2_str = (1/4N) sum_alpha S(k_alpha) of C (..., N)."""
    return np.asarray(C, np.float64) @ stripe_weights(LX, LY, stripe_k(LX, LY, family))


# ------------------------------------------------------------------------ d-wave pair correlation
#
# The terms of main-text Eq. (2) as the reference measures them (``code/measure/dwave_packed.py``,
# ``code/measure/observables.py:1-13``): A+(bond, o) = c+_{a up} c+_{b dn} with (a, b) the (lower,
# upper) ends of a nearest-neighbour bond for o = 0 and swapped for o = 1;
#   F[bl, oL, br, oR] = <A+(bl, oL) A(br, oR)>  for the +x and +y bonds bl at R = (0, 0),
#   G[i, j, s]        = <c+_{i s} c_{j s}>       for the ends i of those bonds,
# and C_p(r) = (1/16) sum over the directions of the two bonds of h h' [ (1/2) sum_{oL, oR} F
#   - (1/2) sum of the four Wick products G G ] (``code/measure/assemble.py:103-175``).

DIRS = ((1, 0), (-1, 0), (0, 1), (0, -1))
FORM_FACTOR = {(1, 0): 1.0, (-1, 0): 1.0, (0, 1): -1.0, (0, -1): -1.0}


def _canon_bond(site, d, LX, LY):
    """(x, y, axis) of the bond from `site` along the unit vector d, keyed by its lower end."""
    x, y = site
    if d[0] + d[1] < 0:
        x, y = x + d[0], y + d[1]
    return (x % LX, y % LY, 0 if d[0] != 0 else 1)


def _bond_sites(bond, LX, LY):
    x, y, ax = bond
    return (x, y), (((x + 1) % LX, y) if ax == 0 else (x, (y + 1) % LY))


def _left_bonds():
    return [(0, 0, 0), (0, 0, 1)]  # the +x and +y bonds at R = (0, 0)


def _left_sites(LY):
    return [0, 1, LY]  # (0, 0), (0, 1), (1, 0): the ends of the left bonds in sorted order


def _apply_ops_sign(occ, ops):
    """Jordan-Wigner sign of applying `ops` [(mode tensor (I,), create)] in order (the first acts
    first) to occupations occ (I, 2N) bool; every step is known to be allowed."""
    torch = _torch()
    occ = occ.clone()
    sign = torch.ones(occ.shape[0], dtype=torch.float64, device=occ.device)
    rows = torch.arange(occ.shape[0], device=occ.device)
    for m, create in ops:
        below = occ.cumsum(1, dtype=torch.int32).gather(1, m[:, None])[:, 0] - occ[rows, m].int()
        sign = sign * (1 - 2 * (below & 1)).double()
        occ[rows, m] = create
    return sign


def dwave_local(state, spins, psi, LX: int, LY: int):
    """Per-sample local values (F (B, 4, 4N), G (B, 3, N, 2)) of the d-wave terms at R = (0, 0):
    F[b, 2 l + oL, 2 (2 site + axis) + oR] for left bond l (0: +x, 1: +y) and every right bond
    (site, axis), G[b, li, j, s] for the left sites li = (0,0), (0,1), (1,0); diagonal terms from the
    occupations (``dwave_packed.PackedDwave.measure``).  Flattened, F and G follow the reference's
    ``F_keys`` and ``G_keys`` (`dwave_keys`)."""
    torch = _torch()
    N = LX * LY
    s = spins
    B = s.shape[0]
    dev = s.device
    occ = s > 0
    up, dn = occ[:, :N], occ[:, N:]
    F = torch.zeros((B, 4, 4 * N), dtype=torch.float64, device=dev)
    G = torch.zeros((B, 3, N, 2), dtype=torch.float64, device=dev)
    # right terms: (c, d) = string sites of (bond k = 2 site + axis, oR)
    k = torch.arange(2 * N, device=dev)
    site, axis = k // 2, k % 2
    xs, ys = site // LY, site % LY
    upper = torch.where(axis == 0, ((xs + 1) % LX) * LY + ys, xs * LY + (ys + 1) % LY)
    c_of = torch.stack([site, upper], 1).reshape(-1)  # right term t = 2 k + oR: up site
    d_of = torch.stack([upper, site], 1).reshape(-1)  # down site

    def evaluate(b, sp, ops):
        me = _apply_ops_sign(occ[b], ops)
        p2 = state(sp)
        return me * p2.sign.double() * psi.sign[b].double() * torch.exp(p2.logabs - psi.logabs[b])

    for l, bl in enumerate(_left_bonds()):
        (ax_, ay_), (bx_, by_) = _bond_sites(bl, LX, LY)
        lo, hi = ax_ * LY + ay_, bx_ * LY + by_
        for oL in (0, 1):
            a, bsite = (lo, hi) if oL == 0 else (hi, lo)
            row = 2 * l + oL
            diag_t = 2 * (2 * (ax_ * LY + ay_) + bl[2]) + oL
            F[:, row, diag_t] = (up[:, a] & dn[:, bsite]).double()
            # O+ = A+(br, oR) A(bl, oL) = c+_{c up} c+_{d dn} c_{b dn} c_{a up} applied to s
            ok = (up[:, a] & dn[:, bsite])[:, None] & ((~up[:, c_of]) | (c_of == a)[None]) & \
                 ((~dn[:, d_of]) | (d_of == bsite)[None])
            ok[:, diag_t] = False
            b, t = torch.nonzero(ok, as_tuple=True)
            if b.numel():
                c, d = c_of[t], d_of[t]
                sp = s[b].clone()
                r_ = torch.arange(b.numel(), device=dev)
                sp[r_, a] = -1
                sp[r_, N + bsite] = -1
                sp[r_, c] = 1
                sp[r_, N + d] = 1
                am, bm = torch.full_like(c, a), torch.full_like(c, N + bsite)
                ops = ((am, False), (bm, False), (N + d, True), (c, True))
                F.index_put_((b, torch.full_like(b, row), t), evaluate(b, sp, ops), accumulate=True)
    for li, i in enumerate(_left_sites(LY)):
        for spin in (0, 1):
            o = occ[:, spin * N:(spin + 1) * N]
            G[:, li, i, spin] = o[:, i].double()
            # O+ = c+_{j s} c_{i s} applied to s: the electron of i moves to j
            ok = o[:, i:i + 1] & ~o
            ok[:, i] = False
            b, j = torch.nonzero(ok, as_tuple=True)
            if b.numel():
                sp = s[b].clone()
                r_ = torch.arange(b.numel(), device=dev)
                sp[r_, spin * N + i] = -1
                sp[r_, spin * N + j] = 1
                ops = ((torch.full_like(j, spin * N + i), False), (spin * N + j, True))
                G.index_put_((b, torch.full_like(b, li), j, torch.full_like(b, spin)), evaluate(b, sp, ops),
                             accumulate=True)
    return F, G


def dwave_measure(state, spins, psi, LX: int, LY: int, chunk: int = 512):
    """Sample means (F (4, 4N), G (3, N, 2)) of the d-wave terms, in chunks of `chunk` samples
    (``dwave_packed.PackedDwave.measure``)."""
    from .backflow import LogArray

    B = spins.shape[0]
    if B == 0:
        raise ValueError("no samples to measure")
    F = G = 0.0
    for b0 in range(0, B, chunk):
        sl = slice(b0, min(B, b0 + chunk))
        f, g = dwave_local(state, spins[sl], LogArray(psi.sign[sl], psi.logabs[sl]), LX, LY)
        F = F + f.sum(0)
        G = G + g.sum(0)
    return F / B, G / B


def _bonds_from(x, y, d, LX: int, LY: int):
    """Canonical bonds (lower end x, y; axis) from the sites (x, y) along the unit vector d, and
    their upper ends; x, y arrays."""
    if d[0] + d[1] < 0:  # the bond points down or left: its lower end is the neighbour
        x, y = (x + d[0]) % LX, (y + d[1]) % LY
    ax = 0 if d[0] else 1
    return x, y, ax, ((x + 1 - ax) % LX, (y + ax) % LY)


class DwaveIndex:
    """C_p(r) from the measured term means: F (the pair terms of the +x and +y bonds at
    R = (0, 0) with every right bond, dwave_keys' order) and G (<c+_i c_j> for the three ends of
    those bonds).  Terms of other left bonds and sites come from the measured ones by translation
    (exact for translation-invariant states).  For each of the 16 pairs of bond directions
    (d_l, d_r) and every r the columns are computed once, vectorised over r:
      F column (2 l + o_l) 4N + 4 site(r_bond) + 2 axis + o_r, the right bond shifted with the
      left one onto the measured bond l of the left bond's axis;
      G column (3 l_i ... ) of (i, j, s): (index of i among the measured sites, j) or, for any
      other i, (0, j - i)."""

    def __init__(self, LX: int, LY: int):
        self.LX, self.LY = LX, LY
        N = self.N = LX * LY
        rx, ry = np.divmod(np.arange(N), LY)
        measured_sites = {0: 0, 1: 1, LY: 2}  # (0, 0), (0, 1), (1, 0)

        def gcol(ix, iy, jx, jy, spin):
            isite = ix * LY + iy
            li = np.vectorize(lambda v: measured_sites.get(int(v), -1))(np.atleast_1d(isite))
            moved = li < 0
            jx = np.where(moved, (jx - ix) % LX, jx)
            jy = np.where(moved, (jy - iy) % LY, jy)
            li = np.where(moved, 0, li)
            return (li * N + jx * LY + jy) * 2 + spin

        self.pairs = []  # (form factor, F columns (4, N), G columns (4, 2, N)) per direction pair
        for dl in DIRS:
            lx, ly, lax, (ux, uy) = _bonds_from(np.zeros(1, int), np.zeros(1, int), dl, LX, LY)
            left = ((lx, ly), (ux, uy))  # ends (a, b) of the left bond
            for dr in DIRS:
                cx, cy, rax, (dx_, dy_) = _bonds_from(rx, ry, dr, LX, LY)
                right = ((cx, cy), (dx_, dy_))
                tx, ty = (cx - lx) % LX, (cy - ly) % LY  # right bond translated with the left one
                fcols, gcols = [], []
                for ol in (0, 1):
                    for orr in (0, 1):
                        fcols.append((2 * lax + ol) * 4 * N + 4 * (tx * LY + ty) + 2 * rax + orr)
                        # Wick pairing of A+(bl, ol) A(br, or): the up electron created at the first
                        # site of the orientation meets the up electron of the right pair, likewise down
                        (x1, y1), (x2, y2) = left if ol == 0 else left[::-1]
                        (z1, w1), (z2, w2) = right if orr == 0 else right[::-1]
                        gcols.append(np.stack([gcol(x1, y1, z1, w1, 0), gcol(x2, y2, z2, w2, 1)]))
                self.pairs.append((FORM_FACTOR[dl] * FORM_FACTOR[dr], np.stack(fcols), np.stack(gcols)))

    def pair_correlation(self, F, G, connected=True):
        """C_p(r) (..., N) from term means F (..., 4, 4N) and G (..., 3, N, 2), or flattened
        (..., 16N) and (..., 6N): (1/16) sum over direction pairs of h_l h_r [1/2 sum over the four
        orientations of F - 1/2 sum of their Wick products G_up G_dn] (the second part with
        ``connected`` only)."""
        N = self.N
        F = np.asarray(F, np.float64)
        if F.shape[-1] != 16 * N:
            if F.shape[-2:] != (4, 4 * N):
                raise ValueError(f"F of shape {F.shape}: expected (..., 4, {4 * N}) or (..., {16 * N})")
            F = F.reshape(*F.shape[:-2], 16 * N)
        if connected:
            G = np.asarray(G, np.float64)
            if G.shape[-1] != 6 * N:
                if G.shape[-3:] != (3, N, 2):
                    raise ValueError(f"G of shape {G.shape}: expected (..., 3, {N}, 2) or (..., {6 * N})")
                G = G.reshape(*G.shape[:-3], 6 * N)
        total = 0.0
        for h, fcols, gcols in self.pairs:
            term = F[..., fcols].sum(axis=-2)
            if connected:
                term = term - (G[..., gcols[:, 0]] * G[..., gcols[:, 1]]).sum(axis=-2)
            total = total + 0.5 * h * term
        return total / 16


def dwave_keys(LX: int, LY: int):
    """(F_keys (16N, 8), G_keys (6N, 5)) int32 in the reference's series format
    (``observables.F_keys_to_int`` / ``G_keys_to_int`` of ``dwave_keys(geo, (0, 0))``): rows
    (bx, by, bax, oL, cx, cy, cax, oR) and (ix, iy, jx, jy, spin), spin 0 up and 1 down."""
    F = [(*bl, oL, x, y, ax, oR) for bl in _left_bonds() for oL in (0, 1)
         for x in range(LX) for y in range(LY) for ax in (0, 1) for oR in (0, 1)]
    G = [(*divmod(i, LY), *divmod(j, LY), sp) for i in _left_sites(LY) for j in range(LX * LY) for sp in (0, 1)]
    return np.array(F, np.int32), np.array(G, np.int32)


def far_mask(LX: int, LY: int):
    """(mask of the displacements with |r| >= r_max, r_max) with |r| the minimum-image length and
    r_max half the largest of them."""
    rx, ry = np.divmod(np.arange(LX * LY), LY)
    length = np.sqrt(np.minimum(rx, LX - rx) ** 2 + np.minimum(ry, LY - ry) ** 2)
    r_max = 0.5 * length.max()
    return length >= r_max - 1e-9, float(r_max)


def dwave(F, G, LX: int, LY: int) -> dict:
    """C_p(r) with its delete-one-sweep error, Delta^2_SC (the mean of C_p over the far
    displacements) and Delta_SC = sqrt(Delta^2_SC) with block-jackknife errors, from per-sweep term
    means F (S, 4, 4N) and G (S, 3, N, 2)."""
    from .measure import delete_one_means, jackknife

    N = LX * LY
    F = np.asarray(F, np.float64).reshape(-1, 16 * N)
    G = np.asarray(G, np.float64).reshape(-1, 6 * N)
    n = F.shape[0]
    idx = DwaveIndex(LX, LY)
    far, r_max = far_mask(LX, LY)
    Cp = idx.pair_correlation(F.mean(axis=0), G.mean(axis=0))
    if n > 1:
        reps = idx.pair_correlation(delete_one_means(F, 1), delete_one_means(G, 1))
        Cp_err = np.sqrt((n - 1) / n * ((reps - reps.mean(axis=0)) ** 2).sum(axis=0))
    else:
        Cp_err = np.full_like(Cp, np.nan)

    def far_mean(f, g):
        return idx.pair_correlation(f, g)[far].mean()

    def root(f, g):
        v = far_mean(f, g)
        return np.sqrt(v) if v > 0 else np.nan

    d2, d2_err, d2_blk = jackknife(far_mean, F, G)
    _, d_err, _ = jackknife(root, F, G)
    d2 = float(d2)
    return {"C_p": Cp, "C_p_err": Cp_err, "Delta2_SC": d2, "Delta2_SC_err": float(d2_err),
            "Delta_SC": float(np.sqrt(d2)) if d2 > 0 else float("nan"), "Delta_SC_err": float(d_err),
            "Delta2_SC_err_blocks": {int(b): float(e) for b, e in d2_blk.items()}, "r_max": r_max,
            "n_far": int(far.sum()), "nsweeps": int(n)}
