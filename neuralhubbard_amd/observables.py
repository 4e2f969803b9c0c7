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
    """nref reference sites spread along the diagonal (``observables.reference_sites``)."""
    return [(x, x % LY) for x in range(0, LX, max(LX // nref, 1))][:nref]


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


FAMILIES = {"pi8": (1, 8), "pi4": (1, 4)}  # stripe wavevector q / pi (``observables.paper_family``)


def stripe_k(LX: int, LY: int, family: str):
    """Integer momenta of k_alpha = (pi -+ q, pi), (pi, pi -+ q) (``observables.stripe_k``)."""
    num, den = FAMILIES[family]
    if (num * LX) % (2 * den) or (num * LY) % (2 * den):
        raise ValueError(f"q = {num}/{den} pi is not on the momentum grid of {LX}x{LY}")
    nx, ny = num * LX // (2 * den), num * LY // (2 * den)
    hx, hy = LX // 2, LY // 2
    return ((hx - nx, hy), (hx + nx, hy), (hx, hy - ny), (hx, hy + ny))


def charge_q(LX: int, LY: int, family: str):
    """Integer momenta of the charge wavevectors +-(2q, 0), +-(0, 2q) (``observables.charge_q``)."""
    num, den = FAMILIES[family]
    if (num * LX) % (2 * den) or (num * LY) % (2 * den):
        raise ValueError(f"q = {num}/{den} pi is not on the momentum grid of {LX}x{LY}")
    nx, ny = 2 * (num * LX // (2 * den)), 2 * (num * LY // (2 * den))
    return ((nx, 0), (LX - nx, 0), (0, ny), (0, LY - ny))


def structure_factor(C, LX: int, LY: int):
    """S(k) = sum_r e^{ik.r} C(r) on the momentum grid, (..., LX, LY) (``assemble.structure_factor``)."""
    C = np.asarray(C, dtype=np.float64)
    return (np.fft.ifft2(C.reshape(*C.shape[:-1], LX, LY), axes=(-2, -1)) * (LX * LY)).real


def m2_str(C, LX: int, LY: int, family: str):
    """M^2_str = (1/4N) sum_alpha S(k_alpha) (``assemble.family_mean`` of ``assemble.spin``)."""
    Sk = structure_factor(C, LX, LY)
    return np.mean([Sk[..., kx % LX, ky % LY] for kx, ky in stripe_k(LX, LY, family)], axis=0) / (LX * LY)


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


class DwaveIndex:
    """Index tables of C_p(r) over the measured terms of `dwave_local` (the role of
    ``assemble.DwaveIndex``): terms of left bonds or sites other than the measured ones come from
    the measured ones by translation (exact for translation-invariant states)."""

    def __init__(self, LX: int, LY: int):
        N = self.N = LX * LY
        lefts = _left_bonds()
        lsites = _left_sites(LY)

        def right_index(bond, o):
            x, y, ax = bond
            return 2 * (2 * (x * LY + y) + ax) + o

        def fcol(bl, oL, br, oR):
            l = next(q for q, b in enumerate(lefts) if b[2] == bl[2])
            tx, ty = lefts[l][0] - bl[0], lefts[l][1] - bl[1]
            brt = ((br[0] + tx) % LX, (br[1] + ty) % LY, br[2])
            return (2 * l + oL) * (4 * N) + right_index(brt, oR)

        def gcol(i, j, s):
            ii = i[0] * LY + i[1]
            if ii in lsites:
                return (lsites.index(ii) * N + j[0] * LY + j[1]) * 2 + s
            jt = ((j[0] - i[0]) % LX, (j[1] - i[1]) % LY)
            return (0 * N + jt[0] * LY + jt[1]) * 2 + s

        self.fidx = np.empty((N, 16, 4), np.int64)
        self.gidx = np.empty((N, 16, 4, 2), np.int64)
        self.ff = np.empty(16)
        for n in range(N):
            r2 = divmod(n, LY)
            m = 0
            for dl in DIRS:
                bl = _canon_bond((0, 0), dl, LX, LY)
                a, b = _bond_sites(bl, LX, LY)
                for d2 in DIRS:
                    br = _canon_bond(r2, d2, LX, LY)
                    c, d = _bond_sites(br, LX, LY)
                    self.ff[m] = FORM_FACTOR[dl] * FORM_FACTOR[d2]
                    self.fidx[n, m] = [fcol(bl, oL, br, oR) for oL in (0, 1) for oR in (0, 1)]
                    self.gidx[n, m] = [[gcol(a, c, 0), gcol(b, d, 1)], [gcol(a, d, 0), gcol(b, c, 1)],
                                       [gcol(b, c, 0), gcol(a, d, 1)], [gcol(b, d, 0), gcol(a, c, 1)]]
                    m += 1

    def pair_correlation(self, F, G, connected=True):
        """C_p(r) (..., N) from term means F (..., 4, 4N) and G (..., 3, N, 2), or flattened
        (..., 16N) and (..., 6N)."""
        N = self.N
        F = np.asarray(F, np.float64)
        if F.shape[-1] != 16 * N:
            if F.shape[-2:] != (4, 4 * N):
                raise ValueError(f"F of shape {F.shape}: expected (..., 4, {4 * N}) or (..., {16 * N})")
            F = F.reshape(*F.shape[:-2], -1)
        out = 0.5 * F[..., self.fidx].sum(-1)
        if connected:
            G = np.asarray(G, np.float64)
            if G.shape[-1] != 6 * N:
                if G.shape[-3:] != (3, N, 2):
                    raise ValueError(f"G of shape {G.shape}: expected (..., 3, {N}, 2) or (..., {6 * N})")
                G = G.reshape(*G.shape[:-3], -1)
            out = out - 0.5 * (G[..., self.gidx[..., 0]] * G[..., self.gidx[..., 1]]).sum(-1)
        return (out * self.ff).sum(-1) / 16


def dwave_keys(LX: int, LY: int):
    """(F_keys (16N, 8), G_keys (6N, 5)) int32 in the reference's series format
    (``observables.F_keys_to_int`` / ``G_keys_to_int`` of ``dwave_keys(geo, (0, 0))``): rows
    (bx, by, bax, oL, cx, cy, cax, oR) and (ix, iy, jx, jy, spin), spin 0 up and 1 down."""
    F = [(*bl, oL, x, y, ax, oR) for bl in _left_bonds() for oL in (0, 1)
         for x in range(LX) for y in range(LY) for ax in (0, 1) for oR in (0, 1)]
    G = [(*divmod(i, LY), *divmod(j, LY), sp) for i in _left_sites(LY) for j in range(LX * LY) for sp in (0, 1)]
    return np.array(F, np.int32), np.array(G, np.int32)


def far_mask(LX: int, LY: int):
    """(mask of |r| >= r_max, r_max = d_max / 2) over displacements (``assemble.far_mask``)."""
    dx, dy = np.divmod(np.arange(LX * LY), LY)
    dist = np.hypot(np.minimum(dx, LX - dx), np.minimum(dy, LY - dy))
    r_max = dist.max() / 2
    return dist >= r_max - 1e-9, float(r_max)


def dwave(F, G, LX: int, LY: int) -> dict:
    """C_p(r) with leave-one-out errors, Delta^2_SC = mean of C_p over the far displacements and
    Delta_SC with block jackknife errors, from per-sweep term means F (S, 4, 4N) and G (S, 3, N, 2)
    (``assemble.dwave``)."""
    from .measure import BLOCK_SIZES, MIN_BLOCKS

    N = LX * LY
    F = np.asarray(F, np.float64).reshape(-1, 16 * N)
    G = np.asarray(G, np.float64).reshape(-1, 6 * N)
    idx = DwaveIndex(LX, LY)
    far, r_max = far_mask(LX, LY)
    n = F.shape[0]
    Cp = idx.pair_correlation(F.mean(0), G.mean(0))
    if n > 1:
        loo = idx.pair_correlation((F.sum(0) - F) / (n - 1), (G.sum(0) - G) / (n - 1))
        Cp_err = np.sqrt(n - 1) * loo.std(axis=0)
    else:
        Cp_err = np.full_like(Cp, np.nan)
    d2 = float(Cp[far].mean())
    d2_errs, d_errs = {}, {}
    for b in BLOCK_SIZES:
        nb = n // b
        if nb < MIN_BLOCKS:
            continue
        Fb = F[:nb * b].reshape(nb, b, -1).mean(1)
        Gb = G[:nb * b].reshape(nb, b, -1).mean(1)
        reps = idx.pair_correlation((Fb.sum(0) - Fb) / (nb - 1), (Gb.sum(0) - Gb) / (nb - 1))[:, far].mean(-1)
        d2_errs[b] = float(np.sqrt(nb - 1) * reps.std())
        d_errs[b] = float(np.sqrt(nb - 1) * np.sqrt(reps).std()) if reps.min() > 0 else float("nan")
    return {"C_p": Cp, "C_p_err": Cp_err, "Delta2_SC": d2, "Delta2_SC_err": max(d2_errs.values(), default=float("nan")),
            "Delta_SC": float(np.sqrt(d2)) if d2 > 0 else float("nan"),
            "Delta_SC_err": max(d_errs.values(), default=float("nan")), "Delta2_SC_err_blocks": d2_errs,
            "r_max": r_max, "n_far": int(far.sum()), "nsweeps": n}
