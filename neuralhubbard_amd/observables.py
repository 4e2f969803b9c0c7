"""Spin correlations of a state: the diagonal S^z S^z part and the transverse S^+ S^- terms.

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


def structure_factor(C, LX: int, LY: int):
    """S(k) = sum_r e^{ik.r} C(r) on the momentum grid, (..., LX, LY) (``assemble.structure_factor``)."""
    C = np.asarray(C, dtype=np.float64)
    return (np.fft.ifft2(C.reshape(*C.shape[:-1], LX, LY), axes=(-2, -1)) * (LX * LY)).real


def m2_str(C, LX: int, LY: int, family: str):
    """M^2_str = (1/4N) sum_alpha S(k_alpha) (``assemble.family_mean`` of ``assemble.spin``)."""
    Sk = structure_factor(C, LX, LY)
    return np.mean([Sk[..., kx % LX, ky % LY] for kx, ky in stripe_k(LX, LY, family)], axis=0) / (LX * LY)
