r"""Lattice conventions of the Pfaffian/BCS backflow path (host side, integer work only).

Conventions held bit-exact against the reference text:

* A configuration ``s`` is int8 of length 2N, +1 occupied / -1 empty: the N spin-up
  occupations, then the N spin-down ones (``README.md:85-86``; ``code/states.py:16-17``).
* Site (x, y) has index ``x * LY + y`` (row major, checked at ``code/backflow/models.py:96-97``).
* Mode m = spin * N + site, so modes are spin major (up block, then down block).
* Displacement index ``r(p, q) = ((x_p - x_q) mod LX) * LY + (y_p - y_q) mod LY``
  (``code/backflow/pf_kernel.py:83-96, 144-153``).
* Projection group ``Flip(axis=0) @ Flip(axis=1) @ SpinInverse`` of 8 elements with trivial
  characters (``code/states.py:274-279``).  Element g = fx | fy << 1 | si << 2 acts on modes by
  ``(spin, x, y) -> (spin ^ si, fx ? -x : x, fy ? -y : y)`` (mod L).  Every element is an
  involution.  The reflection is taken about the origin; a reflection about another axis differs
  by a translation, and the amplitudes here are translation invariant in the fermionic sense,
  so the projected amplitude does not depend on that choice.
* Hubbard hopping bonds: nearest neighbours (t) and, with ``n_neighbor=[1, 2]``, next-nearest
  (diagonal) neighbours (t'), periodic, spin preserving (``code/states.py:244-247``).  Every
  distinct unordered site pair at that distance is one bond.
"""

from __future__ import annotations

import functools
from dataclasses import dataclass

import numpy as np

GROUP_SIZE = 8


@dataclass(frozen=True)
class Geometry:
    """Static (LX, LY, n_up, n_dn) of one periodic lattice; mirrors ``PairGeometry``
    (``code/backflow/pf_kernel.py:58-77``) without the feature width."""

    LX: int
    LY: int
    n_up: int
    n_dn: int

    @property
    def N(self) -> int:
        return self.LX * self.LY

    @property
    def n_occ(self) -> int:
        return self.n_up + self.n_dn


@functools.lru_cache(maxsize=None)
def displacement_tables(LX: int, LY: int):
    """``(disp_idx (N, N) int32, neg_idx (N,) int32, dist_r (N,) float64)``.

    Same tables as ``displacement_tables`` (``code/backflow/pf_kernel.py:83-96``):
    ``disp_idx[i, j]`` is the displacement index of i - j, ``neg_idx[r]`` that of -r and
    ``dist_r[r]`` the minimum-image length of r.
    """
    N = LX * LY
    x, y = np.divmod(np.arange(N), LY)
    dx = (x[:, None] - x[None, :]) % LX
    dy = (y[:, None] - y[None, :]) % LY
    disp = (dx * LY + dy).astype(np.int32)
    neg = (((-x) % LX) * LY + ((-y) % LY)).astype(np.int32)
    mx = np.minimum(x, LX - x)
    my = np.minimum(y, LY - y)
    dist = np.sqrt((mx * mx + my * my).astype(np.float64))
    return disp, neg, dist


@functools.lru_cache(maxsize=None)
def dwave_form_factor(LX: int, LY: int) -> np.ndarray:
    """(N, N) F^d: -1 on x bonds, +1 on y bonds (``code/backflow/models.py:69-83``)."""
    N = LX * LY
    x, y = np.divmod(np.arange(N), LY)
    dx = (x[:, None] - x[None, :]) % LX
    dy = (y[:, None] - y[None, :]) % LY
    mx = np.minimum(dx, LX - dx)
    my = np.minimum(dy, LY - dy)
    fd = np.zeros((N, N), np.float64)
    fd[(mx == 1) & (my == 0)] = -1.0
    fd[(mx == 0) & (my == 1)] = 1.0
    return fd


@functools.lru_cache(maxsize=None)
def group_mode_perms(LX: int, LY: int) -> np.ndarray:
    """(8, 2N) int32: ``perm[g, m]`` is the image of mode m under group element g."""
    N = LX * LY
    m = np.arange(2 * N)
    spin, site = np.divmod(m, N)
    x, y = np.divmod(site, LY)
    out = np.empty((GROUP_SIZE, 2 * N), np.int32)
    for g in range(GROUP_SIZE):
        fx, fy, si = g & 1, (g >> 1) & 1, (g >> 2) & 1
        gx = (-x) % LX if fx else x
        gy = (-y) % LY if fy else y
        out[g] = (spin ^ si) * N + gx * LY + gy
    return out


def apply_group(s: np.ndarray, g: int, LX: int, LY: int) -> np.ndarray:
    """Configuration g.s: ``(g.s)[perm_g[m]] = s[m]`` (batched over leading axes)."""
    perm = group_mode_perms(LX, LY)[g]
    out = np.empty_like(s)
    out[..., perm] = s
    return out


@functools.lru_cache(maxsize=None)
def neighbour_table(LX: int, LY: int, n_neighbor: int):
    """Bond partners of every site, as ``(nbr (N, z) int32, cls (N, z) int8)``.

    ``cls`` is 0 for a nearest-neighbour bond and 1 for a next-nearest one; ``z`` is the
    largest partner count, unused slots hold -1.  Duplicated partners (lattices with an
    extent of 2) and the site itself are dropped, so every unordered bond appears once per
    end point.
    """
    N = LX * LY
    shells = [[(1, 0), (-1, 0), (0, 1), (0, -1)]]
    if n_neighbor >= 2:
        shells.append([(1, 1), (1, -1), (-1, 1), (-1, -1)])
    rows, classes = [], []
    for i in range(N):
        x, y = divmod(i, LY)
        seen, row, cl = set(), [], []
        for c, shell in enumerate(shells):
            for dx, dy in shell:
                j = ((x + dx) % LX) * LY + (y + dy) % LY
                if j == i or j in seen:
                    continue
                seen.add(j)
                row.append(j)
                cl.append(c)
        rows.append(row)
        classes.append(cl)
    z = max(len(r) for r in rows)
    nbr = np.full((N, z), -1, np.int32)
    cls = np.full((N, z), -1, np.int8)
    for i, (r, c) in enumerate(zip(rows, classes)):
        nbr[i, : len(r)] = r
        cls[i, : len(c)] = c
    return nbr, cls


def fermion_idx(s: np.ndarray, n_occ: int | None = None) -> np.ndarray:
    """Ascending occupied modes (up block first), as ``quantax.nn.fermion_idx`` is used at
    ``code/backflow/models.py:200``.  Batched over the leading axes."""
    s = np.asarray(s)
    occ = s > 0
    if n_occ is None:
        n_occ = int(occ.reshape(-1, s.shape[-1])[0].sum())
    flat = occ.reshape(-1, s.shape[-1])
    idx = np.stack([np.flatnonzero(row) for row in flat]).astype(np.int32)
    if idx.shape[1] != n_occ:
        raise ValueError("configurations with different particle numbers")
    return idx.reshape(*s.shape[:-1], n_occ)


def random_configurations(geom: Geometry, batch: int, seed: int) -> np.ndarray:
    """Seeded fixed-count configurations: n_up distinct up sites and n_dn distinct down
    sites drawn uniformly per configuration (PCG64), +1 occupied / -1 empty, int8."""
    rng = np.random.Generator(np.random.PCG64(seed))
    N = geom.N
    s = -np.ones((batch, 2 * N), np.int8)
    for b in range(batch):
        s[b, rng.choice(N, geom.n_up, replace=False)] = 1
        s[b, N + rng.choice(N, geom.n_dn, replace=False)] = 1
    return s


def pack_spins(spins: np.ndarray) -> dict:
    """Chain file layout of ``weights/<state>_spins.npz``: ``bits = packbits(spins > 0)``
    along the last axis plus the shape (``code/measure_orderparams.py:170-172``)."""
    spins = np.asarray(spins)
    return dict(bits=np.packbits(spins > 0, axis=-1), shape=np.array(spins.shape))


def unpack_spins(bits: np.ndarray, shape) -> np.ndarray:
    """Inverse of `pack_spins` (``code/states.py:328-332``)."""
    shape = tuple(int(x) for x in shape)
    s = np.unpackbits(bits, axis=-1, count=shape[-1]).astype(np.int8)
    return (2 * s - 1).reshape(shape)
