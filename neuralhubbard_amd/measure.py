"""Energy statistics of a run of sweeps (host side, float64).

Mirrors ``code/measure/assemble.py:48-97`` (``BLOCK_SIZES``, ``MIN_BLOCKS``, ``jackknife``,
``energy``): the per-sweep means of E_loc and E_loc^2 are grouped in blocks of 1, 2, 4 and 8
consecutive sweeps (an incomplete last block is dropped); every block size that leaves at least
8 blocks gives a leave-one-block-out jackknife error, and the quoted error is the largest.
"""

from __future__ import annotations

import numpy as np

BLOCK_SIZES = (1, 2, 4, 8)
MIN_BLOCKS = 8


def jackknife(f, *series):
    """(f(means), max error over block sizes, {block size: error}) of per-sweep series."""
    xs = [np.asarray(s, dtype=np.float64) for s in series]
    value = f(*(x.mean(axis=0) for x in xs))
    errs = {}
    n = xs[0].shape[0]
    for b in BLOCK_SIZES:
        nb = n // b
        if nb < MIN_BLOCKS:
            continue
        blocks = [x[:nb * b].reshape((nb, b) + x.shape[1:]).mean(axis=1) for x in xs]
        # leave-one-block-out means of every series, shape (nb, ...)
        loo = [(blk.sum(axis=0)[None] - blk) / (nb - 1) for blk in blocks]
        reps = np.array([f(*(l[i] for l in loo)) for i in range(nb)])
        errs[b] = np.sqrt(nb - 1) * reps.std(axis=0)
    error = np.max(np.stack(list(errs.values())), axis=0) if errs else np.full(np.shape(value), np.nan)
    return value, error, errs


def energy(E_mean, E_sqr, N: int) -> dict:
    """E/N and sigma^2/N with jackknife errors from per-sweep means of E_loc and E_loc^2."""
    e, e_err, e_errs = jackknife(lambda m: m / N, E_mean)
    v, v_err, v_errs = jackknife(lambda m, q: (q - m * m) / N, E_mean, E_sqr)
    return {"E_per_site": float(e), "E_err": float(e_err), "var_per_site": float(v), "var_err": float(v_err),
            "E_err_blocks": {int(b): float(x) for b, x in e_errs.items()},
            "var_err_blocks": {int(b): float(x) for b, x in v_errs.items()}, "nsweeps": int(len(E_mean))}


def within_error_bars(a: dict, b: dict, k: float = 2.0) -> bool:
    """|E_a - E_b| <= k * sqrt(err_a^2 + err_b^2)."""
    return abs(a["E_per_site"] - b["E_per_site"]) <= k * float(np.hypot(a["E_err"], b["E_err"]))
