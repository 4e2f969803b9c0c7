"""Energy statistics of a run of sweeps (host side, float64), and the energy channel of the
reference's measurement script with its file formats.

Mirrors ``code/measure/assemble.py:48-97`` (``BLOCK_SIZES``, ``MIN_BLOCKS``, ``jackknife``,
``energy``): the per-sweep means of E_loc and E_loc^2 are grouped in blocks of 1, 2, 4 and 8
consecutive sweeps (an incomplete last block is dropped); every block size that leaves at least
8 blocks gives a leave-one-block-out jackknife error, and the quoted error is the largest.
`run_energy` writes ``series.npz``, ``spins_last.npz`` and ``results.json`` as
``code/measure_orderparams.py`` and ``code/measure/assemble.py`` do for the energy channel (the
spin, charge and d-wave channels are not built).
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


# ------------------------------------------------------------------------ energy channel of a run

SAVE_EVERY = 4  # sweeps between rewrites of series.npz (code/measure_orderparams.py:28)


def paper_family(tprime: float) -> str:
    """Stripe family label the reference stores with a series (``observables.paper_family``)."""
    return "pi8" if tprime == 0 else "pi4"


def _json_value(v):
    if isinstance(v, dict):
        return {str(k): _json_value(x) for k, x in v.items()}
    if isinstance(v, np.generic):
        v = v.item()
    if isinstance(v, float) and not np.isfinite(v):
        return None
    return v


def assemble_energy(series: dict) -> dict:
    """The scalars and ``energy`` entry that ``assemble`` derives from a series
    (``code/measure/assemble.py:301-311``)."""
    LX, LY = int(series["LX"]), int(series["LY"])
    res = {"LX": LX, "LY": LY, "family": str(series["family"])}
    for k in ("state", "tprime", "nsamples"):
        if k in series:
            res[k] = np.asarray(series[k]).item()
    if "E_mean" in series and len(series["E_mean"]):
        res["energy"] = energy(series["E_mean"], series["E_sqr"], LX * LY)
    return res


def summary(res: dict) -> str:
    """First lines of the reference's summary (``code/measure/assemble.py:330-342``)."""
    lines = [f"{res.get('state', '')} {res['LX']}x{res['LY']}, stripe family {res['family']}"]
    if "energy" in res:
        e = res["energy"]
        lines.append(f"E/N = {e['E_per_site']:.6f} +- {e['E_err']:.1e}, "
                     f"sigma^2/N = {e['var_per_site']:.5f} +- {e['var_err']:.1e} ({e['nsweeps']} sweeps)")
    return "\n".join(lines)


def save_results(res: dict, out_dir: str) -> None:
    """results.json with non-finite floats as null (``code/measure/assemble.py:383-399``)."""
    import json
    import os

    with open(os.path.join(out_dir, "results.json"), "w") as fh:
        json.dump(_json_value(res), fh, indent=1, allow_nan=False)


def run_energy(state, H, sampler, nsweeps: int, out_dir: str, name: str = "", tprime: float = 0.0,
               log=print) -> dict:
    """The energy channel of ``code/measure_orderparams.py:108-174``: every sweep's samples get
    psi once and E_loc; per-sweep means of E_loc and E_loc^2 go to ``series.npz`` (rewritten every
    SAVE_EVERY sweeps; sweeps with non-finite values excluded), the final chains to
    ``spins_last.npz`` in the released-chain format, the assembled energy to ``results.json``.
    `sampler` is a `sampler.MetropolisSampler`; returns the assembled results."""
    import os

    from .backflow import Samples
    from .released import pack_spins

    os.makedirs(out_dir, exist_ok=True)
    w = state.model.workload
    series = {"state": name, "tprime": float(tprime), "LX": w.LX, "LY": w.LY, "family": paper_family(tprime),
              "nsamples": int(sampler.s.shape[0])}
    rec = {"E_mean": [], "E_sqr": []}
    kept, excluded = [], []
    path = os.path.join(out_dir, "series.npz")

    def save():
        tmp = os.path.join(out_dir, "series_tmp.npz")
        np.savez(tmp, **series, **{k: np.asarray(v, np.float64) for k, v in rec.items()},
                 kept_sweeps=np.array(kept, dtype=np.int32), excluded_sweeps=np.array(excluded, dtype=np.int32))
        os.replace(tmp, path)

    s = None
    for n in range(nsweeps):
        smp = sampler.sweep()
        s = smp.spins
        e = H.Oloc(state, Samples(s, smp.psi)).detach().cpu().numpy().astype(np.float64)
        m, q = float(e.mean()), float((e * e).mean())
        if not (np.isfinite(m) and np.isfinite(q)):
            excluded.append(n)
            log(f"sweep {n}: non-finite values, excluded")
            continue
        kept.append(n)
        rec["E_mean"].append(m)
        rec["E_sqr"].append(q)
        log(f"sweep {n}: E/N = {m / w.N:.6f}, sigma^2/N = {(q - m * m) / w.N:.4f}")
        if (n + 1) % SAVE_EVERY == 0:
            save()
    save()
    if s is not None:
        spins = s.detach().cpu().numpy()
        np.savez(os.path.join(out_dir, "spins_last.npz"), **pack_spins(spins))
    res = assemble_energy({**series, **{k: np.asarray(v) for k, v in rec.items()}})
    if kept:
        log(summary(res))
        save_results(res, out_dir)
    return res
