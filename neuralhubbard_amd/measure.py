"""Statistics of a run of sweeps (host side, float64) and the measurement driver with the file
formats of the reference's measurement script.

Error bars are the reference's definition (``code/measure/assemble.py:48-97``): per-sweep
means are grouped into blocks of 1, 2, 4 and 8 consecutive sweeps (an incomplete last block is
dropped); each block size leaving at least 8 blocks gives a delete-one-block jackknife error, and
the quoted error is the largest of them.  `run_measure` writes ``series.npz``,
``spins_last.npz``, ``results.json`` and ``results.npz`` with the reference's keys
(``code/measure_orderparams.py``, ``code/measure/assemble.py:296-411``) for the energy, spin,
charge and d-wave channels; `assemble` turns a series into E/N, M^2_str, N2(q_c), C_p(r) and
Delta^2_SC.
"""

from __future__ import annotations

import numpy as np

BLOCK_SIZES = (1, 2, 4, 8)
MIN_BLOCKS = 8


def delete_one_means(x, b: int):
    """Means over all blocks but one, one row per deleted block, for blocks of b consecutive rows
    of x (the incomplete tail dropped): (nb, ...)."""
    x = np.asarray(x, np.float64)
    nb = x.shape[0] // b
    used = x[: nb * b]
    block_sums = np.add.reduceat(used, np.arange(0, nb * b, b), axis=0)
    return (used.sum(axis=0) - block_sums) / ((nb - 1) * b)


def jackknife(f, *series):
    """(f of the sweep means, error, {block size: error}) of per-sweep series (leading axis =
    sweeps); f may return arrays, the error is the element-wise maximum over the admissible block
    sizes (NaN where none is admissible)."""
    xs = [np.asarray(x, np.float64) for x in series]
    value = f(*(x.mean(axis=0) for x in xs))
    errs = {}
    n = xs[0].shape[0]
    for b in BLOCK_SIZES:
        nb = n // b
        if nb < MIN_BLOCKS:
            continue
        reps = np.stack([np.asarray(f(*cols), np.float64)
                         for cols in zip(*(delete_one_means(x, b) for x in xs))])
        dev = reps - reps.mean(axis=0)
        errs[b] = np.sqrt((nb - 1) / nb * np.sum(dev * dev, axis=0))
    error = np.maximum.reduce(list(errs.values())) if errs else np.full(np.shape(value), np.nan)
    return value, error, errs


def _per_block(errs: dict) -> dict:
    return {int(b): float(np.asarray(e)) for b, e in errs.items()}


def energy(E_mean, E_sqr, N: int) -> dict:
    """E/N and sigma^2/N = (<E^2> - <E>^2) / N with their errors, from per-sweep means of E_loc and
    E_loc^2 (keys of the reference's results.json)."""
    E_mean, E_sqr = np.asarray(E_mean, np.float64), np.asarray(E_sqr, np.float64)
    e, e_err, e_blk = jackknife(lambda m: m / N, E_mean)
    var, var_err, var_blk = jackknife(lambda m, q: (q - m * m) / N, E_mean, E_sqr)
    out = {"E_per_site": e, "E_err": e_err, "var_per_site": var, "var_err": var_err}
    out = {k: float(v) for k, v in out.items()}
    out.update(E_err_blocks=_per_block(e_blk), var_err_blocks=_per_block(var_blk), nsweeps=int(E_mean.shape[0]))
    return out


def within_error_bars(a: dict, b: dict, k: float = 2.0) -> bool:
    """|E_a - E_b| <= k * sqrt(err_a^2 + err_b^2)."""
    return abs(a["E_per_site"] - b["E_per_site"]) <= k * float(np.hypot(a["E_err"], b["E_err"]))


# ------------------------------------------------------------------------ a run of sweeps

CHANNELS = ("energy", "spin", "charge", "dwave")
NREF = 8  # reference sites of the transverse spin terms (code/measure_orderparams.py:27)
SAVE_EVERY = 4  # sweeps between rewrites of series.npz (code/measure_orderparams.py:28)


def paper_family(tprime: float) -> str:
    """Stripe family label stored with a series: pi/8 stripes at t' = 0, pi/4 otherwise."""
    return "pi8" if tprime == 0 else "pi4"


def spin(Czz, Tpm, Tmp, LX: int, LY: int, family: str) -> dict:
    """Spin correlation C_s(r) = C_zz(r) + (T+-(r) + T-+(r)) / 2 per sweep (transverse terms of
    several reference sites averaged), its structure factor, and the stripe order
    M^2_str = (1/4N) sum_alpha S(k_alpha) with the zz part alone as M2_zz."""
    from .observables import stripe_k, stripe_weights, structure_factor

    Czz = np.asarray(Czz, np.float64)
    Tpm, Tmp = (np.asarray(t, np.float64) for t in (Tpm, Tmp))
    if Tpm.ndim == 3:  # (sweeps, reference sites, N)
        Tpm, Tmp = Tpm.mean(axis=1), Tmp.mean(axis=1)
    Cs = Czz + 0.5 * (Tpm + Tmp)
    w = stripe_weights(LX, LY, stripe_k(LX, LY, family))  # M^2 = w . C, linear in C
    m2, m2_err, m2_blk = jackknife(lambda c: c @ w, Cs)
    mzz, mzz_err, _ = jackknife(lambda c: c @ w, Czz)
    C, C_err, _ = jackknife(lambda c: c, Cs)
    return {"C_s": C, "C_s_err": C_err, "S_k": structure_factor(C, LX, LY), "k_alpha": np.array(stripe_k(LX, LY, family)),
            "family": family, "M2_str": float(m2), "M2_str_err": float(m2_err), "M2_str_err_blocks": _per_block(m2_blk),
            "M2_zz": float(mzz), "M2_zz_err": float(mzz_err), "nsweeps": int(Cs.shape[0])}


def charge(Cnn, LX: int, LY: int, family: str) -> dict:
    """Charge correlation C_c(r) per sweep, its structure factor N(q) and
    N2 = (1/4N) sum_alpha N(q_alpha) at the charge wavevectors."""
    from .observables import charge_q, stripe_weights, structure_factor

    Cnn = np.asarray(Cnn, np.float64)
    w = stripe_weights(LX, LY, charge_q(LX, LY, family))
    n2, n2_err, _ = jackknife(lambda c: c @ w, Cnn)
    C, C_err, _ = jackknife(lambda c: c, Cnn)
    return {"C_c": C, "C_c_err": C_err, "N_q": structure_factor(C, LX, LY), "q_c": np.array(charge_q(LX, LY, family)),
            "N2": float(n2), "N2_err": float(n2_err), "nsweeps": int(Cnn.shape[0])}


def _local(series, LX, LY, family):
    sz = np.asarray(series["sz_site"], np.float64).mean(axis=0)
    n = np.asarray(series["n_site"], np.float64).mean(axis=0)
    return {"max_abs_sz": float(np.abs(sz).max()), "n_min": float(n.min()), "n_max": float(n.max())}


def _dwave(series, LX, LY, family):
    from .observables import dwave

    dx, dy = np.divmod(np.arange(LX * LY), LY)
    return {"dr": np.stack([dx, dy], 1), **dwave(series["F"], series["G"], LX, LY), "translation_fill": True}


# result key, the series keys it needs, how it is assembled
_ASSEMBLERS = (
    ("energy", ("E_mean", "E_sqr"), lambda s, LX, LY, fam: energy(s["E_mean"], s["E_sqr"], LX * LY)),
    ("spin", ("Czz", "Tpm", "Tmp"), lambda s, LX, LY, fam: spin(s["Czz"], s["Tpm"], s["Tmp"], LX, LY, fam)),
    ("charge", ("Cnn",), lambda s, LX, LY, fam: charge(s["Cnn"], LX, LY, fam)),
    ("dwave", ("F", "G"), _dwave),
    ("local", ("sz_site", "n_site"), _local),
)


def assemble(series: dict) -> dict:
    """Every observable a series (the content of series.npz) holds."""
    LX, LY = int(series["LX"]), int(series["LY"])
    family = str(series["family"])
    res = {"LX": LX, "LY": LY, "family": family}
    res.update({k: np.asarray(series[k]).item() for k in ("state", "tprime", "nsamples") if k in series})
    for key, needs, make in _ASSEMBLERS:
        if all(k in series for k in needs) and len(series[needs[0]]):
            res[key] = make(series, LX, LY, family)
    return res


assemble_energy = assemble


def summary(res: dict) -> str:
    """Human-readable summary of assembled results, one line per observable."""
    def pm(d, k, err, fmt):
        return f"{d[k]:{fmt}} +- {d[err]:.1g}"

    out = [f"[{res.get('state', '') or 'state'}] {res['LX']}x{res['LY']} lattice, stripe family {res['family']}"]
    if "energy" in res:
        e = res["energy"]
        out.append(f"  energy   E/N {pm(e, 'E_per_site', 'E_err', '.6f')}  var/N {pm(e, 'var_per_site', 'var_err', '.5f')}  "
                   f"[{e['nsweeps']} sweeps]")
    if "spin" in res:
        sp = res["spin"]
        out.append(f"  spin     M2_str {pm(sp, 'M2_str', 'M2_str_err', '.6f')}  M2_zz {sp['M2_zz']:.6f}  C_s(0) {sp['C_s'][0]:.5f}  "
                   f"k_alpha {sp['k_alpha'].tolist()}")
    if "charge" in res:
        c = res["charge"]
        out.append(f"  charge   N2 {pm(c, 'N2', 'N2_err', '.6f')}  sum_r C_c {c['C_c'].sum():.1e}  q_c {c['q_c'].tolist()}")
    if "dwave" in res:
        d = res["dwave"]
        out.append(f"  d-wave   Delta2_SC {pm(d, 'Delta2_SC', 'Delta2_SC_err', '.6f')}  Delta_SC {d['Delta_SC']:.6f}  "
                   f"[{d['n_far']} displacements |r| >= {d['r_max']:.4f}, {d['nsweeps']} sweeps]")
    if "local" in res:
        loc = res["local"]
        out.append(f"  local    max|<S^z_i>| {loc['max_abs_sz']:.1e}  <n_i> in [{loc['n_min']:.4f}, {loc['n_max']:.4f}]")
    return "\n".join(out)


def _to_json(v):
    """JSON-ready copy: numpy scalars to Python numbers, non-finite floats to None, keys to str."""
    if isinstance(v, dict):
        return {str(k): _to_json(x) for k, x in v.items()}
    v = v.item() if isinstance(v, np.generic) else v
    return None if isinstance(v, float) and not np.isfinite(v) else v


def split_results(res: dict):
    """(scalars, maps) of assembled results: the maps are the arrays of rank >= 1 one level down
    (results.npz keys are their inner names), everything else stays in the scalar tree."""
    scalars, maps = {}, {}
    for key, val in res.items():
        if not isinstance(val, dict):
            scalars[key] = _to_json(val)
            continue
        inner = scalars.setdefault(key, {})
        for k, v in val.items():
            if isinstance(v, np.ndarray) and v.ndim:
                maps[k] = v
            else:
                inner[k] = _to_json(v)
    return scalars, maps


def save_results(res: dict, out_dir: str) -> None:
    """results.json (scalars; NaN and inf written as null) and results.npz (the maps) in out_dir."""
    import json
    import os

    scalars, maps = split_results(res)
    text = json.dumps(scalars, indent=1)  # every non-finite float is None by now
    with open(os.path.join(out_dir, "results.json"), "w") as fh:
        fh.write(text)
    np.savez(os.path.join(out_dir, "results.npz"), **maps)


def run_measure(state, H, sampler, nsweeps: int, out_dir: str, name: str = "", tprime: float = 0.0,
                channels=CHANNELS, nbar: float | None = None, log=print) -> dict:
    """``code/measure_orderparams.py:108-174``: every sweep's samples get psi once; the selected
    channels add their per-sweep means to ``series.npz`` (rewritten every SAVE_EVERY sweeps; a sweep
    with any non-finite value is excluded):

    * energy: E_mean, E_sqr of E_loc;
    * spin / charge: sz_site, n_site (LX, LY), Czz, Cnn (N,) and the transverse terms Tpm, Tmp
      (NREF, N) of the reference sites `observables.reference_sites` (NREF of them, fewer on
      lattices narrower than NREF) — sample k serves reference site k mod NREF, so that each site's terms are means over B / NREF samples (the reference
      evaluates every site on every sample; both estimate the same expectation);
    * dwave: F (16N,), G (6N,) in the order of F_keys, G_keys, which the series holds.

    The final chains go to ``spins_last.npz`` in the released-chain format, the assembled
    observables to ``results.json`` / ``results.npz``.  `sampler` is a `sampler.MetropolisSampler`
    (or anything whose ``sweep()`` returns spins and their LogArray psi); returns the assembled
    results."""
    import os
    import time

    import torch

    from . import observables as obs
    from .backflow import LogArray, Samples
    from .released import pack_spins

    channels = tuple(channels)
    if not set(channels) <= set(CHANNELS):
        raise ValueError(f"channels must be a subset of {CHANNELS}, got {channels}")
    os.makedirs(out_dir, exist_ok=True)
    w = state.model.workload
    LX, LY, N = w.LX, w.LY, w.N
    nbar = (w.n_up + w.n_dn) / N if nbar is None else float(nbar)
    family = paper_family(tprime)
    series = {"state": name, "tprime": float(tprime), "LX": LX, "LY": LY, "family": family,
              "nsamples": int(sampler.s.shape[0])}
    if {"spin", "charge"} & set(channels):
        obs.stripe_k(LX, LY, family)  # fail early if q is not on the momentum grid
    if "spin" in channels:
        refs = obs.reference_sites(LX, LY, NREF)
        series["REF_SITES"] = np.array(refs)
    if "charge" in channels:
        series["nbar"] = nbar
    if "dwave" in channels:
        series["F_keys"], series["G_keys"] = obs.dwave_keys(LX, LY)
    rec, kept, excluded = {}, [], []
    path = os.path.join(out_dir, "series.npz")

    def save():
        tmp = os.path.join(out_dir, "series_tmp.npz")
        np.savez(tmp, **series, **{k: np.stack(v) for k, v in rec.items()},
                 kept_sweeps=np.array(kept, dtype=np.int32), excluded_sweeps=np.array(excluded, dtype=np.int32))
        os.replace(tmp, path)

    def np64(t):
        return t.detach().cpu().numpy().astype(np.float64)

    t0 = time.time()
    s = None
    for n in range(nsweeps):
        smp = sampler.sweep()
        s = smp.spins
        psi = smp.psi
        if psi is None and {"spin", "dwave"} & set(channels):
            psi = state(s)
        cur = {}
        if "energy" in channels:
            e = np64(H.Oloc(state, Samples(s, psi)))
            cur["E_mean"], cur["E_sqr"] = e.mean(), (e * e).mean()
        if {"spin", "charge"} & set(channels):
            gz, gn = obs.sz_grid(s, N), obs.dn_grid(s, N, nbar)
            cur["sz_site"] = np64(gz.mean(0)).reshape(LX, LY)
            cur["n_site"] = np64(gn.mean(0)).reshape(LX, LY) + nbar
            if "spin" in channels:
                cur["Czz"] = np64(obs.autocorr(gz, LX, LY))
            if "charge" in channels:
                cur["Cnn"] = np64(obs.autocorr(gn, LX, LY))
        if "spin" in channels:
            B = s.shape[0]
            T = np.zeros((len(refs), 2, N))
            for i, (ax, ay) in enumerate(refs):
                sel = np.arange(i, B, len(refs))
                if len(sel) == 0:
                    raise ValueError(f"{B} samples leave reference site {i} without samples")
                sel_t = torch.as_tensor(sel, device=s.device)
                Ti = obs.transverse_local(state, s[sel_t], LogArray(psi.sign[sel_t], psi.logabs[sel_t]),
                                          ax * LY + ay, LX, LY)
                T[i] = np64(Ti.mean(0))
            cur["Tpm"], cur["Tmp"] = T[:, 0], T[:, 1]
        if "dwave" in channels:
            F, G = obs.dwave_measure(state, s, psi, LX, LY)
            cur["F"], cur["G"] = np64(F).reshape(-1), np64(G).reshape(-1)  # flat, in F_keys / G_keys order
        if not all(np.isfinite(v).all() for v in cur.values()):
            excluded.append(n)
            log(f"sweep {n}: non-finite values, excluded")
            continue
        kept.append(n)
        for k, v in cur.items():
            rec.setdefault(k, []).append(v)
        msg = f"sweep {n}: {time.time() - t0:.0f} s"
        if "energy" in channels:
            msg += f", E/N = {cur['E_mean'] / N:.6f}, sigma^2/N = {(cur['E_sqr'] - cur['E_mean'] ** 2) / N:.4f}"
        log(msg)
        if (n + 1) % SAVE_EVERY == 0:
            save()
    save()
    if s is not None:
        np.savez(os.path.join(out_dir, "spins_last.npz"), **pack_spins(s.detach().cpu().numpy()))
    res = assemble({**series, **{k: np.stack(v) for k, v in rec.items()}}) if kept else {}
    if kept:
        log(summary(res))
        save_results(res, out_dir)
    return res


def run_energy(state, H, sampler, nsweeps: int, out_dir: str, name: str = "", tprime: float = 0.0,
               log=print) -> dict:
    """`run_measure` with the energy channel alone."""
    return run_measure(state, H, sampler, nsweeps, out_dir, name, tprime, ("energy",), log=log)


# ------------------------------------------------------------------------ command line

def _load_state_files(name: str, weights: str, spins: str | None):
    """(workload, params, chains) of a released state from its ``.eqx`` weight file (and
    ``_spins.npz`` chains), or from a parameter ``.npz`` holding the same arrays and the packed
    chains (``conv_w``, ``conv_b``, ``W``, ``theta``, ``emb``, optional ``head`` / ``phi0``,
    ``bits``, ``shape``)."""
    from .configs import Params
    from .released import STATES, load_released, load_spins, unpack_spins

    info = STATES[name]
    if weights.endswith(".eqx"):
        wk, params = load_released(name, weights)
        chains = None
    else:
        with np.load(weights, allow_pickle=False) as f:
            g = {k: f[k] for k in f.files}
        wk = info.workload()
        params = Params(list(g["conv_w"]), list(g["conv_b"]), g["W"], float(g["theta"]), None, g.get("phi0"),
                        g["emb"], g.get("head"))
        chains = unpack_spins(g["bits"], g["shape"]) if "bits" in g else None
    if spins is not None:
        chains = load_spins(spins, info)
    if chains is None:
        raise ValueError("no initial chains: give --spins (weights/<state>_spins.npz)")
    return wk, params, chains


def main(argv=None) -> dict:
    """The command line of ``code/measure_orderparams.py`` on a released state:

        python -m neuralhubbard_amd.measure --state tpm02_L16_pfaffian \\
            --weights weights/tpm02_L16_pfaffian.eqx --spins weights/tpm02_L16_pfaffian_spins.npz \\
            --out runs/tpm02_L16_pfaffian

    The chains start at the released chains (repeated or truncated to --nsamples) and move by
    the paper's sampler (`released.make_sampler`: particle hops and site exchanges with the
    state's neighbour ranges; --moves exchange keeps the plain same-spin exchange), thermalised
    for --thermal x N steps and measured every --sweep-steps x N steps (defaults: the state's
    values, ``code/states.py:105-119``).  Writes series.npz, spins_last.npz, results.json and
    results.npz into --out."""
    import argparse

    import torch

    from .backflow import Hubbard, Variational, build_backflow
    from .released import STATES, make_sampler
    from .sampler import MetropolisSampler

    p = argparse.ArgumentParser(description="Measure energy, spin, charge and d-wave correlations of a released state.")
    p.add_argument("--state", required=True, choices=sorted(STATES))
    p.add_argument("--weights", required=True, help="weights/<state>.eqx, or a parameter .npz with packed chains")
    p.add_argument("--spins", help="weights/<state>_spins.npz (default: the chains of a parameter .npz)")
    p.add_argument("--out", required=True, help="output folder")
    p.add_argument("--nsamples", type=int, default=8192, help="chains, one sample each per sweep")
    p.add_argument("--nsweeps", type=int, default=64)
    p.add_argument("--thermal", type=float, help="thermalisation steps per chain / N (default: the state's)")
    p.add_argument("--sweep-steps", type=float, help="steps per chain between sweeps / N (default: the state's)")
    p.add_argument("--moves", default="mix", choices=["mix", "exchange"],
                   help="mix: the paper's particle hops + site exchanges; exchange: same-spin NN exchange")
    p.add_argument("--channels", default=",".join(CHANNELS), help="comma-separated subset of " + ",".join(CHANNELS))
    p.add_argument("--precision", default="mixed", choices=["fp32", "mixed"])
    p.add_argument("--seed", type=int, default=42)
    args = p.parse_args(argv)
    channels = tuple(c for c in args.channels.split(",") if c)
    if not set(channels) <= set(CHANNELS):
        p.error(f"--channels must be a subset of {','.join(CHANNELS)}")
    info = STATES[args.state]
    wk, params, chains = _load_state_files(args.state, args.weights, args.spins)
    state = Variational(build_backflow(wk, params, args.precision))
    start = np.resize(chains, (args.nsamples, 2 * wk.N))
    sweep = int(round((info.sweeps[0] if args.sweep_steps is None else args.sweep_steps) * wk.N))
    thermal = int(round((info.sweeps[1] if args.thermal is None else args.thermal) * wk.N))
    print(f"{info.name}: {wk.LX}x{wk.LY}, {args.nsamples} chains, {args.precision}, {args.moves} moves; "
          f"thermalising", flush=True)
    if args.moves == "mix":
        sampler = make_sampler(state, info, args.nsamples, start, sweep, thermal, seed=args.seed)
    else:
        sampler = MetropolisSampler(state, torch.as_tensor(start), sweep, seed=args.seed, thermal_steps=thermal)
    H = Hubbard(U=info.U, t=1.0, tprime=info.tprime)
    return run_measure(state, H, sampler, args.nsweeps, args.out, name=info.name, tprime=info.tprime,
                       channels=channels)


if __name__ == "__main__":
    main()
