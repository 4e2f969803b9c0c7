"""Accepted-sample throughput of the 16x16 Pfaffian backflow path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--chains C] [--sweep P] [--slice S]
                    [--form both|plain|mixed-da] [--no-cpu] [--workload pf16|tpm02_L16]

Workload: the 16x16 Pfaffian backflow (t' = -0.2, U = 8, 112 + 112 electrons, 224 x 224
Pfaffians, 6-layer 32-channel conv, 8-image projection) with seeded random-init weights, run as
delayed-acceptance Metropolis with 8192 chains resident in HBM and sweeps of 512 exchange
proposals (``--workload tpm02_L16``: the released tpm02_L16_pfaffian state instead).  The chains
start from an equilibrated ensemble of this workload (tests/golden/pf16_chains_8192.npz, written
by the path's own exact sampler over 210 sweeps; the random-init state relaxes slowly: at the end
<log|psi|> still rises by ~0.02 and D/N by ~4e-5 per sweep, i.e. E/N by ~3e-4 per sweep, far
below the E/N error bars of a run).

One step = a slice of S = 16 exchange proposals for all C chains, then log psi_P (carried by
the chain) and E_loc (every connected configuration evaluated in full) of the M = C * S / P = 256
chains whose turn it is: P / S = 32 steps are one sweep of every chain and one measured sample
per chain, as in a measurement run of 8192 chains x 512-proposal sweeps.  Two forms run from the
same ensemble:

  plain      FP32 feature network and Pfaffians, plain Metropolis;
  mixed-da   split-FP16 MFMA feature network (FP32 pairing / Pfaffian) with delayed acceptance
             screened by the conv surrogate, trained beforehand on the GPU by the library's own
             forward / backward kernels on 2^20 (configuration, exact log|psi|) pairs recorded
             from the chains (recording and training time reported apart).

value = samples/s of mixed-da (or of the one form run), inputs resident in HBM.  The line also
carries per form: samples/s, accepted moves/s, an autocorrelation-aware effective sample rate,
acceptance statistics, E/N with block-jackknife errors, the drift of the ensemble over the timed
steps; the ratio to plain FP32; energy_within_error_bars (flagged on stderr when false); the
surrogate's RMS log-ratio error; log|psi| / sign / E_loc against the FP64 C restatement on
sampled configurations, and

  roofline      the dominant kernel class of the headline form, its algorithmic flops per
                launch over the average launch time measured with HIP events on its stream,
                against the MI355X dense peak of its arithmetic type; traffic from the
                committed rocprofv3 counter summary;
  cpu_baseline  the FP32 C restatement of the reference CPU path (oracle/nh_cpu.c, OpenMP over
                all host cores) timed on a bounded sample of the same workload: E_loc plus the
                proposals' amplitude evaluations of one measured sample.
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "samples/s (log-psi+E_loc, 16x16 Pfaffian): mixed+surrogate vs plain FP32"
PEAK = {"fp32": 157.3e12, "fp16": 2.5e15, "bf16": 2.5e15}
CHAINS_FILE = os.path.join(ROOT, "tests", "golden", "pf16_chains_8192.npz")
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "k_image_pf_cm_pf16_hbm_traffic.md")
TRAFFIC_SUMMARY_RESCONV = os.path.join(ROOT, "profiles", "k_resconv_bf16_tpm02_L16_hbm_traffic.md")


def parse():
    p = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=8)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--chains", type=int, default=8192)
    p.add_argument("--sweep", type=int, default=512, help="exchange proposals per chain between its samples")
    p.add_argument("--slice", type=int, default=16, help="exchange proposals per chain per step")
    p.add_argument("--form", default="both", choices=["both", "plain", "mixed-da"])
    p.add_argument("--pairs", type=int, default=1 << 20, help="(configuration, exact log|psi|) pairs for the surrogate")
    p.add_argument("--record-gap", type=int, default=4, help="proposals between recorded states")
    p.add_argument("--epochs", type=int, default=3)
    p.add_argument("--chunk", type=int, default=16, help="E_loc samples per chunk")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and FP64 checks")
    p.add_argument("--cpu-samples", type=int, default=12)
    p.add_argument("--workload", default="pf16", choices=["pf16", "tpm02_L16"],
                   help="pf16: the synthetic 6-layer 32-channel workload (default); tpm02_L16: the "
                        "released tpm02_L16_pfaffian state (16-block 64-channel residual CNN) from --weights, "
                        "chains started from its released chains")
    p.add_argument("--weights", default=os.path.join(ROOT, "tests", "golden", "tpm02_L16_pfaffian.npz"),
                   help="parameters and chains of the released state (written by tests/golden/make_released.py)")
    p.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                   help="only copy BYTES between two device buffers (reference dispatch for HBM counters)")
    return p.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def read_traffic(images_per_launch, summary=TRAFFIC_SUMMARY):
    """HBM bytes per launch of the dominant kernel: the committed rocprofv3 counter summary's
    bytes per image (FETCH_SIZE + WRITE_SIZE, copy-calibrated) times this run's images per launch."""
    try:
        with open(summary) as f:
            m = re.search(r"traffic_bytes_per_image:\s*([0-9.eE+]+)", f.read())
        return float(m.group(1)) * images_per_launch if m else None
    except OSError:
        return None


def decorrelation(series, slice_props, sweep):
    """Autocorrelation of a per-chain observable sampled every `slice_props` proposals (series
    [steps, chains]) over an equilibrated ensemble: deviations from the ensemble mean of each step
    (so a chain that stays high or low for long contributes at every lag), rho(lag) averaged over
    chains and steps; the slow mode rho ~ A exp(-lag / tau) fitted on the upper half of the lags
    (tau in proposals); and the factor by which measuring every `sweep` proposals inflates the
    variance of a mean, 1 + 2 A r / (1 - r) with r = exp(-sweep / tau), whose inverse turns
    samples/s into effective samples/s."""
    import numpy as np

    x = np.asarray(series, np.float64)
    x = x - x.mean(axis=1, keepdims=True)
    var = (x * x).mean()
    K = x.shape[0] // 2
    if K < 2 or var == 0:
        return {"rho": [], "tau_proposals": None, "amplitude": None, "variance_inflation": None}
    rho = np.array([(x[:-k] * x[k:]).mean() / var for k in range(1, K + 1)])
    lags = np.arange(1, K + 1) * slice_props
    tail = slice(K // 2, K)
    ok = rho[tail] > 0.02
    if ok.sum() >= 2:
        slope, icpt = np.polyfit(lags[tail][ok], np.log(rho[tail][ok]), 1)
        tau = float(-1.0 / slope) if slope < 0 else float("inf")
        amp = float(min(1.0, np.exp(icpt)))
    else:  # decorrelated within the measured lags
        tau, amp = float(lags[K // 2]) / 3.0, 0.0
    r = float(np.exp(-sweep / tau)) if np.isfinite(tau) else 1.0
    infl = 1.0 + 2.0 * amp * r / (1.0 - r) if r < 1 else float("inf")
    return {"rho": [float(v) for v in rho[:8]], "tau_proposals": tau, "amplitude": amp, "variance_inflation": infl}


def cpu_leg(wk, params, spins, gpu_psi, gpu_eloc, props_exact):
    """CPU baseline (FP32 C port, timed on all host cores) and the FP64 oracle as checker."""
    import numpy as np

    from oracle.cpu_port import CpuPath

    cores = len(os.sched_getaffinity(0))
    out = {}
    cp = CpuPath(wk, params, bits=32, threads=cores)
    x = spins
    cp.psi(x[:1])  # warm-up (page-in, thread pool)
    # E_loc of sampled configurations, two at a time, for about 15 s
    t0 = time.perf_counter()
    n_done = 0
    while n_done < len(x) and (time.perf_counter() - t0) < 15.0:
        xi = x[n_done:n_done + 2]
        sg, la = cp.psi(xi)
        cp.local_energy(xi, sg, la)
        n_done += len(xi)
    t_eloc = (time.perf_counter() - t0) / n_done
    # projected amplitudes of proposal configurations (the chain's exact evaluations)
    m = 4 * cores
    t0 = time.perf_counter()
    cp.psi(np.resize(x, (m, x.shape[1])))
    t_psi = (time.perf_counter() - t0) / m
    per_sample = t_eloc + props_exact * t_psi
    out["cpu_baseline"] = {"value": 1.0 / per_sample, "unit": "samples/s", "cores": cores, "kind": "port",
                           "sample": f"E_loc of {n_done} sampled configurations ({t_eloc:.2f} s each) plus "
                                     f"{props_exact:.0f} projected amplitudes per sample for the sweep's exact "
                                     f"evaluations ({t_psi * 1e3:.1f} ms each, timed on {m}), oracle/nh_cpu.c "
                                     f"in FP32 on {cores} threads"}
    ref = CpuPath(wk, params, bits=64, threads=cores)
    sg64, la64 = ref.psi(x)
    acc = {}
    ne = len(next(iter(gpu_eloc.values())))
    e64 = ref.local_energy(x[:ne], sg64[:ne], la64[:ne]) if ne else None
    for form, (sg, la) in gpu_psi.items():
        err = np.abs(np.asarray(la) - la64)
        acc[form] = {"configs": int(len(x)), "max_abs_logpsi_err": float(err.max()),
                     "median_abs_logpsi_err": float(np.median(err)),
                     "sign_agreement": float(np.mean(np.asarray(sg) == sg64)), "eloc_configs": int(ne)}
        if ne:
            acc[form]["max_abs_eloc_per_site_err"] = float((np.abs(np.asarray(gpu_eloc[form]) - e64) / wk.N).max())
    out["accuracy_vs_fp64"] = acc
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: --gpus must be 1 (how the path would shard "
                                   "across GPUs is described in the design notes, not launched)"}))
        sys.exit(2)
    import numpy as np
    import torch

    if args.calibrate_copy:
        n = args.calibrate_copy // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copied_bytes": n * 4}))
        return

    from neuralhubbard_amd import _lib
    from neuralhubbard_amd.backflow import Hubbard, LogArray, Samples, Variational, build_backflow, hops
    from neuralhubbard_amd.configs import WORKLOADS, make_params
    from neuralhubbard_amd.measure import energy, within_error_bars
    from neuralhubbard_amd.released import unpack_spins
    from neuralhubbard_amd.sampler import MetropolisSampler, Surrogate

    lib = _lib.load()
    C, S, P = args.chains, args.slice, args.sweep
    if P % S or C % (P // S):
        sys.exit("--sweep must be a multiple of --slice, and --chains a multiple of sweep / slice")
    M = C // (P // S)  # chains measured per step
    if args.workload == "tpm02_L16":
        from neuralhubbard_amd.configs import Params
        from neuralhubbard_amd.released import STATES

        wk = STATES["tpm02_L16_pfaffian"].workload()
        with np.load(args.weights, allow_pickle=False) as f:
            params = Params(list(f["conv_w"]), list(f["conv_b"]), f["W"], float(f["theta"]), None, None, f["emb"])
            released = unpack_spins(f["bits"], f["shape"])
        if C > len(released):
            sys.exit(f"--chains {C} exceeds the {len(released)} released chains")
        start0 = released[:C]
        data = "released tpm02_L16_pfaffian parameters; chains started from its released equilibrated chains"
        net = "16-block 64-channel residual CNN (released tpm02_L16_pfaffian)"
    else:
        wk = WORKLOADS["pf16"]
        params = make_params(wk)
        with np.load(CHAINS_FILE, allow_pickle=False) as f:
            ens = unpack_spins(f["bits"], f["shape"])
            eq_sweeps = len(f["acceptance"])
        if C > len(ens):
            sys.exit(f"--chains {C} exceeds the {len(ens)} chains of {CHAINS_FILE}")
        start0 = ens[:C]
        data = (f"synthetic: random-init weights of the workload's architecture (seeded); chains from an "
                f"ensemble equilibrated by the path's own exact sampler over {eq_sweeps} sweeps")
        net = "6-layer 32-channel conv"
    resconv = wk.network == "resconv"
    N = wk.N
    forms = ["plain", "mixed-da"] if args.form == "both" else [args.form]
    states = {"plain": Variational(build_backflow(wk, params, "fp32")),
              "mixed-da": Variational(build_backflow(wk, params, "mixed"))}
    H = Hubbard(wk.U, 1.0, wk.tprime, chunk=args.chunk)

    # (configuration, exact log|psi|) pairs for the surrogate: exact Metropolis on the mixed
    # amplitude, one recorded state per chain every record_gap proposals
    sur = None
    sur_info = None
    start = torch.as_tensor(start0, device="cuda")
    if "mixed-da" in forms:
        t0 = time.perf_counter()
        rec = MetropolisSampler(states["mixed-da"], start, args.record_gap, seed=1)
        X, Yl = [], []
        n_rec = max(1, args.pairs // C)
        for _ in range(n_rec):
            smp = rec.sweep()
            X.append(smp.spins)
            Yl.append(smp.psi.logabs)
        X = torch.cat(X).cpu().numpy()
        Yl = torch.cat(Yl).cpu().numpy()
        t_record = time.perf_counter() - t0
        log(f"recorded {len(X)} pairs in {t_record:.1f} s")
        ntr = int(0.95 * len(X))
        sur = Surrogate(wk.LX, wk.LY, 32, seed=2)
        t0 = time.perf_counter()
        hist = sur.train(X[:ntr], Yl[:ntr], epochs=args.epochs, batch=1024, lr=2e-3,
                         log=lambda m: log(m))
        torch.cuda.synchronize()
        t_train = time.perf_counter() - t0
        held = sur(X[ntr:]).cpu().numpy()
        sur_info = {"pairs": int(ntr), "heldout": int(len(X) - ntr), "epochs": args.epochs, "record_s": t_record,
                    "train_s": t_train, "train_mse": hist[-1], "heldout_mse": float(np.mean((held - Yl[ntr:]) ** 2)),
                    "target_var": float(Yl.var()), "arithmetic": "fp32 (VALU kernels)"}
        log(f"surrogate trained in {t_train:.1f} s: {sur_info}")
        del X, Yl
        start = rec.s.clone()

    results = {}
    prof_names = ((_lib.PROF_FEATURES, "features"), (_lib.PROF_PAIRING_LINALG, "pairing+pfaffian"),
                  (_lib.PROF_OTHER, "other"))
    for f in forms:
        st = states[f]
        sm = MetropolisSampler(st, start, S, seed=7, surrogate=sur if f == "mixed-da" else None)
        turn = [0]

        def step():
            sm.run(S)
            j = turn[0] % (P // S)
            turn[0] += 1
            sl = slice(j * M, (j + 1) * M)
            spins = sm.s[sl].clone()
            smp = Samples(spins, LogArray(sm.sign[sl].clone(), sm.logabs[sl].clone()))
            return smp, H.Oloc(st, smp)

        for i in range(args.warmup):
            step()
            log(f"{f}: warmup step {i + 1}/{args.warmup} done")
        sm.reset_stats()
        torch.cuda.synchronize()
        lib.nh_profile_reset()
        lib.nh_profile_enable(1)
        e_mean, e_sqr, spins, dser, lser = [], [], [], [], []
        t0 = time.perf_counter()
        for i in range(args.steps):
            smp, e = step()
            e_mean.append(e.mean())
            e_sqr.append((e * e).mean())
            spins.append(smp.spins)
            dser.append(((sm.s[:, :N] > 0) & (sm.s[:, N:] > 0)).sum(1, dtype=torch.int32))
            lser.append(sm.logabs.clone())
            log(f"{f}: step {i + 1}/{args.steps} issued")
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        lib.nh_profile_enable(0)
        kinds = {}
        for kid, name in prof_names:
            n = ctypes.c_int64()
            ms = ctypes.c_double()
            _lib.check(lib.nh_profile_query(kid, ctypes.byref(n), ctypes.byref(ms)), "nh_profile_query")
            kinds[name] = (n.value, ms.value)
        stats = sm.stats()
        n_conn = sum(int(hops(x, wk.LX, wk.LY, 1.0, wk.tprime, wk.n_up + wk.n_dn)[0].sum().item()) for x in spins)
        e_mean = torch.stack(e_mean).cpu().numpy()
        e_sqr = torch.stack(e_sqr).cpu().numpy()
        dser = torch.stack(dser).double().cpu().numpy()
        lser = torch.stack(lser).cpu().numpy()
        dec = decorrelation(dser, S, P)
        rate = args.steps * M / dt
        infl = dec["variance_inflation"]
        drift = {"D_per_site": [float(dser[0].mean() / N), float(dser[-1].mean() / N)],
                 "mean_logabs": [float(lser[0].mean()), float(lser[-1].mean())]}
        results[f] = {"rate": rate, "ms_per_step": dt / args.steps * 1e3, "stats": stats, "kinds": kinds,
                      "n_conn": n_conn, "energy": energy(e_mean, e_sqr, N), "sampler": sm,
                      "accepted_per_s": stats["accepted"] / dt, "decorrelation": dec,
                      "effective_rate": rate / infl if infl and np.isfinite(infl) else None, "drift": drift}
        log(f"{f}: {rate:.2f} samples/s, {stats}, decorrelation {dec}, kernel ms "
            f"{({k: v[1] for k, v in kinds.items()})}")

    head = "mixed-da" if "mixed-da" in forms else "plain"
    r = results[head]
    images = 8 * (r["stats"]["exact_evaluations"] + r["n_conn"])
    flops = {"features": images * wk.cnn_flops(), "pairing+pfaffian": images * (wk.pair_flops() + wk.linalg_flops())}
    dom = max(flops, key=lambda k: r["kinds"][k][1])
    launches, ms = r["kinds"][dom]
    achieved = flops[dom] / (ms * 1e-3)
    arith = ("bf16" if resconv else "fp16") if (head == "mixed-da" and dom == "features") else "fp32"
    traffic = None
    if dom == "pairing+pfaffian" and not resconv:
        traffic = read_traffic(images / max(1, launches))
    elif dom == "features" and resconv and arith == "bf16":
        traffic = read_traffic(images / max(1, launches), TRAFFIC_SUMMARY_RESCONV)
    cnn_kernel = ("k_resconv_bf16" if arith == "bf16" else "k_resconv_fs") if resconv else "k_cnn_mfma"
    roof = {"bound": "mfma", "kernel": "k_image_pf_cm" if dom == "pairing+pfaffian" else cnn_kernel,
            "achieved": achieved / 1e12, "peak": PEAK[arith] / 1e12, "unit": "TFLOP/s",
            "frac": achieved / PEAK[arith], "traffic": traffic, "launches": launches,
            "avg_launch_ms": ms / max(1, launches), "flops_per_launch": flops[dom] / max(1, launches),
            "algorithmic_flops_per_image": (wk.pair_flops() + wk.linalg_flops()) if dom == "pairing+pfaffian"
            else wk.cnn_flops(),
            "images_timed": images, "kernel_class_ms": {k: v[1] for k, v in r["kinds"].items()}}

    out = {
        "metric": METRIC,
        "value": r["rate"],
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp16+fp32" if head == "mixed-da" else "fp32",
        "data": data,
        "config": {"workload": f"16x16 Pfaffian backflow, t'=-0.2, U=8, 112+112 electrons (224x224 Pfaffians), "
                               f"{net}, 8-image projection; {C} chains resident, sweeps of {P} exchange "
                               f"proposals; one step = {S} proposals for every chain, then log psi_P + full E_loc "
                               f"of the {M} chains whose turn it is ({P // S} steps = one sweep and one "
                               f"measured sample per chain)",
                   "form": head, "chains": C, "sweep": P, "slice": S, "measured_per_step": M,
                   "connected_per_sample": r["n_conn"] / (args.steps * M), "parallelism": "single GPU"},
        "roofline": roof,
    }
    for f in forms:
        q = results[f]
        st_ = q["stats"]
        out[f] = {"samples_per_s": q["rate"], "ms_per_step": q["ms_per_step"],
                  "accepted_moves_per_s": q["accepted_per_s"],
                  "effective_samples_per_s": q["effective_rate"],
                  "decorrelation_D": q["decorrelation"],
                  "acceptance": st_["acceptance"], "exact_evaluations": st_["exact_evaluations"],
                  "stage1_acceptance": st_["stage1_acceptance"], "stage2_acceptance": st_["stage2_acceptance"],
                  "stage2_rejection": (1 - st_["stage2_acceptance"]) if st_["stage2_acceptance"] is not None else None,
                  "E_per_site": q["energy"]["E_per_site"], "E_err": q["energy"]["E_err"],
                  "drift_over_timed_steps": q["drift"]}
    if len(forms) == 2:
        out["ratio_vs_plain_fp32"] = results["mixed-da"]["rate"] / results["plain"]["rate"]
        ep, em = results["plain"]["effective_rate"], results["mixed-da"]["effective_rate"]
        out["effective_ratio_vs_plain_fp32"] = em / ep if ep and em else None
        out["accepted_moves_ratio_vs_plain_fp32"] = (results["mixed-da"]["accepted_per_s"] /
                                                     results["plain"]["accepted_per_s"])
        ok = within_error_bars(results["mixed-da"]["energy"], results["plain"]["energy"])
        out["energy_within_error_bars"] = bool(ok)
        if not ok:
            log("WARNING: E/N of mixed-da and plain differ by more than 2 standard errors: "
                f"{results['mixed-da']['energy']['E_per_site']:.5f} +- {results['mixed-da']['energy']['E_err']:.5f} vs "
                f"{results['plain']['energy']['E_per_site']:.5f} +- {results['plain']['energy']['E_err']:.5f}")
    if sur is not None:
        # RMS error of the surrogate's log-ratio on one random changing exchange hop per chain
        sm = results["mixed-da"]["sampler"]
        s = sm.s[:2048].cpu().numpy()
        rng = np.random.default_rng(5)
        sp = s.copy()
        for c in range(len(s)):
            while True:
                i, dirn, spin = rng.integers(N), rng.integers(4), rng.integers(2)
                x, y = divmod(int(i), wk.LY)
                dx, dy = ((1, 0), (-1, 0), (0, 1), (0, -1))[dirn]
                j = ((x + dx) % wk.LX) * wk.LY + (y + dy) % wk.LY
                a, b = spin * N + i, spin * N + j
                if s[c, a] != s[c, b]:
                    sp[c, a], sp[c, b] = s[c, b], s[c, a]
                    break
        st = states["mixed-da"]
        d_exact = (st(sp).logabs - st(s).logabs).cpu().numpy()
        d_sur = (sur(sp) - sur(s)).double().cpu().numpy()
        ok = np.isfinite(d_exact)
        sur_info["rms_log_ratio_err"] = float(np.sqrt(np.mean((d_sur[ok] - d_exact[ok]) ** 2)))
        sur_info["rms_exact_log_ratio"] = float(np.sqrt(np.mean(d_exact[ok] ** 2)))
        out["surrogate"] = sur_info
    if not args.no_cpu:
        n = max(1, min(args.cpu_samples, M))
        x = results[head]["sampler"].s[:n]
        gpu_psi, gpu_eloc = {}, {}
        for f in forms:
            p = states[f](x)
            gpu_psi[f] = (p.sign.cpu().numpy(), p.logabs.cpu().numpy())
            gpu_eloc[f] = np.zeros(0) if resconv else H.Oloc(states[f], Samples(x[:2])).cpu().numpy()
        plain_exact = results["plain"]["stats"]["exact_evaluations"] if "plain" in results else \
            results[head]["stats"]["changing"]
        props_exact = plain_exact / (args.steps * M)
        out.update(cpu_leg(wk, params, x.cpu().numpy(), gpu_psi, gpu_eloc, props_exact))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
