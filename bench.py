"""Accepted-sample throughput of the 16x16 Pfaffian backflow path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--chains C] [--sweep P]
                    [--form both|plain|mixed-da] [--no-cpu] [--workload pf16|tpm02_L16]

Workload: the 16x16 Pfaffian backflow (t' = -0.2, U = 8, 112 + 112 electrons, 224 x 224
Pfaffians, 6-layer 32-channel conv, 8-image projection) with seeded random-init weights
(``--workload tpm02_L16``: the released tpm02_L16_pfaffian state instead, 16-block 64-channel
residual CNN, chains started from its released chains).  C
Metropolis chains live in HBM.  One step = one sweep of P exchange proposals per chain, then
log psi_P and E_loc (every connected configuration evaluated in full) of every chain's state:
C measured samples per step.  Two forms run from the same thermalised chains:

  plain      FP32 feature network and Pfaffians, plain Metropolis;
  mixed-da   split-BF16 MFMA feature network (FP32 pairing/Pfaffian) with delayed acceptance
             screened by the conv surrogate, trained beforehand on the GPU on (configuration,
             exact log|psi|) pairs recorded while thermalising (training time reported apart).

value = samples/s of mixed-da (with --form both or mixed-da), inputs resident in HBM.  The JSON
line also carries the plain FP32 rate and the ratio, acceptance statistics, the surrogate's RMS
log-ratio error, E/N of both forms with block-jackknife errors, log|psi| and sign agreement
with the FP64 oracle on a few of the sampled configurations, and

  roofline      the dominant kernel class of the headline form, its algorithmic flops per
                launch over the average launch time measured with HIP events on its stream,
                against the MI355X dense peak of its arithmetic type; traffic from the
                committed rocprofv3 counter summary when present;
  cpu_baseline  the FP32 C restatement of the reference CPU path (oracle/nh_cpu.c, OpenMP) on
                a bounded sample of the same workload on this host's cores.
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
PEAK = {"fp32": 157.3e12, "bf16": 2.5e15}
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "k_image_pf_pf16_hbm_traffic.md")
TRAFFIC_SUMMARY_RESCONV = os.path.join(ROOT, "profiles", "k_resconv_bf16_tpm02_L16_hbm_traffic.md")


def parse():
    p = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=8)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--chains", type=int, default=256)
    p.add_argument("--sweep", type=int, default=512, help="exchange proposals per chain per step")
    p.add_argument("--form", default="both", choices=["both", "plain", "mixed-da"])
    p.add_argument("--thermal", type=int, default=256, help="plain proposals per chain before recording")
    p.add_argument("--records", type=int, default=256, help="recorded states per chain for the surrogate")
    p.add_argument("--record-gap", type=int, default=4)
    p.add_argument("--epochs", type=int, default=8)
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


def cpu_leg(wk, params, spins, gpu_psi, gpu_eloc, per_call=2):
    """CPU baseline (FP32 C port, timed) and the FP64 oracle as checker of the GPU amplitudes."""
    import numpy as np

    from oracle.cpu_port import CpuPath

    cores = min(16, len(os.sched_getaffinity(0)))
    out = {}
    cp = CpuPath(wk, params, bits=32, threads=cores)
    x = spins
    cp.psi(x[:1])  # warm-up (page-in, thread pool)
    # per_call configurations at a time (2: 16 images keep 16 threads busy) until about 15 s
    t0 = time.perf_counter()
    n_done = 0
    while n_done < len(x) and (time.perf_counter() - t0) < 15.0:
        xi = x[n_done:n_done + per_call]
        sg, la = cp.psi(xi)
        cp.local_energy(xi, sg, la)
        n_done += len(xi)
    dt = time.perf_counter() - t0
    out["cpu_baseline"] = {"value": n_done / dt, "unit": "samples/s", "cores": cores, "kind": "port",
                           "sample": f"{n_done} sampled configurations of the same workload, {per_call} at a time: "
                                     f"projected psi + full E_loc, no sweep ({dt:.1f} s on {cores} threads, "
                                     f"oracle/nh_cpu.c in FP32)"}
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
    from neuralhubbard_amd.backflow import Hubbard, Samples, Variational, build_backflow, hops
    from neuralhubbard_amd.configs import WORKLOADS, make_configurations, make_params
    from neuralhubbard_amd.measure import energy, within_error_bars
    from neuralhubbard_amd.sampler import MetropolisSampler, Surrogate

    lib = _lib.load()
    C = args.chains
    if args.workload == "tpm02_L16":
        from neuralhubbard_amd.configs import Params
        from neuralhubbard_amd.released import STATES, unpack_spins

        wk = STATES["tpm02_L16_pfaffian"].workload()
        with np.load(args.weights, allow_pickle=False) as f:
            params = Params(list(f["conv_w"]), list(f["conv_b"]), f["W"], float(f["theta"]), None, None, f["emb"])
            released = unpack_spins(f["bits"], f["shape"])
        start0 = np.resize(released, (C, 2 * wk.N))
        data = "released tpm02_L16_pfaffian parameters; chains started from its released equilibrated chains"
        net = "16-block 64-channel residual CNN (released tpm02_L16_pfaffian)"
    else:
        wk = WORKLOADS["pf16"]
        params = make_params(wk)
        start0 = make_configurations(wk, C)
        data = ("synthetic: seeded random configurations thermalised by the path's own Metropolis chains, "
                "random-init weights of the workload's architecture")
        net = "6-layer 32-channel conv"
    resconv = wk.network == "resconv"
    forms = ["plain", "mixed-da"] if args.form == "both" else [args.form]
    states = {"plain": Variational(build_backflow(wk, params, "fp32")),
              "mixed-da": Variational(build_backflow(wk, params, "mixed"))}
    H = Hubbard(wk.U, 1.0, wk.tprime, chunk=args.chunk)

    # thermalise with exact plain Metropolis on the mixed-precision amplitude and record the
    # (configuration, exact log|psi|) pairs the surrogate learns from
    t0 = time.perf_counter()
    rec = MetropolisSampler(states["mixed-da"], start0, args.record_gap, seed=1,
                            thermal_steps=args.thermal)
    X, Y = [], []
    for _ in range(args.records):
        smp = rec.sweep()
        X.append(smp.spins)
        Y.append(smp.psi.logabs)
    X = torch.cat(X).cpu().numpy()
    Y = torch.cat(Y).cpu().numpy()
    t_record = time.perf_counter() - t0
    log(f"thermalised + recorded {len(X)} pairs in {t_record:.1f} s")
    sur = None
    sur_info = None
    if "mixed-da" in forms:
        ntr = int(0.9 * len(X))
        sur = Surrogate(wk.LX, wk.LY, 32, seed=2)
        t0 = time.perf_counter()
        hist = sur.train(X[:ntr], Y[:ntr], epochs=args.epochs, batch=512, lr=2e-3)
        torch.cuda.synchronize()
        t_train = time.perf_counter() - t0
        held = sur(X[ntr:]).cpu().numpy()
        sur_info = {"pairs": int(ntr), "epochs": args.epochs, "train_s": t_train, "train_mse": hist[-1],
                    "heldout_mse": float(np.mean((held - Y[ntr:]) ** 2)), "target_var": float(Y.var())}
        log(f"surrogate trained in {t_train:.1f} s: {sur_info}")
    start = rec.s.clone()

    results = {}
    prof_names = ((_lib.PROF_FEATURES, "features"), (_lib.PROF_PAIRING_LINALG, "pairing+pfaffian"),
                  (_lib.PROF_OTHER, "other"))
    for f in forms:
        st = states[f]
        sm = MetropolisSampler(st, start, args.sweep, seed=7, surrogate=sur if f == "mixed-da" else None)

        def step():
            smp = sm.sweep()
            e = H.Oloc(st, smp)
            return smp, e

        for i in range(args.warmup):
            step()
            log(f"{f}: warmup step {i + 1}/{args.warmup} done")
        sm.reset_stats()
        torch.cuda.synchronize()
        lib.nh_profile_reset()
        lib.nh_profile_enable(1)
        e_mean, e_sqr, spins = [], [], []
        t0 = time.perf_counter()
        for i in range(args.steps):
            smp, e = step()
            e_mean.append(e.mean())
            e_sqr.append((e * e).mean())
            spins.append(smp.spins)
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
        results[f] = {"rate": args.steps * C / dt, "ms_per_step": dt / args.steps * 1e3, "stats": stats,
                      "kinds": kinds, "n_conn": n_conn, "energy": energy(e_mean, e_sqr, wk.N), "sampler": sm}
        log(f"{f}: {results[f]['rate']:.2f} samples/s, {stats}, kernel ms {({k: v[1] for k, v in kinds.items()})}")

    head = "mixed-da" if "mixed-da" in forms else "plain"
    r = results[head]
    images = 8 * (r["stats"]["exact_evaluations"] + r["n_conn"])
    flops = {"features": images * wk.cnn_flops(), "pairing+pfaffian": images * (wk.pair_flops() + wk.linalg_flops())}
    dom = max(flops, key=lambda k: r["kinds"][k][1])
    launches, ms = r["kinds"][dom]
    achieved = flops[dom] / (ms * 1e-3)
    arith = "bf16" if (head == "mixed-da" and dom == "features") else "fp32"
    traffic = None
    if dom == "pairing+pfaffian" and not resconv:
        traffic = read_traffic(images / max(1, launches))
    elif dom == "features" and resconv and arith == "bf16":
        traffic = read_traffic(images / max(1, launches), TRAFFIC_SUMMARY_RESCONV)
    cnn_kernel = ("k_resconv_bf16" if arith == "bf16" else "k_resconv_fs") if resconv else "k_cnn_mfma"
    roof = {"bound": "mfma", "kernel": "k_image_pf" if dom == "pairing+pfaffian" else cnn_kernel,
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
        "dtype": "bf16+fp32" if head == "mixed-da" else "fp32",
        "data": data,
        "config": {"workload": f"16x16 Pfaffian backflow, t'=-0.2, U=8, 112+112 electrons (224x224 Pfaffians), "
                               f"{net}, 8-image projection; {C} chains, {args.sweep} exchange "
                               f"proposals per chain per step, then log psi_P + full E_loc of every chain",
                   "form": head, "chains": C, "sweep": args.sweep,
                   "connected_per_sample": r["n_conn"] / (args.steps * C), "parallelism": "single GPU"},
        "roofline": roof,
    }
    for f in forms:
        q = results[f]
        out[f] = {"samples_per_s": q["rate"], "ms_per_step": q["ms_per_step"],
                  "acceptance": q["stats"]["acceptance"], "exact_evaluations": q["stats"]["exact_evaluations"],
                  "stage1_acceptance": q["stats"]["stage1_acceptance"],
                  "stage2_acceptance": q["stats"]["stage2_acceptance"],
                  "stage2_rejection": (1 - q["stats"]["stage2_acceptance"]) if q["stats"]["stage2_acceptance"]
                  is not None else None,
                  "E_per_site": q["energy"]["E_per_site"], "E_err": q["energy"]["E_err"]}
    if len(forms) == 2:
        out["ratio_vs_plain_fp32"] = results["mixed-da"]["rate"] / results["plain"]["rate"]
        out["energy_within_error_bars"] = within_error_bars(results["mixed-da"]["energy"], results["plain"]["energy"])
    if sur is not None:
        # RMS error of the surrogate's log-ratio on one random exchange hop per chain
        sm = results["mixed-da"]["sampler"]
        s = sm.s.cpu().numpy()
        rng = np.random.default_rng(5)
        sp = s.copy()
        N = wk.N
        for c in range(C):
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
        # the residual CNN costs ~30 s of host time per sample's E_loc: one at a time, and its FP64
        # E_loc parity is left to tests/test_gpu_released.py
        n = max(1, min(args.cpu_samples, C))
        x = results[head]["sampler"].s[:n]
        gpu_psi, gpu_eloc = {}, {}
        for f in forms:
            p = states[f](x)
            gpu_psi[f] = (p.sign.cpu().numpy(), p.logabs.cpu().numpy())
            gpu_eloc[f] = np.zeros(0) if resconv else H.Oloc(states[f], Samples(x[:2])).cpu().numpy()
        out.update(cpu_leg(wk, params, x.cpu().numpy(), gpu_psi, gpu_eloc, per_call=1 if resconv else 2))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
