"""Throughput of the projected amplitude + local energy on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload pf16] [--batch B]
                    [--precision fp32|mixed] [--no-cpu]

One step evaluates, for B configurations of the workload's seeded synthetic batch (a
different slice each step), the projected log psi (8 images) and E_loc (every connected
configuration by a full projected evaluation).  value = configurations per second with
log psi and E_loc, the whole job on one GPU with inputs resident in HBM.

The JSON line also carries:
  roofline      the dominant kernel class, timed by HIP events on its launch stream over the
                timed steps, against the MI355X dense peak for its arithmetic type;
  cpu_baseline  the C restatement of the reference CPU path (oracle/nh_cpu.c, FP32, OpenMP)
                on a bounded sample of the same workload on this host's cores.
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "samples/s (log-psi+E_loc, 16x16 Pfaffian): mixed+surrogate vs plain FP32"
PEAK = {"fp32": 157.3e12, "bf16": 2.5e15, "fp16": 2.5e15}


def parse():
    p = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=4)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--workload", default="pf16")
    p.add_argument("--batch", type=int, default=64, help="configurations per step")
    p.add_argument("--precision", default="fp32", choices=["fp32", "mixed"])
    p.add_argument("--chunk", type=int, default=16, help="E_loc samples per chunk")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--cpu-samples", type=int, default=2)
    return p.parse_args()


def cpu_baseline(wk, params, s, samples):
    """Time the FP32 C port on `samples` configurations: projected psi + E_loc."""
    from oracle.cpu_port import CpuPath

    cores = min(16, len(os.sched_getaffinity(0)))
    cp = CpuPath(wk, params, bits=32, threads=cores)
    x = s[:samples]
    cp.psi(x[:1])  # warm-up (page-in, thread pool)
    t0 = time.perf_counter()
    sg, la = cp.psi(x)
    cp.local_energy(x, sg, la)
    dt = time.perf_counter() - t0
    return {"value": samples / dt, "unit": "samples/s", "cores": cores, "kind": "port",
            "sample": f"{samples} configuration(s) of the same workload: projected psi + full E_loc "
                      f"({dt:.1f} s on {cores} threads, oracle/nh_cpu.c FP32)"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (multi-GPU sharding is "
                                   "described in DESIGN.md, not launched)"}))
        sys.exit(2)
    import numpy as np
    import torch

    from neuralhubbard_amd import _lib
    from neuralhubbard_amd.backflow import Hubbard, Samples, Variational, build_backflow, hops
    from neuralhubbard_amd.configs import WORKLOADS, make_configurations, make_params

    wk = WORKLOADS[args.workload]
    params = make_params(wk)
    s_host = make_configurations(wk)  # the seeded synthetic batch
    s_all = torch.tensor(s_host, device="cuda")
    model = build_backflow(wk, params, args.precision)
    state = Variational(model)
    H = Hubbard(wk.U, 1.0, wk.tprime, chunk=args.chunk)
    B = args.batch
    nslices = max(1, s_all.shape[0] // B)
    lib = _lib.load()

    def step(i):
        x = s_all[(i % nslices) * B:(i % nslices + 1) * B]
        psi = state(x)
        return H.Oloc(state, Samples(x, psi))

    for i in range(args.warmup):
        step(i)
    torch.cuda.synchronize()
    # connected configurations of the timed slices, for the algorithmic work count
    n_conn = 0
    for i in range(args.warmup, args.warmup + args.steps):
        x = s_all[(i % nslices) * B:(i % nslices + 1) * B]
        n_conn += int(hops(x, wk.LX, wk.LY, 1.0, wk.tprime, wk.n_up + wk.n_dn)[0].sum().item())
    lib.nh_profile_reset()
    lib.nh_profile_enable(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.warmup, args.warmup + args.steps):
        e = step(i)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    lib.nh_profile_enable(0)
    assert np.isfinite(e.cpu().numpy()).all()

    kinds = {}
    for kid, name in ((_lib.PROF_FEATURES, "features"), (_lib.PROF_PAIRING_LINALG, "pairing+pfaffian")):
        n = ctypes.c_int64()
        ms = ctypes.c_double()
        _lib.check(lib.nh_profile_query(kid, ctypes.byref(n), ctypes.byref(ms)), "nh_profile_query")
        kinds[name] = (n.value, ms.value)
    lib.nh_profile_reset()
    images = 8 * (args.steps * B + n_conn)
    flops = {"features": images * wk.cnn_flops(), "pairing+pfaffian": images * (wk.pair_flops() + wk.linalg_flops())}
    dom = max(kinds, key=lambda k: kinds[k][1])
    launches, ms = kinds[dom]
    achieved = flops[dom] / (ms * 1e-3)
    arith = "bf16" if (args.precision == "mixed" and dom == "features") else "fp32"
    roof = {"bound": "mfma", "kernel": dom, "achieved": achieved / 1e12, "peak": PEAK[arith] / 1e12,
            "unit": "TFLOP/s", "frac": achieved / PEAK[arith], "traffic": None,
            "launches": launches, "avg_launch_ms": ms / max(1, launches),
            "flops_per_launch": flops[dom] / max(1, launches),
            "kernel_ms": {k: v[1] for k, v in kinds.items()}}

    samples_per_s = args.steps * B / dt
    out = {
        "metric": METRIC,
        "value": samples_per_s,
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32" if args.precision == "fp32" else "bf16+fp32",
        "data": "synthetic (seeded configurations and random-init weights of the workload's architecture)",
        "config": {"workload": f"{wk.name}: {wk.notes}; {wk.n_layers}-layer {wk.channels}-channel conv backflow, "
                               f"{wk.n_up + wk.n_dn}x{wk.n_up + wk.n_dn} Pfaffians, 8-image projection, "
                               f"E_loc by full evaluation of every connected configuration",
                   "configurations_per_step": B, "connected_per_configuration": n_conn / (args.steps * B),
                   "form": "plain FP32" if args.precision == "fp32" else "mixed precision",
                   "parallelism": "single GPU"},
        "roofline": roof,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(wk, params, s_host, args.cpu_samples)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
