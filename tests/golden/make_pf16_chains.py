"""Equilibrates 8192 Markov chains of the 16x16 Pfaffian workload (t' = -0.2, U = 8, 112 + 112
electrons, 6-layer 32-channel network with the seeded random-init parameters) with the path's own
exact Metropolis sampler, and writes them as pf16_chains_8192.npz: the thermalised starting
ensemble of bench.py and of the delayed-acceptance GPU test, as the reference's measurement script
starts from its released equilibrated chains (code/measure_orderparams.py, code/states.py:335-352).

Runs on the GPU, resumable (the random-init state mixes slowly: <log|psi|> relaxes with a time
constant of ~60 sweeps of 512 exchange proposals):

    python tests/golden/make_pf16_chains.py --sweeps 100 [--resume FILE] --out FILE

Sampler: nh_metropolis_sweep without surrogate (min(1, |psi'/psi|^2)) on the mixed-precision
amplitude (split-FP16 features, FP32 Pfaffians; |d log psi| <= 1e-3 of FP64 on sampled
configurations), seed 2025, Philox step counter continued across resumes.  The file stores the
packed configurations, the step counter and per-sweep means of D/N (double occupancy), log|psi|
and the acceptance, the record from which stationarity is judged.
"""

import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

CHAINS = 8192
SWEEP = 512
SEED = 2025


def main():
    import torch

    from neuralhubbard_amd.backflow import Variational, build_backflow
    from neuralhubbard_amd.configs import WORKLOADS, make_configurations, make_params
    from neuralhubbard_amd.released import pack_spins, unpack_spins
    from neuralhubbard_amd.sampler import MetropolisSampler

    ap = argparse.ArgumentParser()
    ap.add_argument("--sweeps", type=int, default=100)
    ap.add_argument("--resume", default=None)
    ap.add_argument("--out", default=os.path.join(HERE, "pf16_chains_8192.npz"))
    args = ap.parse_args()
    wk = WORKLOADS["pf16"]
    st = Variational(build_backflow(wk, make_params(wk), "mixed"))
    hist = {"D_per_site": [], "mean_logabs": [], "acceptance": []}
    step0 = 0
    if args.resume:
        with np.load(args.resume, allow_pickle=False) as f:
            s0 = unpack_spins(f["bits"], f["shape"])
            step0 = int(f["step"])
            for k in hist:
                hist[k] = list(f[k])
    else:
        s0 = make_configurations(wk, CHAINS, seed=SEED)
    sm = MetropolisSampler(st, s0, SWEEP, seed=SEED)
    sm.step = step0
    N = wk.N
    t0 = time.time()

    def save():
        rec = pack_spins(sm.s.cpu().numpy())
        np.savez_compressed(args.out, bits=rec["bits"], shape=rec["shape"], step=sm.step, seed=SEED, sweep=SWEEP,
                            **{k: np.asarray(v) for k, v in hist.items()})

    for k in range(args.sweeps):
        sm.reset_stats()
        smp = sm.sweep()
        s = smp.spins
        hist["D_per_site"].append(((s[:, :N] > 0) & (s[:, N:] > 0)).sum(1).double().mean().item() / N)
        hist["mean_logabs"].append(smp.psi.logabs.mean().item())
        hist["acceptance"].append(sm.stats()["acceptance"])
        print(f"sweep {len(hist['acceptance'])}: D/N {hist['D_per_site'][-1]:.5f} <log|psi|> "
              f"{hist['mean_logabs'][-1]:.3f} acceptance {hist['acceptance'][-1]:.4f} ({time.time() - t0:.0f} s)",
              flush=True)
        if (k + 1) % 10 == 0 or k + 1 == args.sweeps:
            save()
    torch.cuda.synchronize()


if __name__ == "__main__":
    main()
