"""Writes pf16_random_10k.npz: FP64 (and FP32) amplitudes of 10 000 seeded random configurations
of the 16x16 Pfaffian workload (112 + 112 electrons, 6-layer 32-channel network, 8-image
projection), computed by the C restatement of the path (oracle/nh_cpu.c, libnhcpu64.so in FP64,
libnhcpu32.so in FP32).  Run in the build container: ``python tests/golden/make_pf16_random.py``.

The configurations themselves are not stored: they are make_configurations(pf16, 10000,
seed=SEED), which the GPU test regenerates.  The FP32 values record the precision floor of FP32
arithmetic on these inputs (their error against FP64 is what an exact FP32 evaluation of the
same formula loses), so a test can separate that floor from errors of the GPU kernels.
"""

import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from neuralhubbard_amd.configs import WORKLOADS, make_configurations, make_params  # noqa: E402
from oracle.cpu_port import CpuPath  # noqa: E402

SEED = 2024
COUNT = 10000
OUT = os.path.join(HERE, "pf16_random_10k.npz")


def main():
    w = WORKLOADS["pf16"]
    p = make_params(w)
    S = make_configurations(w, COUNT, seed=SEED)
    s64, l64 = CpuPath(w, p, bits=64).psi(S)
    s32, l32 = CpuPath(w, p, bits=32).psi(S)
    np.savez_compressed(OUT, seed=SEED, count=COUNT, s_sha256=hashlib.sha256(S.tobytes()).hexdigest(),
                        sign64=s64.astype(np.int8), logabs64=l64, sign32=s32.astype(np.int8),
                        logabs32=l32.astype(np.float64))
    d = np.abs(l32 - l64)
    print(f"wrote {OUT}: FP32 vs FP64 median {np.median(d):.2e}, p99 {np.quantile(d, 0.99):.2e}, "
          f"max {d.max():.2e}, > 1e-3 on {(d > 1e-3).sum()}, sign mismatches {(s32 != s64).sum()}")


if __name__ == "__main__":
    main()
