"""Writes the golden fixtures of this directory from the FP64 oracle (run in the build
container: ``python tests/golden/make_golden.py``).

The reference ships no golden vectors for this path (its third-party Pfaffian, projection and
Hubbard code are not vendored), so these fixtures are the oracle's own outputs on seeded
inputs, pinned by the oracle's closed-form and brute-force tests.  They let the GPU tests check
the HIP path without recomputing the oracle and guard the oracle itself against drift.

Each case stores: the seeded configurations s, the workload's parameters' checksum, the
identity image's occupied modes and features, its pairing matrix F, (sign, log|Pf| or log|det|)
of the identity image, the projected (sign, log|psi_P|), and E_loc (t = 1, t', U of the case).
"""

import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from neuralhubbard_amd.configs import WORKLOADS, make_configurations, make_params, small_workload  # noqa: E402

CASES = {
    "pf_4x4_3p3": dict(kind="pfaffian", LX=4, LY=4, n_up=3, n_dn=3, channels=8, n_layers=2, base_f0=True,
                       tprime=-0.2),
    "bcs_4x4_2p2": dict(kind="bcs", LX=4, LY=4, n_up=2, n_dn=2, channels=8, n_layers=2, tprime=0.0),
    "pf_3x5_2p2": dict(kind="pfaffian", LX=3, LY=5, n_up=2, n_dn=2, channels=8, n_layers=1, tprime=0.3),
    "pf8": "pf8",
}
BATCH = {"pf8": 3}


def workload(name):
    spec = CASES[name]
    return WORKLOADS[spec] if isinstance(spec, str) else small_workload(**spec)


def params_checksum(p) -> str:
    h = hashlib.sha256()
    for a in list(p.conv_w) + list(p.conv_b) + [p.W, np.float32(p.theta)]:
        h.update(np.ascontiguousarray(a, np.float32).tobytes())
    for a in (p.F0, p.phi0):
        if a is not None:
            h.update(np.ascontiguousarray(a, np.float32).tobytes())
    return h.hexdigest()


def make_case(name):
    wk = workload(name)
    p = make_params(wk)
    m = O.Model(wk, p)
    s = make_configurations(wk, BATCH.get(name, 4))
    modes = np.stack([np.flatnonzero(x > 0) for x in s]).astype(np.int32)
    y = np.stack([m.features(x) for x in s])
    F = np.stack([m.image_matrix(x, 0) for x in s])
    img = np.array([m.image_amplitude(x, 0) for x in s])
    proj = np.array([m.psi(x) for x in s])
    el = np.array([O.eloc(m, x, wk.U, 1.0, wk.tprime) for x in s])
    return dict(s=s, modes=modes, features=y, F=F, image_sign=img[:, 0], image_logabs=img[:, 1],
                psi_sign=proj[:, 0], psi_logabs=proj[:, 1], eloc=el,
                params_sha256=np.array(params_checksum(p)))


def main():
    for name in CASES:
        out = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(out, **make_case(name))
        print(out, os.path.getsize(out))


if __name__ == "__main__":
    main()
