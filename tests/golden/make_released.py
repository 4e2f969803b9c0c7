"""Writes ``<name>.npz`` for the released tpm02_L16_pfaffian, tp0_L16_pfaffian and tp0_L16_bcs
states as fixtures (run in the build container, where the reference checkout is readable:
``python tests/golden/make_released.py [reference_root] [name ...]``).

A fixture holds data only: the parameters read from ``weights/<name>.eqx`` (leaf stream read
with ``np.load(allow_pickle=False)``; for BCS also the linear head and delta_d F^d as the base
table phi0), the first CHAINS released chains of ``weights/<name>_spins.npz`` in their packed
form, and the FP64 oracle's values on a few of them: features of image 0, projected and bare
(sign, log|psi|), and E_loc (t = 1, the state's t', U = 8; the C restatement in FP64, every
connected configuration in full), for the Pfaffian state also with the up and down embedding
vectors exchanged.  The paper's E/N and sigma^2/N are in `neuralhubbard_amd.released.STATES`.
"""

import copy
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from neuralhubbard_amd import released as R  # noqa: E402
from oracle import oracle as O  # noqa: E402
from oracle.cpu_port import CpuPath  # noqa: E402

NAMES = ("tpm02_L16_pfaffian", "tp0_L16_pfaffian", "tp0_L16_bcs", "tpm02_L24_pfaffian", "tp0_L32_bcs")
CHAINS = 2048  # chains kept for the E/N check
N_PSI = 8      # chains with oracle amplitudes
N_FEAT = 2     # chains with oracle features
N_ELOC = 2     # chains with oracle E_loc
# the 24x24 and 32x32 states: fewer chains (their E/N checks are short) and no oracle E_loc (one
# chain's ~2000 connected configurations x 8 images of a 1.4-2.4 GFLOP network take the C
# restatement many minutes)
SIZES = {"tpm02_L24_pfaffian": dict(chains=256, n_psi=2, n_feat=1, n_eloc=0),
         "tp0_L32_bcs": dict(chains=64, n_psi=2, n_feat=1, n_eloc=0)}


def make(NAME, ref_root):
    wdir = os.path.join(ref_root, "weights")
    info = R.STATES[NAME]
    sz = {**dict(chains=CHAINS, n_psi=N_PSI, n_feat=N_FEAT, n_eloc=N_ELOC), **SIZES.get(NAME, {})}
    wk, p = R.load_released(NAME, os.path.join(wdir, f"{NAME}.eqx"))
    S = R.load_spins(os.path.join(wdir, f"{NAME}_spins.npz"), info)[:sz["chains"]]
    m = O.Model(wk, p)
    feats = m.features_batch(S[:sz["n_feat"]])
    cp = CpuPath(wk, p, bits=64, threads=os.cpu_count())
    sg, la = cp.psi(S[:sz["n_psi"]])
    bare = CpuPath(wk, p, bits=64, project=False, threads=os.cpu_count())
    bsg, bla = bare.psi(S[:sz["n_psi"]])
    t0 = time.time()
    n_el = sz["n_eloc"]
    el = cp.local_energy(S[:n_el], sg[:n_el], la[:n_el]) if n_el else np.zeros(0)
    print(f"oracle E_loc of {n_el} chains: {el} ({time.time() - t0:.0f} s)", flush=True)
    extra = {}
    if info.ansatz == "pfaffian" and n_el:
        # the same with the up and down embedding vectors exchanged (a plausible convention
        # mistake), to show that E/N against the paper discriminates it
        q = copy.copy(p)
        q.emb = np.ascontiguousarray(p.emb[:, [0, 2, 1, 3]])
        cq = CpuPath(wk, q, bits=64, threads=os.cpu_count())
        qs, ql = cq.psi(S[:n_el])
        extra["eloc_swapped_embedding"] = cq.local_energy(S[:n_el], qs, ql)
        print(f"... with up/down embeddings exchanged: {extra['eloc_swapped_embedding']}", flush=True)
    elif info.ansatz == "bcs":
        extra["head"], extra["phi0"] = p.head, p.phi0
    out = os.path.join(HERE, f"{NAME}.npz")
    np.savez_compressed(
        out,
        emb=p.emb, conv_w=np.stack(p.conv_w), conv_b=np.stack(p.conv_b), W=p.W, theta=np.float32(p.theta),
        **R.pack_spins(S),
        features=feats.astype(np.float32), psi_sign=sg, psi_logabs=la, bare_sign=bsg, bare_logabs=bla,
        eloc=el, **extra,
    )
    print(out, os.path.getsize(out))


def main(ref_root="/root/reference", *names):
    for name in names or NAMES:
        make(name, ref_root)


if __name__ == "__main__":
    main(*sys.argv[1:])
