"""Writes ``real_space_corr_tpm02_L16.npz``: the spin and charge correlation maps that the
reference publishes for the released tpm02_L16_pfaffian state
(``data/supplement/real_space_corr/figS17_real_space_corr_tpm02_16x16_pfaffian/maps.npz``: ``Cs``,
``Cs_err``, ``Cc``, ``Cc_err``, indexed by r = (rx, ry) and symmetrized as [C(r) + C(-r)] / 2,
128 sweeps of 8192 samples) and its stripe order M^2_str of Fig. 4 panel c
(``data/fig4_order_params_tpm02/panel_c_orders.csv``).  Run in the build container:
``python tests/golden/make_real_space.py [reference_root]``.  Data only, copied as stored."""

import csv
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main(ref_root="/root/reference"):
    src = os.path.join(ref_root, "data", "supplement", "real_space_corr",
                       "figS17_real_space_corr_tpm02_16x16_pfaffian", "maps.npz")
    with np.load(src, allow_pickle=False) as f:
        out = {k: f[k] for k in ("Cs", "Cs_err", "Cc", "Cc_err", "state", "n_sweeps_cs", "samples_per_sweep")}
    # M^2_str of the same state, Fig. 4 panel c (data/fig4_order_params_tpm02/panel_c_orders.csv)
    with open(os.path.join(ref_root, "data", "fig4_order_params_tpm02", "panel_c_orders.csv"), newline="") as fh:
        for row in csv.DictReader(fh):
            if row["state"] == "tpm02_L16_pfaffian":
                out["M2_str"] = np.float64(row["M2_str"])
                out["M2_str_err"] = np.float64(row["M2_str_err"])
    dst = os.path.join(HERE, "real_space_corr_tpm02_L16.npz")
    np.savez_compressed(dst, **out)
    print(dst, os.path.getsize(dst))


if __name__ == "__main__":
    main(*sys.argv[1:])
