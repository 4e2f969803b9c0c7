"""Writes ``real_space_corr_tpm02_L16.npz`` and ``real_space_corr_tp0_L16_bcs.npz``: the spin, charge
and pair correlation maps that the reference publishes for the released tpm02_L16_pfaffian and
tp0_L16_bcs states (``data/supplement/real_space_corr/figS17_real_space_corr_tpm02_16x16_pfaffian/
maps.npz`` and ``figS12_real_space_corr_tp0_16x16_bcs/maps.npz``: ``Cs``, ``Cc``, ``Cp`` and their
``*_err``, indexed by r = (rx, ry) and symmetrized as [C(r) + C(-r)] / 2, 128 sweeps of 8192
samples; ``Delta_SC2_plateau``, the mean of ``Cp`` over |r| >= ``r_max``, over ``Delta_SC2_nvec``
displacements; for tpm02 also the per-sweep ``Cs`` and ``Cc`` of ``series.npz`` as ``Cs_series``,
``Cc_series``) and their stripe order M^2_str and Delta^2_SC of Fig. 4 / Fig. 3 panel c
(``data/fig4_order_params_tpm02/panel_c_orders.csv``, ``data/fig3_order_params_tp0/
panel_c_orders.csv``).  Run in the build container:
``python tests/golden/make_real_space.py [reference_root]``.  Data only, copied as stored."""

import csv
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

SOURCES = (
    # (output, figure folder, order-parameter csv, state label in the csv)
    ("real_space_corr_tpm02_L16.npz", "figS17_real_space_corr_tpm02_16x16_pfaffian",
     os.path.join("fig4_order_params_tpm02", "panel_c_orders.csv"), "tpm02_L16_pfaffian"),
    ("real_space_corr_tp0_L16_bcs.npz", "figS12_real_space_corr_tp0_16x16_bcs",
     os.path.join("fig3_order_params_tp0", "panel_c_orders.csv"), "tp0_L16_bcs_step3999"),
)
MAP_KEYS = ("Cs", "Cs_err", "Cc", "Cc_err", "Cp", "Cp_err", "Delta_SC2_plateau", "Delta_SC2_nvec", "r_max", "state",
            "n_sweeps_cs", "samples_per_sweep")


def main(ref_root="/root/reference"):
    for out_name, fig, orders, label in SOURCES:
        src = os.path.join(ref_root, "data", "supplement", "real_space_corr", fig, "maps.npz")
        with np.load(src, allow_pickle=False) as f:
            out = {k: f[k] for k in MAP_KEYS}
        if out_name.startswith("real_space_corr_tpm02"):  # the per-sweep series behind the maps
            with np.load(os.path.join(os.path.dirname(src), "series.npz"), allow_pickle=False) as f:
                out["Cs_series"], out["Cc_series"] = f["Cs"], f["Cc"]
        with open(os.path.join(ref_root, "data", orders), newline="") as fh:
            for row in csv.DictReader(fh):
                if row["state"] == label:
                    for k in ("M2_str", "M2_str_err", "Delta_SC2", "Delta_SC2_err"):
                        out[k] = np.float64(row[k])
        dst = os.path.join(HERE, out_name)
        np.savez_compressed(dst, **out)
        print(dst, os.path.getsize(dst))


if __name__ == "__main__":
    main(*sys.argv[1:])
