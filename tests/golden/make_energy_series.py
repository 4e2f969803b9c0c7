"""Writes ``energy_series.npz``: the reference's per-sweep energy series of the released and
parent states (``data/supplement/tableS4_energy_benchmarks/energy_series/*.npz``: ``E_mean``,
``E_sqr`` per sweep, ``Lx``, ``Ly``) with the values Table S4 prints for them
(``tableS4.csv``, columns ``E_*``/``var_*`` of the rows whose ``data_*`` names the file).
Run in the build container: ``python tests/golden/make_energy_series.py [reference_root]``.
The fixture is data only: the arrays as the reference stores them and the printed strings.
"""

import csv
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main(ref_root="/root/reference"):
    base = os.path.join(ref_root, "data", "supplement", "tableS4_energy_benchmarks")
    out = {}
    names = []
    with open(os.path.join(base, "tableS4.csv"), newline="") as fh:
        for row in csv.DictReader(fh):
            for kind in ("projected", "unprojected"):
                rel = row[f"data_{kind}"]
                if not rel.startswith("energy_series/"):
                    continue
                name = os.path.splitext(os.path.basename(rel))[0]
                with np.load(os.path.join(base, rel), allow_pickle=False) as f:
                    out[f"{name}/E_mean"] = f["E_mean"]
                    out[f"{name}/E_sqr"] = f["E_sqr"]
                    out[f"{name}/N"] = np.int64(f["Lx"] * f["Ly"])
                out[f"{name}/E_printed"] = np.array(row[f"E_{kind}"])
                out[f"{name}/var_printed"] = np.array(row[f"var_{kind}"])
                out[f"{name}/training_average"] = np.array(row["training_average"])
                out[f"{name}/kind"] = np.array(kind)
                names.append(name)
    out["names"] = np.array(names)
    dst = os.path.join(HERE, "energy_series.npz")
    np.savez_compressed(dst, **out)
    print(dst, len(names), os.path.getsize(dst))


if __name__ == "__main__":
    main(*sys.argv[1:])
