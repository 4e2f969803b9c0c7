"""File formats of the energy channel (code/measure_orderparams.py:108-174, code/measure/
assemble.py:296-411, code/states.py:328-352), exercised with host-side stand-ins for the state,
operator and sampler (no GPU): series.npz keys and per-sweep means, the exclusion of a sweep with
non-finite E_loc, spins_last.npz in the released-chain format, and results.json equal to the
energy statistics of the kept sweeps."""

import json
import os
from types import SimpleNamespace

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from neuralhubbard_amd import measure as M  # noqa: E402
from neuralhubbard_amd.configs import small_workload  # noqa: E402
from neuralhubbard_amd.released import unpack_spins  # noqa: E402


class _Sampler:
    def __init__(self, wk, chains, seed=0):
        self.rng = np.random.default_rng(seed)
        self.wk = wk
        self.s = torch.tensor(self._draw(chains))

    def _draw(self, c):
        N = self.wk.N
        out = -np.ones((c, 2 * N), np.int8)
        for i in range(c):
            out[i, self.rng.choice(N, self.wk.n_up, replace=False)] = 1
            out[i, N + self.rng.choice(N, self.wk.n_dn, replace=False)] = 1
        return out

    def sweep(self):
        self.s = torch.tensor(self._draw(self.s.shape[0]))
        return SimpleNamespace(spins=self.s, psi=None)


class _H:
    def __init__(self, bad_sweep):
        self.calls, self.bad = 0, bad_sweep
        self.rng = np.random.default_rng(1)

    def Oloc(self, state, samples):
        e = -12.0 + self.rng.standard_normal(samples.spins.shape[0])
        if self.calls == self.bad:
            e[3] = np.nan
        self.calls += 1
        return torch.tensor(e)


def test_energy_channel_files(tmp_path):
    wk = small_workload("pfaffian", LX=4, LY=4, n_up=3, n_dn=3)
    state = SimpleNamespace(model=SimpleNamespace(workload=wk))
    sampler, H = _Sampler(wk, 16), _H(bad_sweep=5)
    out = str(tmp_path / "run")
    res = M.run_energy(state, H, sampler, 12, out, name="small", tprime=wk.tprime, log=lambda m: None)
    with np.load(os.path.join(out, "series.npz"), allow_pickle=False) as f:
        series = {k: f[k] for k in f.files}
    assert set(series) >= {"state", "tprime", "LX", "LY", "family", "nsamples", "E_mean", "E_sqr",
                           "kept_sweeps", "excluded_sweeps"}
    assert series["excluded_sweeps"].tolist() == [5]
    assert series["kept_sweeps"].tolist() == [n for n in range(12) if n != 5]
    assert series["E_mean"].shape == (11,) and str(series["family"]) == "pi4"
    with np.load(os.path.join(out, "spins_last.npz"), allow_pickle=False) as f:
        last = unpack_spins(f["bits"], f["shape"])
    np.testing.assert_array_equal(last, sampler.s.numpy())
    with open(os.path.join(out, "results.json")) as fh:
        js = json.load(fh)
    ref = M.energy(series["E_mean"], series["E_sqr"], wk.N)
    assert js["energy"]["E_per_site"] == pytest.approx(ref["E_per_site"], rel=1e-15)
    assert js["energy"]["E_err"] == pytest.approx(ref["E_err"], rel=1e-15)
    assert js["energy"]["E_err_blocks"] == {str(b): v for b, v in ref["E_err_blocks"].items()}
    assert js["LX"] == 4 and js["state"] == "small" and js["nsamples"] == 16
    assert res["energy"]["nsweeps"] == 11


def test_results_json_writes_nan_as_null(tmp_path):
    series = {"LX": 4, "LY": 4, "family": "pi8", "E_mean": np.full(3, -12.0), "E_sqr": np.full(3, 150.0)}
    res = M.assemble_energy(series)
    assert np.isnan(res["energy"]["E_err"])  # 3 sweeps leave no block size with 8 blocks
    M.save_results(res, str(tmp_path))
    with open(tmp_path / "results.json") as fh:
        assert json.load(fh)["energy"]["E_err"] is None
