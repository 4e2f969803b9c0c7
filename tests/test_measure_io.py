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


class _CheapState:
    """A host amplitude (log|psi| linear in the occupations, sign from their parity pattern) for
    file-format tests: any amplitude exercises the layout."""

    def __init__(self, wk):
        self.model = SimpleNamespace(workload=wk)
        self.w = torch.linspace(-0.3, 0.4, 2 * wk.N, dtype=torch.float64)

    def __call__(self, spins):
        from neuralhubbard_amd.backflow import LogArray

        occ = (spins > 0).double()
        sign = 1.0 - 2.0 * ((occ[:, ::3].sum(1) % 2))
        return LogArray(sign.float(), occ @ self.w)


def test_all_channel_files(tmp_path):
    """series.npz of the four channels carries the reference's keys and per-sweep shapes
    (code/measure_orderparams.py:90-160); results.json / results.npz hold what `assemble` makes
    of it (code/measure/assemble.py:301-399); the spin channel's Tpm / Tmp per reference site are
    the means of `transverse_local` over the samples that site serves."""
    from neuralhubbard_amd import observables as obs
    from neuralhubbard_amd.backflow import LogArray

    wk = small_workload("pfaffian", LX=8, LY=8, n_up=28, n_dn=28, tprime=-0.2)  # q = pi/4 needs L % 8 == 0
    N, L = wk.N, wk.LX
    state = _CheapState(wk)
    sampler, H = _Sampler(wk, 16), _H(bad_sweep=-1)
    out = str(tmp_path / "run")
    logs = []
    res = M.run_measure(state, H, sampler, 9, out, name="small", tprime=wk.tprime, log=logs.append)
    with np.load(os.path.join(out, "series.npz"), allow_pickle=False) as f:
        series = {k: f[k] for k in f.files}
    S = 9
    shapes = {"E_mean": (S,), "E_sqr": (S,), "sz_site": (S, L, L), "n_site": (S, L, L), "Czz": (S, N), "Cnn": (S, N),
              "Tpm": (S, M.NREF, N), "Tmp": (S, M.NREF, N), "F": (S, 16 * N), "G": (S, 6 * N),
              "F_keys": (16 * N, 8), "G_keys": (6 * N, 5), "REF_SITES": (M.NREF, 2)}
    for k, shp in shapes.items():
        assert series[k].shape == shp, (k, series[k].shape)
    assert float(series["nbar"]) == 56 / 64
    # the last sweep's spin and d-wave rows from the estimators on the final chains
    s = sampler.s
    psi = state(s)
    refs = obs.reference_sites(L, L, M.NREF)
    for i, (ax, ay) in enumerate(refs):
        sel = torch.arange(i, s.shape[0], len(refs))
        T = obs.transverse_local(state, s[sel], LogArray(psi.sign[sel], psi.logabs[sel]), ax * L + ay, L, L)
        np.testing.assert_allclose(series["Tpm"][-1, i], T[:, 0].mean(0).numpy(), atol=1e-12)
        np.testing.assert_allclose(series["Tmp"][-1, i], T[:, 1].mean(0).numpy(), atol=1e-12)
    F, G = obs.dwave_measure(state, s, psi, L, L)
    np.testing.assert_allclose(series["F"][-1], F.numpy().reshape(-1), atol=1e-12)
    np.testing.assert_allclose(series["G"][-1], G.numpy().reshape(-1), atol=1e-12)
    np.testing.assert_allclose(series["Czz"][-1], obs.autocorr(obs.sz_grid(s, N), L, L).numpy(), atol=1e-12)
    # results: the assembly of the series
    again = M.assemble(series)
    with open(os.path.join(out, "results.json")) as fh:
        js = json.load(fh)
    with np.load(os.path.join(out, "results.npz"), allow_pickle=False) as f:
        maps = {k: f[k] for k in f.files}
    assert set(maps) >= {"C_s", "C_s_err", "S_k", "k_alpha", "C_c", "C_c_err", "N_q", "q_c", "dr", "C_p", "C_p_err"}
    np.testing.assert_allclose(maps["C_p"], again["dwave"]["C_p"], atol=1e-15)
    np.testing.assert_allclose(maps["C_s"], again["spin"]["C_s"], atol=1e-15)
    assert js["spin"]["M2_str"] == pytest.approx(obs.m2_str(again["spin"]["C_s"], L, L, "pi4"), rel=1e-12)
    assert js["dwave"]["Delta2_SC"] == pytest.approx(again["dwave"]["Delta2_SC"], rel=1e-12)
    assert js["dwave"]["n_far"] == int(obs.far_mask(L, L)[0].sum())
    assert js["charge"]["N2"] == pytest.approx(again["charge"]["N2"], rel=1e-12)
    assert js["local"]["n_min"] <= 56 / 64 <= js["local"]["n_max"]
    assert res["energy"]["nsweeps"] == S
    assert any(line.strip().startswith("d-wave   Delta2_SC") for line in logs[-1].splitlines())


def test_spin_assembly_matches_direct_sums():
    """`measure.spin` on per-sweep maps: C_s = C_zz + (T+- + T-+)/2 averaged over the reference
    sites, M^2_str the stripe-momentum mean of its structure factor, M2_zz that of C_zz alone."""
    from neuralhubbard_amd import observables as obs

    rng = np.random.default_rng(3)
    S, N = 10, 64
    Czz, Tpm, Tmp = rng.standard_normal((S, N)), rng.standard_normal((S, 8, N)), rng.standard_normal((S, 8, N))
    r = M.spin(Czz, Tpm, Tmp, 8, 8, "pi4")
    Cs = Czz + 0.5 * (Tpm.mean(1) + Tmp.mean(1))
    np.testing.assert_allclose(r["C_s"], Cs.mean(0), atol=1e-12)
    assert r["M2_str"] == pytest.approx(obs.m2_str(Cs.mean(0), 8, 8, "pi4"), rel=1e-10)
    assert r["M2_zz"] == pytest.approx(obs.m2_str(Czz.mean(0), 8, 8, "pi4"), rel=1e-10)
    # N(q) at +-(2q, 0), +-(0, 2q): q = pi/4 on 8x8 is momentum index 1 (k = 2 pi n / L), 2q index 2
    c = M.charge(Czz, 8, 8, "pi4")
    assert [tuple(q) for q in c["q_c"]] == [(2, 0), (6, 0), (0, 2), (0, 6)]
    Sk = obs.structure_factor(Czz.mean(0), 8, 8)
    assert c["N2"] == pytest.approx((Sk[2, 0] + Sk[6, 0] + Sk[0, 2] + Sk[0, 6]) / 4 / 64, rel=1e-10)
    assert [tuple(k) for k in r["k_alpha"]] == [(3, 4), (5, 4), (4, 3), (4, 5)]


def test_command_line_loads_released_fixture_and_rejects_bad_channels():
    """`measure.main`'s loader reads a parameter .npz with packed chains (the fixture format) and
    the argument check of code/measure_orderparams.py:44-48 rejects unknown channels."""
    from neuralhubbard_amd.released import STATES

    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "tpm02_L16_pfaffian.npz")
    wk, params, chains = M._load_state_files("tpm02_L16_pfaffian", path, None)
    assert wk.network == "resconv" and len(params.conv_w) == 32 and chains.shape == (2048, 512)
    assert np.all((chains[:, :256] > 0).sum(1) == STATES["tpm02_L16_pfaffian"].n_up)
    with pytest.raises(SystemExit):
        M.main(["--state", "tpm02_L16_pfaffian", "--weights", path, "--out", "x", "--channels", "energy,magic"])
