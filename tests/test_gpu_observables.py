"""Transverse spin correlations on the GPU (marked gpu).

* small lattice: the local values <s|S+_a S-_j|s'> psi(s')/psi(s) per sample and term from
  `observables.transverse_local` (amplitudes by nh_log_psi) against the FP64 oracle's operator
  algebra (itself pinned to a Jordan-Wigner Fock-space computation in tests/test_oracle.py);
* the released tpm02_L16_pfaffian state: C_s(r) = C_zz(r) + (C_+-(r) + C_-+(r)) / 2 over its
  2048 released chains against the map the reference publishes
  (data/supplement/real_space_corr/figS17_real_space_corr_tpm02_16x16_pfaffian/maps.npz, 128 sweeps
  x 8192 samples), and the stripe order M^2_str against Fig. 4 panel c
  (data/fig4_order_params_tpm02/panel_c_orders.csv): four standard errors of our 2048-sample mean
  plus three of the published one;
* the d-wave terms F and G per sample from `observables.dwave_local` against the FP64 oracle
  (pinned to Jordan-Wigner Fock matrices and the direct <Delta+_0 Delta_r> in tests/test_oracle.py),
  and the pair correlation C_p(r) of both released 16x16 states over their 2048 released chains
  against the published maps (figS17 for tpm02_L16_pfaffian, figS12 for tp0_L16_bcs) and
  Delta^2_SC of Fig. 3 / Fig. 4 panel c.
"""

import os

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible MI355X")


def test_transverse_local_matches_oracle():
    from neuralhubbard_amd.backflow import Variational, build_backflow
    from neuralhubbard_amd.configs import make_configurations, make_params, small_workload
    from neuralhubbard_amd.observables import transverse_local

    wk = small_workload(LX=4, LY=4, n_up=5, n_dn=5, tprime=-0.2, channels=8, n_layers=2, base_f0=True, seed=3)
    p = make_params(wk)
    ref = O.Model(wk, p)
    state = Variational(build_backflow(wk, p))
    S = make_configurations(wk, 6)
    s = torch.tensor(S, device="cuda")
    psi = state(s)
    for a in (0, 5, 11):
        print(f"transverse terms, reference site {a}", flush=True)
        T = transverse_local(state, s, psi, a, 4, 4).cpu().numpy()
        for b in range(len(S)):
            want = O.transverse_local(S[b], a, 4, 4, ref.psi)
            np.testing.assert_allclose(T[b], want, rtol=1e-4, atol=1e-6 * max(1.0, np.abs(want).max()))


def test_released_spin_correlation_matches_published_map():
    from neuralhubbard_amd.backflow import Variational, build_backflow
    from neuralhubbard_amd.observables import m2_str, spin_correlation_samples
    from tests.test_gpu_released import _released

    info, wk, p, S, _ = _released("tpm02_L16_pfaffian")
    L = wk.LX
    with np.load(os.path.join(GOLDEN, "real_space_corr_tpm02_L16.npz"), allow_pickle=False) as f:
        Cs_pub, Cs_err, M2_pub, M2_err = f["Cs"], f["Cs_err"], float(f["M2_str"]), float(f["M2_str_err"])
    state = Variational(build_backflow(wk, p, "mixed"))
    s = torch.tensor(S, device="cuda")
    psi = state(s)
    print(f"C_s over {len(S)} released chains, one reference site per chain", flush=True)
    per = spin_correlation_samples(state, s, psi, L, L).cpu().numpy()
    mean = per.mean(0).reshape(L, L)
    mean = 0.5 * (mean + np.roll(mean[::-1, ::-1], (1, 1), axis=(0, 1)))  # [C(r) + C(-r)] / 2
    err = per.std(0, ddof=1).reshape(L, L) / np.sqrt(len(S))
    for r in ((0, 0), (0, 1), (1, 0), (1, 1), (0, 2), (2, 0), (1, 2), (2, 1)):
        print(f"C_s{r}: {mean[r]:+.5f} +- {err[r]:.5f}, published {Cs_pub[r]:+.5f} +- {Cs_err[r]:.5f}", flush=True)
        assert abs(mean[r] - Cs_pub[r]) < 4 * err[r] + 3 * Cs_err[r]
    # stripe order M^2_str = (1/4N) sum_alpha S(k_alpha), q = pi/4 (Fig. 4 panel c)
    m2 = m2_str(per, L, L, "pi4")
    m2_mean, m2_err = m2.mean(), m2.std(ddof=1) / np.sqrt(len(m2))
    print(f"M2_str {m2_mean:.5f} +- {m2_err:.5f}, published {M2_pub:.6f} +- {M2_err:.6f}", flush=True)
    assert abs(m2_mean - M2_pub) < 4 * m2_err + 3 * M2_err


def test_dwave_local_matches_oracle():
    from neuralhubbard_amd.backflow import Variational, build_backflow
    from neuralhubbard_amd.configs import make_configurations, make_params, small_workload
    from neuralhubbard_amd.observables import dwave_local

    for kind in ("pfaffian", "bcs"):
        wk = small_workload(kind=kind, LX=4, LY=4, n_up=5, n_dn=5, tprime=-0.2, channels=8, n_layers=2,
                            base_f0=True, seed=3)
        p = make_params(wk)
        ref = O.Model(wk, p)
        state = Variational(build_backflow(wk, p))
        S = make_configurations(wk, 6)
        s = torch.tensor(S, device="cuda")
        psi = state(s)
        print(f"d-wave terms, {kind} 4x4, {len(S)} samples", flush=True)
        F, G = dwave_local(state, s, psi, 4, 4)
        F, G = F.cpu().numpy(), G.cpu().numpy()
        # every value is psi(s') / psi(s) times a sign: two FP32 log|psi| within the 1e-3 bound of
        # FP64 leave the ratio within exp(2e-3) - 1
        for b in range(len(S)):
            f, g = O.dwave_local(S[b], 4, 4, ref.psi)
            np.testing.assert_allclose(F[b], f, rtol=2e-3, atol=1e-6 * max(1.0, np.abs(f).max()))
            np.testing.assert_allclose(G[b], g, rtol=2e-3, atol=1e-6 * max(1.0, np.abs(g).max()))
            np.testing.assert_array_equal(np.sign(F[b][np.abs(f) > 1e-6]), np.sign(f[np.abs(f) > 1e-6]))
            np.testing.assert_array_equal(np.sign(G[b][np.abs(g) > 1e-6]), np.sign(g[np.abs(g) > 1e-6]))


@pytest.mark.parametrize("name,fixture", [("tpm02_L16_pfaffian", "real_space_corr_tpm02_L16.npz"),
                                          ("tp0_L16_bcs", "real_space_corr_tp0_L16_bcs.npz")])
def test_released_pair_correlation_matches_published_map(name, fixture):
    """C_p(r) from 16 blocks of 128 released chains (leave-one-block-out errors) against the
    published map at the shortest displacements and Delta^2_SC against the published value: four
    of our standard errors plus three of the published ones."""
    from neuralhubbard_amd.backflow import Variational, build_backflow
    from neuralhubbard_amd.observables import dwave, dwave_measure
    from tests.test_gpu_released import _released

    info, wk, p, S, _ = _released(name)
    L = wk.LX
    with np.load(os.path.join(GOLDEN, fixture), allow_pickle=False) as f:
        Cp_pub, Cp_err, d2_pub, d2_err = f["Cp"], f["Cp_err"], float(f["Delta_SC2"]), float(f["Delta_SC2_err"])
    state = Variational(build_backflow(wk, p, "mixed"))
    s = torch.tensor(S, device="cuda")
    psi = state(s)
    nblk = 16
    blk = len(S) // nblk
    Fs, Gs = [], []
    for k in range(nblk):
        sl = slice(k * blk, (k + 1) * blk)
        F, G = dwave_measure(state, s[sl], type(psi)(psi.sign[sl], psi.logabs[sl]), L, L)
        Fs.append(F.cpu().numpy())
        Gs.append(G.cpu().numpy())
        print(f"{name}: d-wave block {k + 1}/{nblk}", flush=True)
    res = dwave(np.stack(Fs), np.stack(Gs), L, L)
    Cp = res["C_p"].reshape(L, L)
    Cp = 0.5 * (Cp + np.roll(Cp[::-1, ::-1], (1, 1), axis=(0, 1)))  # [C(r) + C(-r)] / 2
    err = res["C_p_err"].reshape(L, L)
    for r in ((0, 0), (0, 1), (1, 0), (1, 1), (0, 2), (2, 0), (2, 2), (0, 4)):
        print(f"C_p{r}: {Cp[r]:+.6f} +- {err[r]:.6f}, published {Cp_pub[r]:+.6f} +- {Cp_err[r]:.6f}", flush=True)
        assert abs(Cp[r] - Cp_pub[r]) < 4 * err[r] + 3 * Cp_err[r], r
    d2, e2 = res["Delta2_SC"], res["Delta2_SC_err"]
    print(f"Delta2_SC {d2:.6f} +- {e2:.6f}, published {d2_pub:.6f} +- {d2_err:.6f}", flush=True)
    assert abs(d2 - d2_pub) < 4 * e2 + 3 * d2_err


def test_run_measure_all_channels(tmp_path):
    """`measure.run_measure` end to end on the GPU (8x8, 28 + 28 electrons, 64 chains, 3 sweeps of
    the Metropolis sampler): the series and results files of all four channels, every value
    finite, <n_i> = 7/8 on average and E/N in the range of a random-weight state."""
    import json

    from neuralhubbard_amd import lattice as Lat
    from neuralhubbard_amd import measure as M
    from neuralhubbard_amd.backflow import Hubbard, Variational, build_backflow
    from neuralhubbard_amd.configs import make_params, small_workload
    from neuralhubbard_amd.sampler import MetropolisSampler

    wk = small_workload(LX=8, LY=8, n_up=28, n_dn=28, tprime=-0.2, channels=8, n_layers=2, seed=5)
    p = make_params(wk)
    state = Variational(build_backflow(wk, p))
    sampler = MetropolisSampler(state, Lat.random_configurations(wk.geom, 64, 3), sweep_steps=64, seed=2,
                                thermal_steps=128)
    out = str(tmp_path / "run")
    print("run_measure: 8x8, 64 chains, 3 sweeps, all channels", flush=True)
    res = M.run_measure(state, Hubbard(U=wk.U, tprime=wk.tprime), sampler, 3, out, name="small8",
                        tprime=wk.tprime, log=lambda m: print(m, flush=True))
    with np.load(os.path.join(out, "series.npz"), allow_pickle=False) as f:
        series = {k: f[k] for k in f.files}
    for k in ("E_mean", "E_sqr", "Czz", "Cnn", "Tpm", "Tmp", "F", "G", "sz_site", "n_site"):
        assert series[k].shape[0] == 3 and np.isfinite(series[k]).all(), k
    assert abs(series["n_site"].mean() - 0.875) < 1e-12
    assert -3.0 < res["energy"]["E_per_site"] < 3.0
    with open(os.path.join(out, "results.json")) as fh:
        js = json.load(fh)
    assert set(js) >= {"energy", "spin", "charge", "dwave", "local"}
    with np.load(os.path.join(out, "results.npz"), allow_pickle=False) as f:
        assert f["C_p"].shape == (64,) and np.isfinite(f["C_p"]).all()


def test_measure_command_line_on_released_state(tmp_path):
    """`python -m neuralhubbard_amd.measure` on the released tpm02_L16_pfaffian state (fixture
    parameters and chains; 16 chains, 2 sweeps of N / 2 proposals, every channel): the files of
    code/measure_orderparams.py and an E/N near the paper's -0.74533."""
    import json

    from neuralhubbard_amd import measure as M

    path = os.path.join(GOLDEN, "tpm02_L16_pfaffian.npz")
    out = str(tmp_path / "run")
    print("measure command line: tpm02_L16_pfaffian, 16 chains, 2 sweeps", flush=True)
    res = M.main(["--state", "tpm02_L16_pfaffian", "--weights", path, "--out", out, "--nsamples", "16",
                  "--nsweeps", "2", "--thermal", "0.5", "--sweep-steps", "0.5"])
    for f in ("series.npz", "spins_last.npz", "results.json", "results.npz"):
        assert os.path.exists(os.path.join(out, f)), f
    with open(os.path.join(out, "results.json")) as fh:
        js = json.load(fh)
    assert set(js) >= {"energy", "spin", "charge", "dwave", "local"}
    assert -0.80 < res["energy"]["E_per_site"] < -0.70
