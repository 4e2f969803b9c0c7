"""Spin and charge correlations on the CPU: the diagonal parts against the reference's published
maps for the released tpm02_L16_pfaffian chains (data/supplement/real_space_corr/
figS17_real_space_corr_tpm02_16x16_pfaffian/maps.npz, copied to tests/golden/
real_space_corr_tpm02_L16.npz); the transverse part needs amplitudes and is in
tests/test_gpu_observables.py."""

import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from neuralhubbard_amd import observables as obs  # noqa: E402
from neuralhubbard_amd.released import STATES, unpack_spins  # noqa: E402
from oracle import oracle as O  # noqa: E402

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _symm(c, L):
    m = c.reshape(L, L)
    return 0.5 * (m + np.roll(m[::-1, ::-1], (1, 1), axis=(0, 1)))  # [C(r) + C(-r)] / 2


def test_charge_correlation_of_released_chains_matches_published_map():
    """C_c from 2048 released chains against the published map (96 sweeps x 8192 samples), four
    standard errors of the 2048-sample mean at the shortest distances."""
    with np.load(os.path.join(GOLDEN, "tpm02_L16_pfaffian.npz"), allow_pickle=False) as f:
        S = unpack_spins(f["bits"], f["shape"])
    with np.load(os.path.join(GOLDEN, "real_space_corr_tpm02_L16.npz"), allow_pickle=False) as f:
        Cc_pub, Cc_err = f["Cc"], f["Cc_err"]
    L = STATES["tpm02_L16_pfaffian"].L
    t = torch.tensor(S)
    per = torch.fft.ifft2(torch.fft.fft2(obs.dn_grid(t, L * L, 0.875).reshape(-1, L, L)).abs() ** 2).real
    per = per.reshape(len(S), -1).numpy() / (L * L)
    mean = _symm(per.mean(0), L)
    err = per.std(0, ddof=1).reshape(L, L) / np.sqrt(len(S))
    np.testing.assert_allclose(obs.charge_correlation(t, L, L, 0.875).numpy(), per.mean(0), atol=1e-12)
    for r in ((0, 0), (0, 1), (1, 0), (1, 1), (0, 2), (2, 0)):
        assert abs(mean[r] - Cc_pub[r]) < 4 * err[r] + 3 * Cc_err[r], (r, mean[r], Cc_pub[r], err[r])


def test_sz_autocorr_matches_direct_sums():
    rng = np.random.default_rng(0)
    S = np.where(rng.random((5, 2 * 24)) < 0.4, 1, -1).astype(np.int8)
    got = obs.autocorr(obs.sz_grid(torch.tensor(S), 24), 4, 6).numpy()
    np.testing.assert_allclose(got, O.sz_autocorr(S, 4, 6), atol=1e-12)


def test_m2_str_of_published_map_matches_published_value():
    """The structure factor and stripe momenta (assemble.py:241-276): M^2_str of the published
    C_s map is the published M^2_str of Fig. 4 panel c (0.011789(9)); pi/8 stripes would give
    another value."""
    with np.load(os.path.join(GOLDEN, "real_space_corr_tpm02_L16.npz"), allow_pickle=False) as f:
        Cs, M2 = f["Cs"].reshape(-1), float(f["M2_str"])
    assert abs(obs.m2_str(Cs, 16, 16, "pi4") - M2) < 1e-6
    assert abs(obs.m2_str(Cs, 16, 16, "pi8") - M2) > 1e-3


# ------------------------------------------------------------------------ d-wave pair correlation


@pytest.mark.parametrize("name", ["real_space_corr_tpm02_L16.npz", "real_space_corr_tp0_L16_bcs.npz"])
def test_far_mask_reproduces_published_plateau(name):
    """Delta^2_SC is the mean of C_p over |r| >= r_max = d_max / 2 (``assemble.far_mask``): on the
    published maps the mask gives the published plateau, its number of vectors and r_max."""
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as f:
        Cp, plateau, nvec, r_max = f["Cp"].reshape(-1), float(f["Delta_SC2_plateau"]), int(f["Delta_SC2_nvec"]), \
            float(f["r_max"])
        d2, d2_err = float(f["Delta_SC2"]), float(f["Delta_SC2_err"])
    far, r = obs.far_mask(16, 16)
    assert int(far.sum()) == nvec and abs(r - r_max) < 1e-12
    assert abs(Cp[far].mean() - plateau) < 1e-15
    assert abs(Cp[far].mean() - d2) < 0.1 * d2_err  # the csv rounds the same number


def test_m2_str_of_published_bcs_map_matches_published_value():
    with np.load(os.path.join(GOLDEN, "real_space_corr_tp0_L16_bcs.npz"), allow_pickle=False) as f:
        Cs, M2 = f["Cs"].reshape(-1), float(f["M2_str"])
    assert abs(obs.m2_str(Cs, 16, 16, "pi8") - M2) < 1e-6


def test_dwave_keys_follow_the_measured_layout():
    """Row k of F_keys (G_keys) names the term at flat position k of `dwave_local`'s F (G)."""
    LX, LY = 4, 6
    N = LX * LY
    Fk, Gk = obs.dwave_keys(LX, LY)
    assert Fk.shape == (16 * N, 8) and Gk.shape == (6 * N, 5)
    for k, (bx, by, bax, oL, cx, cy, cax, oR) in enumerate(Fk):
        assert (bx, by) == (0, 0)
        assert k == (2 * bax + oL) * 4 * N + 2 * (2 * (cx * LY + cy) + cax) + oR
    lefts = [0, 1, LY]
    for k, (ix, iy, jx, jy, sp) in enumerate(Gk):
        assert k == (lefts.index(ix * LY + iy) * N + jx * LY + jy) * 2 + sp


class _OracleState:
    """Amplitudes of the FP64 oracle behind the call interface of `backflow.Variational`, for host
    tests of the estimators' combinatorics."""

    def __init__(self, model):
        self.model = model

    def __call__(self, spins):
        from neuralhubbard_amd.backflow import LogArray

        rows = [self.model.psi(np.asarray(s)) for s in spins.cpu().numpy()]
        return LogArray(torch.tensor([r[0] for r in rows], dtype=torch.float32),
                        torch.tensor([r[1] for r in rows], dtype=torch.float64))


def test_dwave_local_matches_oracle_on_host():
    """`observables.dwave_local` (torch, vectorized) against the oracle's term-by-term FP64 loop
    (itself pinned to Jordan-Wigner Fock matrices), with oracle amplitudes on both sides."""
    from neuralhubbard_amd.configs import make_configurations, make_params, small_workload

    wk = small_workload(LX=4, LY=4, n_up=5, n_dn=5, tprime=-0.2, channels=4, n_layers=1, base_f0=True, seed=2)
    m = O.Model(wk, make_params(wk))
    state = _OracleState(m)
    S = make_configurations(wk, 6)
    s = torch.tensor(S)
    psi = state(s)
    F, G = obs.dwave_local(state, s, psi, 4, 4)
    nonzero = 0
    for b in range(len(S)):
        f, g = O.dwave_local(S[b], 4, 4, m.psi)
        np.testing.assert_allclose(F[b].numpy(), f, rtol=1e-9, atol=1e-12)
        np.testing.assert_allclose(G[b].numpy(), g, rtol=1e-9, atol=1e-12)
        nonzero += np.count_nonzero(f)
    assert nonzero > 50  # off-diagonal pair terms are exercised


def test_dwave_assembly_statistics():
    """`dwave` on per-sweep terms: C_p of the mean terms, Delta^2_SC the far mean, its block error
    zero for identical sweeps and the leave-one-out error of C_p that of the sweep values when C_p
    is linear in them (G = 0)."""
    rng = np.random.default_rng(1)
    LX = LY = 4
    N = 16
    F = rng.standard_normal((16, 4, 4 * N))
    G = np.zeros((16, 3, N, 2))
    res = obs.dwave(F, G, LX, LY)
    idx = obs.DwaveIndex(LX, LY)
    per = idx.pair_correlation(F, G)
    np.testing.assert_allclose(res["C_p"], per.mean(0), atol=1e-12)
    np.testing.assert_allclose(res["C_p_err"], per.std(0, ddof=1) / 4, rtol=1e-10)
    far, _ = obs.far_mask(LX, LY)
    assert abs(res["Delta2_SC"] - per.mean(0)[far].mean()) < 1e-12
    same = obs.dwave(np.repeat(F[:1], 16, 0), np.repeat(G[:1], 16, 0), LX, LY)
    assert same["Delta2_SC_err"] < 1e-12


def test_assembly_of_published_series_reproduces_published_maps():
    """`measure.jackknife` and `measure.spin` on the reference's own per-sweep C_s (128 sweeps) and
    C_c (96 sweeps) of tpm02_L16_pfaffian (figS17 series.npz): the symmetrized sweep means are the
    published maps, the 4-sweep block errors the published error maps (data/supplement/
    real_space_corr/README.md: "standard errors over blocks of 4 sweeps"), and M^2_str with its
    jackknife error the published 0.011789(9) of Fig. 4 panel c."""
    from neuralhubbard_amd import measure as M

    with np.load(os.path.join(GOLDEN, "real_space_corr_tpm02_L16.npz"), allow_pickle=False) as f:
        pub = {k: f[k] for k in f.files}
    L = 16
    for key in ("Cs", "Cc"):
        X = pub[f"{key}_series"].astype(np.float64)
        X = 0.5 * (X + np.roll(X[:, ::-1, ::-1], (1, 1), axis=(1, 2)))
        v, _, errs = M.jackknife(lambda x: x, X.reshape(len(X), -1))
        np.testing.assert_allclose(v.reshape(L, L), pub[key], atol=1e-9)
        np.testing.assert_allclose(errs[4].reshape(L, L), pub[f"{key}_err"], rtol=1e-6, atol=1e-12)
    Cs = pub["Cs_series"].astype(np.float64).reshape(128, -1)
    sp = M.spin(Cs, np.zeros_like(Cs), np.zeros_like(Cs), L, L, "pi4")
    assert abs(sp["M2_str"] - float(pub["M2_str"])) < 1e-6
    assert abs(sp["M2_str_err"] - float(pub["M2_str_err"])) < 1e-6
