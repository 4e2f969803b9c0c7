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
