"""The golden fixtures of tests/golden against the oracle and the seeded generators (CPU).

The fixtures were written by tests/golden/make_golden.py from the FP64 oracle; recomputing them
here catches drift of the oracle or of the seeded parameter / configuration generators."""

import importlib.util
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
spec = importlib.util.spec_from_file_location("make_golden", os.path.join(HERE, "golden", "make_golden.py"))
G = importlib.util.module_from_spec(spec)
spec.loader.exec_module(G)


def load(name):
    return dict(np.load(os.path.join(HERE, "golden", f"{name}.npz"), allow_pickle=False))


@pytest.mark.parametrize("name", list(G.CASES))
def test_fixture_inputs_match_generators(name):
    from neuralhubbard_amd.configs import make_configurations, make_params

    wk = G.workload(name)
    f = load(name)
    assert str(f["params_sha256"]) == G.params_checksum(make_params(wk))
    assert np.array_equal(f["s"], make_configurations(wk, len(f["s"])))


@pytest.mark.parametrize("name", ["pf_4x4_3p3", "bcs_4x4_2p2", "pf_3x5_2p2"])
def test_oracle_reproduces_fixture(name):
    f = load(name)
    g = G.make_case(name)
    for k in ("modes", "image_sign", "psi_sign"):
        assert np.array_equal(f[k], g[k]), k
    for k in ("features", "F", "image_logabs", "psi_logabs", "eloc"):
        np.testing.assert_allclose(g[k], f[k], rtol=1e-12, atol=1e-12, err_msg=k)


def test_pf8_fixture_consistency():
    """pf8 (8x8, 28 + 28): Pf(F)^2 = det(F) for the stored pairing matrices, and the stored
    identity-image amplitude is the Pfaffian of the stored F."""
    from oracle import oracle as O

    f = load("pf8")
    for b in range(len(f["s"])):
        F = f["F"][b]
        assert np.abs(F + F.T).max() == 0.0
        sg, la = O.slogpf(F)
        assert abs(la - f["image_logabs"][b]) < 1e-12
        s2, l2 = np.linalg.slogdet(F)
        assert s2 > 0 and abs(2 * la - l2) < 1e-8
