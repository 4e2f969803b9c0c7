"""Reverse mode of the pairing blocks, the Pfaffian and the determinant on the GPU (marked gpu),
against the FP64 oracle's VJPs (pinned by finite differences in tests/test_oracle.py):
nh_pairing_occ_vjp (dY, dWt per sample), nh_pairing_ud_vjp (dYu, dYd, dWt per sample),
nh_slogpf_vjp ((1/2) F^{-T}) and nh_slogdet_vjp (F^{-T}), and their chains, the gradients of
log|Pf F(Y, Wt)| and of log|det F_ud(Yu, Yd, Wt)| with respect to the features and the table,
against FP64 finite differences.
Tolerances: 1e-5 of the largest magnitude for the pairing VJP (FP32 sums of n terms), 5e-4 for
the FP32 Gauss-Jordan inverse of random antisymmetric matrices (condition numbers up to ~1e3
here; an FP32 emulation of the same elimination reaches 9e-5 at n = 300)."""

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible MI355X")


def _case(LX, LY, n, d, B, seed):
    rng = np.random.default_rng(seed)
    N = LX * LY
    Wt = O.build_wtilde(rng.standard_normal((3, N, d)) * 30.0, 0.4, LX, LY)
    idx = np.stack([np.sort(rng.choice(2 * N, n, replace=False)) for _ in range(B)]).astype(np.int32)
    Y = rng.standard_normal((B, n, d)).astype(np.float32)
    G = rng.standard_normal((B, n, n)).astype(np.float32)
    return Wt, idx, Y, G


@pytest.mark.parametrize("LX,LY,n,d,B", [(4, 3, 10, 4, 3), (8, 8, 56, 16, 4), (16, 16, 224, 32, 2)])
def test_pairing_vjp_matches_oracle(LX, LY, n, d, B):
    from neuralhubbard_amd.backflow import pairing_occ_vjp

    print(f"pairing vjp {LX}x{LY} n={n} d={d} batch={B}", flush=True)
    Wt, idx, Y, G = _case(LX, LY, n, d, B, seed=LX + n)
    dY, dWt = pairing_occ_vjp(torch.tensor(Y, device="cuda"), torch.tensor(idx, device="cuda"),
                              torch.tensor(Wt.astype(np.float32), device="cuda"), torch.tensor(G, device="cuda"), LX, LY)
    dY, dWt = dY.cpu().numpy(), dWt.cpu().numpy()
    for b in range(B):
        rY, rW = O.pairing_occ_vjp(Y[b], idx[b], Wt.astype(np.float32), G[b], LX, LY)
        np.testing.assert_allclose(dY[b], rY, rtol=0, atol=1e-5 * np.abs(rY).max())
        np.testing.assert_allclose(dWt[b], rW, rtol=0, atol=1e-5 * np.abs(rW).max())


@pytest.mark.parametrize("n,B", [(8, 5), (56, 4), (224, 3), (300, 2)])
def test_pfaffian_vjp_matches_oracle(n, B):
    from neuralhubbard_amd.backflow import slogpf_vjp

    print(f"pfaffian vjp n={n} batch={B}", flush=True)
    rng = np.random.default_rng(n)
    A = rng.standard_normal((B, n, n))
    A = (A - np.swapaxes(A, 1, 2)).astype(np.float32)
    cot = rng.standard_normal(B).astype(np.float32)
    G = slogpf_vjp(torch.tensor(A, device="cuda"), torch.tensor(cot, device="cuda")).cpu().numpy()
    for b in range(B):
        ref = cot[b] * O.slogpf_vjp(A[b])
        np.testing.assert_allclose(G[b], ref, rtol=0, atol=5e-4 * np.abs(ref).max())


def test_singular_pfaffian_vjp_is_nan():
    from neuralhubbard_amd.backflow import slogpf_vjp

    A = np.zeros((1, 6, 6), np.float32)
    A[0, 0, 1], A[0, 1, 0] = 1.0, -1.0  # rank 2
    assert np.all(np.isnan(slogpf_vjp(torch.tensor(A, device="cuda")).cpu().numpy()))


def test_log_pfaffian_gradient_chain():
    """GPU F = pairing_occ(Y, idx, Wt), G = slogpf_vjp(F), (dY, dWt) = pairing_occ_vjp(G): the
    directional derivative of log|Pf F| against an FP64 central difference along random (vY, vW)
    with vW antisymmetrised like a table."""
    from neuralhubbard_amd.backflow import pairing_occ, pairing_occ_vjp, slogpf_vjp

    LX, LY, n, d = 8, 8, 56, 16
    N = LX * LY
    rng = np.random.default_rng(9)
    Wt, idx, Y, _ = _case(LX, LY, n, d, 1, seed=21)
    Wt32 = Wt.astype(np.float32)
    tY, ti, tW = (torch.tensor(Y, device="cuda"), torch.tensor(idx, device="cuda"), torch.tensor(Wt32, device="cuda"))
    F = pairing_occ(tY, ti, tW, LX, LY)
    G = slogpf_vjp(F)
    dY, dWt = pairing_occ_vjp(tY, ti, tW, G, LX, LY)
    dY, dWt = dY.cpu().numpy()[0], dWt.cpu().numpy()[0]
    V = O.build_wtilde(rng.standard_normal((3, N, d)), 0.4, LX, LY)  # a direction inside the table space
    vY = rng.standard_normal((n, d))
    h = 1e-5

    def logpf(t):
        Yfull = np.zeros((2 * N, d))
        Yfull[idx[0]] = Y[0].astype(np.float64) + t * vY
        return O.slogpf(O.pairing_occ(Yfull, idx[0], Wt32.astype(np.float64) + t * V, LX, LY))[1]

    fd = (logpf(h) - logpf(-h)) / (2 * h)
    got = float(np.sum(dY * vY) + np.sum(dWt * V))
    print(f"directional derivative: GPU {got:.6f}, FP64 central difference {fd:.6f}", flush=True)
    assert got == pytest.approx(fd, rel=1e-3, abs=1e-3)


def _ud_case(LX, LY, nu, nd, d, B, seed):
    rng = np.random.default_rng(seed)
    N = LX * LY
    Wt = O.build_wt_ud(rng.standard_normal((N, d)) * 30.0, 0.4, LX, LY).astype(np.float32)
    iu = np.stack([np.sort(rng.choice(N, nu, replace=False)) for _ in range(B)]).astype(np.int32)
    idn = np.stack([np.sort(rng.choice(N, nd, replace=False)) for _ in range(B)]).astype(np.int32)
    Yu = rng.standard_normal((B, nu, d)).astype(np.float32)
    Yd = rng.standard_normal((B, nd, d)).astype(np.float32)
    G = rng.standard_normal((B, nu, nd)).astype(np.float32)
    return Wt, iu, idn, Yu, Yd, G


@pytest.mark.parametrize("LX,LY,nu,nd,d,B", [(4, 3, 5, 4, 4, 3), (8, 8, 28, 28, 16, 3), (16, 16, 112, 112, 64, 2),
                                             (32, 32, 448, 448, 64, 1)])
def test_bcs_pairing_vjp_matches_oracle(LX, LY, nu, nd, d, B):
    from neuralhubbard_amd.backflow import pairing_ud_vjp

    print(f"bcs pairing vjp {LX}x{LY} {nu}+{nd} d={d} batch={B}", flush=True)
    Wt, iu, idn, Yu, Yd, G = _ud_case(LX, LY, nu, nd, d, B, seed=LX + nu)
    t = lambda x: torch.tensor(x, device="cuda")  # noqa: E731
    dYu, dYd, dWt = (x.cpu().numpy() for x in pairing_ud_vjp(t(Yu), t(Yd), t(iu), t(idn), t(Wt), t(G), LX, LY))
    for b in range(B):
        rYu, rYd, rW = O.pairing_ud_vjp(Yu[b], Yd[b], iu[b], idn[b], Wt, G[b], LX, LY)
        for got, ref in ((dYu[b], rYu), (dYd[b], rYd), (dWt[b], rW)):
            np.testing.assert_allclose(got, ref, rtol=0, atol=1e-5 * np.abs(ref).max())


@pytest.mark.parametrize("n,B", [(5, 4), (56, 3), (112, 3), (448, 1)])
def test_determinant_vjp_matches_oracle(n, B):
    from neuralhubbard_amd.backflow import slogdet_vjp

    rng = np.random.default_rng(100 + n)
    A = rng.standard_normal((B, n, n)).astype(np.float32)
    cot = rng.standard_normal(B).astype(np.float32)
    G = slogdet_vjp(torch.tensor(A, device="cuda"), torch.tensor(cot, device="cuda")).cpu().numpy()
    for b in range(B):
        ref = cot[b] * O.slogdet_vjp(A[b])
        print(f"det vjp n={n} b={b}: max |d| / max |ref| {np.abs(G[b] - ref).max() / np.abs(ref).max():.2e}", flush=True)
        np.testing.assert_allclose(G[b], ref, rtol=0, atol=5e-4 * np.abs(ref).max())


def test_log_determinant_gradient_chain():
    """GPU F = pairing_ud(Yu, Yd, iu, idn, Wt), G = slogdet_vjp(F), (dYu, dYd, dWt) =
    pairing_ud_vjp(G): the directional derivative of log|det F| along random (vu, vd, vW) against
    an FP64 central difference."""
    from neuralhubbard_amd.backflow import pairing_ud, pairing_ud_vjp, slogdet_vjp

    LX, LY, n, d = 8, 8, 28, 16
    rng = np.random.default_rng(12)
    Wt, iu, idn, Yu, Yd, _ = _ud_case(LX, LY, n, n, d, 1, seed=31)
    t = lambda x: torch.tensor(x, device="cuda")  # noqa: E731
    F = pairing_ud(t(Yu), t(Yd), t(iu), t(idn), t(Wt), LX, LY)
    G = slogdet_vjp(F)
    dYu, dYd, dWt = (x.cpu().numpy()[0] for x in pairing_ud_vjp(t(Yu), t(Yd), t(iu), t(idn), t(Wt), G, LX, LY))
    vu, vd, vW = rng.standard_normal((n, d)), rng.standard_normal((n, d)), rng.standard_normal(Wt.shape)
    h = 1e-5

    def logdet(s):
        F64 = O.pairing_ud(Yu[0].astype(np.float64) + s * vu, Yd[0].astype(np.float64) + s * vd, iu[0], idn[0],
                           Wt.astype(np.float64) + s * vW, LX, LY)
        return O.slogdet(F64)[1]

    fd = (logdet(h) - logdet(-h)) / (2 * h)
    got = float(np.sum(dYu * vu) + np.sum(dYd * vd) + np.sum(dWt * vW))
    print(f"directional derivative of log|det|: GPU {got:.6f}, FP64 central difference {fd:.6f}", flush=True)
    assert got == pytest.approx(fd, rel=1e-3, abs=1e-3)

