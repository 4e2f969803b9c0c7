"""HIP path vs the FP64 oracle, through the C ABI (marked gpu).

Tolerances: log|psi| within 1e-4 absolute on the small lattices and 1e-3 at 16x16 (the
target bound for FP32 against FP64); signs, hop lists and configurations bit-exact;
E_loc within 1e-6 relative + 1e-5 absolute on small lattices.
"""

import math

import numpy as np
import pytest

from oracle import oracle as O
from tests._util import cancellation, log_tol
from neuralhubbard_amd import lattice as L
from neuralhubbard_amd.configs import WORKLOADS, make_params, make_configurations, small_workload

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible MI355X")
    from neuralhubbard_amd import _lib

    _lib.load()


def rand_skew(n, rng, scale=1.0):
    A = rng.standard_normal((n, n)) * scale
    return np.triu(A, 1) - np.triu(A, 1).T


def test_slogpf_matches_oracle():
    from neuralhubbard_amd.backflow import slogpf

    rng = np.random.default_rng(0)
    for n, B in [(2, 8), (8, 16), (56, 32), (64, 8), (224, 8), (256, 4)]:
        print(f"slogpf n={n} B={B}", flush=True)
        A = np.stack([rand_skew(n, rng, 0.3) for _ in range(B)])
        A32 = A.astype(np.float32).astype(np.float64)
        for ws in (False, True):  # the LDS form and the workspace form
            sg, la = slogpf(torch.tensor(A, dtype=torch.float32, device="cuda"), workspace=ws)
            sg, la = sg.cpu().numpy(), la.cpu().numpy()
            for b in range(B):
                s_ref, l_ref = O.slogpf(A32[b])
                assert sg[b] == s_ref, (n, b, ws)
                assert abs(la[b] - l_ref) < 2e-4 * max(1.0, n / 56), (n, b, ws, la[b], l_ref)


def _cond_tol(A64, base):
    """base + 1e-6 x condition number: FP32 elimination loses about eps * cond in log|det|."""
    sv = np.linalg.svd(A64, compute_uv=False)
    return base + 1e-6 * sv[0] / max(sv[-1], 1e-300)


@pytest.mark.parametrize("n", [66, 70, 96, 130, 176, 178, 192, 222, 240, 242, 250, 256])
def test_slogpf_blocked_sizes(n):
    """Blocked Pfaffian of nh_slogpf at sizes around its boundaries (256-thread workgroups up to
    n = 176, 1024 threads above; 8-step panels up to n = 240, 4-step panels above, partial last
    panels): sign against the FP64 oracle, log|Pf| against
    LAPACK's log|det| / 2 within a condition-scaled tolerance."""
    from neuralhubbard_amd.backflow import slogpf

    rng = np.random.default_rng(n)
    B = 12
    A = np.stack([rand_skew(n, rng, 0.3) for _ in range(B)]).astype(np.float32)
    sg, la = slogpf(torch.tensor(A, device="cuda"), workspace=False)
    sg, la = sg.cpu().numpy(), la.cpu().numpy()
    for b in range(B):
        A64 = A[b].astype(np.float64)
        s_ref, l_ref = O.slogpf(A64)
        _, ld = np.linalg.slogdet(A64)
        assert abs(l_ref - ld / 2) < 1e-8
        assert sg[b] == s_ref, (n, b)
        assert abs(la[b] - l_ref) < _cond_tol(A64, 2e-4 * n / 56), (n, b, la[b], l_ref)


@pytest.mark.parametrize("n", [66, 96, 98, 100, 130, 144, 146, 194, 224, 256, 258, 320, 352, 354, 384, 460, 512, 514,
                               896, 1024])
def test_slogpf_workspace_form(n):
    """nh_slogpf_ws (column-major squares in the workspace for 100 <= n <= 352, packed triangles
    at the other sizes, n <= 1024; n = 896 is the 32x32 Pfaffian):
    sign against the FP64 oracle, log|Pf| against LAPACK's log|det| / 2 within the
    condition-scaled tolerance of the LDS form, and a workspace of two matrices (the batch in
    chunks) gives bit-identical results to one of the whole batch."""
    import ctypes

    from neuralhubbard_amd import _lib
    from neuralhubbard_amd.backflow import slogpf

    rng = np.random.default_rng(1000 + n)
    B = 5 if n <= 512 else 3
    A = np.stack([rand_skew(n, rng, 0.3) for _ in range(B)]).astype(np.float32)
    At = torch.tensor(A, device="cuda")
    sg, la = slogpf(At, workspace=True)
    sg, la = sg.cpu().numpy(), la.cpu().numpy()
    print(f"slogpf_ws n={n}", flush=True)
    tol = []
    for b in range(B):
        A64 = A[b].astype(np.float64)
        _, ld = np.linalg.slogdet(A64)
        s_ref, _ = O.slogpf(A64) if n <= 512 else (None, None)
        if s_ref is not None:
            assert sg[b] == s_ref, (n, b)
        tol.append(_cond_tol(A64, 2e-4 * n / 56))
        print(f"  b={b}: |d log| vs LAPACK {abs(la[b] - ld / 2):.2e} (tolerance {tol[-1]:.2e})", flush=True)
        assert abs(la[b] - ld / 2) < tol[-1], (n, b, la[b], ld / 2)
    if n > 512:  # sign (the FP64 oracle is slow here): swapping two rows and columns flips it
        P = np.arange(n)
        P[[0, 1]] = P[[1, 0]]
        sg2, la2 = slogpf(torch.tensor(A[:, P][:, :, P].copy(), device="cuda"), workspace=True)
        assert (sg2.cpu().numpy() == -sg).all()
        assert (np.abs(la2.cpu().numpy() - la) < 2 * np.array(tol)).all()
    lib = _lib.load()
    two = lib.nh_slogpf_workspace_size(2, n)
    assert two < lib.nh_slogpf_workspace_size(B, n)
    ws = torch.empty(two, dtype=torch.uint8, device="cuda")
    sgc = torch.empty(B, dtype=torch.float32, device="cuda")
    lac = torch.empty(B, dtype=torch.float64, device="cuda")
    _lib.check(lib.nh_slogpf_ws(_lib.stream_handle(), B, n, _lib.ptr(At), _lib.ptr(sgc), _lib.ptr(lac),
                                _lib.ptr(ws), ctypes.c_size_t(two)), "nh_slogpf_ws")
    np.testing.assert_array_equal(sgc.cpu().numpy(), sg)
    np.testing.assert_array_equal(lac.cpu().numpy(), la)
    # too small a workspace is refused before any launch
    assert lib.nh_slogpf_ws(_lib.stream_handle(), B, n, _lib.ptr(At), _lib.ptr(sgc), _lib.ptr(lac), _lib.ptr(ws),
                            ctypes.c_size_t(lib.nh_slogpf_workspace_size(1, n) // 2)) == _lib.NH_ERR_WORKSPACE


@pytest.mark.parametrize("n", [65, 100, 112, 127, 128, 129, 160, 192])
def test_slogdet_blocked_sizes(n):
    """Blocked LU at sizes around its 16-column panels and thread-count switch (n = 128)."""
    from neuralhubbard_amd.backflow import slogdet

    rng = np.random.default_rng(n)
    B = 12
    A = rng.standard_normal((B, n, n)).astype(np.float32)
    sg, la = slogdet(torch.tensor(A, device="cuda"))
    sg, la = sg.cpu().numpy(), la.cpu().numpy()
    for b in range(B):
        A64 = A[b].astype(np.float64)
        s_ref, l_ref = np.linalg.slogdet(A64)
        assert sg[b] == s_ref, (n, b)
        assert abs(la[b] - l_ref) < _cond_tol(A64, 1e-4 * n / 28), (n, b, la[b], l_ref)


def test_slogpf_singular_and_signs():
    from neuralhubbard_amd.backflow import slogpf

    A = np.zeros((3, 6, 6), np.float32)
    a = np.array([0.5, -2.0, 3.0])
    for k in range(3):
        A[1, 2 * k, 2 * k + 1], A[1, 2 * k + 1, 2 * k] = a[k], -a[k]
    A[2] = A[1][[1, 0, 2, 3, 4, 5]][:, [1, 0, 2, 3, 4, 5]]
    sg, la = slogpf(torch.tensor(A, device="cuda"))
    sg, la = sg.cpu().numpy(), la.cpu().numpy()
    assert sg[0] == 0 and la[0] == -np.inf
    assert sg[1] == np.sign(a.prod()) and abs(la[1] - np.log(np.abs(a)).sum()) < 1e-6
    assert sg[2] == -sg[1]


def test_slogdet_matches_oracle():
    from neuralhubbard_amd.backflow import slogdet

    rng = np.random.default_rng(1)
    for n, B in [(1, 4), (5, 8), (28, 16), (112, 8)]:
        A = rng.standard_normal((B, n, n)).astype(np.float32)
        sg, la = slogdet(torch.tensor(A, device="cuda"))
        for b in range(B):
            s_ref, l_ref = O.slogdet(A[b].astype(np.float64))
            assert sg[b].item() == s_ref
            assert abs(la[b].item() - l_ref) < 1e-4 * max(1, n / 28)


@pytest.mark.parametrize("LX,LY,d,n", [(8, 8, 8, 40), (16, 16, 16, 40), (5, 7, 4, 40), (16, 16, 32, 224),
                                        (4, 4, 12, 20)])
def test_pairing_occ_matches_oracle(LX, LY, d, n):
    """Tiled kernel for d in {4, 8, 16, 32, 64}; d = 12 takes the one-thread-per-entry kernel."""
    from neuralhubbard_amd.backflow import build_wtilde, pairing_occ

    rng = np.random.default_rng(2)
    N = LX * LY
    W = (rng.standard_normal((3, N, d)) / np.sqrt(d)).astype(np.float32)
    Wt = build_wtilde(W, 0.55, LX, LY)
    B = 3
    idx = np.stack([np.sort(rng.choice(2 * N, n, replace=False)) for _ in range(B)]).astype(np.int32)
    Y = rng.standard_normal((B, n, d)).astype(np.float32)
    F = pairing_occ(torch.tensor(Y, device="cuda"), torch.tensor(idx, device="cuda"),
                    torch.tensor(Wt, device="cuda"), LX, LY).cpu().numpy()
    assert np.array_equal(F, -np.transpose(F, (0, 2, 1)))
    Wt_ref = O.build_wtilde(W, 0.55, LX, LY)
    for b in range(B):
        Yfull = np.zeros((2 * N, d))
        Yfull[idx[b]] = Y[b]
        ref = O.pairing_occ(Yfull, idx[b], Wt_ref, LX, LY)
        np.testing.assert_allclose(F[b], ref, rtol=1e-4, atol=1e-6 * np.abs(ref).max())


def test_pairing_ud_matches_oracle():
    from neuralhubbard_amd.backflow import build_wt_ud, pairing_ud

    rng = np.random.default_rng(3)
    LX = LY = 8
    N, d, nu, nd = 64, 8, 10, 12
    W = rng.standard_normal((N, d)).astype(np.float32)
    Wt = build_wt_ud(W, 0.55, LX, LY)
    iu = np.sort(rng.choice(N, nu, replace=False)).astype(np.int32)[None]
    idn = np.sort(rng.choice(N, nd, replace=False)).astype(np.int32)[None]
    Yu = rng.standard_normal((1, nu, d)).astype(np.float32)
    Yd = rng.standard_normal((1, nd, d)).astype(np.float32)
    F = pairing_ud(*(torch.tensor(x, device="cuda") for x in (Yu, Yd, iu, idn, Wt)), LX, LY).cpu().numpy()
    ref = O.pairing_ud(Yu[0], Yd[0], iu[0], idn[0], O.build_wt_ud(W, 0.55, LX, LY), LX, LY)
    np.testing.assert_allclose(F[0], ref, rtol=1e-4, atol=1e-6 * np.abs(ref).max())


def _model(wk, precision="fp32"):
    from neuralhubbard_amd.backflow import build_backflow

    p = make_params(wk)
    return build_backflow(wk, p, precision), O.Model(wk, p)


@pytest.mark.parametrize("name", ["small", "pf8"])
def test_features_match_oracle(name):
    wk = small_workload(channels=8, n_layers=2) if name == "small" else WORKLOADS["pf8"]
    dev, ref = _model(wk)
    s = make_configurations(wk, 4)
    y = dev.features(torch.tensor(s, device="cuda")).cpu().numpy()
    for b in range(4):
        np.testing.assert_allclose(y[b], ref.features(s[b]), rtol=1e-4, atol=1e-4)


@pytest.mark.parametrize("kind,symm", [("pfaffian", True), ("pfaffian", False), ("bcs", True), ("bcs", False)])
def test_log_psi_small_matches_oracle(kind, symm):
    from neuralhubbard_amd.backflow import Variational

    wk = small_workload(kind=kind, LX=4, LY=4, n_up=3, n_dn=3 if kind == "bcs" else 1, channels=8,
                        n_layers=2, base_f0=(kind == "pfaffian"), tprime=-0.3)
    dev, ref = _model(wk)
    s = make_configurations(wk, 16)
    psi = Variational(dev, symm=symm)(torch.tensor(s, device="cuda"))
    sg, la = psi.numpy()
    for b in range(16):
        rs, rl = ref.psi(s[b], project=symm)
        c = cancellation(ref, s[b]) if symm else 0.0
        if c > 12:  # exactly cancelling symmetric configuration: psi_P = 0 in FP64
            continue
        assert sg[b] == rs, b
        assert abs(la[b] - rl) < log_tol(1e-4, c), (b, la[b], rl, c)


def test_log_psi_pf8_matches_oracle():
    from neuralhubbard_amd.backflow import Variational

    wk = WORKLOADS["pf8"]
    dev, ref = _model(wk)
    s = make_configurations(wk, 64)
    sg, la = Variational(dev)(torch.tensor(s, device="cuda")).numpy()
    for b in range(0, 64, 8):
        rs, rl = ref.psi(s[b])
        c = cancellation(ref, s[b])
        assert sg[b] == rs and abs(la[b] - rl) < log_tol(1e-4, c), (b, la[b], rl, c)


def test_log_psi_pf16_matches_oracle():
    """16x16 Pfaffian backflow (224 x 224 Pfaffians): |dlog psi| < 1e-3 vs FP64, same sign."""
    from neuralhubbard_amd.backflow import Variational

    wk = WORKLOADS["pf16"]
    dev, ref = _model(wk)
    s = make_configurations(wk, 32)
    sg, la = Variational(dev)(torch.tensor(s, device="cuda")).numpy()
    for b in range(0, 32, 11):
        rs, rl = ref.psi(s[b])
        c = cancellation(ref, s[b])
        print(f"pf16 b={b} gpu={la[b]:.6f} fp64={rl:.6f} cancellation={c:.2f}", flush=True)
        assert sg[b] == rs and abs(la[b] - rl) < log_tol(1e-3, c), (b, la[b], rl)


@pytest.mark.parametrize("LX,LY,n_up,n_dn,channels", [
    (12, 13, 72, 74, 16),  # column form, n = 146: the last 4-column group reaches the padding rows
    (12, 13, 76, 76, 8),   # column form at d = 4 features
    (14, 14, 96, 98, 32),  # column form, n = 194, d = 16, seven 16-step panels
    (10, 10, 48, 52, 16),  # column form at its smallest n = 100 (four panels, the last of 2 steps)
    (12, 12, 60, 62, 32),  # column form, n = 122: the last 4-column group reaches the padding rows
    (13, 13, 74, 76, 16),  # n_up % 4 != 0: the packed-triangle kernel at n = 150
    (10, 10, 48, 50, 16),  # n = 98, below the column form's sizes: the packed-triangle kernel
    (10, 10, 37, 37, 16),  # n = 74, packed triangle
])
def test_log_psi_image_forms_match_oracle(LX, LY, n_up, n_dn, channels):
    """Both Pfaffian image kernels above n = 64 (the column-major lazy-pivot form, used for
    100 <= n <= 256 when n_up % 4 == 0, and the packed-triangle form otherwise) against the FP64
    oracle's projected amplitude, with t' != 0 and a base matrix F0, on sizes where their edge
    cases arise."""
    from neuralhubbard_amd.backflow import Variational

    wk = small_workload(kind="pfaffian", LX=LX, LY=LY, n_up=n_up, n_dn=n_dn, channels=channels,
                        n_layers=2, base_f0=True, tprime=-0.2)
    dev, ref = _model(wk)
    s = make_configurations(wk, 6)
    sg, la = Variational(dev)(torch.tensor(s, device="cuda")).numpy()
    for b in range(6):
        rs, rl = ref.psi(s[b])
        c = cancellation(ref, s[b])
        assert sg[b] == rs and abs(la[b] - rl) < log_tol(2e-4, c), (b, la[b], rl, c)


@pytest.mark.parametrize("precision", ["fp32", "mixed"])
def test_log_psi_pf32_matches_oracle(precision):
    """32x32 Pfaffian backflow (896 x 896 Pfaffians, triangles in the global workspace; the
    feature network on VALU in FP32, on split-FP16 MFMA with 32-site rows in the mixed form) on 8
    configurations against the FP64 C restatement: same sign and |d log psi| < 1e-3, the 16x16
    bar held at 32x32 too (measured up to 1.1e-4; FP32 elimination over 448 pivot steps)."""
    from neuralhubbard_amd.backflow import Variational
    from oracle.cpu_port import CpuPath

    wk = WORKLOADS["pf32"]
    p = make_params(wk)
    dev, _ = _model(wk, precision)
    s = make_configurations(wk, 8)
    sg, la = Variational(dev)(torch.tensor(s, device="cuda")).numpy()
    rs, rl = CpuPath(wk, p, bits=64).psi(s)
    for b in range(len(s)):
        print(f"pf32 {precision} b={b} gpu={la[b]:.6f} fp64={rl[b]:.6f} |d|={abs(la[b] - rl[b]):.2e}", flush=True)
    assert np.array_equal(sg, rs)
    assert np.abs(la - rl).max() < 1e-3


@pytest.mark.parametrize("precision", ["fp32", "mixed"])
def test_local_energy_pf32_matches_materialized_hops(precision):
    """E_loc at 32x32 Pfaffian (n = 896) through nh_local_energy (hop items, chunked images,
    triangles in the workspace) equals U D(s) + sum over the oracle's hop list (Jordan-Wigner signs
    included) of h psi(s') / psi(s) with every psi(s') from nh_log_psi on the materialised
    configurations; 16 of those psi(s') against the FP64 C restatement (its full E_loc, 4036
    amplitudes of 8 896 x 896 Pfaffians each, takes minutes on the host)."""
    from neuralhubbard_amd.backflow import Hubbard, Samples, Variational
    from oracle.cpu_port import CpuPath

    wk = WORKLOADS["pf32"]
    dev, _ = _model(wk, precision)
    st = Variational(dev)
    s = make_configurations(wk, 1)
    ts = torch.tensor(s, device="cuda")
    psi = st(ts)
    e = Hubbard(wk.U, 1.0, wk.tprime, chunk=1).Oloc(st, Samples(ts, psi)).cpu().numpy()[0]
    hl = list(O.hops(s[0], wk.LX, wk.LY, 1.0, wk.tprime))
    sp = np.repeat(s[:1], len(hl), axis=0)
    for k, (a, b2, _) in enumerate(hl):
        sp[k, a], sp[k, b2] = -1, 1
    sg2, la2 = st(torch.tensor(sp, device="cuda")).numpy()
    sg0, la0 = float(psi.sign[0]), float(psi.logabs[0])
    want = O.diag_energy(s[0], wk.U) + sum(h * sg2[k] * sg0 * math.exp(la2[k] - la0) for k, (_, _, h) in enumerate(hl))
    print(f"pf32 {precision} eloc gpu={e / wk.N:.8f} materialised={want / wk.N:.8f} per site ({len(hl)} hops)",
          flush=True)
    assert abs(e - want) < 1e-5 * abs(want) + 1e-4
    pick = np.linspace(0, len(hl) - 1, 16).astype(int)
    r_s, r_l = CpuPath(wk, make_params(wk), bits=64).psi(sp[pick])
    assert np.array_equal(sg2[pick], r_s)
    assert np.abs(la2[pick] - r_l).max() < 1e-2


def test_log_psi_bcs16_matches_oracle():
    from neuralhubbard_amd.backflow import Variational

    wk = WORKLOADS["bcs16"]
    dev, ref = _model(wk)
    s = make_configurations(wk, 16)
    sg, la = Variational(dev)(torch.tensor(s, device="cuda")).numpy()
    for b in range(0, 16, 7):
        rs, rl = ref.psi(s[b])
        c = cancellation(ref, s[b])
        assert sg[b] == rs and abs(la[b] - rl) < log_tol(1e-3, c), (b, la[b], rl, c)


@pytest.mark.parametrize("n", [193, 208, 230, 256])
def test_log_psi_bcs_global_matrix_sizes(n):
    """BCS amplitudes whose n x n determinants leave LDS (n > 192: matrix in the workspace, 9-bit
    pivot keys, 512-thread LU) at sizes around the 16-column panels, 16x16 lattice, against the
    FP64 oracle: same sign, |d log psi| within the condition-aware bound."""
    from neuralhubbard_amd.backflow import Variational

    wk = small_workload(kind="bcs", LX=16, LY=16, n_up=n, n_dn=n, tprime=0.0, channels=8, n_layers=2, seed=n)
    dev, ref = _model(wk)
    s = make_configurations(wk, 3)
    sg, la = Variational(dev)(torch.tensor(s, device="cuda")).numpy()
    for b in range(3):
        rs, rl = ref.psi(s[b])
        c = cancellation(ref, s[b])
        print(f"bcs n={n} b={b} gpu={la[b]:.6f} fp64={rl:.6f} |d|={abs(la[b] - rl):.2e} cancellation={c:.2f}",
              flush=True)
        assert sg[b] == rs and abs(la[b] - rl) < log_tol(1e-3, c), (n, b, la[b], rl, c)


def test_log_psi_bcs32_matches_oracle():
    """32x32 BCS backflow (448 x 448 determinants in the workspace, VALU feature network): same
    sign as FP64 and |d log psi| < 1e-2 (the measured error is printed)."""
    from neuralhubbard_amd.backflow import Variational

    wk = WORKLOADS["bcs32"]
    dev, ref = _model(wk)
    s = make_configurations(wk, 2)
    sg, la = Variational(dev)(torch.tensor(s, device="cuda")).numpy()
    for b in range(2):
        rs, rl = ref.psi(s[b])
        print(f"bcs32 b={b} gpu={la[b]:.6f} fp64={rl:.6f} |d|={abs(la[b] - rl):.2e}", flush=True)
        assert sg[b] == rs and abs(la[b] - rl) < 1e-2, (b, la[b], rl)


def test_local_energy_bcs32_matches_materialized_hops():
    """E_loc at 32x32 BCS through nh_local_energy (hop items, chunked images, matrices in the
    workspace) equals U D(s) + sum over the oracle's hop list (Jordan-Wigner signs included) of
    h psi(s') / psi(s) with every psi(s') from nh_log_psi on the materialised configurations."""
    from neuralhubbard_amd.backflow import Hubbard, Samples, Variational

    wk = WORKLOADS["bcs32"]
    dev, _ = _model(wk)
    st = Variational(dev)
    s = make_configurations(wk, 1)
    ts = torch.tensor(s, device="cuda")
    psi = st(ts)
    e = Hubbard(wk.U, 1.0, wk.tprime, chunk=1).Oloc(st, Samples(ts, psi)).cpu().numpy()[0]
    hl = list(O.hops(s[0], wk.LX, wk.LY, 1.0, wk.tprime))
    sp = np.repeat(s[:1], len(hl), axis=0)
    for k, (a, b2, _) in enumerate(hl):
        sp[k, a], sp[k, b2] = -1, 1
    sg2, la2 = st(torch.tensor(sp, device="cuda")).numpy()
    sg0, la0 = float(psi.sign[0]), float(psi.logabs[0])
    want = O.diag_energy(s[0], wk.U) + sum(h * sg2[k] * sg0 * math.exp(la2[k] - la0) for k, (_, _, h) in enumerate(hl))
    print(f"bcs32 eloc gpu={e:.8f} materialised={want:.8f} ({len(hl)} hops)", flush=True)
    assert abs(e - want) < 1e-5 * abs(want) + 1e-4


@pytest.mark.parametrize("LX,LY,tp", [(4, 4, -0.3), (8, 8, 0.0), (16, 16, -0.2), (3, 5, 0.4)])
def test_hops_bit_exact(LX, LY, tp):
    from neuralhubbard_amd.backflow import hops

    geom = L.Geometry(LX, LY, (LX * LY) // 3, (LX * LY) // 4)
    s = L.random_configurations(geom, 8, 5)
    count, fr, to, co = (x.cpu().numpy() for x in hops(torch.tensor(s, device="cuda"), LX, LY, 1.0, tp,
                                                        geom.n_occ))
    for b in range(8):
        ref = O.hops(s[b], LX, LY, 1.0, tp)
        assert count[b] == len(ref)
        assert [(int(f), int(t)) for f, t in zip(fr[b, : count[b]], to[b, : count[b]])] == [(a, c) for a, c, _ in ref]
        np.testing.assert_array_equal(co[b, : count[b]], np.array([h for _, _, h in ref], np.float32))


@pytest.mark.parametrize("kind", ["pfaffian", "bcs"])
def test_local_energy_small_matches_oracle(kind):
    from neuralhubbard_amd.backflow import Hubbard, Samples, Variational

    wk = small_workload(kind=kind, LX=4, LY=4, n_up=3, n_dn=3 if kind == "bcs" else 1, channels=8,
                        n_layers=2, base_f0=(kind == "pfaffian"), tprime=-0.3)
    dev, ref = _model(wk)
    s = make_configurations(wk, 8)
    st = Variational(dev)
    ts = torch.tensor(s, device="cuda")
    psi = st(ts)
    e = Hubbard(8.0, 1.0, -0.3, chunk=3).Oloc(st, Samples(ts, psi)).cpu().numpy()
    for b in range(8):
        c = cancellation(ref, s[b])
        if c > 4:  # strongly cancelling psi_P(s): ratios psi(s')/psi(s) are ill-conditioned
            continue
        er = O.eloc(ref, s[b], 8.0, 1.0, -0.3)
        assert abs(e[b] - er) < log_tol(1e-6, c) * abs(er) + 1e-5, (b, e[b], er)


def test_local_energy_of_malformed_configuration_is_nan():
    """A configuration with more electrons than the model (its hop count exceeds the item slots of
    nh_local_energy) gives NaN, as its amplitude does, and leaves the other samples' E_loc intact."""
    from neuralhubbard_amd.backflow import Hubbard, LogArray, Samples, Variational

    wk = small_workload(LX=4, LY=4, n_up=3, n_dn=3, channels=8, n_layers=2, tprime=-0.3)
    dev, ref = _model(wk)
    s = make_configurations(wk, 3)
    bad = s[1].copy()
    bad[np.flatnonzero(bad[:16] < 0)[:4]] = 1  # 7 + 3 electrons
    s[1] = bad
    st = Variational(dev)
    ts = torch.tensor(s, device="cuda")
    psi = st(ts)
    sg, la = psi.numpy()
    assert np.isnan(la[1]) and sg[1] == 0
    ok = LogArray(torch.where(torch.isnan(psi.logabs), 0, psi.sign), torch.nan_to_num(psi.logabs, nan=0.0))
    e = Hubbard(8.0, 1.0, -0.3, chunk=3).Oloc(st, Samples(ts, ok)).cpu().numpy()
    assert np.isnan(e[1])
    for b in (0, 2):
        er = O.eloc(ref, s[b], 8.0, 1.0, -0.3)
        assert abs(e[b] - er) < log_tol(1e-6, cancellation(ref, s[b])) * abs(er) + 1e-5, (b, e[b], er)


def test_local_energy_pf8_matches_oracle():
    from neuralhubbard_amd.backflow import Hubbard, Samples, Variational

    wk = WORKLOADS["pf8"]
    dev, ref = _model(wk)
    s = make_configurations(wk, 4)
    st = Variational(dev)
    ts = torch.tensor(s, device="cuda")
    e = Hubbard(wk.U, 1.0, wk.tprime).Oloc(st, Samples(ts, st(ts))).cpu().numpy()
    for b in range(2):
        assert cancellation(ref, s[b]) < 2
        er = O.eloc(ref, s[b], wk.U, 1.0, wk.tprime)
        print(f"pf8 eloc b={b} gpu={e[b]:.8f} fp64={er:.8f}", flush=True)
        assert abs(e[b] - er) < 1e-5 * abs(er) + 1e-4


def test_features_mfma_fp32_pf16():
    """16x16, 32 channels: the FP32 MFMA feature network against the FP64 oracle."""
    wk = WORKLOADS["pf16"]
    dev, ref = _model(wk)
    assert dev.conv_w_mfma is not None
    s = make_configurations(wk, 3)
    y = dev.features(torch.tensor(s, device="cuda")).cpu().numpy()
    for b in range(3):
        r = ref.features(s[b])
        np.testing.assert_allclose(y[b], r, rtol=1e-4, atol=1e-4 * np.abs(r).max())


@pytest.mark.parametrize("LX", [4, 10, 32])
def test_features_mfma_f16_32_site_rows(LX):
    """Split-FP16 MFMA features on lattices with 32-site rows (two 16-site segments per row, the
    y-taps across the segment edge from the other segment; 1024 threads at 32x32) against the
    FP64 oracle: within 1e-6 of their largest magnitude, as on 16-site rows."""
    wk = small_workload(LX=LX, LY=32, n_up=LX * 14 // 2, n_dn=LX * 14 // 2, channels=32, n_layers=6)
    dev, ref = _model(wk, "mixed")
    s = make_configurations(wk, 2)
    y = dev.features(torch.tensor(s, device="cuda")).cpu().numpy()
    for b in range(2):
        r = ref.features(s[b])
        err = np.abs(y[b] - r).max() / np.abs(r).max()
        print(f"split-fp16 features {LX}x32: max |dy| / max |y| = {err:.2e}", flush=True)
        assert err < 1e-6


def test_features_mfma_f16_pf16():
    """Split-FP16 operands / FP32 accumulation (weights pre-scaled by 2^6): features within 1e-6
    of their largest magnitude (measured ~2e-7; split BF16 gave ~6e-6)."""
    wk = WORKLOADS["pf16"]
    dev, ref = _model(wk, "mixed")
    s = make_configurations(wk, 3)
    y = dev.features(torch.tensor(s, device="cuda")).cpu().numpy()
    for b in range(3):
        r = ref.features(s[b])
        err = np.abs(y[b] - r).max() / np.abs(r).max()
        print(f"split-fp16 features: max |dy| / max |y| = {err:.2e}", flush=True)
        assert err < 1e-6
