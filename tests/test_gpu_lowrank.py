"""Local energy with the hop ratios as low-rank updates (nh_local_energy_lowrank): every one of a
sample's ~200 (8x8), ~1000 (16x16) or ~4000 (32x32) hop ratios enters E_loc with its sign, so a
wrong sign, cone or ordering in any of them shows.  On random-init states E_loc cancels between
terms of large magnitude (sum_h |coef_h psi(s_h) / psi(s)| up to 1e4), so the bars are stated
against that scale: at 8x8 and 16x16 against the FP64 C restatement of the whole path, the
low-rank error within 3x the full-recompute path's own error or 3e-5 of the scale; at 32x32
(where the FP64 restatement takes minutes per sample) against the full-recompute path
(nh_local_energy, pinned against materialised connected configurations), within 2e-4 of the
scale."""

import numpy as np
import pytest

from neuralhubbard_amd.configs import WORKLOADS, make_params, make_configurations

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible MI355X")
    from neuralhubbard_amd import _lib

    _lib.load()


def _abs_scale(st, H, S):
    """sum over hops of |coef psi(s') / psi(s)| per sample, from the materialised hops."""
    from neuralhubbard_amd.backflow import hops

    w = st.model.workload
    n = w.n_up + w.n_dn
    psi = st(S)
    count, fr, to, co = hops(S, w.LX, w.LY, H.t, H.tprime, n)
    out = []
    for b in range(S.shape[0]):
        c = int(count[b])
        sp = S[b].repeat(c, 1).clone()
        idx = torch.arange(c, device="cuda")
        sp[idx, fr[b, :c].long()] = -1
        sp[idx, to[b, :c].long()] = 1
        p = st(sp)
        out.append(float((co[b, :c].double().abs() * torch.exp(p.logabs - psi.logabs[b])).sum()))
    return np.array(out)


def _workload(name):
    if name == "cone12":  # 12x12, 3 convolutions: the cone covers 7x7 sites, a strict subset
        from neuralhubbard_amd.configs import small_workload

        return small_workload(kind="pfaffian", LX=12, LY=12, n_up=40, n_dn=40, channels=8, n_layers=2,
                              tprime=-0.2, base_f0=True)
    return WORKLOADS[name]


@pytest.mark.parametrize("name,precision,nconf", [("cone12", "fp32", 4), ("pf8", "fp32", 6), ("pf16", "fp32", 3),
                                                  ("pf16", "mixed", 2), ("pf32", "fp32", 2), ("pf32", "mixed", 1)])
def test_lowrank_local_energy_matches_full(name, precision, nconf):
    from neuralhubbard_amd.backflow import Hubbard, Samples, Variational, build_backflow

    wk = _workload(name)
    st = Variational(build_backflow(wk, make_params(wk), precision))
    S = torch.tensor(make_configurations(wk, nconf, seed=21), device="cuda")
    psi = st(S)
    full = Hubbard(U=wk.U, t=wk.t, tprime=wk.tprime, chunk=nconf).Oloc(st, Samples(S, psi)).cpu().numpy()
    H = Hubbard(U=wk.U, t=wk.t, tprime=wk.tprime, chunk=nconf, lowrank=True)
    low = H.Oloc(st, Samples(S, psi)).cpu().numpy()
    scale = _abs_scale(st, H, S)
    ref = None
    if name != "pf32":
        from oracle.cpu_port import CpuPath

        cp = CpuPath(wk, make_params(wk), bits=64)
        Sn = S.cpu().numpy()
        rs, rl = cp.psi(Sn)
        ref = cp.local_energy(Sn, rs, rl, t=wk.t)
    for b in range(nconf):
        msg = (f"{name} {precision} b={b}: E_loc full {full[b]:+.8f} low-rank {low[b]:+.8f}"
               f" |d| {abs(full[b] - low[b]):.2e} (sum |terms| {scale[b]:.3e})")
        if ref is not None:
            ef, el = abs(full[b] - ref[b]), abs(low[b] - ref[b])
            print(msg + f"; FP64 {ref[b]:+.8f}: full off by {ef:.2e}, low-rank by {el:.2e}", flush=True)
            assert el <= max(3 * ef, 3e-5 * scale[b]), (b, full[b], low[b], ref[b])
        else:
            print(msg, flush=True)
            assert abs(full[b] - low[b]) < 2e-4 * scale[b] + 1e-9, (b, full[b], low[b])


def test_lowrank_refuses_unsupported_models():
    """BCS states, the residual CNN (receptive field over the whole lattice) and the bare network
    are refused before any launch."""
    import ctypes

    from neuralhubbard_amd import _lib
    from neuralhubbard_amd.backflow import Variational, build_backflow

    lib = _lib.load()
    for name, symm in (("bcs16", True), ("tpm02_L16", True), ("pf16", False)):
        wk = WORKLOADS[name]
        st = Variational(build_backflow(wk, make_params(wk), "fp32"), symm=symm)
        m = st.model.struct(symm)
        assert lib.nh_local_energy_lowrank_workspace_size(ctypes.byref(m), wk.tprime, 4) == 0, name
