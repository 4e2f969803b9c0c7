r"""Host side of the Pfaffian/BCS backflow path, mirroring the reference's interfaces.

Reference call surface (ChenAo-Phys/NeuralHubbard) and what replaces it here:

=========================================================  ======================================
reference                                                  here (all device work in libnh.so)
=========================================================  ======================================
``pf_kernel.build_wtilde`` (``pf_kernel.py:107-124``)        `build_wtilde` (host, parameters only)
``bcs_kernel.build_wt_ud`` (``bcs_kernel.py:63-70``)        `build_wt_ud`
``pf_kernel.pairing_occ`` (``pf_kernel.py:335-351``)        `pairing_occ` -> ``nh_pairing_occ``
``bcs_kernel.pairing_ud`` (``bcs_kernel.py:257-271``)       `pairing_ud` -> ``nh_pairing_ud``
``fermix.slogpf`` / ``slogdet`` (``models.py:210, 297``)    `slogpf` / `slogdet` -> ``nh_slogpf``...
``pairing_occ`` backward (``pf_kernel.py:354-365``)          `pairing_occ_vjp` -> ``nh_pairing_occ_vjp``
``fermix.slogpf`` reverse mode                              `slogpf_vjp` -> ``nh_slogpf_vjp``
``PfaffianBackflow`` / ``BCSBackflow`` (``models.py``)      `PfaffianBackflow` / `BCSBackflow`
``quantax.state.Variational`` (``states.py:318-320``)       `Variational` -> ``nh_log_psi``
``quantax.operator.Hubbard.Oloc`` (``quick_start.py:33``)   `Hubbard.Oloc` -> ``nh_local_energy``
``quantax.utils.LogArray`` (``models.py:211``)              `LogArray`
=========================================================  ======================================

Tensors are torch tensors on the current CUDA (HIP) device; kernels run on the current
stream.  Errors raise `NhError`; non-finite amplitudes propagate as sign 0 / logabs -inf (a
singular matrix) or NaN (a configuration with the wrong particle number), as the reference
lets them propagate to its callers (``code/measure_orderparams.py:153-156``).
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np

from . import _lib
from .configs import Params, Workload
from .lattice import displacement_tables


def _torch():
    import torch

    return torch


# ------------------------------------------------------------------------ value type


@dataclass
class LogArray:
    """sign * exp(logabs), the quantax ``LogArray`` (``README.md:90-91``)."""

    sign: object
    logabs: object

    def __mul__(self, other):
        if isinstance(other, LogArray):
            return LogArray(self.sign * other.sign, self.logabs + other.logabs)
        return LogArray(self.sign * np.sign(other), self.logabs + np.log(np.abs(other)))

    def numpy(self):
        s = self.sign.detach().cpu().numpy() if hasattr(self.sign, "detach") else np.asarray(self.sign)
        l = self.logabs.detach().cpu().numpy() if hasattr(self.logabs, "detach") else np.asarray(self.logabs)
        return s, l


# ------------------------------------------------------------------------ parameter tables


def softplus(t: float) -> float:
    return math.log1p(math.exp(-abs(t))) + max(t, 0.0)


def build_wtilde(W, theta, LX, LY) -> np.ndarray:
    """(3, N, d) raw W -> (4, N, d) table [uu, ud, du, dd] with the distance envelope;
    uu/dd antisymmetrised in r and du(r) = -ud(-r) (``code/backflow/pf_kernel.py:107-124``).
    Computed in float64, returned as float32."""
    _, neg, dist = displacement_tables(LX, LY)
    W = np.asarray(W, np.float64)
    N = LX * LY
    blocks = [(W[1] - W[1][neg]) / 2, W[0], -W[0][neg], (W[2] - W[2][neg]) / 2]
    pref = np.exp(-softplus(float(theta)) * dist)
    pref /= N * pref.sum()
    return (np.stack(blocks) * pref[None, :, None]).astype(np.float32)


def build_wt_ud(W, theta, LX, LY) -> np.ndarray:
    """(N, d) -> W * pref (``code/backflow/bcs_kernel.py:63-70``)."""
    _, _, dist = displacement_tables(LX, LY)
    pref = np.exp(-softplus(float(theta)) * dist)
    pref /= LX * LY * pref.sum()
    return (np.asarray(W, np.float64) * pref[:, None]).astype(np.float32)


def pack_conv(params: Params, dtype=np.float32):
    """Conv weights as [conv][tap = (dx+1)*3 + (dy+1)][c_in][c_out] and biases [conv][C]; a
    residual CNN's occupation embedding goes first as [class][C], a linear head last as
    [k][c][d]."""
    w = [np.ascontiguousarray(np.transpose(np.asarray(cw, dtype), (2, 3, 1, 0))).reshape(-1)
         for cw in params.conv_w]
    if params.emb is not None:
        w.insert(0, np.ascontiguousarray(np.asarray(params.emb, dtype).T).reshape(-1))
    if params.head is not None:  # (2, d, C) -> [k][c][d]
        w.append(np.ascontiguousarray(np.transpose(np.asarray(params.head, dtype), (0, 2, 1))).reshape(-1))
    b = [np.asarray(cb, dtype).reshape(-1) for cb in params.conv_b]
    return np.concatenate(w), np.concatenate(b)


# ------------------------------------------------------------------------ kernel wrappers


def _dev(x, dtype):
    torch = _torch()
    return torch.as_tensor(np.ascontiguousarray(x), dtype=dtype, device="cuda")


def pairing_occ(Y, idx, Wt, LX: int, LY: int):
    """Occupied pairing block F[b, p, q] (``pf_kernel.pairing_occ``); Y (B, n, d) float32,
    idx (B, n) int32 mode indices, Wt (4, N, d).  Returns (B, n, n) float32, exactly
    antisymmetric."""
    torch = _torch()
    _lib.require_gpu()
    Y = Y.contiguous().float()
    idx = idx.contiguous().to(torch.int32)
    Wt = Wt.contiguous().float()
    B, n, d = Y.shape
    F = torch.empty((B, n, n), dtype=torch.float32, device=Y.device)
    lib = _lib.load()
    _lib.check(lib.nh_pairing_occ(_lib.stream_handle(), LX, LY, d, B, n, _lib.ptr(Y), _lib.ptr(idx),
                                  _lib.ptr(Wt), _lib.ptr(F)), "nh_pairing_occ")
    return F


def pairing_ud(Yu, Yd, iu, idn, Wt, LX: int, LY: int):
    """Up x down pairing block (``bcs_kernel.pairing_ud``)."""
    torch = _torch()
    _lib.require_gpu()
    Yu, Yd = Yu.contiguous().float(), Yd.contiguous().float()
    iu, idn = iu.contiguous().to(torch.int32), idn.contiguous().to(torch.int32)
    Wt = Wt.contiguous().float()
    B, nu, d = Yu.shape
    nd = Yd.shape[1]
    F = torch.empty((B, nu, nd), dtype=torch.float32, device=Yu.device)
    lib = _lib.load()
    _lib.check(lib.nh_pairing_ud(_lib.stream_handle(), LX, LY, d, B, nu, nd, _lib.ptr(Yu), _lib.ptr(Yd),
                                 _lib.ptr(iu), _lib.ptr(idn), _lib.ptr(Wt), _lib.ptr(F)), "nh_pairing_ud")
    return F


_SLOGPF_WS = None
_SLOGPF_WS_CAP = 1 << 30  # bytes of workspace per launch; larger batches run in chunks


def slogpf(F, workspace: bool | None = None):
    """Batched (sign, log|Pf|) of antisymmetric (B, n, n) float32 (``fermix.slogpf``), n <= 1024.
    workspace: use `nh_slogpf_ws` (the matrices in a workspace); default for n >= 100, where
    its column-major form is faster than the LDS form (0.116 against 0.128 us per matrix at
    n = 130, 0.297 against 0.510 at n = 224), and above n = 256, where the LDS form does not
    reach; below n = 100 the LDS form is the faster (0.068 against 0.075 us at n = 98)."""
    global _SLOGPF_WS
    torch = _torch()
    _lib.require_gpu()
    F = F.contiguous().float()
    B, n, _ = F.shape
    sign = torch.empty(B, dtype=torch.float32, device=F.device)
    logabs = torch.empty(B, dtype=torch.float64, device=F.device)
    lib = _lib.load()
    if workspace is None:
        workspace = n >= 100
    if not workspace:
        _lib.check(lib.nh_slogpf(_lib.stream_handle(), B, n, _lib.ptr(F), _lib.ptr(sign), _lib.ptr(logabs)),
                   "nh_slogpf")
        return sign, logabs
    per = max(1, lib.nh_slogpf_workspace_size(1, n) - 256)
    nb = max(1, min(B, _SLOGPF_WS_CAP // per))
    if _SLOGPF_WS is None:
        _SLOGPF_WS = _Workspace()
    ws = _SLOGPF_WS.get(lib.nh_slogpf_workspace_size(nb, n))
    _lib.check(lib.nh_slogpf_ws(_lib.stream_handle(), B, n, _lib.ptr(F), _lib.ptr(sign), _lib.ptr(logabs),
                                _lib.ptr(ws), ws.numel()), "nh_slogpf_ws")
    return sign, logabs


def slogdet(F):
    """Batched (sign, log|det|) of (B, n, n) float32 (``fermix.slogdet``)."""
    torch = _torch()
    _lib.require_gpu()
    F = F.contiguous().float()
    B, n, _ = F.shape
    sign = torch.empty(B, dtype=torch.float32, device=F.device)
    logabs = torch.empty(B, dtype=torch.float64, device=F.device)
    _lib.check(_lib.load().nh_slogdet(_lib.stream_handle(), B, n, _lib.ptr(F), _lib.ptr(sign), _lib.ptr(logabs)),
               "nh_slogdet")
    return sign, logabs


def pairing_occ_vjp(Y, idx, Wt, G, LX: int, LY: int):
    """(dY (B, n, d), dWt (B, 4, N, d)): the VJP of `pairing_occ` for the cotangent G (B, n, n),
    per sample (the backward of ``pairing_occ``'s custom_vjp, ``pf_kernel.py:354-365``)."""
    torch = _torch()
    _lib.require_gpu()
    Y = Y.contiguous().float()
    idx = idx.contiguous().to(torch.int32)
    Wt = Wt.contiguous().float()
    G = G.contiguous().float()
    B, n, d = Y.shape
    N = LX * LY
    dY = torch.empty((B, n, d), dtype=torch.float32, device=Y.device)
    dWt = torch.empty((B, 4, N, d), dtype=torch.float32, device=Y.device)
    _lib.check(_lib.load().nh_pairing_occ_vjp(_lib.stream_handle(), LX, LY, d, B, n, _lib.ptr(Y), _lib.ptr(idx),
                                              _lib.ptr(Wt), _lib.ptr(G), _lib.ptr(dY), _lib.ptr(dWt)),
               "nh_pairing_occ_vjp")
    return dY, dWt


def slogpf_vjp(F, cot=None):
    """cot * d log|Pf F| / dF = cot * F^{-T} / 2 for antisymmetric (B, n, n) float32."""
    torch = _torch()
    _lib.require_gpu()
    F = F.contiguous().float()
    B, n, _ = F.shape
    G = torch.empty_like(F)
    c = None if cot is None else torch.as_tensor(cot, dtype=torch.float32, device=F.device).contiguous()
    _lib.check(_lib.load().nh_slogpf_vjp(_lib.stream_handle(), B, n, _lib.ptr(F), 0 if c is None else _lib.ptr(c),
                                         _lib.ptr(G)), "nh_slogpf_vjp")
    return G


def pairing_ud_vjp(Yu, Yd, iu, idn, Wt, G, LX: int, LY: int):
    """(dYu (B, n_up, d), dYd (B, n_dn, d), dWt (B, N, d)): the VJP of `pairing_ud` for the
    cotangent G (B, n_up, n_dn), per sample (the backward of ``pairing_ud``'s custom_vjp,
    ``bcs_kernel.py:274-288``)."""
    torch = _torch()
    _lib.require_gpu()
    Yu, Yd = Yu.contiguous().float(), Yd.contiguous().float()
    iu, idn = iu.contiguous().to(torch.int32), idn.contiguous().to(torch.int32)
    Wt, G = Wt.contiguous().float(), G.contiguous().float()
    B, nu, d = Yu.shape
    nd = Yd.shape[1]
    dYu = torch.empty((B, nu, d), dtype=torch.float32, device=Yu.device)
    dYd = torch.empty((B, nd, d), dtype=torch.float32, device=Yu.device)
    dWt = torch.empty((B, LX * LY, d), dtype=torch.float32, device=Yu.device)
    _lib.check(_lib.load().nh_pairing_ud_vjp(_lib.stream_handle(), LX, LY, d, B, nu, nd, _lib.ptr(Yu), _lib.ptr(Yd),
                                             _lib.ptr(iu), _lib.ptr(idn), _lib.ptr(Wt), _lib.ptr(G), _lib.ptr(dYu),
                                             _lib.ptr(dYd), _lib.ptr(dWt)), "nh_pairing_ud_vjp")
    return dYu, dYd, dWt


def slogdet_vjp(F, cot=None):
    """cot * d log|det F| / dF = cot * F^{-T} for (B, n, n) float32."""
    torch = _torch()
    _lib.require_gpu()
    F = F.contiguous().float()
    B, n, _ = F.shape
    G = torch.empty_like(F)
    c = None if cot is None else torch.as_tensor(cot, dtype=torch.float32, device=F.device).contiguous()
    _lib.check(_lib.load().nh_slogdet_vjp(_lib.stream_handle(), B, n, _lib.ptr(F), 0 if c is None else _lib.ptr(c),
                                          _lib.ptr(G)), "nh_slogdet_vjp")
    return G


# ------------------------------------------------------------------------ models


class _Backflow:
    kind = None

    def __init__(self, workload: Workload, params: Params, precision: str = "fp32"):
        torch = _torch()
        _lib.require_gpu()
        self.workload = workload
        self.params = params
        self.precision = precision
        w, b = pack_conv(params)
        self.conv_w = _dev(w, torch.float32)
        self.conv_b = _dev(b, torch.float32)
        LX, LY = workload.LX, workload.LY
        if self.kind == _lib.NH_KIND_PFAFFIAN:
            self.Wt = _dev(build_wtilde(params.W, params.theta, LX, LY), torch.float32)
            self.base = None if params.F0 is None else _dev(params.F0, torch.float32)
        else:
            self.Wt = _dev(build_wt_ud(params.W, params.theta, LX, LY), torch.float32)
            self.base = None if params.phi0 is None else _dev(params.phi0, torch.float32)
        self.conv_w_lp = None
        self.conv_w_mfma = None
        if workload.network == "resconv":
            self.network = _lib.NH_NET_RESCONV_HEAD if params.head is not None else _lib.NH_NET_RESCONV
        else:
            self.network = _lib.NH_NET_CONV
        if self.network != _lib.NH_NET_CONV:  # residual CNN: FP32 weight slabs
            from .mixed import pack_resconv_f32

            self.conv_w_mfma = _dev(pack_resconv_f32(params), torch.float32)
        elif workload.LY == 16 and workload.channels == 32:
            from .mixed import pack_mfma_f32

            self.conv_w_mfma = _dev(pack_mfma_f32(params), torch.float32)
        if precision == "mixed":
            from .mixed import pack_lowp

            self.conv_w_lp = pack_lowp(self)

    @property
    def d(self):
        return self.workload.d

    def struct(self, project: bool) -> _lib.NhModel:
        w = self.workload
        return _lib.NhModel(
            self.kind, w.LX, w.LY, w.n_up, w.n_dn, w.channels, w.n_layers,
            _lib.NH_PREC_MIXED if self.precision == "mixed" else _lib.NH_PREC_FP32,
            1 if project else 0, self.network,
            self.conv_w.data_ptr(), self.conv_b.data_ptr(), self.Wt.data_ptr(),
            0 if self.base is None else self.base.data_ptr(),
            0 if self.conv_w_lp is None else self.conv_w_lp.data_ptr(),
            0 if self.conv_w_mfma is None else self.conv_w_mfma.data_ptr(),
        )

    def features(self, s):
        """(B, 2, N, d) features of configurations s (B, 2N) int8 (``models.py:163-167``)."""
        torch = _torch()
        s = torch.as_tensor(s, dtype=torch.int8, device="cuda").contiguous()
        if s.ndim == 1:
            s = s[None]
        w = self.workload
        y = torch.empty((s.shape[0], 2, w.N, w.d), dtype=torch.float32, device="cuda")
        m = self.struct(False)
        lib = _lib.load()
        need = lib.nh_features_workspace_size(ctypes.byref(m), s.shape[0])
        if need == 0:
            raise _lib.NhError("invalid model description")
        ws = torch.empty(int(need), dtype=torch.uint8, device="cuda")
        _lib.check(lib.nh_features_ws(_lib.stream_handle(), ctypes.byref(m), s.shape[0], _lib.ptr(s), _lib.ptr(y),
                                      _lib.ptr(ws), ws.numel()), "nh_features_ws")
        return y

    def __call__(self, s) -> LogArray:
        """Unprojected amplitude of configurations s (``models.py:208-211``)."""
        return Variational(self, symm=False)(s)


class PfaffianBackflow(_Backflow):
    """psi(s) = Pf[n * f(n) * n] (``code/backflow/models.py:121-211``)."""

    kind = _lib.NH_KIND_PFAFFIAN


class BCSBackflow(_Backflow):
    """psi(s) = (-1)^{n(n-1)/2} det[n_up * f_ud(n) * n_dn] (``code/backflow/models.py:214-302``)."""

    kind = _lib.NH_KIND_BCS


def build_backflow(workload: Workload, params: Params, precision: str = "fp32"):
    """The backflow of a workload (``build_backflow``, ``code/backflow/models.py:305-332``)."""
    if workload.kind == "pfaffian":
        return PfaffianBackflow(workload, params, precision)
    if workload.kind == "bcs":
        return BCSBackflow(workload, params, precision)
    raise ValueError(f"Unknown backflow kind {workload.kind!r}; use 'bcs' or 'pfaffian'.")


# ------------------------------------------------------------------------ state and operator


class _Workspace:
    def __init__(self):
        self.buf = None

    def get(self, nbytes: int):
        torch = _torch()
        if self.buf is None or self.buf.numel() < nbytes:
            self.buf = None
            self.buf = torch.empty(int(nbytes), dtype=torch.uint8, device="cuda")
        return self.buf


class Variational:
    """Projected state (``quantax.state.Variational(model, symm=paper_symmetry())``,
    ``code/states.py:318-320``): ``state(spins)`` returns the LogArray of
    sum_{g in G} eta_g(s) psi(g.s) over Flip(0) @ Flip(1) @ SpinInverse (``symm=True``), or the
    bare network (``symm=False``).  ``max_parallel`` bounds the configurations per launch."""

    def __init__(self, model: _Backflow, symm: bool = True, max_parallel: int = 8192):
        self.model = model
        self.symm = bool(symm)
        self.max_parallel = int(max_parallel)
        self._ws = _Workspace()

    def __call__(self, spins) -> LogArray:
        torch = _torch()
        s = torch.as_tensor(spins, dtype=torch.int8, device="cuda").contiguous()
        if s.ndim == 1:
            s = s[None]
        B = s.shape[0]
        sign = torch.empty(B, dtype=torch.float32, device="cuda")
        logabs = torch.empty(B, dtype=torch.float64, device="cuda")
        m = self.model.struct(self.symm)
        lib = _lib.load()
        for b0 in range(0, B, self.max_parallel):
            nb = min(self.max_parallel, B - b0)
            need = lib.nh_log_psi_workspace_size(ctypes.byref(m), nb)
            if need == 0:
                raise _lib.NhError("invalid model description")
            ws = self._ws.get(need)
            _lib.check(lib.nh_log_psi(_lib.stream_handle(), ctypes.byref(m), nb, _lib.ptr(s[b0:]),
                                      _lib.ptr(sign[b0:]), _lib.ptr(logabs[b0:]), _lib.ptr(ws), ws.numel()),
                       "nh_log_psi")
        return LogArray(sign, logabs)

    def jacobian(self, spins, workspace_cap: int = 4 << 30):
        """(LogArray of psi, J (B, P) float32): J[b, k] = d log|psi(s_b)| / d theta_k, the
        per-sample Jacobian SR training takes (``AdamSR``, ``train_pfaffian_16x16_tpm02.py:100-101``),
        columns in the order of `param_vector`.  Batches run in chunks whose workspace fits
        `workspace_cap` bytes."""
        torch = _torch()
        s = torch.as_tensor(spins, dtype=torch.int8, device="cuda").contiguous()
        if s.ndim == 1:
            s = s[None]
        B = s.shape[0]
        m = self.model.struct(self.symm)
        lib = _lib.load()
        P = lib.nh_log_psi_jacobian_size(ctypes.byref(m))
        if P == 0:
            raise _lib.NhError("model not supported by nh_log_psi_jacobian")
        per = max(1, lib.nh_log_psi_jacobian_workspace_size(ctypes.byref(m), 1))
        nb_max = max(1, min(B, workspace_cap // per, 65535 // (8 if self.symm else 1)))
        sign = torch.empty(B, dtype=torch.float32, device="cuda")
        logabs = torch.empty(B, dtype=torch.float64, device="cuda")
        J = torch.empty((B, P), dtype=torch.float32, device="cuda")
        theta = float(self.model.params.theta)
        for b0 in range(0, B, nb_max):
            nb = min(nb_max, B - b0)
            ws = self._ws.get(lib.nh_log_psi_jacobian_workspace_size(ctypes.byref(m), nb))
            _lib.check(lib.nh_log_psi_jacobian(_lib.stream_handle(), ctypes.byref(m), theta, nb, _lib.ptr(s[b0:]),
                                               _lib.ptr(sign[b0:]), _lib.ptr(logabs[b0:]), _lib.ptr(J[b0:]),
                                               _lib.ptr(ws), ws.numel()), "nh_log_psi_jacobian")
        return LogArray(sign, logabs), J


def param_vector(params: Params) -> np.ndarray:
    """The parameters as one float64 vector in the column order of `Variational.jacobian`
    (``nh_log_psi_jacobian``): packed conv weights, conv biases, W, theta, then F0 or phi0
    when the state has a base table."""
    w, b = pack_conv(params, np.float64)
    parts = [w, b, np.asarray(params.W).ravel(), [params.theta]]
    if params.F0 is not None:
        parts.append(np.asarray(params.F0).ravel())
    elif params.phi0 is not None:
        parts.append(np.asarray(params.phi0).ravel())
    return np.concatenate([np.asarray(p, np.float64).ravel() for p in parts])


@dataclass
class Samples:
    """Configurations with their amplitudes (``quantax.sampler.Samples``)."""

    spins: object
    psi: LogArray | None = None


class Hubbard:
    """Hubbard model (``quantax.operator.Hubbard``; built at ``code/states.py:244-247``):
    H = -t sum_<ij>,s (c+_is c_js + h.c.) - t' sum_<<ij>>,s (...) + U sum_i n_iu n_id."""

    def __init__(self, U: float = 8.0, t: float = 1.0, tprime: float = 0.0, chunk: int = 16,
                 lowrank: bool = False):
        """lowrank: the hop ratios as low-rank updates of each sample's image matrices
        (``nh_local_energy_lowrank``: Pfaffian states on the synthetic network, e.g. 32x32)."""
        self.U, self.t, self.tprime = float(U), float(t), float(tprime)
        self.chunk = int(chunk)
        self.lowrank = bool(lowrank)
        self._ws = _Workspace()

    def Oloc(self, state: Variational, samples) -> "object":
        """E_loc(s) for every sample (float64 tensor)."""
        torch = _torch()
        if not isinstance(samples, Samples):
            samples = Samples(samples)
        s = torch.as_tensor(samples.spins, dtype=torch.int8, device="cuda").contiguous()
        if s.ndim == 1:
            s = s[None]
        psi = samples.psi if samples.psi is not None else state(s)
        B = s.shape[0]
        eloc = torch.empty(B, dtype=torch.float64, device="cuda")
        m = state.model.struct(state.symm)
        lib = _lib.load()
        chunk = max(1, min(self.chunk, B))
        size = lib.nh_local_energy_lowrank_workspace_size if self.lowrank else lib.nh_local_energy_workspace_size
        need = size(ctypes.byref(m), self.tprime, chunk)
        if need == 0:
            raise _lib.NhError("model or chunk not supported" if self.lowrank else "invalid model description")
        ws = self._ws.get(need)
        sign = psi.sign.contiguous().float()
        logabs = psi.logabs.contiguous().double()
        fn = lib.nh_local_energy_lowrank if self.lowrank else lib.nh_local_energy
        _lib.check(fn(_lib.stream_handle(), ctypes.byref(m), self.t, self.tprime, self.U, B, _lib.ptr(s),
                      _lib.ptr(sign), _lib.ptr(logabs), _lib.ptr(eloc), chunk, _lib.ptr(ws), ws.numel()),
                   "nh_local_energy_lowrank" if self.lowrank else "nh_local_energy")
        return eloc


def hops(s, LX, LY, t, tprime, n_occ):
    """Hop enumeration of ``nh_hops``: (count (B,), from (B, H), to (B, H), coef (B, H))."""
    torch = _torch()
    s = torch.as_tensor(s, dtype=torch.int8, device="cuda").contiguous()
    B = s.shape[0]
    H = n_occ * (8 if tprime != 0 else 4)
    count = torch.empty(B, dtype=torch.int32, device="cuda")
    fr = torch.empty((B, H), dtype=torch.int32, device="cuda")
    to = torch.empty((B, H), dtype=torch.int32, device="cuda")
    co = torch.empty((B, H), dtype=torch.float32, device="cuda")
    _lib.check(_lib.load().nh_hops(_lib.stream_handle(), LX, LY, t, tprime, B, n_occ, _lib.ptr(s), H,
                                   _lib.ptr(count), _lib.ptr(fr), _lib.ptr(to), _lib.ptr(co)), "nh_hops")
    return count, fr, to, co
