"""Metropolis exchange sampler and the learned surrogate of delayed acceptance (host side).

`MetropolisSampler` plays the role of the reference's ``make_sampler`` / quantax
``MixSampler.sweep()`` (``code/states.py:355-408``, ``code/measure_orderparams.py:128``):
chains of configurations resident in HBM, ``sweep()`` runs ``sweep_steps`` proposals per chain
and returns `Samples` with the projected amplitudes.  With a `Surrogate` it runs delayed
acceptance: the surrogate screens every proposal and only survivors reach the exact amplitude,
which keeps |psi|^2 exactly stationary (Christen & Fox 2005).

`Surrogate` is a small translation-equivariant conv network predicting log|psi| (layout in
``include/nh.h``), trained on the GPU by the library's own forward/backward kernels with Adam
on (configuration, exact log|psi|) pairs the exact path produced.
"""

from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _lib
from .backflow import Samples, Variational, _Workspace


def _torch():
    import torch

    return torch


class Surrogate:
    """f(s) = sum_sites conv3(gelu(conv2(gelu(conv1(x0))))) with `channels` channels."""

    def __init__(self, LX: int, LY: int, channels: int = 32, seed: int = 0):
        torch = _torch()
        _lib.require_gpu()
        self.LX, self.LY, self.channels = LX, LY, channels
        lib = _lib.load()
        self.nparams = lib.nh_sur_nparams(channels)
        rng = np.random.Generator(np.random.PCG64(seed))
        C = channels
        parts = [
            rng.standard_normal(9 * 2 * C) / math.sqrt(18),  # w1
            np.zeros(C),  # b1
            rng.standard_normal(9 * C * C) / math.sqrt(9 * C),  # w2
            np.zeros(C),  # b2
            rng.standard_normal(9 * C) / math.sqrt(9 * C) * 0.1,  # w3
            np.zeros(1),  # b3
        ]
        p = np.concatenate(parts).astype(np.float32)
        assert p.size == self.nparams
        self.params = torch.as_tensor(p, device="cuda")
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.adam_step = 0
        self.offset = 0.0  # f(s) + offset approximates log|psi(s)|
        self._ws = _Workspace()

    def struct(self) -> _lib.NhSurrogate:
        return _lib.NhSurrogate(self.LX, self.LY, self.channels, 0, self.params.data_ptr())

    def __call__(self, s):
        """f(s) + offset for s (B, 2N) int8."""
        torch = _torch()
        s = torch.as_tensor(s, dtype=torch.int8, device="cuda").contiguous()
        out = torch.empty(s.shape[0], dtype=torch.float32, device="cuda")
        q = self.struct()
        _lib.check(_lib.load().nh_sur_forward(_lib.stream_handle(), ctypes.byref(q), s.shape[0], _lib.ptr(s),
                                              _lib.ptr(out)), "nh_sur_forward")
        return out + self.offset

    def grad(self, s, target, weight: float):
        """(gradient of sum 0.5 weight (f(s) - target)^2, that loss)."""
        torch = _torch()
        s = torch.as_tensor(s, dtype=torch.int8, device="cuda").contiguous()
        target = torch.as_tensor(target, dtype=torch.float32, device="cuda").contiguous()
        B = s.shape[0]
        q = self.struct()
        lib = _lib.load()
        ws = self._ws.get(lib.nh_sur_grad_workspace_size(ctypes.byref(q), B))
        g = torch.empty(self.nparams, dtype=torch.float32, device="cuda")
        loss = torch.empty(1, dtype=torch.float32, device="cuda")
        _lib.check(lib.nh_sur_grad(_lib.stream_handle(), ctypes.byref(q), B, _lib.ptr(s), _lib.ptr(target),
                                   float(weight), _lib.ptr(g), _lib.ptr(loss), _lib.ptr(ws), ws.numel()),
                   "nh_sur_grad")
        return g, loss

    def adam(self, g, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        self.adam_step += 1
        _lib.check(_lib.load().nh_adam(_lib.stream_handle(), self.nparams, _lib.ptr(self.params), _lib.ptr(g),
                                       _lib.ptr(self.m), _lib.ptr(self.v), lr, beta1, beta2, eps, self.adam_step),
                   "nh_adam")

    def train(self, spins, logabs, epochs=10, batch=1024, lr=2e-3, seed=0, augment=True, log=None):
        """Fit f + offset to log|psi| by mean squared error with Adam.  The pairs stay on the GPU;
        with ``augment`` each configuration enters every batch in a random image of the
        projection group (|psi_P| is invariant under it; every group element is an involution, so
        the image is a gather of the modes).  Returns the per-epoch mean squared error."""
        torch = _torch()
        from .lattice import group_mode_perms

        x_all = torch.as_tensor(spins, dtype=torch.int8, device="cuda")
        y = torch.as_tensor(logabs, dtype=torch.float64, device="cuda")
        self.offset = float(y.mean().item())
        target = (y - self.offset).float()
        perms = torch.as_tensor(np.asarray(group_mode_perms(self.LX, self.LY)), dtype=torch.int64, device="cuda")
        gen = torch.Generator(device="cuda")
        gen.manual_seed(seed)
        n = x_all.shape[0]
        hist = []
        for ep in range(epochs):
            order = torch.randperm(n, generator=gen, device="cuda")
            tot = torch.zeros((), dtype=torch.float64, device="cuda")
            nb = 0
            for b0 in range(0, n - batch + 1, batch):
                idx = order[b0:b0 + batch]
                x = x_all[idx]
                if augment:
                    gs = torch.randint(0, 8, (batch,), generator=gen, device="cuda")
                    x = torch.gather(x, 1, perms[gs])
                g, loss = self.grad(x, target[idx], 1.0 / batch)
                self.adam(g, lr=lr)
                tot += loss[0].double()
                nb += 1
            hist.append(2.0 * float(tot.item()) / max(1, nb))
            if log:
                log(f"surrogate epoch {ep}: mse {hist[-1]:.4f}")
        return hist


class MetropolisSampler:
    """Chains resident on the GPU; ``sweep()`` = ``sweep_steps`` exchange proposals per chain
    (quantax ``Sampler.sweep``), optionally screened by a surrogate (delayed acceptance)."""

    def __init__(self, state: Variational, initial_spins, sweep_steps: int, seed: int = 42,
                 surrogate: Surrogate | None = None, thermal_steps: int = 0):
        torch = _torch()
        self.state = state
        self.sweep_steps = int(sweep_steps)
        self.seed = int(seed)
        self.surrogate = surrogate
        if not isinstance(initial_spins, torch.Tensor):
            initial_spins = np.asarray(initial_spins, np.int8)
        self.s = torch.as_tensor(initial_spins, dtype=torch.int8, device="cuda").contiguous().clone()
        n = self.s.shape[0]
        psi = state(self.s)
        self.sign = psi.sign.contiguous().clone()
        self.logabs = psi.logabs.contiguous().clone()
        self.f_cur = surrogate(self.s).contiguous().clone() if surrogate is not None else None
        self.counters = torch.zeros((4, n), dtype=torch.int32, device="cuda")
        self.step = 0
        self._stats_from = 0
        self._ws = _Workspace()
        if thermal_steps:
            self.run(thermal_steps)

    @property
    def n_chains(self) -> int:
        return self.s.shape[0]

    def run(self, n_steps: int):
        m = self.state.model.struct(self.state.symm)
        lib = _lib.load()
        ws = self._ws.get(lib.nh_metropolis_workspace_size(ctypes.byref(m), self.n_chains))
        sur = self.surrogate.struct() if self.surrogate is not None else None
        if self.surrogate is not None:
            # the kernels compare raw network outputs; the offset cancels in differences
            f_raw = self.f_cur - self.surrogate.offset
        _lib.check(lib.nh_metropolis_sweep(
            _lib.stream_handle(), ctypes.byref(m), ctypes.byref(sur) if sur is not None else None, self.n_chains,
            _lib.ptr(self.s), _lib.ptr(self.sign), _lib.ptr(self.logabs),
            _lib.ptr(f_raw) if sur is not None else None, self.seed, self.step, int(n_steps),
            _lib.ptr(self.counters), _lib.ptr(ws), ws.numel()), "nh_metropolis_sweep")
        if sur is not None:
            self.f_cur = f_raw + self.surrogate.offset
        self.step += int(n_steps)

    def sweep(self) -> Samples:
        from .backflow import LogArray

        self.run(self.sweep_steps)
        return Samples(self.s.clone(), LogArray(self.sign.clone(), self.logabs.clone()))

    def refresh_surrogate(self):
        """Re-evaluate f at the current states; needed after the surrogate was retrained, since
        delayed acceptance compares f(s') with the stored f(s)."""
        if self.surrogate is not None:
            self.f_cur = self.surrogate(self.s).contiguous().clone()

    def reset_stats(self):
        """Start the acceptance counters afresh (for example after thermalisation)."""
        self.counters.zero_()
        self._stats_from = self.step

    def stats(self) -> dict:
        c = self.counters.sum(dim=1).cpu().numpy().astype(np.int64)
        total = (self.step - self._stats_from) * self.n_chains
        return {"proposals": int(total), "changing": int(c[0]), "stage1_passed": int(c[1]),
                "accepted": int(c[2]), "exact_evaluations": int(c[3]),
                "acceptance": float(c[2] / max(1, c[0])),
                "stage1_acceptance": float(c[1] / max(1, c[0])) if self.surrogate is not None else None,
                "stage2_acceptance": float(c[2] / max(1, c[1])) if self.surrogate is not None else None}
