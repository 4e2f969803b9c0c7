"""Metropolis samplers and the learned surrogate of delayed acceptance (host side).

`MetropolisSampler` plays the role of quantax's samplers behind the reference's
``make_sampler`` (``code/states.py:355-408``, ``code/measure_orderparams.py:128``): chains of
configurations resident in HBM, ``sweep()`` runs ``sweep_steps`` proposals per chain and returns
`Samples` with the projected amplitudes.  The move set is a `Proposal`: the default exchange of
two nearest-neighbour modes of one spin, or the reference's `ParticleHop`, `SiteExchange` and
their `MixSampler` with n_neighbor 1 (nearest) or (1, 2) (and next-nearest).  With a `Surrogate`
it runs delayed acceptance: the surrogate screens every proposal and only survivors reach the
exact amplitude, which keeps |psi|^2 exactly stationary (Christen & Fox 2005).

`Surrogate` is a small translation-equivariant conv network predicting log|psi| (layout in
``include/nh.h``), trained on the GPU by the library's own forward/backward kernels with Adam
on (configuration, exact log|psi|) pairs the exact path produced.
"""

from __future__ import annotations

import ctypes
import math

import numpy as np

from dataclasses import dataclass

from . import _lib
from .backflow import Samples, Variational, _Workspace


@dataclass(frozen=True)
class Proposal:
    """Move set of a sweep (``nh_proposal`` of include/nh.h): kind `_lib.MOVE_*` and the number of
    partner sites of a particle hop / site exchange (4 nearest, or 8 with the next-nearest)."""

    kind: int = _lib.MOVE_EXCHANGE
    ph_neighbors: int = 4
    se_neighbors: int = 4

    def struct(self) -> _lib.NhProposal:
        return _lib.NhProposal(self.kind, self.ph_neighbors, self.se_neighbors)


def neighbor_count(n_neighbor) -> int:
    """quantax's ``n_neighbor``: 1 -> the 4 nearest neighbours, (1, 2) -> 8 with the next-nearest
    (``code/states.py:95-96``)."""
    t = tuple(int(v) for v in np.atleast_1d(n_neighbor))
    if t == (1,):
        return 4
    if t == (1, 2):
        return 8
    raise ValueError(f"n_neighbor {n_neighbor!r}: expected 1 or (1, 2)")


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
    """Chains resident on the GPU; ``sweep()`` = ``sweep_steps`` proposals per chain (quantax
    ``Sampler.sweep``) drawn from `proposal`, optionally screened by a surrogate (delayed
    acceptance)."""

    def __init__(self, state: Variational, initial_spins, sweep_steps: int, seed: int = 42,
                 surrogate: Surrogate | None = None, thermal_steps: int = 0, proposal: Proposal | None = None):
        torch = _torch()
        self.state = state
        self.sweep_steps = int(sweep_steps)
        self.seed = int(seed)
        self.surrogate = surrogate
        self.proposal = proposal if proposal is not None else Proposal()
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
        prop = self.proposal.struct()
        _lib.check(lib.nh_metropolis_sweep(
            _lib.stream_handle(), ctypes.byref(m), ctypes.byref(sur) if sur is not None else None,
            ctypes.byref(prop), self.n_chains,
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


def _start_spins(state: Variational, nsamples: int, initial_spins, seed: int):
    """initial_spins repeated or truncated to nsamples chains, or random configurations with the
    state's particle numbers."""
    if initial_spins is None:
        from .configs import make_configurations

        return make_configurations(state.model.workload, nsamples, seed=seed)
    torch = _torch()
    if isinstance(initial_spins, torch.Tensor):
        initial_spins = initial_spins.cpu().numpy()
    spins = np.asarray(initial_spins, np.int8)
    return np.resize(spins, (nsamples, spins.shape[-1])) if spins.shape[0] != nsamples else spins


class _MoveSampler(MetropolisSampler):
    """A sampler with one move kind, in the form of quantax's constructors.  Inside a
    `MixSampler` it only names its move, neighbour range and chain count; its own chains are
    made on first use."""

    kind = _lib.MOVE_EXCHANGE

    def __init__(self, state: Variational, nsamples: int, thermal_steps: int = 0, sweep_steps: int | None = None,
                 initial_spins=None, n_neighbor=1, seed: int = 42, surrogate: Surrogate | None = None):
        self.state = state
        self.nsamples = int(nsamples)
        self.z = neighbor_count(n_neighbor)
        self._pending = (initial_spins, sweep_steps, seed, surrogate, thermal_steps)

    def _start(self):
        if getattr(self, "_pending", None) is None:
            return
        init, sweep_steps, seed, surrogate, thermal = self._pending
        self._pending = None
        N = self.state.model.workload.N
        MetropolisSampler.__init__(self, self.state, _start_spins(self.state, self.nsamples, init, seed),
                                   N if sweep_steps is None else sweep_steps, seed, surrogate, thermal,
                                   Proposal(self.kind, self.z, self.z))

    def run(self, n_steps: int):
        self._start()
        super().run(n_steps)

    def __getattr__(self, name):
        if name in ("s", "sign", "logabs", "counters", "f_cur", "step", "_stats_from", "_ws", "proposal",
                    "sweep_steps", "seed", "surrogate"):
            self._start()
            return self.__dict__[name]
        raise AttributeError(name)


class ParticleHop(_MoveSampler):
    """quantax ``sampler.ParticleHop``: a uniformly drawn particle hops to one of its 4 (or 8)
    neighbouring sites of the same spin; no move when that mode is occupied."""

    kind = _lib.MOVE_PARTICLE_HOP


class SiteExchange(_MoveSampler):
    """quantax ``sampler.SiteExchange``: a uniformly drawn site trades its (up, down) occupation
    with one of its 4 (or 8) neighbouring sites."""

    kind = _lib.MOVE_SITE_EXCHANGE


class MixSampler(MetropolisSampler):
    """quantax ``sampler.MixSampler`` of `ParticleHop` and/or `SiteExchange` as the reference
    builds it (``code/states.py:397-408``): sum(nsamples) chains; with both moves each chain draws
    particle hop or site exchange with probability 1/2 at every step, a symmetric mixture (the
    reference gives each move its share of the chains, nsamples // 2 each)."""

    def __init__(self, samplers, thermal_steps: int = 0, sweep_steps: int | None = None, initial_spins=None,
                 seed: int = 42, surrogate: Surrogate | None = None):
        samplers = list(samplers)
        if not samplers or not all(isinstance(x, (ParticleHop, SiteExchange)) for x in samplers):
            raise TypeError("MixSampler takes ParticleHop and SiteExchange samplers")
        state = samplers[0].state
        ph = [x for x in samplers if isinstance(x, ParticleHop)]
        se = [x for x in samplers if isinstance(x, SiteExchange)]
        if len({x.z for x in ph}) > 1 or len({x.z for x in se}) > 1:
            raise ValueError("one neighbour range per move kind")
        if ph and se:
            prop = Proposal(_lib.MOVE_MIX, ph[0].z, se[0].z)
        elif ph:
            prop = Proposal(_lib.MOVE_PARTICLE_HOP, ph[0].z, 4)
        else:
            prop = Proposal(_lib.MOVE_SITE_EXCHANGE, 4, se[0].z)
        nsamples = sum(x.nsamples for x in samplers)
        N = state.model.workload.N
        super().__init__(state, _start_spins(state, nsamples, initial_spins, seed),
                         N if sweep_steps is None else sweep_steps, seed, surrogate, thermal_steps, prop)
