"""Released states of the reference: weight files, chain files and the paper's energies.

Reference interfaces mirrored here (ChenAo-Phys/NeuralHubbard):

* ``weights/<name>.eqx`` written by ``equinox.tree_serialise_leaves`` (``weights/README.md:6-7``):
  the leaves of the model pytree, one ``np.save`` record after another.  `load_eqx_leaves` reads
  them with ``np.load(allow_pickle=False)``, so nothing in the file is executed.
* ``weights/<name>_spins.npz`` (``code/states.py:328-352``): ``bits = np.packbits(spins > 0,
  axis=-1)`` and ``shape = (nchains, 2N)``; `unpack_spins` / `pack_spins` / `load_spins`.
* the state table of ``code/states.py:99-128`` for the backflow states this library evaluates
  (`STATES`: the three 16x16 states, the 24x24 Pfaffian state and the 32x32 BCS state), with the
  projected E/N and sigma^2/N of the paper.

`released_params` maps the leaves of a ``PfaffianBackflow`` (``code/backflow/models.py:121-151``:
``net`` = quantax ResConv with 16 blocks of 64 channels and a final bias, ``W`` (3, N, 32),
``theta``, ``pad``) or a ``BCSBackflow`` (``models.py:214-249``: the ResConv trunk without the
final bias, ``head`` (2, 64, 64), ``W`` (N, 64), ``theta``, ``delta_d``, ``pad``) onto this
library's `Params` of a ``network="resconv"`` workload.
"""

from __future__ import annotations

import os
from dataclasses import dataclass

import numpy as np

from .configs import Params, Workload


@dataclass(frozen=True)
class StateInfo:
    """A released backflow state (``code/states.py:39-92``, rows at ``:105-119``)."""

    name: str
    ansatz: str  # "pfaffian" or "bcs"
    L: int
    tprime: float
    md5: str
    nparams: int
    E_per_site: float
    E_err: float
    var_per_site: float
    var_err: float
    U: float = 8.0
    # n_neighbor of (ParticleHop, SiteExchange), and (measurement sweep, thermalisation,
    # optimisation sweep) in Metropolis steps per chain / N (code/states.py:95-103,105-119)
    neighbors: tuple = ((1,), (1,))
    sweeps: tuple = (4, 50, 2)

    @property
    def N(self) -> int:
        return self.L * self.L

    @property
    def n_up(self) -> int:  # 1/8 hole doping, code/states.py:184
        return 7 * self.N // 16

    def workload(self, batch: int = 8192) -> Workload:
        # BCS: d = 64 features per spin from a linear head on the 64 channels (models.py:246-270)
        return Workload(self.name, self.ansatz, self.L, self.L, self.n_up, self.n_up, self.tprime, self.U,
                        64, 16, batch, network="resconv", features=64 if self.ansatz == "bcs" else 0,
                        notes=f"released state {self.name}")


_NN, _NNN = (1,), (1, 2)

STATES = {
    s.name: s
    for s in (
        StateInfo("tp0_L16_bcs", "bcs", 16, 0.0, "cdfdb8e6065b904232bf10dec6ca2227", 1_208_320,
                  -0.75878, 2e-5, 0.0255, 0.0020, neighbors=(_NN, _NN), sweeps=(4, 50, 2)),
        StateInfo("tp0_L16_pfaffian", "pfaffian", 16, 0.0, "17b770fe57367b09b27ad37538842ba0", 1_208_320,
                  -0.75813, 1e-5, 0.0259, 0.0002, neighbors=(_NN, _NN), sweeps=(4, 50, 2)),
        StateInfo("tpm02_L16_pfaffian", "pfaffian", 16, -0.2, "aef48e8b7c10cfa183f8c0e8dcf84d06", 1_208_320,
                  -0.74533, 1e-5, 0.0272, 0.0005, neighbors=(_NNN, _NNN), sweeps=(4, 50, 3)),
        StateInfo("tpm02_L24_pfaffian", "pfaffian", 24, -0.2, "17653481d380416e6df7033ba7902aee", 1_241_088,
                  -0.74585, 1e-5, 0.0257, 0.0003, neighbors=(_NNN, _NNN), sweeps=(4, 50, 3)),
        StateInfo("tp0_L32_bcs", "bcs", 32, 0.0, "e45a997c3202fae9df7c4f38bd7e6acc", 1_257_472,
                  -0.75617, 1e-5, 0.046, 0.010, neighbors=(_NN, _NN), sweeps=(5, 40, 4)),
    )
}


def load_eqx_leaves(path: str) -> list:
    """All leaves of an ``equinox.tree_serialise_leaves`` file, in pytree order."""
    leaves = []
    size = os.path.getsize(path)
    with open(path, "rb") as f:
        while f.tell() < size:
            leaves.append(np.load(f, allow_pickle=False))
    return leaves


def dwave_table(LX: int, LY: int) -> np.ndarray:
    """F^d as a function of the displacement r = dx * LY + dy: -1 on x bonds, +1 on y bonds
    (minimum image), 0 elsewhere (``code/backflow/models.py:69-83``)."""
    dx, dy = np.divmod(np.arange(LX * LY), LY)
    mx, my = np.minimum(dx, LX - dx), np.minimum(dy, LY - dy)
    return np.where((mx == 1) & (my == 0), -1.0, np.where((mx == 0) & (my == 1), 1.0, 0.0)).astype(np.float32)


def released_params(leaves: list, LX: int, LY: int) -> Params:
    """`Params` of a released backflow from its leaves: the (C, 4, 1, 1) occupation embedding, then
    2 x blocks convolutions (C, C, 3, 3), each followed by its (C, 1, 1) bias (the BCS trunk's last
    convolution has none: zero here), the integer and boolean fields of the network, and the
    pairing leaves: Pfaffian W (3, N, d), theta, pad (``models.py:121-151``); BCS head (2, d, C),
    W (N, d), theta, delta_d, pad (``models.py:214-249``), delta_d F^d going into the
    displacement-indexed base table phi0."""
    arrays = [a for a in leaves if a.dtype.kind == "f"]
    emb = arrays[0]
    if emb.ndim != 4 or emb.shape[1:] != (4, 1, 1):
        raise ValueError(f"expected the (C, 4, 1, 1) embedding first, got {emb.shape}")
    C = emb.shape[0]
    conv_w, conv_b = [], []
    i = 1
    no_bias = []  # convolutions without a bias leaf
    while i < len(arrays) and arrays[i].shape == (C, C, 3, 3):
        conv_w.append(np.ascontiguousarray(arrays[i], np.float32))
        i += 1
        if i < len(arrays) and arrays[i].shape == (C, 1, 1):
            conv_b.append(np.ascontiguousarray(arrays[i].reshape(C), np.float32))
            i += 1
        else:
            conv_b.append(np.zeros(C, np.float32))
            no_bias.append(len(conv_w) - 1)
    if not conv_w or len(conv_w) % 2:
        raise ValueError(f"expected an even number of convolutions, got {len(conv_w)}")
    if no_bias and no_bias != [len(conv_w) - 1]:
        raise ValueError(f"only the last convolution may lack a bias, not {no_bias}")
    missing_bias = len(no_bias)
    N = LX * LY
    tail = arrays[i:]
    if len(tail) == 3:  # Pfaffian: W, theta, pad
        W, theta, pad = tail
        if missing_bias or W.shape != (3, N, C // 2) or theta.shape != () or pad.ndim != 1:
            raise ValueError(f"unexpected Pfaffian tail: W {W.shape}, theta {theta.shape}, pad {pad.shape}")
        head = phi0 = None
    elif len(tail) == 5:  # BCS: head, W, theta, delta_d, pad
        head, W, theta, delta_d, pad = tail
        d = head.shape[1] if head.ndim == 3 else -1
        if (head.shape != (2, d, C) or W.shape != (N, d) or theta.shape != () or delta_d.shape != ()
                or pad.ndim != 1 or missing_bias != 1):
            raise ValueError(f"unexpected BCS tail: head {head.shape}, W {W.shape}, pad {pad.shape}")
        head = np.ascontiguousarray(head, np.float32)
        phi0 = (np.float32(delta_d) * dwave_table(LX, LY)).astype(np.float32)
    else:
        raise ValueError(f"unexpected number of pairing leaves: {len(tail)}")
    if np.any(pad != 0):
        raise ValueError("the pad leaf is not zero")
    return Params(conv_w, conv_b, np.ascontiguousarray(W, np.float32), float(theta), None, phi0,
                  np.ascontiguousarray(emb.reshape(C, 4), np.float32), head)


def n_params(params: Params, pad: int = 0) -> int:
    """Float parameter count as quantax counts it (``code/states.py:321-324``): the BCS trunk's last
    convolution has no bias, and its delta_d is one more scalar."""
    arrays = [params.emb, params.W, *params.conv_w, *params.conv_b]
    n = sum(np.asarray(a).size for a in arrays) + 1 + pad
    if params.head is not None:
        n += params.head.size - params.conv_b[-1].size + 1
    return int(n)


def load_released(name: str, path: str):
    """(workload, params) of the released state `name` from its ``.eqx`` file."""
    info = STATES[name]
    wk = info.workload()
    return wk, released_params(load_eqx_leaves(path), wk.LX, wk.LY)


def unpack_spins(bits: np.ndarray, shape) -> np.ndarray:
    """int8 +-1 array of `shape` from packed occupation bits (``code/states.py:328-332``)."""
    shape = tuple(int(x) for x in shape)
    s = np.unpackbits(np.asarray(bits, np.uint8), axis=-1, count=shape[-1]).astype(np.int8)
    return (2 * s - 1).reshape(shape)


def pack_spins(spins: np.ndarray) -> dict:
    """The ``_spins.npz`` record of int8 +-1 chains: ``bits`` and ``shape``."""
    spins = np.asarray(spins)
    return {"bits": np.packbits(spins > 0, axis=-1), "shape": np.array(spins.shape, np.int64)}


def load_spins(path: str, info: StateInfo | None = None) -> np.ndarray:
    """Chains of a ``_spins.npz`` file, checked against the state's particle numbers
    (``code/states.py:335-352``)."""
    with np.load(path, allow_pickle=False) as f:
        spins = unpack_spins(f["bits"], f["shape"])
    if info is not None:
        N = info.N
        if spins.ndim != 2 or spins.shape[1] != 2 * N:
            raise ValueError(f"{path}: shape {spins.shape}, expected (nchains, {2 * N}).")
        if np.any((spins[:, :N] > 0).sum(1) != info.n_up) or np.any((spins[:, N:] > 0).sum(1) != info.n_up):
            raise ValueError(f"{path}: wrong particle numbers.")
    return spins


def make_sampler(state, info: StateInfo, nsamples: int, initial_spins=None, sweep_steps: int | None = None,
                 thermal_steps: int | None = None, seed: int = 42, spins_path: str | None = None, surrogate=None):
    """The sampler of the paper (``code/states.py:355-408``): `sampler.MixSampler` of
    ``ParticleHop(nsamples // 2)`` and ``SiteExchange(nsamples // 2)`` with the state's neighbour
    ranges, chains from `initial_spins` (or the released chains at `spins_path`) repeated or
    truncated to nsamples, thermalised for ``thermal x N`` steps, ``sweep x N`` steps per sweep."""
    from .sampler import MixSampler, ParticleHop, SiteExchange

    if nsamples % 2:
        raise ValueError("nsamples must be even.")
    if initial_spins is None:
        if spins_path is None:
            raise ValueError("make_sampler needs initial_spins or spins_path")
        initial_spins = load_spins(spins_path, info)
    N = info.N
    sweep_steps = int(round(info.sweeps[0] * N)) if sweep_steps is None else int(sweep_steps)
    thermal_steps = int(round(info.sweeps[1] * N)) if thermal_steps is None else int(thermal_steps)
    ph = ParticleHop(state, nsamples // 2, n_neighbor=info.neighbors[0])
    se = SiteExchange(state, nsamples // 2, n_neighbor=info.neighbors[1])
    return MixSampler([ph, se], thermal_steps=thermal_steps, sweep_steps=sweep_steps, initial_spins=initial_spins,
                      seed=seed, surrogate=surrogate)

