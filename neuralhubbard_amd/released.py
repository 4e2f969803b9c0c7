"""Released states of the reference: weight files, chain files and the paper's energies.

Reference interfaces mirrored here (ChenAo-Phys/NeuralHubbard):

* ``weights/<name>.eqx`` written by ``equinox.tree_serialise_leaves`` (``weights/README.md:6-7``):
  the leaves of the model pytree, one ``np.save`` record after another.  `load_eqx_leaves` reads
  them with ``np.load(allow_pickle=False)``, so nothing in the file is executed.
* ``weights/<name>_spins.npz`` (``code/states.py:328-352``): ``bits = np.packbits(spins > 0,
  axis=-1)`` and ``shape = (nchains, 2N)``; `unpack_spins` / `pack_spins` / `load_spins`.
* the state table of ``code/states.py:99-128`` for the Pfaffian backflow states this library
  evaluates (`STATES`), with the projected E/N and sigma^2/N of the paper.

`released_params` maps the leaves of a ``PfaffianBackflow`` (``code/backflow/models.py:121-151``:
``net`` = quantax ResConv with 16 blocks of 64 channels and a final bias, ``W`` (3, N, 32),
``theta``, ``pad``) onto this library's `Params` of a ``network="resconv"`` workload.
"""

from __future__ import annotations

import os
from dataclasses import dataclass

import numpy as np

from .configs import Params, Workload


@dataclass(frozen=True)
class StateInfo:
    """A released Pfaffian backflow state (``code/states.py:39-92``, rows at ``:105-119``)."""

    name: str
    L: int
    tprime: float
    md5: str
    nparams: int
    E_per_site: float
    E_err: float
    var_per_site: float
    var_err: float
    U: float = 8.0

    @property
    def N(self) -> int:
        return self.L * self.L

    @property
    def n_up(self) -> int:  # 1/8 hole doping, code/states.py:184
        return 7 * self.N // 16

    def workload(self, batch: int = 8192) -> Workload:
        return Workload(self.name, "pfaffian", self.L, self.L, self.n_up, self.n_up, self.tprime, self.U,
                        64, 16, batch, network="resconv", notes=f"released state {self.name}")


STATES = {
    s.name: s
    for s in (
        StateInfo("tp0_L16_pfaffian", 16, 0.0, "17b770fe57367b09b27ad37538842ba0", 1_208_320,
                  -0.75813, 1e-5, 0.0259, 0.0002),
        StateInfo("tpm02_L16_pfaffian", 16, -0.2, "aef48e8b7c10cfa183f8c0e8dcf84d06", 1_208_320,
                  -0.74533, 1e-5, 0.0272, 0.0005),
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


def released_params(leaves: list, LX: int, LY: int) -> Params:
    """`Params` of a released Pfaffian backflow from its leaves: the (C, 4, 1, 1) occupation
    embedding, then 2 x blocks convolutions (C, C, 3, 3) each followed by its (C, 1, 1) bias,
    the integer and boolean fields of the network, W (3, N, d), theta, and the zero pad."""
    arrays = [a for a in leaves if a.dtype.kind == "f"]
    emb = arrays[0]
    if emb.ndim != 4 or emb.shape[1:] != (4, 1, 1):
        raise ValueError(f"expected the (C, 4, 1, 1) embedding first, got {emb.shape}")
    C = emb.shape[0]
    conv_w, conv_b = [], []
    i = 1
    while i + 1 < len(arrays) and arrays[i].shape == (C, C, 3, 3):
        if arrays[i + 1].shape != (C, 1, 1):
            raise ValueError(f"convolution {len(conv_w)} has no (C, 1, 1) bias")
        conv_w.append(np.ascontiguousarray(arrays[i], np.float32))
        conv_b.append(np.ascontiguousarray(arrays[i + 1].reshape(C), np.float32))
        i += 2
    if not conv_w or len(conv_w) % 2:
        raise ValueError(f"expected an even number of convolutions, got {len(conv_w)}")
    W, theta, pad = arrays[i], arrays[i + 1], arrays[i + 2]
    N = LX * LY
    if W.shape != (3, N, C // 2) or theta.shape != () or pad.ndim != 1 or i + 3 != len(arrays):
        raise ValueError(f"unexpected tail of leaves: W {W.shape}, theta {theta.shape}, pad {pad.shape}")
    if np.any(pad != 0):
        raise ValueError("the pad leaf is not zero")
    return Params(conv_w, conv_b, np.ascontiguousarray(W, np.float32), float(theta), None, None,
                  np.ascontiguousarray(emb.reshape(C, 4), np.float32))


def n_params(params: Params, pad: int = 0) -> int:
    """Float parameter count as quantax counts it (``code/states.py:321-324``)."""
    arrays = [params.emb, params.W, *params.conv_w, *params.conv_b]
    return int(sum(np.asarray(a).size for a in arrays) + 1 + pad)


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
