"""The five workloads of the Pfaffian/BCS backflow path, with seeded synthetic parameters.

Each workload is a periodic LX x LY Hubbard lattice (t = 1, t', U), a backflow network
(a 3x3 periodic conv embedding of the 2-channel +-1 occupancy, then ``n_layers`` residual
3x3 conv layers ``h <- h + conv(gelu(h))`` of ``channels`` channels, no final activation),
and a pairing form.  The released states use instead the reference's residual CNN
(``network="resconv"``, ``data/supplement/tableS1_network/network_settings.json:21-56``): an
embedding vector per site occupation (empty, up, down, double), ``n_layers`` blocks
``h <- h + conv(gelu(conv(gelu(rmsnorm(h)))))`` and the output divided by sqrt(n_layers + 1).
Pairing forms:

* ``pfaffian``: F[p, q] = sum_a Y_pa Y_qa Wt[2 s_p + s_q, r(p, q), a] (+ F0[m_p, m_q]) as in
  ``code/backflow/pf_kernel.py:5`` and ``code/backflow/models.py:178-211``;
* ``bcs``: F[p, q] = sum_a Yu_pa Wt[r(p, q), a] Yd_qa + phi0[r(p, q)] as in
  ``code/backflow/bcs_kernel.py:5`` and ``code/backflow/models.py:272-302``, with the
  displacement-indexed base pair function ``phi0`` in the slot of the reference's
  ``delta_d * F^d`` term.

The feature width per spin is d = channels / 2 (``code/backflow/models.py:145, 163-167``).
Parameters follow the reference initialisers where it has them: W ~ N(0, 1) / sqrt(d)
(``models.py:145-148``), theta = 0.55 (``models.py:149``); conv kernels are LeCun normal and
biases N(0, 0.1^2).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .lattice import Geometry, random_configurations


@dataclass(frozen=True)
class Workload:
    name: str
    kind: str  # "pfaffian" or "bcs"
    LX: int
    LY: int
    n_up: int
    n_dn: int
    tprime: float
    U: float
    channels: int
    n_layers: int
    batch: int
    base_f0: bool = False  # config 0: additive antisymmetric base orbital matrix F0
    seed: int = 1234
    network: str = "conv"  # "conv" (above) or "resconv" (the released states' residual CNN)
    features: int = 0  # d of a linear head on the network output (BCS states); 0: d = channels / 2
    t: float = 1.0
    notes: str = ""
    extra: dict = field(default_factory=dict)

    @property
    def geom(self) -> Geometry:
        return Geometry(self.LX, self.LY, self.n_up, self.n_dn)

    @property
    def N(self) -> int:
        return self.LX * self.LY

    @property
    def d(self) -> int:
        return self.features or self.channels // 2

    @property
    def n_neighbor(self) -> int:
        return 2 if self.tprime != 0 else 1

    def conv_shapes(self):
        C = self.channels
        if self.network == "resconv":  # two convs per residual block, embedding by lookup
            return [(C, C)] * (2 * self.n_layers)
        return [(C, 2)] + [(C, C)] * self.n_layers

    def cnn_flops(self) -> float:
        """2 * 9 * N * sum_l C_in C_out, the convention of the reference's cost table."""
        head = 2.0 * self.N * self.channels * 2 * self.features  # linear head, BCS states
        return 2.0 * 9 * self.N * sum(co * ci for co, ci in self.conv_shapes()) + head

    def pair_flops(self) -> float:
        n = self.n_up + self.n_dn
        if self.kind == "bcs":
            return 3.0 * self.n_up * self.n_dn * self.d
        return 3.0 * n * n * self.d

    def linalg_flops(self) -> float:
        """Pf = n^3 / 3, det = 2 n^3 / 3 (``amplitude_cost.csv`` conventions)."""
        if self.kind == "bcs":
            return 2.0 * self.n_up**3 / 3
        n = self.n_up + self.n_dn
        return n**3 / 3.0

    def image_flops(self) -> float:
        return self.cnn_flops() + self.pair_flops() + self.linalg_flops()


WORKLOADS = {
    "pf8": Workload("pf8", "pfaffian", 8, 8, 28, 28, 0.0, 8.0, 16, 4, 1024, base_f0=True,
                    notes="8x8 Pfaffian backflow with base F0, 1024 configurations"),
    "pf16": Workload("pf16", "pfaffian", 16, 16, 112, 112, -0.2, 8.0, 32, 6, 8192,
                     notes="16x16 Pfaffian backflow, t'=-0.2, 8192 configurations"),
    "bcs16": Workload("bcs16", "bcs", 16, 16, 112, 112, 0.0, 8.0, 32, 6, 8192,
                      notes="16x16 BCS backflow with base pair function, 8192 configurations"),
    "pf32": Workload("pf32", "pfaffian", 32, 32, 448, 448, -0.2, 8.0, 32, 6, 1024,
                     notes="32x32 Pfaffian backflow, 1024 configurations"),
    "bcs32": Workload("bcs32", "bcs", 32, 32, 448, 448, 0.0, 8.0, 32, 6, 1024,
                      notes="32x32 BCS backflow with base pair function (448 x 448 determinants)"),
    # the released tpm02_L16_pfaffian state (code/states.py:114-115): 16 residual blocks of 64
    # channels, d = 32, U = 8, t' = -0.2, 112 + 112 electrons
    "tpm02_L16": Workload("tpm02_L16", "pfaffian", 16, 16, 112, 112, -0.2, 8.0, 64, 16, 8192,
                          network="resconv", notes="released 16x16 Pfaffian backflow, t'=-0.2"),
}


def small_workload(kind="pfaffian", LX=4, LY=4, n_up=2, n_dn=2, tprime=-0.2, channels=4,
                   n_layers=2, base_f0=False, seed=7, U=8.0, network="conv", features=0) -> Workload:
    """A lattice small enough for exact enumeration (tests)."""
    return Workload(f"small_{kind}_{LX}x{LY}", kind, LX, LY, n_up, n_dn, tprime, U, channels,
                    n_layers, 16, base_f0=base_f0, seed=seed, network=network, features=features)


@dataclass
class Params:
    """Host (numpy, float32) parameters of one backflow network."""

    conv_w: list  # [(C_out, C_in, 3, 3)]
    conv_b: list  # [(C_out,)]
    W: np.ndarray  # pfaffian (3, N, d); bcs (N, d)
    theta: float
    F0: np.ndarray | None = None  # pfaffian base (2N, 2N), antisymmetric
    phi0: np.ndarray | None = None  # bcs base pair function (N,)
    emb: np.ndarray | None = None  # resconv: (C, 4) vectors of the occupations (empty, up, down, double)
    head: np.ndarray | None = None  # resconv with features > 0: (2, d, C) linear head y_k = head[k] x


def make_params(w: Workload, seed: int | None = None) -> Params:
    rng = np.random.Generator(np.random.PCG64(w.seed if seed is None else seed))
    conv_w, conv_b = [], []
    emb = None
    if w.network == "resconv":
        emb = (rng.standard_normal((w.channels, 4)) / 2.0).astype(np.float32)
    for co, ci in w.conv_shapes():
        conv_w.append((rng.standard_normal((co, ci, 3, 3)) / np.sqrt(9 * ci)).astype(np.float32))
        conv_b.append((0.1 * rng.standard_normal(co)).astype(np.float32))
    N, d = w.N, w.d
    F0 = phi0 = None
    if w.kind == "pfaffian":
        W = (rng.standard_normal((3, N, d)) / np.sqrt(d)).astype(np.float32)
        if w.base_f0:
            A = rng.standard_normal((2 * N, 2 * N)) / np.sqrt(N)
            F0 = np.triu(A, 1)
            F0 = (F0 - F0.T).astype(np.float32)
    else:
        W = (rng.standard_normal((N, d)) / np.sqrt(d)).astype(np.float32)
        phi0 = rng.standard_normal(N).astype(np.float32)
    head = None
    if w.network == "resconv" and w.features:
        head = (rng.standard_normal((2, w.features, w.channels)) / np.sqrt(w.channels)).astype(np.float32)
    return Params(conv_w, conv_b, W, 0.55, F0, phi0, emb, head)


def make_configurations(w: Workload, batch: int | None = None, seed: int | None = None):
    return random_configurations(w.geom, w.batch if batch is None else batch,
                                 (w.seed + 1) if seed is None else seed)
