"""FP64 CPU oracle of the Pfaffian/BCS backflow amplitude and local-energy path.

TEST INFRASTRUCTURE ONLY.  Nothing in ``neuralhubbard_amd`` imports, links or runs this
module: it is imported by ``tests/``, by ``__graft_entry__.smoke()`` and by the
``cpu_baseline`` leg of ``bench.py``, always as the checker of the HIP path.

It restates the reference algorithm (ChenAo-Phys/NeuralHubbard) in float64 numpy; each
function cites the reference lines it follows.  The reference's third-party pieces
(``fermix.slogpf``/``slogdet`` 0.1.0, ``quantax`` 0.4.0 projection, ``Hubbard.Oloc``,
samplers) are not vendored, so those functions restate the published algorithms (Parlett-Reid
Pfaffian of Wimmer, ACM TOMS 38 (2012) 30; Gaussian elimination with partial pivoting;
Metropolis; Christen & Fox (2005) delayed acceptance) and are pinned by the closed forms and
brute-force checks in ``tests/test_oracle.py``.

Pinning status per function (see the tests named in brackets):
  lattice tables            closed form + brute force                [test_oracle_tables]
  build_wtilde / wt_ud      antisymmetry identity + normalisation    [test_oracle_wtilde]
  conv / features           scipy.ndimage.correlate(mode="wrap")     [test_oracle_conv]
  pairing_occ               full 2N x 2N einsum path, sliced         [test_oracle_pairing]
  apply_fermion_ops /       exact expectation over |psi|^2 of the local
  transverse_local          values = <psi|O|psi> with O built from
                            Jordan-Wigner matrices on the full Fock
                            space; hermiticity <S+_a S-_j> =
                            <S-_a S+_j>                              [test_oracle_transverse_*]
  sz_autocorr               FFT form                                 [test_oracle_transverse_*]
  dwave_local               exact |psi|^2 averages = <psi|A+A|psi>,
                            <psi|c+c|psi> from Jordan-Wigner Fock
                            matrices, every left bond, orientation and
                            right term; C_p of the assembled means =
                            <Delta+_0 Delta_r> built directly on the
                            Fock space over all bonds                 [test_oracle_dwave_*]
  pairing_occ_vjp /         finite differences of <G, F(Y, Wt)> and of
  slogpf_vjp /              log|Pf| / log|det| along random directions
  pairing_ud_vjp /          (skew ones for the Pfaffian); the BCS dW by
  slogdet_vjp               a direct loop over displacements          [test_oracle_vjp_*]
  slogpf                    Pf^2 = det, exact rational expansion,
                            Pf(B A B^T) = det(B) Pf(A), block form    [test_oracle_pfaffian]
  slogdet                   numpy (LAPACK) slogdet                   [test_oracle_det]
  projection                U_g psi_P = psi_P with U_g built on bit
                            strings; explicit inversion-count signs   [test_oracle_projection]
  eloc                      sparse JW Hamiltonian built on bit
                            strings over the full Hilbert space       [test_oracle_eloc]
  philox4x32                published known-answer vectors           [test_oracle_philox]
  jackknife / energy        closed forms for i.i.d. blocks           [test_oracle_jackknife]
                            the reference's published Table S4 values
                            from its own energy series (17 states)   [test_energy_series]
  resconv_features          translation equivariance, embedding order
  (+ rmsnorm, conv batch)   with zero convs, RMSNorm closed form, the
                            scipy-pinned conv3x3; the composite by the
                            released states' E/N against the paper
                            (a swapped embedding fails it); the BCS
                            head by tp0_L16_bcs's E/N                [test_released, test_gpu_released]
  sur_forward / sur_loss    the same network in torch conv2d (circular
                            padding), translation invariance, f = N b3
                            for zero weights                         [test_oracle_surrogate_forward]
  sur_grad_fd               torch autograd                           [test_oracle_surrogate_gradient]
  philox_draw /             chi-square stationarity of |psi|^2 over an
  metropolis_chain          enumerated 3x3 space, plain and delayed;
                            a |psi|-targeting chain fails it          [test_oracle_metropolis_*]
  propose                   the exact proposal matrix over enumerated
                            3x3 / 4x3 spaces is symmetric; the moves
                            are the neighbour hops / site swaps drawn
                            by brute force; site exchange alone keeps
                            the doublon count, the mix connects the
                            space; chi-square of the mixed chain     [test_oracle_proposals_*]
"""

from __future__ import annotations

import math

import numpy as np

# ------------------------------------------------------------------------ lattice tables


def lattice_tables(LX: int, LY: int):
    """(disp (N,N), neg (N,), dist_r (N,)) as ``displacement_tables``
    (``code/backflow/pf_kernel.py:83-96``), dist in float64."""
    N = LX * LY
    x, y = np.divmod(np.arange(N), LY)
    disp = ((x[:, None] - x[None, :]) % LX) * LY + (y[:, None] - y[None, :]) % LY
    neg = ((-x) % LX) * LY + (-y) % LY
    mx, my = np.minimum(x, LX - x), np.minimum(y, LY - y)
    return disp, neg, np.sqrt(mx**2 + my**2.0)


def group_perm(g: int, LX: int, LY: int) -> np.ndarray:
    """Mode map of group element g = fx | fy << 1 | si << 2 (Flip(0) @ Flip(1) @ SpinInverse,
    ``code/states.py:274-279``)."""
    N = LX * LY
    m = np.arange(2 * N)
    spin, site = np.divmod(m, N)
    x, y = np.divmod(site, LY)
    if g & 1:
        x = (-x) % LX
    if g & 2:
        y = (-y) % LY
    return ((spin ^ ((g >> 2) & 1)) * N + x * LY + y).astype(np.int64)


# ------------------------------------------------------------------------ network


def gelu(x):
    """tanh form of GELU (the default of ``jax.nn.gelu``)."""
    return 0.5 * x * (1.0 + np.tanh(math.sqrt(2.0 / math.pi) * (x + 0.044715 * x**3)))


def conv3x3(h, w, b):
    """Periodic 3x3 cross-correlation: out[o,x,y] = b[o] + sum w[o,c,1+dx,1+dy] h[c,x+dx,y+dy]."""
    out = np.zeros((w.shape[0],) + h.shape[1:])
    for dx in (-1, 0, 1):
        for dy in (-1, 0, 1):
            sh = np.roll(h, shift=(-dx, -dy), axis=(1, 2))
            out += np.einsum("oc,cxy->oxy", w[:, :, dx + 1, dy + 1].astype(np.float64), sh)
    return out + np.asarray(b, np.float64)[:, None, None]


def features(s, conv_w, conv_b, LX, LY):
    """(2, N, d) per-spin features of one configuration: embedding conv, residual layers
    ``h + conv(gelu(h))``, channels split as y_up | y_dn (``code/backflow/models.py:163-167``)."""
    x = np.asarray(s, np.float64).reshape(2, LX, LY)
    h = conv3x3(x, conv_w[0], conv_b[0])
    for w, b in zip(conv_w[1:], conv_b[1:]):
        h = h + conv3x3(gelu(h), w, b)
    C = h.shape[0]
    return np.swapaxes(h.reshape(2, C // 2, LX * LY), 1, 2)


RMS_EPS = 1e-5  # eqx.nn.RMSNorm default, added to the mean square


def occupation_class(s, N):
    """Site occupation 0 empty, 1 up, 2 down, 3 double (the embedding order of
    ``data/supplement/tableS1_network/network_settings.json:50``); s (..., 2N)."""
    s = np.asarray(s)
    return (s[..., :N] > 0).astype(np.int64) + 2 * (s[..., N:] > 0).astype(np.int64)


def rmsnorm(h):
    """h / sqrt(mean(h^2) + eps) over all channels and sites of each image, no affine
    parameters (``network_settings.json:53``); h (B, C, LX, LY)."""
    return h / np.sqrt(np.mean(h * h, axis=(1, 2, 3), keepdims=True) + RMS_EPS)


def conv3x3_batch(h, w, b):
    """`conv3x3` over a batch: h (B, C, LX, LY)."""
    w = np.asarray(w, np.float64)
    out = np.zeros((h.shape[0], w.shape[0]) + h.shape[2:])
    for dx in (-1, 0, 1):
        for dy in (-1, 0, 1):
            sh = np.roll(h, shift=(-dx, -dy), axis=(2, 3))
            out += np.einsum("oc,bcxy->boxy", w[:, :, dx + 1, dy + 1], sh, optimize=True)
    return out + np.asarray(b, np.float64)[None, :, None, None]


def resconv_features(S, emb, conv_w, conv_b, LX, LY, head=None):
    """(B, 2, N, d) features of the released states' residual CNN (quantax ResConv as built at
    ``code/backflow/models.py:319-329``, architecture ``network_settings.json:21-56``): the
    embedding vector of each site's occupation, blocks h + conv(gelu(conv(gelu(rmsnorm(h))))),
    output / sqrt(blocks + 1), channels split y_up | y_dn (``models.py:163-167``), or with a
    (2, d, C) `head` the BCS features y_k[n] = head[k] x[:, n] (``models.py:264-270``).
    S (B, 2N) or (2N,); emb (C, 4); conv_w / conv_b: the 2 * blocks convs."""
    S = np.atleast_2d(np.asarray(S))
    B, N = S.shape[0], LX * LY
    emb = np.asarray(emb, np.float64).reshape(-1, 4)
    C = emb.shape[0]
    h = np.transpose(emb[:, occupation_class(S, N)], (1, 0, 2)).reshape(B, C, LX, LY)
    nb = len(conv_w) // 2
    for k in range(nb):
        a = conv3x3_batch(gelu(rmsnorm(h)), conv_w[2 * k], conv_b[2 * k])
        h = h + conv3x3_batch(gelu(a), conv_w[2 * k + 1], conv_b[2 * k + 1])
    h = h / math.sqrt(nb + 1)
    if head is not None:  # BCS: y_k = head[k] x (code/backflow/models.py:264-270)
        return np.einsum("kdc,bcn->bknd", np.asarray(head, np.float64), h.reshape(B, C, N))
    return np.swapaxes(h.reshape(B, 2, C // 2, N), 2, 3)


def softplus(t):
    return math.log1p(math.exp(-abs(t))) + max(t, 0.0)


def build_wtilde(W, theta, LX, LY):
    """(3,N,d) -> (4,N,d) [uu, ud, du, dd] with the distance envelope
    (``code/backflow/pf_kernel.py:107-124``)."""
    _, neg, dist = lattice_tables(LX, LY)
    W = np.asarray(W, np.float64)
    N = LX * LY
    W_ud = W[0]
    W_uu = (W[1] - W[1][neg]) / 2
    W_dd = (W[2] - W[2][neg]) / 2
    W_du = -W_ud[neg]
    pref = np.exp(-softplus(theta) * dist)
    pref = pref / (N * pref.sum())
    return np.stack([W_uu, W_ud, W_du, W_dd]) * pref[None, :, None]


def build_wt_ud(W, theta, LX, LY):
    """(N,d) -> W * pref (``code/backflow/bcs_kernel.py:63-70``)."""
    _, _, dist = lattice_tables(LX, LY)
    pref = np.exp(-softplus(theta) * dist)
    return np.asarray(W, np.float64) * (pref / (LX * LY * pref.sum()))[:, None]


# ------------------------------------------------------------------------ pairing


def pairing_occ(Y, modes, Wt, LX, LY):
    """F[p,q] = sum_a Y_pa Y_qa Wt[2 s_p + s_q, r(p,q), a] over the listed modes
    (``code/backflow/pf_kernel.py:371-382``); Y (2N, d) features by mode."""
    disp, _, _ = lattice_tables(LX, LY)
    N = LX * LY
    m = np.asarray(modes)
    sp, site = np.divmod(m, N)
    r = disp[site[:, None], site[None, :]]
    blk = 2 * sp[:, None] + sp[None, :]
    Yo = np.asarray(Y, np.float64)[m]
    # (Y_p * Y_q) first, then the table: commutative, so F is exactly antisymmetric
    # (the order of ``code/backflow/pf_kernel.py:164-165``)
    return ((Yo[:, None, :] * Yo[None, :, :]) * Wt[blk, r]).sum(-1)


def pairing_occ_vjp(Y, modes, Wt, G, LX, LY):
    """Cotangents (dY (n, d), dWt (4, N, d)) of F = `pairing_occ` for the cotangent G (n, n),
    Y (n, d) the features of the listed modes (the reference's kernel-level argument), every
    entry of F (diagonal included) a function of Y and Wt: the VJP that
    ``code/backflow/pf_kernel.py:354-365`` computes with ``_dy_kernel`` and ``_dw_kernel``.
    dY[p] = sum_q G[p, q] Wt[b(p, q), r(p, q)] Y[q] + G[q, p] Wt[b(q, p), r(q, p)] Y[q];
    dWt[b, r] = sum over pairs (p, q) with b(p, q) = b, r(p, q) = r of G[p, q] Y[p] Y[q]."""
    disp, _, _ = lattice_tables(LX, LY)
    N = LX * LY
    m = np.asarray(modes)
    sp, site = np.divmod(m, N)
    r = disp[site[:, None], site[None, :]]
    blk = 2 * sp[:, None] + sp[None, :]
    Yo = np.asarray(Y, np.float64)
    G = np.asarray(G, np.float64)
    W = np.asarray(Wt, np.float64)[blk, r]  # (n, n, d): Wt of the pair (p, q)
    dY = np.einsum("pq,pqa,qa->pa", G, W, Yo) + np.einsum("qp,qpa,qa->pa", G, W, Yo)
    dWt = np.zeros((4, N, Yo.shape[1]))
    np.add.at(dWt, (blk.ravel(), r.ravel()), (G[:, :, None] * Yo[:, None, :] * Yo[None, :, :]).reshape(-1, Yo.shape[1]))
    return dY, dWt


def slogpf_vjp(A):
    """d log|Pf A| / dA with every entry independent: (1/2) A^{-T} (from d Pf = (1/2) Pf
    tr(A^{-1} dA)), the Pfaffian VJP behind ``fermix.slogpf`` at ``code/backflow/models.py:210``."""
    return 0.5 * np.linalg.inv(np.asarray(A, np.float64)).T


def pairing_ud_vjp(Yu, Yd, sites_u, sites_d, Wt, G, LX, LY):
    """Cotangents (dYu (n_up, d), dYd (n_dn, d), dWt (N, d)) of F = `pairing_ud` for the cotangent
    G (n_up, n_dn): the VJP that ``code/backflow/bcs_kernel.py:274-288`` computes with its
    ``_dy_kernel`` (both sides) and ``_dw_kernel``.
    dYu[p] = sum_q G[p, q] Wt[r(p, q)] Yd[q];  dYd[q] = sum_p G[p, q] Wt[r(p, q)] Yu[p];
    dWt[r] = sum over pairs (p, q) with r(p, q) = r of G[p, q] Yu[p] Yd[q]."""
    disp, _, _ = lattice_tables(LX, LY)
    N = LX * LY
    r = disp[np.asarray(sites_u)[:, None], np.asarray(sites_d)[None, :]]
    Yu = np.asarray(Yu, np.float64)
    Yd = np.asarray(Yd, np.float64)
    G = np.asarray(G, np.float64)
    W = np.asarray(Wt, np.float64)[r]  # (n_up, n_dn, d)
    dYu = np.einsum("pq,pqa,qa->pa", G, W, Yd)
    dYd = np.einsum("pq,pqa,pa->qa", G, W, Yu)
    dWt = np.zeros((N, Yu.shape[1]))
    np.add.at(dWt, r.ravel(), (G[:, :, None] * Yu[:, None, :] * Yd[None, :, :]).reshape(-1, Yu.shape[1]))
    return dYu, dYd, dWt


def slogdet_vjp(A):
    """d log|det A| / dA = A^{-T} (Jacobi's formula), the VJP behind ``fermix.slogdet`` at
    ``code/backflow/models.py:297``."""
    return np.linalg.inv(np.asarray(A, np.float64)).T


def pairing_full(y, W, theta, LX, LY):
    """The reference's einsum path: full (2N, 2N) f, then (f - f^T) / 2
    (``code/backflow/models.py:169-196``)."""
    disp, neg, dist = lattice_tables(LX, LY)
    W = np.asarray(W, np.float64)
    y = np.asarray(y, np.float64)
    W_ud = W[0]
    W_uu = (W[1] - W[1][neg]) / 2
    W_dd = (W[2] - W[2][neg]) / 2
    f_ud = np.einsum("ia,ija,ja->ij", y[0], W_ud[disp], y[1])
    f_uu = np.einsum("ia,ija,ja->ij", y[0], W_uu[disp], y[0])
    f_dd = np.einsum("ia,ija,ja->ij", y[1], W_dd[disp], y[1])
    full_dist = dist[disp]
    pref = np.exp(-softplus(theta) * full_dist)
    pref = pref / pref.sum()
    f = np.block([[f_uu, f_ud], [-f_ud.T, f_dd]]) * np.tile(pref, (2, 2))
    return (f - f.T) / 2


def pairing_ud(Yu, Yd, sites_u, sites_d, Wt, LX, LY):
    """F[p,q] = sum_a Yu_pa Wt[r(p,q), a] Yd_qa (``code/backflow/bcs_kernel.py:296-300``)."""
    disp, _, _ = lattice_tables(LX, LY)
    r = disp[np.asarray(sites_u)[:, None], np.asarray(sites_d)[None, :]]
    return np.einsum("pa,pqa,qa->pq", np.asarray(Yu, np.float64), Wt[r], np.asarray(Yd, np.float64))


# ------------------------------------------------------------------------ linear algebra


def slogpf(A):
    """(sign, log|Pf A|) of a real antisymmetric matrix by Parlett-Reid elimination with
    partial pivoting (Wimmer 2012, algorithm ``pfaffian_LTL``), the role of
    ``fermix.slogpf`` at ``code/backflow/models.py:210``.  sign 0, logabs -inf if singular."""
    A = np.array(A, dtype=np.float64)
    n = A.shape[0]
    if n % 2:
        return 0.0, -math.inf
    sign, logabs = 1.0, 0.0
    for k in range(0, n - 1, 2):
        kp = k + 1 + int(np.argmax(np.abs(A[k + 1 :, k])))
        if kp != k + 1:
            A[[k + 1, kp], :] = A[[kp, k + 1], :]
            A[:, [k + 1, kp]] = A[:, [kp, k + 1]]
            sign = -sign
        piv = A[k, k + 1]
        if piv == 0.0:
            return 0.0, -math.inf
        sign *= math.copysign(1.0, piv)
        logabs += math.log(abs(piv))
        if k + 2 < n:
            tau = A[k, k + 2 :] / piv
            v = A[k + 2 :, k + 1].copy()
            A[k + 2 :, k + 2 :] += np.outer(tau, v) - np.outer(v, tau)
    return sign, logabs


def slogdet(A):
    """(sign, log|det A|) by Gaussian elimination with partial pivoting, the role of
    ``fermix.slogdet`` at ``code/backflow/models.py:297``."""
    A = np.array(A, dtype=np.float64)
    n = A.shape[0]
    sign, logabs = 1.0, 0.0
    for k in range(n):
        p = k + int(np.argmax(np.abs(A[k:, k])))
        if A[p, k] == 0.0:
            return 0.0, -math.inf
        if p != k:
            A[[k, p]] = A[[p, k]]
            sign = -sign
        piv = A[k, k]
        sign *= math.copysign(1.0, piv)
        logabs += math.log(abs(piv))
        A[k + 1 :, k + 1 :] -= np.outer(A[k + 1 :, k] / piv, A[k, k + 1 :])
    return sign, logabs


def signed_logsumexp(signs, logs):
    """LogArray sum: (sign, log|sum_i sign_i exp(log_i)|)."""
    logs = np.asarray(logs, np.float64)
    signs = np.asarray(signs, np.float64)
    ok = signs != 0
    if not ok.any():
        return 0.0, -math.inf
    m = logs[ok].max()
    tot = float(np.sum(signs[ok] * np.exp(logs[ok] - m)))
    if tot == 0.0:
        return 0.0, -math.inf
    return math.copysign(1.0, tot), m + math.log(abs(tot))


# ------------------------------------------------------------------------ amplitudes


class Model:
    """FP64 restatement of ``PfaffianBackflow`` / ``BCSBackflow`` (``code/backflow/models.py``)
    with the workload's synthetic network."""

    def __init__(self, workload, params):
        self.w = workload
        self.p = params
        LX, LY = workload.LX, workload.LY
        if workload.kind == "pfaffian":
            self.Wt = build_wtilde(params.W, params.theta, LX, LY)
        else:
            self.Wt = build_wt_ud(params.W, params.theta, LX, LY)

    def features(self, s):
        w = self.w
        if getattr(w, "network", "conv") == "resconv":
            return resconv_features(s, self.p.emb, self.p.conv_w, self.p.conv_b, w.LX, w.LY, self.p.head)[0]
        return features(s, self.p.conv_w, self.p.conv_b, w.LX, w.LY)

    def features_batch(self, S):
        """(B, 2, N, d) features of a batch of configurations."""
        w = self.w
        if getattr(w, "network", "conv") == "resconv":
            return resconv_features(S, self.p.emb, self.p.conv_w, self.p.conv_b, w.LX, w.LY, self.p.head)
        return np.stack([self.features(s) for s in np.asarray(S)])

    def image_matrix(self, s, g):
        """Matrix whose Pfaffian (determinant for BCS, times the block sign) is
        eta_g(s) * psi(g.s): the occupied modes of s mapped by g, kept in the ascending
        order of s (the permutation sign of the reordering is then inside the Pfaffian)."""
        w = self.w
        LX, LY, N = w.LX, w.LY, w.N
        perm = group_perm(g, LX, LY)
        gs = np.empty_like(np.asarray(s))
        gs[perm] = s
        y = self.features(gs)  # (2, N, d)
        idx = np.flatnonzero(np.asarray(s) > 0)
        modes = perm[idx]
        if w.kind == "pfaffian":
            F = pairing_occ(y.reshape(2 * N, -1), modes, self.Wt, LX, LY)
            if self.p.F0 is not None:
                F = F + np.asarray(self.p.F0, np.float64)[np.ix_(modes, modes)]
            return F
        up = modes[modes < N]
        dn = modes[modes >= N] - N
        disp, _, _ = lattice_tables(LX, LY)
        F = pairing_ud(y[0][up], y[1][dn], up, dn, self.Wt, LX, LY)
        F = F + np.asarray(self.p.phi0, np.float64)[disp[np.ix_(up, dn)]]
        return F

    def image_amplitude(self, s, g):
        w = self.w
        F = self.image_matrix(s, g)
        if w.kind == "pfaffian":
            sg, la = slogpf(F)
            n = F.shape[0]
        else:
            sg, la = slogdet(F)
            n = F.shape[0]
            # Pf [[0, Phi], [-Phi^T, 0]] = (-1)^{n(n-1)/2} det Phi (code/backflow/models.py:300);
            # an image whose first block is spin down carries det(-Phi^T) = (-1)^n det Phi
            sg *= (-1) ** (n * (n - 1) // 2)
            if (g >> 2) & 1:
                sg *= (-1) ** n
        # fermion_inverse_sign: reversal of the Ne creation operators, (-1)^{Ne(Ne-1)/2}
        ne = w.n_up + w.n_dn
        sg *= (-1) ** (ne * (ne - 1) // 2)
        return sg, la

    def psi(self, s, project=True):
        """Projected amplitude sum_g eta_g(s) psi(g.s) (trivial characters), as (sign, log)."""
        if not project:
            return self.image_amplitude(s, 0)
        vals = [self.image_amplitude(s, g) for g in range(8)]
        return signed_logsumexp([v[0] for v in vals], [v[1] for v in vals])

    def psi_batch(self, S, project=True):
        out = np.array([self.psi(s, project) for s in np.asarray(S)])
        return out[:, 0], out[:, 1]


# ------------------------------------------------------------------------ Hubbard local energy


def neighbours(LX, LY, n_neighbor):
    """Distinct bond partners per site and their shell (0: NN, 1: NNN)."""
    shells = [[(1, 0), (-1, 0), (0, 1), (0, -1)]]
    if n_neighbor >= 2:
        shells.append([(1, 1), (1, -1), (-1, 1), (-1, -1)])
    out = []
    for i in range(LX * LY):
        x, y = divmod(i, LY)
        seen, row = set(), []
        for c, shell in enumerate(shells):
            for dx, dy in shell:
                j = ((x + dx) % LX) * LY + (y + dy) % LY
                if j != i and j not in seen:
                    seen.add(j)
                    row.append((j, c))
        out.append(row)
    return out


def hops(s, LX, LY, t, tprime):
    """Connected configurations of the Hubbard hopping term
    H = -sum_{<ij>,sigma} t_ij (c+_i c_j + h.c.) (``code/states.py:244-247``):
    list of (mode_from, mode_to, <s'|H|s>) in enumeration order (electron by ascending mode,
    partner by table order).  The Jordan-Wigner sign is (-1)^(occupied modes strictly
    between the two modes) in spin-major mode order."""
    s = np.asarray(s)
    N = LX * LY
    nb = neighbours(LX, LY, 2 if tprime != 0 else 1)
    tv = (t, tprime)
    occ = s > 0
    out = []
    for m in np.flatnonzero(occ):
        sp, i = divmod(int(m), N)
        for j, c in nb[i]:
            mt = sp * N + j
            if occ[mt]:
                continue
            lo, hi = min(m, mt), max(m, mt)
            between = int(occ[lo + 1 : hi].sum())
            out.append((int(m), mt, -tv[c] * (-1.0) ** between))
    return out


def diag_energy(s, U):
    """U * sum_i n_up,i n_dn,i."""
    s = np.asarray(s)
    N = s.shape[-1] // 2
    return U * float(np.sum((s[:N] > 0) & (s[N:] > 0)))


def eloc(model, s, U, t, tprime, project=True, psi_s=None):
    """E_loc(s) = U D(s) + sum_{s'} <s|H|s'> psi(s') / psi(s) (``Hubbard.Oloc``,
    ``code/quick_start.py:33``), every psi(s') by a full evaluation."""
    w = model.w
    if psi_s is None:
        psi_s = model.psi(s, project)
    e = diag_energy(s, U)
    for a, b, h in hops(s, w.LX, w.LY, t, tprime):
        sp = np.array(s, copy=True)
        sp[a], sp[b] = -1, 1
        sg, la = model.psi(sp, project)
        e += h * sg * psi_s[0] * math.exp(la - psi_s[1])
    return e


# ------------------------------------------------------------------------ spin correlations


def apply_fermion_ops(occ, ops):
    """Apply c / c+ operators, listed left to right as written and acting right to left, to the
    basis state c+_{m1} ... c+_{mk}|0> of the ascending modes `occ`: each operator contributes
    (-1)^(occupied modes below its mode).  Returns (sign, occ') or None for a zero result."""
    occ = sorted(occ)
    sign = 1
    for m, create in reversed(ops):
        k = sum(1 for x in occ if x < m)
        if create == (m in occ):
            return None
        sign *= (-1) ** k
        if create:
            occ.insert(k, m)
        else:
            occ.remove(m)
    return sign, tuple(occ)


def spin_flip(i, k, N):
    """S+_i S-_k = c+_{i up} c_{i dn} c+_{k dn} c_{k up} as an operator list."""
    return [(i, True), (N + i, False), (N + k, True), (k, False)]


def transverse_local(s, a, LX, LY, psi_fn):
    """(2, N) local values of S+_a S-_{a+r} (row 0) and S-_a S+_{a+r} (row 1): sum over s' of
    <s|O|s'> psi(s') / psi(s), <s|O|s'> = <s'|O^dagger|s> from operator algebra, r = 0 the
    diagonal n_up (1 - n_dn), n_dn (1 - n_up) of site a (``code/measure/transverse_spin.py``).
    psi_fn(s) -> (sign, log|psi|)."""
    N = LX * LY
    s = np.asarray(s)
    occ = tuple(int(m) for m in np.flatnonzero(s > 0))
    ps = psi_fn(s)
    ax, ay = divmod(a, LY)
    T = np.zeros((2, N))
    T[0, 0] = float(s[a] > 0 and s[N + a] < 0)
    T[1, 0] = float(s[N + a] > 0 and s[a] < 0)
    for j in range(N):
        if j == a:
            continue
        r = ((j // LY - ax) % LX) * LY + (j % LY - ay) % LY
        # O^dagger of S+_a S-_j is S+_j S-_a; that of S-_a S+_j is S+_a S-_j
        for chan, ops in ((0, spin_flip(j, a, N)), (1, spin_flip(a, j, N))):
            res = apply_fermion_ops(occ, ops)
            if res is None:
                continue
            sign, occ2 = res
            s2 = -np.ones(2 * N, np.int8)
            s2[list(occ2)] = 1
            sg, la = psi_fn(s2)
            T[chan, r] += sign * sg * ps[0] * math.exp(la - ps[1])
    return T


def _bond_ends(x, y, ax, LX, LY):
    """Site indices (lower, upper) of the bond keyed (x, y, axis) (``observables.py:93-96``)."""
    hi = ((x + 1) % LX, y) if ax == 0 else (x, (y + 1) % LY)
    return x * LY + y, hi[0] * LY + hi[1]


def dwave_local(s, LX, LY, psi_fn):
    """Local values (F (4, 4N), G (3, N, 2)) of the d-wave terms at R = (0, 0)
    (``code/measure/dwave_packed.py``): F[2 l + oL, 2 (2 site + axis) + oR] of
    A+(bl, oL) A(br, oR), A+(bond, o) = c+_{a up} c+_{b dn} with (a, b) = (lower, upper) for o = 0
    and swapped for o = 1, for the left bonds bl = (0, 0, 0), (0, 0, 1); G[li, j, s] of
    c+_{i s} c_{j s} for the left sites i = (0, 0), (0, 1), (1, 0).  Each local value is
    sum_{s'} <s'|O^dagger|s> psi(s') / psi(s); the diagonal terms are the occupations."""
    N = LX * LY
    s = np.asarray(s)
    occ = tuple(int(m) for m in np.flatnonzero(s > 0))
    ps = psi_fn(s)
    F = np.zeros((4, 4 * N))
    G = np.zeros((3, N, 2))

    def amp_ratio(ops):
        res = apply_fermion_ops(occ, ops)
        if res is None:
            return 0.0
        sign, occ2 = res
        s2 = -np.ones(2 * N, np.int8)
        s2[list(occ2)] = 1
        sg, la = psi_fn(s2)
        return sign * sg * ps[0] * math.exp(la - ps[1])

    for l, ax in enumerate((0, 1)):
        lo, hi = _bond_ends(0, 0, ax, LX, LY)
        for oL in (0, 1):
            a, b = (lo, hi) if oL == 0 else (hi, lo)
            for site in range(N):
                for bax in (0, 1):
                    clo, chi = _bond_ends(site // LY, site % LY, bax, LX, LY)
                    for oR in (0, 1):
                        c, d = (clo, chi) if oR == 0 else (chi, clo)
                        t = 2 * (2 * site + bax) + oR
                        if (c, d) == (a, b):
                            F[2 * l + oL, t] = float(s[a] > 0 and s[N + b] > 0)
                        else:  # O^dagger = c+_{c up} c+_{d dn} c_{b dn} c_{a up}
                            F[2 * l + oL, t] += amp_ratio([(c, True), (N + d, True), (N + b, False), (a, False)])
    for li, i in enumerate((0, 1, LY)):
        for sp in (0, 1):
            for j in range(N):
                if j == i:
                    G[li, j, sp] = float(s[sp * N + i] > 0)
                else:  # O = c+_i c_j, O^dagger = c+_j c_i
                    G[li, j, sp] = amp_ratio([(sp * N + j, True), (sp * N + i, False)])
    return F, G


def sz_autocorr(S, LX, LY):
    """Batch mean of (1/N) sum_i Sz_i Sz_{i+r}, Sz = (n_up - n_dn) / 2, by direct sums
    (``code/measure/observables.py:179-206``)."""
    S = np.atleast_2d(np.asarray(S, np.float64))
    N = LX * LY
    sz = ((S[:, :N] - S[:, N:]) / 4).reshape(-1, LX, LY)
    out = np.zeros((LX, LY))
    for rx in range(LX):
        for ry in range(LY):
            out[rx, ry] = np.mean(np.sum(sz * np.roll(sz, (-rx, -ry), axis=(1, 2)), axis=(1, 2))) / N
    return out.reshape(-1)


# ------------------------------------------------------------------------ statistics


BLOCK_SIZES = (1, 2, 4, 8)
MIN_BLOCKS = 8


def jackknife(f, *series):
    """Block jackknife (``code/measure/assemble.py:58-82``)."""
    x = [np.asarray(s, dtype=np.float64) for s in series]
    value = f(*[s.mean(axis=0) for s in x])
    errors = {}
    for b in BLOCK_SIZES:
        nb = x[0].shape[0] // b
        if nb < MIN_BLOCKS:
            continue
        xb = [s[: nb * b].reshape(nb, b, *s.shape[1:]).mean(axis=1) for s in x]
        tot = [s.sum(axis=0) for s in xb]
        reps = np.array([f(*[(tt - s[i]) / (nb - 1) for tt, s in zip(tot, xb)]) for i in range(nb)])
        errors[b] = np.sqrt(nb - 1) * reps.std(axis=0)
    error = np.max(np.stack(list(errors.values())), axis=0) if errors else np.nan
    return value, error, errors


# ------------------------------------------------------------------------ counter RNG


PHILOX_M0, PHILOX_M1 = 0xD2511F53, 0xCD9E8D57
PHILOX_W0, PHILOX_W1 = 0x9E3779B9, 0xBB67AE85


def philox4x32(ctr, key, rounds=10):
    """Philox4x32-10 (Salmon et al., SC'11) on python ints: 4 x u32 counter, 2 x u32 key."""
    c = [int(v) & 0xFFFFFFFF for v in ctr]
    k = [int(v) & 0xFFFFFFFF for v in key]
    for _ in range(rounds):
        p0 = PHILOX_M0 * c[0]
        p1 = PHILOX_M1 * c[2]
        c = [
            ((p1 >> 32) ^ c[1] ^ k[0]) & 0xFFFFFFFF,
            p1 & 0xFFFFFFFF,
            ((p0 >> 32) ^ c[3] ^ k[1]) & 0xFFFFFFFF,
            p0 & 0xFFFFFFFF,
        ]
        k = [(k[0] + PHILOX_W0) & 0xFFFFFFFF, (k[1] + PHILOX_W1) & 0xFFFFFFFF]
    return c


# ------------------------------------------------------------------------ surrogate network


def sur_unpack(p, C):
    p = np.asarray(p, np.float64)
    o = 0
    out = []
    for size in (9 * 2 * C, C, 9 * C * C, C, 9 * C, 1):
        out.append(p[o:o + size])
        o += size
    w1, b1, w2, b2, w3, b3 = out
    # [tap][ci][co] -> (co, ci, 3, 3) for conv3x3
    W1 = w1.reshape(3, 3, 2, C).transpose(3, 2, 0, 1)
    W2 = w2.reshape(3, 3, C, C).transpose(3, 2, 0, 1)
    W3 = w3.reshape(3, 3, C, 1).transpose(3, 2, 0, 1)
    return W1, b1, W2, b2, W3, b3


def sur_forward(p, s, LX, LY, C):
    """f(s) = sum_sites conv3(gelu(conv2(gelu(conv1(x0))))) (layout of include/nh.h)."""
    W1, b1, W2, b2, W3, b3 = sur_unpack(p, C)
    x0 = np.asarray(s, np.float64).reshape(2, LX, LY)
    z1 = conv3x3(x0, W1, b1)
    z2 = conv3x3(gelu(z1), W2, b2)
    z3 = conv3x3(gelu(z2), W3, b3)
    return float(z3.sum())


def sur_loss(p, S, T, LX, LY, C, weight):
    return sum(0.5 * weight * (sur_forward(p, s, LX, LY, C) - t) ** 2 for s, t in zip(S, T))


def sur_grad_fd(p, S, T, LX, LY, C, weight, idx, h=1e-6):
    """Central finite differences of the loss at parameter indices idx."""
    p = np.asarray(p, np.float64)
    out = []
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        out.append((sur_loss(p + e, S, T, LX, LY, C, weight) - sur_loss(p - e, S, T, LX, LY, C, weight)) / (2 * h))
    return np.array(out)


# ------------------------------------------------------------------------ Metropolis chains


def philox_draw(chain, step, seed):
    return philox4x32([chain, step & 0xFFFFFFFF, (step >> 32) & 0xFFFFFFFF, 0], [seed & 0xFFFFFFFF, seed >> 32])


MOVE_EXCHANGE, MOVE_PARTICLE_HOP, MOVE_SITE_EXCHANGE, MOVE_MIX = 0, 1, 2, 3
# neighbour directions: the four nearest, then the four next-nearest (the E_loc hop order)
NEIGHBOUR_DIRS = [(1, 0), (-1, 0), (0, 1), (0, -1), (1, 1), (1, -1), (-1, 1), (-1, -1)]


def _neighbour(i, d, LX, LY):
    dx, dy = NEIGHBOUR_DIRS[d]
    x, y = divmod(i, LY)
    return ((x + dx) % LX) * LY + (y + dy) % LY


def propose(s, r0, r1, LX, LY, kind=MOVE_EXCHANGE, ph_z=4, se_z=4):
    """The configuration a proposal makes from s, or None for no change.  r0, r1 are the first
    two Philox words.  Moves of quantax's samplers behind the reference's MixSampler
    (code/states.py:397-408), as this build draws them:
      exchange:      site r0 mod N, direction r1 & 3, spin r1 >> 2 & 1: modes (spin, i), (spin, j)
                     trade occupations;
      particle hop:  the (r0 mod n_occ)-th occupied mode (ascending) moves to neighbour
                     r1 & (ph_z - 1) of its site in its spin block, unless that mode is occupied;
      site exchange: site r0 mod N and neighbour r1 & (se_z - 1) trade (up, down) occupations;
      mix:           r1 >> 4 & 1 picks particle hop (0) or site exchange (1)."""
    N = LX * LY
    if kind == MOVE_MIX:
        kind = MOVE_SITE_EXCHANGE if (r1 >> 4) & 1 else MOVE_PARTICLE_HOP
    sp = np.array(s, dtype=np.int8, copy=True)
    if kind == MOVE_EXCHANGE:
        i = r0 % N
        spin = (r1 >> 2) & 1
        mi, mj = spin * N + i, spin * N + _neighbour(i, r1 & 3, LX, LY)
        if s[mi] == s[mj]:
            return None
        sp[mi], sp[mj] = s[mj], s[mi]
        return sp
    if kind == MOVE_SITE_EXCHANGE:
        i = r0 % N
        j = _neighbour(i, r1 & (se_z - 1), LX, LY)
        if s[i] == s[j] and s[N + i] == s[N + j]:
            return None
        for off in (0, N):
            sp[off + i], sp[off + j] = s[off + j], s[off + i]
        return sp
    occ = np.flatnonzero(np.asarray(s) > 0)
    n_occ = len(occ)
    if n_occ == 0:
        return None
    m = int(occ[r0 % n_occ])
    off = N if m >= N else 0
    t = off + _neighbour(m - off, r1 & (ph_z - 1), LX, LY)
    if s[t] > 0:
        return None
    sp[m], sp[t] = -1, 1
    return sp


def metropolis_chain(psi, s0, chain, seed, step0, n_steps, LX, LY, sur=None, margin=1e-6, kind=MOVE_EXCHANGE,
                     ph_z=4, se_z=4):
    """Replays chain `chain` of the GPU sweep: proposals from Philox4x32-10 at counter
    (chain, step) by `propose`, acceptance with min(1, |psi'/psi|^2) or, with `sur` (a function
    s -> f(s)), delayed acceptance.  psi is a function s -> (sign, logabs).  Returns (final
    configuration, accepted count, near_tie), near_tie True when some acceptance test fell
    within `margin` of its threshold (then a rounding-level difference can legitimately change
    the trajectory).  The particle hop counts the particles of the configuration itself; the
    kernel uses the model's n_up + n_dn, the same number for a well-formed configuration."""
    s = np.array(s0, dtype=np.int8, copy=True)
    cur = psi(s)
    f_cur = sur(s) if sur is not None else 0.0
    acc = 0
    tie = False
    for t in range(step0, step0 + n_steps):
        r = philox_draw(chain, t, seed)
        sp = propose(s, r[0], r[1], LX, LY, kind, ph_z, se_z)
        if sp is None:
            continue
        u1 = (r[2] >> 8) / 16777216.0
        u2 = (r[3] >> 8) / 16777216.0
        f_new = 0.0
        if sur is not None:
            f_new = sur(sp)
            d1 = 2.0 * (f_new - f_cur)
            if d1 < 0 and u1 > 0 and abs(math.log(u1) - d1) < margin:
                tie = True
            if not (d1 >= 0 or u1 < math.exp(d1)):
                continue
        new = psi(sp)
        d = 2.0 * (new[1] - cur[1]) - (2.0 * (f_new - f_cur) if sur is not None else 0.0)
        u = u2 if sur is not None else u1
        if d < 0 and u > 0 and abs(math.log(u) - d) < margin:
            tie = True
        if new[0] != 0 and (d >= 0 or u < math.exp(d)):
            s, cur, f_cur = sp, new, f_new
            acc += 1
    return s, acc, tie
