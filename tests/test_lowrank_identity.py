"""The low-rank Pfaffian ratio that nh_local_energy_lowrank evaluates, checked in FP64 against the
direct ratio Pf(F') / Pf(F) of the FP64 oracle's Parlett-Reid: F' = F + dF with dF skew and
supported on the rows and columns K (odd and even |K|, K padded with unchanged indices as the
kernels pad it), the 2k x 2k matrix assembled from its strict lower triangle as the kernels
assemble it ([[D^T G D, .], [(G D)_K + I, G_KK]], D = dF[:, K] with its K rows halved), and
the sign (-1)^(k(k+1)/2).  A dropped (-1)^k (the sign of Pf(-J) against Pf(J)) fails the odd
cases."""

import numpy as np
import pytest

from oracle import oracle as O


def _pf(A):
    s, l = O.slogpf(A)
    return s * np.exp(l)


def _lowrank_ratio(F, Fp, K):
    G = np.linalg.inv(F)
    k = len(K)
    D = (Fp - F)[:, K].copy()
    D[K, :] *= 0.5
    X = G @ D
    E = D.T @ X
    M = np.zeros((2 * k, 2 * k))
    for i in range(2 * k):
        for j in range(i):
            if i < k:
                M[i, j] = E[i, j]
            elif j < k:
                M[i, j] = X[K[i - k], j] + (1.0 if i - k == j else 0.0)
            else:
                M[i, j] = G[K[i - k], K[j - k]]
    M = M - M.T
    return (-1) ** (k * (k + 1) // 2) * _pf(M)


@pytest.mark.parametrize("n,k,pad", [(10, 3, 0), (10, 4, 0), (16, 5, 2), (16, 7, 0), (24, 9, 3), (12, 12, 0)])
def test_lowrank_pfaffian_ratio(n, k, pad):
    rng = np.random.default_rng(100 * n + k + pad)
    X = rng.standard_normal((n, n))
    F = np.triu(X, 1) - np.triu(X, 1).T
    changed = np.sort(rng.choice(n, k, replace=False))
    dF = np.zeros((n, n))
    dF[:, changed] = rng.standard_normal((n, k))
    dF = dF - dF.T  # skew, supported on the rows and columns `changed`
    Fp = F + dF
    others = np.setdiff1d(np.arange(n), changed)
    K = np.concatenate([changed, others[:pad]])  # changed indices first, then padding
    r = _pf(Fp) / _pf(F)
    assert abs(_lowrank_ratio(F, Fp, K) - r) <= 1e-9 * max(1.0, abs(r))
