"""Shared test helpers."""

import math


def cancellation(ref_model, s):
    """c = max_g log|image_g| - log|psi_P|: the projection's cancellation (in e-folds), which
    scales the relative rounding error of psi_P by e^c.  inf when psi_P is exactly zero."""
    imgs = [ref_model.image_amplitude(s, g)[1] for g in range(8)]
    p = ref_model.psi(s)
    if p[0] == 0:
        return math.inf
    return max(imgs) - p[1]


def log_tol(base, c):
    return base * max(1.0, math.exp(c))
