"""Host-side helpers of the hot path's set-up (utils/utils.py,
utils/scanGrids.py).  These run once per reconstruction on the host; nothing
here is on the per-position path."""

import numpy as np


def circ(x, y, D):
    """Binary disc of diameter D on a meshgrid (utils.py:50-59)."""
    return (x ** 2 + y ** 2) < (D / 2) ** 2


def GenerateNonUniformFermat(n, radius=1000, power=1):
    """Fermat spiral scan grid (scanGrids.py:7-23): r = sqrt(i/n), golden
    angle 137.508 deg; returns (rows, cols)."""
    r = np.sqrt(np.arange(0, n) / n)
    theta0 = 137.508 / 180 * np.pi
    theta = np.arange(0, n) * theta0
    C = radius * r ** power * np.cos(theta)
    R = radius * r ** power * np.sin(theta)
    return R, C


def bitrev_permutation(n):
    """Index array b with b[i] = bit-reversal of i over log2(n) bits."""
    lg = int(np.log2(n))
    idx = np.arange(n)
    out = np.zeros(n, dtype=np.int64)
    for b in range(lg):
        out |= ((idx >> b) & 1) << (lg - 1 - b)
    return out
