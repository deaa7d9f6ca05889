"""Portable synthetic inputs for the transform workloads.

A counter-based SplitMix64 stream with FP64 Box-Muller, so that the same
(seed, count) gives bit-identical values on any machine and in any language
(numpy's and C++'s normal generators are implementation-defined).

Coefficients are drawn in storage order -- (n, l, m) for Zernike, (l, m) for
spherical harmonics -- cos part then sin part, and the sin part at m = 0 is
set to zero, as zest's forward transform always returns it.
"""
from __future__ import annotations

import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, count: int) -> np.ndarray:
    """The first `count` outputs of SplitMix64 started at state `seed`."""
    with np.errstate(over="ignore"):
        i = np.arange(1, count + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + i * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed: int, count: int) -> np.ndarray:
    """Uniform doubles in (0, 1]."""
    z = splitmix64(seed, count)
    return 1.0 - (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed: int, count: int) -> np.ndarray:
    """Standard normal doubles by Box-Muller on consecutive uniform pairs."""
    npair = (count + 1) // 2
    u = uniform(seed, 2 * npair)
    r = np.sqrt(-2.0 * np.log(u[0::2]))
    th = 2.0 * np.pi * u[1::2]
    out = np.empty(2 * npair)
    out[0::2] = r * np.cos(th)
    out[1::2] = r * np.sin(th)
    return out[:count]


def zernike_nlm(order: int):
    """(n, l, m) index arrays in zest's ZernikeTetrahedralLayout order."""
    ns, ls, ms = [], [], []
    for n in range(order):
        for l in range(n & 1, n + 1, 2):
            for m in range(l + 1):
                ns.append(n)
                ls.append(l)
                ms.append(m)
    return np.array(ns), np.array(ls), np.array(ms)


def sh_lm(order: int):
    ls, ms = [], []
    for l in range(order):
        for m in range(l + 1):
            ls.append(l)
            ms.append(m)
    return np.array(ls), np.array(ms)


def zernike_coeffs(order: int, seed: int) -> np.ndarray:
    """Coefficients N(0,1)/(1+n), shape (count, 2)."""
    n, _, m = zernike_nlm(order)
    c = normal(seed, 2 * n.size).reshape(-1, 2)
    c *= (1.0 / (1.0 + n))[:, None]
    c[m == 0, 1] = 0.0
    return c


def sh_coeffs(order: int, seed: int, power: float = 1.5) -> np.ndarray:
    """Coefficients N(0,1)(1+l)^-power, shape (count, 2)."""
    l, m = sh_lm(order)
    c = normal(seed, 2 * l.size).reshape(-1, 2)
    c *= ((1.0 + l) ** -power)[:, None]
    c[m == 0, 1] = 0.0
    return c
