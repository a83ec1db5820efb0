"""Synthetic inputs: a counter-based RNG and the source modes the workloads use.

The mode formulas restate the reference's mode library for input generation only
(``src/modes/gaussian_modes.jl``); they are not on the propagation path.
"""

from __future__ import annotations

import math

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, count: int, start: int = 0) -> np.ndarray:
    """SplitMix64 outputs for counters start..start+count-1 (counter-based, reproducible)."""
    with np.errstate(over="ignore"):
        ctr = np.arange(start + 1, start + count + 1, dtype=np.uint64)
        x = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + ctr * _GOLDEN
        x = (x ^ (x >> np.uint64(30))) * _M1
        x = (x ^ (x >> np.uint64(27))) * _M2
        return x ^ (x >> np.uint64(31))


def uniform(seed: int, shape, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
    """Uniform doubles in [lo, hi) from the top 53 bits of SplitMix64."""
    n = int(np.prod(shape))
    bits = splitmix64(seed, n) >> np.uint64(11)
    u = bits.astype(np.float64) * (1.0 / 9007199254740992.0)
    return (lo + (hi - lo) * u).reshape(shape)


def normal(seed: int, shape) -> np.ndarray:
    """Standard normal doubles by Box-Muller over two SplitMix64 streams."""
    n = int(np.prod(shape))
    u1 = uniform(seed, (n,))
    u2 = uniform(seed ^ 0x5DEECE66D, (n,))
    r = np.sqrt(-2.0 * np.log1p(-u1))
    return (r * np.cos(2.0 * math.pi * u2)).reshape(shape)


def grid(n: int, d: float) -> np.ndarray:
    """Centred grid (i - (n-1)/2)*d (``src/GridUtils.jl:56-60``)."""
    return (np.arange(n, dtype=np.float64) - (n - 1) / 2.0) * d


def gaussian(nx: int, ny: int, dx: float, dy: float, w0: float, xc: float = 0.0,
             yc: float = 0.0) -> np.ndarray:
    """Normalised Gaussian (sqrt(2/pi)/w0) exp(-r^2/w0^2), shape (ny, nx), complex128.

    ``gaussian_modes.jl:22-36`` (normalisation constant per axis) and ``:216-233``.
    """
    x = grid(nx, dx)[None, :] - xc
    y = grid(ny, dy)[:, None] - yc
    c = math.sqrt(2.0 / math.pi) / w0
    return (c * np.exp(-(x * x + y * y) / (w0 * w0))).astype(np.complex128)


def _laguerre(p: int, a: int, t: np.ndarray) -> np.ndarray:
    # three-term recurrence of the generalised Laguerre polynomial (gaussian_modes.jl:480-491)
    if p == 0:
        return np.ones_like(t)
    l0, l1 = np.ones_like(t), 1.0 + a - t
    for k in range(2, p + 1):
        l0, l1 = l1, ((2 * k + a - 1 - t) * l1 - (k + a - 1) * l0) / k
    return l1


def laguerre_gaussian(nx: int, ny: int, dx: float, dy: float, w0: float, p: int,
                      l: int) -> np.ndarray:
    """Vortex LG_pl mode, shape (ny, nx), complex128 (``gaussian_modes.jl:578-613``)."""
    x = grid(nx, dx)[None, :]
    y = grid(ny, dy)[:, None]
    c = math.sqrt(2.0 * math.factorial(p) / (math.pi * math.factorial(p + abs(l)))) / w0
    r2 = (x * x + y * y) / (w0 * w0)
    rl = (math.sqrt(2.0) * np.sqrt(r2)) ** abs(l)
    return c * rl * _laguerre(p, abs(l), 2.0 * r2) * np.exp(-r2) * np.exp(1j * l * np.arctan2(y, x))


def normalize_power(u: np.ndarray, dx: float, dy: float, target: float = 1.0) -> np.ndarray:
    """Scale each slice to power ``target`` (``src/Fields.jl:843-846,889-892``)."""
    u = np.asarray(u)
    p = (np.abs(u) ** 2).sum(axis=(-2, -1), keepdims=True) * dx * dy
    return u * np.sqrt(target / p)


def smooth_noise(ns: int, ny: int, nx: int, sigma: float, peak: float, seed: int) -> np.ndarray:
    """Gaussian-filtered white noise on an (ns, ny, nx) grid (periodic filter of sigma grid
    points along every axis, applied in the Fourier domain) scaled to max |value| = peak, FP32:
    the graded-index volume of a beam-propagation workload."""
    w = normal(seed, (ns, ny, nx))
    kz = np.fft.fftfreq(ns)[:, None, None]
    ky = np.fft.fftfreq(ny)[None, :, None]
    kx = np.fft.rfftfreq(nx)[None, None, :]
    g = np.exp(-2 * (math.pi * sigma) ** 2 * (kz ** 2 + ky ** 2 + kx ** 2))
    f = np.fft.irfftn(np.fft.rfftn(w) * g, s=w.shape)
    return (peak * f / np.max(np.abs(f))).astype(np.float32)
