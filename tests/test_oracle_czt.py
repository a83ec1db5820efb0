"""Pins the oracle's chirp z-transform (CPU only): the DFT special case, brute-force sums for
general a, w (both centrings, one or both dimensions) and the adjoint plan."""

import cmath
import math

import numpy as np
import pytest

from fluxoptics import inputs
from oracle import fluxoptics_oracle as O


def _cplx(shape, seed):
    return inputs.normal(seed, shape) + 1j * inputs.normal(seed + 1, shape)


def _direct(x, axis, a, w, center):
    """X[m] = sum_k x[k] a^-k w^(m k) along one axis, powers through exp(. log)."""
    n = x.shape[axis]
    k = np.arange(n) - ((n - 1) / 2 if center else 0)
    la = math.log(a[0]) + 1j * a[1]
    lw = math.log(w[0]) + 1j * w[1]
    m = np.exp(lw * np.outer(k, k) - la * k[None, :])  # [m, k]
    return np.moveaxis(np.tensordot(m, np.moveaxis(x, axis, 0), axes=(1, 0)), 0, axis)


def test_default_czt_is_the_dft():
    x = _cplx((2, 12, 20), 131)
    assert np.max(np.abs(O.czt(x, (1, 2)) - np.fft.fft2(x))) <= 1e-11 * np.max(np.abs(x)) * 20
    assert np.max(np.abs(O.czt(x, (1,)) - np.fft.fft(x, axis=-1))) <= 1e-11 * np.max(np.abs(x)) * 20


@pytest.mark.parametrize("center", [False, True])
def test_czt_matches_direct_sum(center):
    x = _cplx((10, 14), 133)
    a = {1: (1.02, 0.3), 2: (0.97, -0.2)}
    w = {1: (1.01, -0.21), 2: (0.995, 0.17)}
    ref = _direct(_direct(x, -1, a[1], w[1], center), -2, a[2], w[2], center)
    v = O.czt(x, (1, 2), a, w, center)
    assert np.max(np.abs(v - ref)) <= 1e-10 * np.max(np.abs(ref))
    ref1 = _direct(x, -2, a[2], w[2], center)
    v1 = O.czt(x, (2,), a, w, center)
    assert np.max(np.abs(v1 - ref1)) <= 1e-10 * np.max(np.abs(ref1))


def test_czt_adjoint_plan():
    x, y = _cplx((3, 8, 16), 135), _cplx((3, 8, 16), 137)
    a = {1: (1.0, 0.1), 2: (1.0, -0.05)}
    w = {1: (1.0, -0.3), 2: (1.0, 0.2)}
    cx = O.czt(x, (1, 2), a, w, True)
    ay = O.czt(y, (1, 2), a, w, True, adjoint=True)
    assert abs(np.vdot(cx, y) - np.vdot(x, ay)) <= 1e-11 * abs(np.vdot(cx, y))
