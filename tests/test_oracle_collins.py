"""Pins the oracle's CollinsProp / FourierLens (CPU only): the FFT-and-chirp restatement of
the reference's sequence against the brute-force Collins integral, a closed-form Fourier lens,
and the adjoint relation of its Backward direction."""

import math

import numpy as np

from fluxoptics import inputs
from oracle import fluxoptics_oracle as O


def _cplx(shape, seed):
    return inputs.normal(seed, shape) + 1j * inputs.normal(seed + 1, shape)


def test_collins_matches_direct_integral():
    """Non-square grid, alpha = (1.7, 0.6), general (A, B, D), two wavelengths, per-field
    untracked tilts: the chirp / padded-convolution / chirp sequence equals the Collins sum."""
    nx, ny, ds = 16, 12, (0.7, 0.9)
    ds_out = (1.19, 0.54)
    abd = (0.8, 900.0, -0.3)
    lams = [0.9, 0.7]
    tilts = ([0.02, -0.01], [0.0, 0.015])
    u = _cplx((2, ny, nx), 101)
    v = O.collinsprop(u, ds, ds_out, abd, lams, 1, tilts)
    ref = O.collins_direct(u, ds, ds_out, abd, lams, tilts)
    assert np.max(np.abs(v - ref)) <= 1e-11 * np.max(np.abs(ref))
    # on-axis, negative B
    v = O.collinsprop(u, ds, ds_out, (1.0, -400.0, 1.0), lams, 1)
    ref = O.collins_direct(u, ds, ds_out, (1.0, -400.0, 1.0), lams)
    assert np.max(np.abs(v - ref)) <= 1e-11 * np.max(np.abs(ref))


def test_fourier_lens_of_a_gaussian():
    """FourierLens(f) of exp(-r^2/w^2): (pi w^2 / (i lam f)) exp(-pi^2 w^2 r'^2 / (lam f)^2)."""
    n, dx, w, lam, f = 256, 0.25, 4.0, 1.0, 100.0
    ds_out = (0.5, 0.5)
    xv, yv = O.spatial_vectors((n, n), (dx, dx))
    u = np.exp(-(xv[None, :] ** 2 + yv[:, None] ** 2) / w ** 2)
    v = O.collinsprop(u, (dx, dx), ds_out, (0.0, f, 0.0), lam)[0]
    xo, yo = O.spatial_vectors((n, n), ds_out)
    ref = math.pi * w ** 2 / (1j * lam * f) * np.exp(-(math.pi * w / (lam * f)) ** 2
                                                     * (xo[None, :] ** 2 + yo[:, None] ** 2))
    assert np.max(np.abs(v - ref)) <= 1e-10 * np.max(np.abs(ref))


def test_collins_backward_is_scaled_adjoint():
    """Backward = alpha_x alpha_y times the adjoint of Forward (the adjoint in the ds-weighted
    inner products of the two grids)."""
    n, ds, ds_out = 16, (0.5, 0.5), (0.8, 0.35)
    abd = (1.2, 300.0, 0.5)
    lams = [1.0, 0.8]
    tilts = ([0.03, 0.0], [-0.02, 0.01])
    u, w = _cplx((2, n, n), 111), _cplx((2, n, n), 113)
    pu = O.collinsprop(u, ds, ds_out, abd, lams, 1, tilts)
    bw = O.collinsprop(w, ds, ds_out, abd, lams, -1, tilts)
    scale = (ds_out[0] / ds[0]) * (ds_out[1] / ds[1])
    lhs, rhs = np.vdot(pu, w), np.vdot(u, bw) / scale
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
