"""Pins the oracle's grids, FFT convention and angular-spectrum path (CPU only).

Each check is independent of the oracle's own formula: brute-force long-double DFTs,
mpmath at 50 digits, exact discrete eigenvectors, and the reference's own property
tests (``test/optical_components_test.jl:85-133``).
"""

import math

import mpmath
import numpy as np
import pytest

from oracle import fluxoptics_oracle as O
from fluxoptics import inputs


def brute_dft2(u, sign=-1):
    """O(n^4) 2-D DFT in long double (sum over both axes of e^{sign*2*pi*i*(jk/n)})."""
    u = np.asarray(u, dtype=np.clongdouble)
    ny, nx = u.shape
    jx = np.arange(nx, dtype=np.longdouble)
    jy = np.arange(ny, dtype=np.longdouble)
    two_pi = np.longdouble(2) * np.longdouble(math.pi)
    ex = np.exp(sign * 1j * two_pi * np.outer(jx, jx) / nx)
    ey = np.exp(sign * 1j * two_pi * np.outer(jy, jy) / ny)
    return ey @ u @ ex.T


def test_spatial_vectors_closed_form():
    x, y = O.spatial_vectors((4, 3), (1.0, 0.5))
    assert np.array_equal(x, [-1.5, -0.5, 0.5, 1.5])  # GridUtils.jl:40-50 doc example
    assert np.array_equal(y, [-0.5, 0.0, 0.5])


def test_fftfreq_closed_form():
    assert np.array_equal(O.fftfreq(4, 0.5), [0.0, 0.5, -1.0, -0.5])
    assert np.array_equal(O.fftfreq(5, 1.0), [0.0, 0.2, 0.4, -0.4, -0.2])


@pytest.mark.parametrize("shape", [(4, 4), (8, 16), (16, 6), (5, 7)])
def test_fft_convention_brute_force(shape):
    u = inputs.normal(11, shape) + 1j * inputs.normal(12, shape)
    ref = brute_dft2(u)
    got = O.ft2(u)
    assert np.max(np.abs(got - ref.astype(np.complex128))) < 1e-12 * np.max(np.abs(ref))
    back = O.ift2(got)
    assert np.max(np.abs(back - u)) < 1e-13


def test_parseval():
    u = inputs.normal(3, (32, 48)) + 1j * inputs.normal(4, (32, 48))
    f = O.ft2(u)
    assert abs((np.abs(f) ** 2).sum() - u.size * (np.abs(u) ** 2).sum()) < 1e-10 * (np.abs(f) ** 2).sum()


def _mp_as(fx, fy, lam, n0, z):
    """H from the physical form exp(i*z*sqrt((2*pi*n0/lam)^2 - (2*pi*fx)^2 - (2*pi*fy)^2))."""
    mpmath.mp.dps = 50
    fx, fy, lam, n0, z = (mpmath.mpf(v) for v in (fx, fy, lam, n0, z))
    kz2 = (2 * mpmath.pi * n0 / lam) ** 2 - (2 * mpmath.pi * fx) ** 2 - (2 * mpmath.pi * fy) ** 2
    if kz2 >= 0:
        ph = abs(z) * mpmath.sqrt(kz2)
        val = mpmath.expj(ph)
        return complex(val if z >= 0 else mpmath.conj(val))
    return complex(mpmath.exp(-abs(z) * mpmath.sqrt(-kz2)))


@pytest.mark.parametrize("lam,n0,z", [(0.8, 1.0, 5000.0), (0.455, 1.0, 1e4), (1.0, 1.5, 1.0),
                                      (1.064, 1.0, -1500.0), (0.8, 1.0, 0.0)])
def test_as_kernel_mpmath(lam, n0, z):
    rng_fx = inputs.uniform(21, (40,), -1.6, 1.6)
    rng_fy = inputs.uniform(22, (40,), -1.6, 1.6)
    fc = n0 / lam
    # points exactly around the evanescent circle too
    ang = inputs.uniform(23, (10,), 0, 2 * math.pi)
    edge = [(fc * (1 + e) * math.cos(a), fc * (1 + e) * math.sin(a))
            for a in ang for e in (-1e-9, 1e-9)]
    pts = list(zip(rng_fx, rng_fy)) + edge
    f2 = (n0 / lam) ** 2
    for fx, fy in pts:
        got = complex(O.as_kernel(fx, fy, lam, n0, z))
        ref = _mp_as(fx, fy, lam, n0, z)
        # FP64 error budget of the formula itself: rounding of a = f^2 - fx^2 - fy^2
        # (~eps*(f^2+fx^2+fy^2)) amplified by d sqrt(a)/da = 1/(2 sqrt|a|), times 2*pi*|z|,
        # plus eps times the phase (up to ~1.4e5 rad).
        a = abs(f2 - fx * fx - fy * fy)
        sq = math.sqrt(a)
        tol = (2 * math.pi * abs(z) * 4.5e-16 * (f2 + fx * fx + fy * fy) / (2 * max(sq, 1e-300))
               + 2 * math.pi * abs(z) * sq * 4e-16 + 1e-14)
        assert abs(got - ref) < tol, (fx, fy, got, ref, tol)


@pytest.mark.parametrize("nx,ny,dx,dy", [(16, 8, 0.5, 1.0), (12, 20, 1.0, 0.7)])
def test_asprop_plane_wave_eigenvector(nx, ny, dx, dy):
    """u = e^{2 pi i (k i/nx + l j/ny)} is an exact eigenvector: v = H(f_k, f_l) u."""
    lam, z, n0 = 0.8, 37.0, 1.2
    i = np.arange(nx)[None, :]
    j = np.arange(ny)[:, None]
    for k, l in [(0, 0), (3, 1), (nx - 2, ny - 3), (nx // 2, ny // 2), (5, ny - 1)]:
        u = np.exp(2j * math.pi * (k * i / nx + l * j / ny))
        fk = (k if k < (nx + 1) // 2 else k - nx) / (nx * dx)
        fl = (l if l < (ny + 1) // 2 else l - ny) / (ny * dy)
        h = _mp_as(fk, fl, lam, n0, z)
        v = O.asprop(u, (dx, dy), lam, z, n0, 1)[0]
        assert np.max(np.abs(v - h * u)) < 1e-12
        vb = O.asprop(u, (dx, dy), lam, z, n0, -1)[0]
        assert np.max(np.abs(vb - np.conj(h) * u)) < 1e-12


def test_asprop_brute_force_small():
    """Direct sum v[j,i] = (1/N) sum_kl H_kl e^{+2 pi i(...)} sum_ab u e^{-2 pi i(...)}."""
    nx, ny, dx, dy, lam, z = 6, 4, 0.6, 0.9, 0.7, 13.0
    u = inputs.normal(5, (ny, nx)) + 1j * inputs.normal(6, (ny, nx))
    hk = np.array([[_mp_as(O.fftfreq(nx, dx)[a], O.fftfreq(ny, dy)[b], lam, 1.0, z)
                    for a in range(nx)] for b in range(ny)])
    ref = brute_dft2(hk * brute_dft2(u, -1), +1) / (nx * ny)
    got = O.asprop(u, (dx, dy), lam, z)[0]
    assert np.max(np.abs(got - ref.astype(np.complex128))) < 1e-12


def test_asprop_semigroup_and_adjoint():
    n, ds, lam = 32, (0.5, 0.5), 0.8
    u = inputs.normal(7, (2, n, n)) + 1j * inputs.normal(8, (2, n, n))
    v = inputs.normal(9, (2, n, n)) + 1j * inputs.normal(10, (2, n, n))
    a = O.asprop(O.asprop(u, ds, lam, 30.0), ds, lam, 45.0)
    b = O.asprop(u, ds, lam, 75.0)
    assert np.max(np.abs(a - b)) < 1e-12
    # Backward is the exact adjoint: <P u, v> = <u, P^H v>
    lhs = np.vdot(O.asprop(u, ds, lam, 75.0), v)
    rhs = np.vdot(u, O.asprop(v, ds, lam, 75.0, 1.0, -1))
    assert abs(lhs - rhs) < 1e-11 * abs(lhs)
    # negative z is the inverse on propagating components
    c = O.asprop(O.asprop(u, ds, lam, -75.0), ds, lam, 75.0)
    fx = O.fftfreq(n, 0.5)
    prop = (fx[None, :] ** 2 + fx[:, None] ** 2) < 1 / lam ** 2
    diff = O.ft2(c - u)[:, ~prop]
    assert np.max(np.abs(O.ft2(c - u)[:, prop])) < 1e-10
    assert diff.size > 0


def test_reference_unitarity_and_energy():
    """The reference's own checks: coupling > 0.9999 and power rtol 1e-6."""
    n, ds, lam = 128, (1.0, 1.0), 1.064
    u0 = inputs.gaussian(n, n, 1.0, 1.0, 20.0)
    v = O.asprop(u0, ds, lam, 1000.0, 1.0, 1)
    back = O.asprop(v, ds, lam, 1000.0, 1.0, -1)
    assert O.coupling_efficiency(u0, back)[0] > 0.9999  # optical_components_test.jl:85-103
    u1 = O.normalize_power(inputs.gaussian(n, n, 0.5, 0.5, 15.0), (0.5, 0.5))
    v1 = O.asprop(u1, (0.5, 0.5), lam, 500.0)
    p0, p1 = O.power(u1, (0.5, 0.5))[0], O.power(v1, (0.5, 0.5))[0]
    assert abs(p1 - p0) <= 1e-6 * p0  # optical_components_test.jl:114-133


def test_asprop_multi_lambda_slices_independent():
    n, ds = 16, (1.0, 1.0)
    lams = [0.64, 0.538, 0.455]
    u = inputs.normal(1, (3, n, n)) + 1j * inputs.normal(2, (3, n, n))
    v = O.asprop(u, ds, lams, 200.0)
    for b, lam in enumerate(lams):
        assert np.max(np.abs(v[b] - O.asprop(u[b], ds, lam, 200.0)[0])) < 1e-14


@pytest.mark.parametrize("z", [35.0, -20.0])
def test_oracle_as_z_gradient_finite_differences(z):
    """ASPropZ's distance gradient against central finite differences of the FP64 propagator:
    L(z) = Re sum conj(c) ASProp(u, z) per field, whose cotangent at the output is c.  The grid
    (ds 0.4 with lambda 0.5 and 0.7) holds evanescent frequencies, so the decaying part and the
    z < 0 conjugation are both exercised."""
    u = inputs.normal(71, (2, 24, 20)) + 1j * inputs.normal(72, (2, 24, 20))
    c = inputs.normal(73, (2, 24, 20)) + 1j * inputs.normal(74, (2, 24, 20))
    ds, lams, n0 = (0.4, 0.4), [0.5, 0.7], 1.2
    loss = lambda zz: np.real(np.sum(np.conj(c) * O.asprop(u, ds, lams, zz, n0), axis=(-2, -1)))
    v = O.asprop(u, ds, lams, z, n0)
    g = O.as_z_gradient(v, c, ds, lams, z, n0)
    h = 1e-4
    fd = (loss(z + h) - loss(z - h)) / (2 * h)
    assert np.allclose(g, fd, rtol=1e-6, atol=1e-8)
