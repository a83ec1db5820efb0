"""CPU pins of the oracle's ASProp frequency filter (angular_spectrum.jl:1-76: the kernel times
Complex(filter(...)), conjugated with the kernel for z <= 0 except in the paraxial form): exact
plane-wave eigenvectors under an asymmetric complex filter, the carrier identity of tilted fields
(the filter sees the shifted frequencies), the exact adjoint, and the reference's own filter_lp
docstring example (angular_spectrum.jl:187-188)."""

import math

import numpy as np
import pytest

from fluxoptics import inputs
from oracle import fluxoptics_oracle as O


def asym_filter(fx, fy):
    """Complex, neither even nor symmetric in fx <-> fy: a transposed grid, a wrong sign of a
    frequency or a missing conjugation changes its samples."""
    return (1.0 + 0.5j * fx) * complex(math.cos(7.0 * fy), math.sin(7.0 * fy)) + (0.25 if fx > 0 else 0.0)


@pytest.mark.parametrize("z", [37.0, -37.0])
@pytest.mark.parametrize("paraxial", [False, True])
def test_filtered_plane_wave_eigenvector(z, paraxial):
    """u = e^{2 pi i (k i/nx + l j/ny)}: v = H(f_k, f_l) F~ u, F~ = F(f_k, f_l), conjugated for
    z <= 0 in the non-paraxial kernel; Backward gives the conjugate factor."""
    nx, ny, dx, dy, lam, n0 = 16, 8, 0.5, 1.0, 0.8, 1.2
    i = np.arange(nx)[None, :]
    j = np.arange(ny)[:, None]
    for k, l in [(0, 0), (3, 1), (nx - 2, ny - 3), (nx // 2, ny // 2), (5, ny - 1)]:
        u = np.exp(2j * math.pi * (k * i / nx + l * j / ny))
        fk = (k if k < nx // 2 else k - nx) / (nx * dx)
        fl = (l if l < ny // 2 else l - ny) / (ny * dy)
        f = asym_filter(fk, fl)
        if paraxial:
            h = O.as_paraxial_kernel(fk, fl, lam, 0.0, 0.0, False, n0, z)
        else:
            h = O.as_kernel(fk, fl, lam, n0, z)
            f = f if z > 0 else np.conj(f)
        v = O.asprop(u, (dx, dy), lam, z, n0, 1, paraxial=paraxial, filt=asym_filter)[0]
        assert np.max(np.abs(v - h * f * u)) < 1e-12, (k, l)
        vb = O.asprop(u, (dx, dy), lam, z, n0, -1, paraxial=paraxial, filt=asym_filter)[0]
        assert np.max(np.abs(vb - np.conj(h * f) * u)) < 1e-12, (k, l)


def test_filtered_tilted_equals_carrier_shift():
    """An untracked tilted field u stands for u * carrier; with f0 on the grid, the filtered
    tilted propagation equals the filtered on-axis propagation of u * carrier divided by the
    carrier: the filter is evaluated at fx + f0x, fy + f0y."""
    n, dx, lam, n0, z = 64, 1.0, 0.8, 1.3, 300.0
    le = lam / n0
    mx, my = 3, -2
    tx, ty = math.asin(mx * le / (n * dx)), math.asin(my * le / (n * dx))
    x = inputs.grid(n, dx)[None, :]
    y = inputs.grid(n, dx)[:, None]
    u = inputs.gaussian(n, n, dx, dx, 6.0) * np.exp(1j * 0.01 * x * y)
    carrier = np.exp(2j * math.pi * (mx * x + my * y) / (n * dx))
    for zz in (z, -z):
        ref = O.asprop(u * carrier, (dx, dx), lam, zz, n0, 1, filt=asym_filter)[0] / carrier
        v = O.asprop(u, (dx, dx), lam, zz, n0, 1, tilts=(tx, ty), filt=asym_filter)[0]
        assert np.linalg.norm(v - ref) <= 1e-10 * np.linalg.norm(ref), zz


def test_filtered_asprop_adjoint_and_unit_filter():
    """<P u, w> = <u, P^H w> with the complex filter; a filter of ones changes nothing."""
    ny, nx, ds, lam, z = 12, 20, (0.7, 1.0), 0.9, 55.0
    u = inputs.normal(1, (2, ny, nx)) + 1j * inputs.normal(2, (2, ny, nx))
    w = inputs.normal(3, (2, ny, nx)) + 1j * inputs.normal(4, (2, ny, nx))
    pu = O.asprop(u, ds, lam, z, 1.0, 1, filt=asym_filter)
    phw = O.asprop(w, ds, lam, z, 1.0, -1, filt=asym_filter)
    assert abs(np.vdot(w, pu) - np.vdot(phw, u)) <= 1e-12 * np.linalg.norm(u) * np.linalg.norm(w)
    one = O.asprop(u, ds, lam, z, 1.0, 1, filt=lambda fx, fy: 1.0)
    assert np.array_equal(one, O.asprop(u, ds, lam, z, 1.0, 1))


def test_reference_filter_lp_example():
    """The docstring's filter_lp = (fx, fy) -> sqrt(fx^2 + fy^2) < 0.3 ? 1.0 : 0.0
    (angular_spectrum.jl:187-188): the output spectrum is the unfiltered output's inside the
    disc and zero outside."""
    n, ds, lam, z = 32, (1.0, 1.0), 1.064, 400.0
    u = inputs.normal(7, (n, n)) + 1j * inputs.normal(8, (n, n))
    lp = lambda fx, fy: 1.0 if math.sqrt(fx * fx + fy * fy) < 0.3 else 0.0
    v = O.asprop(u, ds, lam, z, 1.0, 1, filt=lp)[0]
    ref = O.asprop(u, ds, lam, z, 1.0, 1)[0]
    f = np.fft.fftfreq(n, 1.0)
    inside = np.sqrt(f[None, :] ** 2 + f[:, None] ** 2) < 0.3
    V, R = np.fft.fft2(v), np.fft.fft2(ref)
    assert np.max(np.abs(V[~inside])) <= 1e-10 * np.max(np.abs(R))
    assert np.max(np.abs(V[inside] - R[inside])) <= 1e-10 * np.max(np.abs(R))
    assert 0 < inside.sum() < n * n
