"""CPU tests pinning the oracle's paraxial and tilted transfer functions by identities that
tie them to the on-axis kernels (themselves pinned by mpmath, brute-force DFTs and closed
forms in the other oracle tests)."""

import math

import numpy as np
import pytest

from fluxoptics import inputs
from oracle import fluxoptics_oracle as O


@pytest.mark.parametrize("n0", [1.0, 1.5])
def test_paraxial_is_second_order_expansion(n0):
    """phase(exact) - phase(paraxial) = -pi*lam'^3*z*f^4/4 + O(f^6) for on-axis fields
    (sqrt(1/lam'^2 - f^2) = 1/lam' - lam'*f^2/2 - lam'^3*f^4/8 - ...)."""
    lam, z = 0.8, 2000.0
    le = lam / n0
    f = np.linspace(0.005, 0.04, 8)
    fx, fy = f * 0.6, f * 0.8
    d = np.angle(O.as_kernel(fx, fy, lam, n0, z) * np.conj(O.as_paraxial_kernel(fx, fy, lam, 0, 0, False, n0, z)))
    pred = -math.pi * le ** 3 * z * f ** 4 / 4
    assert np.all(np.abs(d - pred) <= 0.02 * np.abs(pred) + 1e-12)


def test_tilted_as_equals_carrier_shift():
    """A tilted field u stands for u*exp(2*pi*i*f0.x); with f0 on the frequency grid, the
    untracked tilted propagation of u equals the on-axis propagation of u*carrier divided
    by the carrier."""
    n, dx, lam, n0, z = 64, 1.0, 0.8, 1.3, 300.0
    le = lam / n0
    mx, my = 3, -2
    tx, ty = math.asin(mx * le / (n * dx)), math.asin(my * le / (n * dx))
    x = inputs.grid(n, dx)[None, :]
    y = inputs.grid(n, dx)[:, None]
    u = inputs.gaussian(n, n, dx, dx, 6.0) * np.exp(1j * 0.01 * x * y)
    carrier = np.exp(2j * math.pi * (mx * x + my * y) / (n * dx))
    ref = O.asprop(u * carrier, (dx, dx), lam, z, n0, 1)[0] / carrier
    v = O.asprop(u, (dx, dx), lam, z, n0, 1, tilts=(tx, ty))[0]
    assert np.linalg.norm(v - ref) <= 1e-10 * np.linalg.norm(ref)
    # adjoint direction and negative z use conj(H(-z))
    vb = O.asprop(v, (dx, dx), lam, z, n0, -1, tilts=(tx, ty))[0]
    assert np.linalg.norm(vb - u) <= 1e-10 * np.linalg.norm(u)
    vn = O.asprop(u, (dx, dx), lam, -z, n0, 1, tilts=(tx, ty))[0]
    va = O.asprop(u, (dx, dx), lam, z, n0, -1, tilts=(tx, ty))[0]
    assert np.linalg.norm(vn - va) <= 1e-12 * np.linalg.norm(u)


def test_tracked_tilt_shifts_the_output():
    """Tracking adds cis(2*pi*z*lam'*f0.f): the output moves by z*sin(theta); with that an
    integer number of pixels it is a roll of the untracked output."""
    n, dx, lam, n0 = 64, 1.0, 0.8, 1.0
    tx, ty = math.asin(0.05), math.asin(-0.025)
    z = 80.0  # z*sin(tx) = 4 px, z*sin(ty) = -2 px
    u = inputs.gaussian(n, n, dx, dx, 8.0).astype(complex)
    v0 = O.asprop(u, (dx, dx), lam, z, n0, 1, tilts=(tx, ty), track=False)[0]
    v1 = O.asprop(u, (dx, dx), lam, z, n0, 1, tilts=(tx, ty), track=True)[0]
    assert np.linalg.norm(v1 - np.roll(v0, (2, -4), axis=(0, 1))) <= 1e-10 * np.linalg.norm(v0)


def test_paraxial_tracked_offset_is_a_shift():
    """The paraxial kernel's tracked offset, -2*(f0x*fx + f0y*fy) radians, moves the output
    by -f0/pi; with f0/pi a whole number of pixels it is a roll."""
    n, dx, lam, n0, z = 64, 0.05, 0.5, 1.0, 3.0
    f0x = 2 * math.pi * dx     # output(x) = untracked(x - 2 px)
    f0y = -1 * math.pi * dx    # output(y) = untracked(y + 1 px)
    tx, ty = math.asin(f0x * lam), math.asin(f0y * lam)
    u = inputs.gaussian(n, n, dx, dx, 0.4).astype(complex)
    v0 = O.asprop(u, (dx, dx), lam, z, n0, 1, tilts=(tx, ty), paraxial=True, track=False)[0]
    v1 = O.asprop(u, (dx, dx), lam, z, n0, 1, tilts=(tx, ty), paraxial=True, track=True)[0]
    assert np.linalg.norm(v1 - np.roll(v0, (-1, 2), axis=(0, 1))) <= 1e-10 * np.linalg.norm(v0)


def test_tilted_rs_equals_carrier_shift():
    """(u conv (h*conj(c)))(x) = conj(c(x)) * ((u*c) conv h)(x) for c(x) = exp(2*pi*i*f0.x),
    exactly, for any f0; tracking leaves h unchanged."""
    n, dx, lam, z = 32, 1.0, 1.064, 900.0
    tx, ty = 0.013, -0.021
    x = inputs.grid(n, dx)[None, :]
    y = inputs.grid(n, dx)[:, None]
    c = np.exp(2j * math.pi * (math.sin(tx) / lam * x + math.sin(ty) / lam * y))
    u = inputs.normal(3, (n, n)) + 1j * inputs.normal(4, (n, n))
    ref = np.conj(c) * O.rsprop(u * c, (dx, dx), lam, z, 1)[0]
    v = O.rsprop(u, (dx, dx), lam, z, 1, tilts=(tx, ty))[0]
    assert np.linalg.norm(v - ref) <= 1e-12 * np.linalg.norm(ref)
    vt = O.rsprop(u, (dx, dx), lam, z, 1, tilts=(tx, ty), track=True)[0]
    assert np.linalg.norm(vt - O.rsprop(u, (dx, dx), lam, z, 1)[0]) == 0.0


def test_per_field_distances():
    """One z per field equals separate propagations."""
    n, dx, lam = 32, 1.0, 0.8
    u = inputs.normal(5, (3, n, n)) + 1j * inputs.normal(6, (3, n, n))
    zs = [100.0, 250.0, -40.0]
    v = O.asprop(u, (dx, dx), lam, zs)
    for k in range(3):
        assert np.array_equal(v[k], O.asprop(u[k], (dx, dx), lam, zs[k])[0])


def test_shiftprop_is_a_whole_pixel_roll():
    """``ShiftProp`` moves a field by z tan(theta); with that a whole number of pixels it is
    exactly a periodic roll (both axes, per field), and Backward undoes it."""
    n, dx = 32, 0.5
    u = inputs.normal(61, (2, n, n)) + 1j * inputs.normal(62, (2, n, n))
    z = 10.0
    tilts = ([math.atan(3 * dx / z), math.atan(-2 * dx / z)], [math.atan(-1 * dx / z), 0.0])
    v = O.shiftprop(u, (dx, dx), z, tilts)
    assert np.linalg.norm(v[0] - np.roll(u[0], (-1, 3), axis=(0, 1))) <= 1e-12 * np.linalg.norm(u[0])
    assert np.linalg.norm(v[1] - np.roll(u[1], (0, -2), axis=(0, 1))) <= 1e-12 * np.linalg.norm(u[1])
    assert np.linalg.norm(O.shiftprop(v, (dx, dx), z, tilts, -1) - u) <= 1e-12 * np.linalg.norm(u)
