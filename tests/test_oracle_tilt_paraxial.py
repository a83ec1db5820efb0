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


def test_offset_tilts_shifts_the_spectrum():
    """``offset_tilts!`` from theta to 0 multiplies by exp(+i 2 pi sin(theta)/lam x): with
    sin(theta)/lam = k/(n dx) the spectrum moves by +k bins (magnitudes exactly), and offsetting
    back restores the field."""
    n, dx, lam = 64, 0.5, 0.8
    u = inputs.normal(71, (n, n)) + 1j * inputs.normal(72, (n, n))
    kx, ky = 5, -3
    th = (math.asin(kx * lam / (n * dx)), math.asin(ky * lam / (n * dx)))
    v = O.offset_tilts(u, (dx, dx), lam, th, (0.0, 0.0))[0]
    ref = np.roll(np.abs(np.fft.fft2(u)), (ky, kx), axis=(0, 1))
    assert np.max(np.abs(np.abs(np.fft.fft2(v)) - ref)) <= 1e-10 * np.max(ref)
    back = O.offset_tilts(v, (dx, dx), lam, (0.0, 0.0), th)[0]
    assert np.max(np.abs(back - u)) <= 1e-12 * np.max(np.abs(u))


def test_tilted_asprop_is_anchored_onaxis_asprop():
    """For tilts on the frequency grid (n0 = 1): the untracked tilted ASProp equals anchoring
    the field to zero tilt, on-axis ASProp, and anchoring back (a second route through the
    tilted kernel); per-field tilts and wavelengths."""
    n, dx, z = 64, 0.5, 150.0
    lams = [0.8, 0.64]
    # band-limited to |k| < n/4 so that the carrier's cyclic spectral shift wraps nothing
    spec = np.fft.fft2(inputs.normal(73, (2, n, n)) + 1j * inputs.normal(74, (2, n, n)))
    kk = np.abs(np.fft.fftfreq(n, 1.0 / n))
    spec[:, kk >= n // 4, :] = 0
    spec[:, :, kk >= n // 4] = 0
    u = np.fft.ifft2(spec)
    k = [(4, 0), (-2, 3)]
    tx = [math.asin(k[b][0] * lams[b] / (n * dx)) for b in range(2)]
    ty = [math.asin(k[b][1] * lams[b] / (n * dx)) for b in range(2)]
    direct = O.asprop(u, (dx, dx), lams, z, 1.0, tilts=(tx, ty))
    a = O.offset_tilts(u, (dx, dx), lams, (tx, ty), (0.0, 0.0))
    a = O.asprop(a, (dx, dx), lams, z, 1.0)
    a = O.offset_tilts(a, (dx, dx), lams, (0.0, 0.0), (tx, ty))
    assert np.max(np.abs(a - direct)) <= 1e-11 * np.max(np.abs(direct))


def test_pipeline_with_tilt_anchor_adjoint_and_gradient():
    """A tilted ASProp -> TiltAnchor(0) -> Phase -> ASProp chain: the reverse pass is the exact
    adjoint, and the Phase gradient through the anchor matches central differences."""
    n, ds, lam = 32, (0.5, 0.5), 0.8
    th = ([0.05, -0.02], [0.01, 0.03])
    u = inputs.normal(75, (2, n, n)) + 1j * inputs.normal(76, (2, n, n))
    phi = inputs.normal(77, (n, n))
    stages = lambda p: [("as", 40.0, 1.0, th), ("anchor", th, (0.0, 0.0)), ("phase", p), ("as", 25.0)]
    v, saved = O.pipeline_forward(u, stages(phi), ds, lam)
    w = inputs.normal(78, (2, n, n)) + 1j * inputs.normal(79, (2, n, n))
    du, _ = O.pipeline_backward(w, stages(phi), ds, lam, saved)
    assert abs(np.vdot(v, w) - np.vdot(u, du)) <= 1e-11 * abs(np.vdot(v, w))
    t = inputs.uniform(80, (2, n, n))
    loss = lambda p: O.sq_intensity_diff(O.pipeline_forward(u, stages(p), ds, lam)[0], t)[0]
    _, dv = O.sq_intensity_diff(v, t)
    _, (g,) = O.pipeline_backward(dv, stages(phi), ds, lam, saved)
    for i in (3, 200, 517, 1000):
        h = 1e-6
        e = np.zeros(n * n)
        e[i] = h
        fd = (loss(phi + e.reshape(n, n)) - loss(phi - e.reshape(n, n))) / (2 * h)
        assert abs(fd - g.reshape(-1)[i]) <= 1e-5 * max(1.0, abs(fd))
