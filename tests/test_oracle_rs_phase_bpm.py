"""Pins the oracle's RSProp, Phase and AS_BPM paths (CPU only).

Checks: the RS kernel against an mpmath derivative of the free-space Green's function,
delta responses and brute-force linear convolutions, the 2n-1 / 2n padding identity,
exact adjointness, the reference's Phase property tests, BPM closed forms and central
finite differences of every gradient.
"""

import math

import mpmath
import numpy as np
import pytest

from oracle import fluxoptics_oracle as O
from fluxoptics import inputs


def _cplx(shape, seed):
    return inputs.normal(seed, shape) + 1j * inputs.normal(seed + 1000, shape)


def _mp_rs(x, y, lam, z, dxdy):
    """h = dx*dy * (-d/dz) [e^{ikr} / (2 pi r)], differentiated numerically by mpmath."""
    mpmath.mp.dps = 40
    k = 2 * mpmath.pi / mpmath.mpf(lam)
    x, y = mpmath.mpf(x), mpmath.mpf(y)

    def g(zz):
        r = mpmath.sqrt(x * x + y * y + zz * zz)
        return mpmath.expj(k * r) / (2 * mpmath.pi * r)

    return complex(-mpmath.mpf(dxdy) * mpmath.diff(g, mpmath.mpf(z)))


@pytest.mark.parametrize("x,y,lam,z", [(0.0, 0.0, 1.064, 1500.0), (37.0, -12.0, 0.8, 200.0),
                                       (-300.0, 250.0, 1.064, 2000.0), (5.0, 3.0, 0.5, 10.0)])
def test_rs_kernel_mpmath(x, y, lam, z):
    dxdy = 0.75
    got = complex(O.rs_kernel(x, y, lam, z, dxdy / (2 * math.pi)))
    ref = _mp_rs(x, y, lam, z, dxdy)
    kr = 2 * math.pi / lam * math.sqrt(x * x + y * y + z * z)
    # FP64 phase k*r of up to ~1.2e4 rad: absolute phase error ~ kr * eps
    assert abs(got - ref) < (kr * 4e-16 + 1e-13) * abs(ref)
    # z < 0 is the conjugate of +|z| (rayleigh_sommerfeld.jl:9-12)
    assert complex(O.rs_kernel(x, y, lam, -z, dxdy / (2 * math.pi))) == pytest.approx(np.conj(got))


@pytest.mark.parametrize("pad", ["2n-1", "2n"])
def test_rsprop_delta_response(pad):
    nx, ny, dx, dy, lam, z = 10, 7, 1.0, 0.8, 1.064, 30.0
    i0, j0 = 3, 5
    u = np.zeros((ny, nx), complex)
    u[j0, i0] = 1.0
    v = O.rsprop(u, (dx, dy), lam, z, 1, pad)[0]
    i = np.arange(nx)[None, :]
    j = np.arange(ny)[:, None]
    h = np.vectorize(lambda a, b: _mp_rs(a, b, lam, z, dx * dy))((i - i0) * dx, (j - j0) * dy)
    assert np.max(np.abs(v - h)) < 1e-13 * np.max(np.abs(h))


def test_rsprop_brute_force_convolution():
    nx, ny, dx, dy, lam, z = 6, 5, 0.9, 1.1, 0.9, 25.0
    u = _cplx((ny, nx), 31)
    hh = lambda a, b: O.rs_kernel(a, b, lam, z, dx * dy / (2 * math.pi))
    ref = np.zeros((ny, nx), complex)
    for j in range(ny):
        for i in range(nx):
            for jj in range(ny):
                for ii in range(nx):
                    ref[j, i] += u[jj, ii] * hh((i - ii) * dx, (j - jj) * dy)
    got = O.rsprop(u, (dx, dy), lam, z)[0]
    assert np.max(np.abs(got - ref)) < 1e-13 * np.max(np.abs(ref))


def test_rsprop_pow2_padding_identity_and_adjoint():
    n, ds, lam, z = 24, (1.0, 1.0), 1.064, 300.0
    u = _cplx((2, n, n), 41)
    v = _cplx((2, n, n), 43)
    a = O.rsprop(u, ds, lam, z, 1, "2n-1")
    b = O.rsprop(u, ds, lam, z, 1, "2n")
    assert np.linalg.norm(a - b) <= 1e-14 * np.linalg.norm(a)
    lhs = np.vdot(a, v)
    rhs = np.vdot(u, O.rsprop(v, ds, lam, z, -1, "2n-1"))
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    ab = O.rsprop(v, ds, lam, z, -1, "2n-1")
    bb = O.rsprop(v, ds, lam, z, -1, "2n")
    assert np.linalg.norm(ab - bb) <= 1e-14 * np.linalg.norm(ab)


def test_rs_reference_unitarity():
    """``test/optical_components_test.jl:105-111``: fwd->bwd coupling > 0.999."""
    n = 128
    u0 = inputs.gaussian(n, n, 1.0, 1.0, 20.0)
    v = O.rsprop(u0, (1.0, 1.0), 1.064, 1000.0, 1)
    back = O.rsprop(v, (1.0, 1.0), 1.064, 1000.0, -1)
    assert O.coupling_efficiency(u0, back)[0] > 0.999


def test_phase_reference_properties():
    """``test/optical_components_test.jl:22-44``: constant pi/2 and ramp x+y."""
    n = 64
    x, y = O.spatial_vectors((n, n), (1.0, 1.0))
    u = np.ones((n, n), complex)
    r = O.phase_apply(u, np.full((n, n), math.pi / 2))
    assert np.allclose(np.angle(r), math.pi / 2)
    ramp = x[None, :] + y[:, None]
    r2 = O.phase_apply(u, ramp)
    assert np.allclose(np.angle(r2), np.angle(np.exp(1j * ramp)))
    # Backward undoes Forward
    assert np.allclose(O.phase_apply(r2, ramp, -1), u)


def _fd_check(loss, theta, idx, analytic, h=1e-4, rtol=1e-6):
    """Central differences; the bound adds the FP64 cancellation noise 8*eps*|L|/h."""
    noise = 8 * 2.2e-16 * abs(loss(theta)) / h
    for p in idx:
        t1 = theta.copy()
        t2 = theta.copy()
        t1.flat[p] += h
        t2.flat[p] -= h
        fd = (loss(t1) - loss(t2)) / (2 * h)
        scale = max(abs(fd), np.max(np.abs(analytic)) * 1e-3)
        assert abs(fd - analytic.flat[p]) <= rtol * scale + noise, (p, fd, analytic.flat[p])


def test_phase_gradient_finite_difference():
    """Phase -> AS -> 1/2|v - t|^2: dphi vs central differences on random pixels."""
    n, ds, lam, z = 24, (1.0, 1.0), 0.8, 150.0
    u = _cplx((3, n, n), 51)
    t = _cplx((3, n, n), 53)
    phi = inputs.uniform(55, (n, n), 0, 2 * math.pi)

    def loss(ph):
        v, _ = O.pipeline_forward(u, [("phase", ph), ("as", z)], ds, lam)
        return O.half_sq_field_diff(v, t)[0]

    v, saved = O.pipeline_forward(u, [("phase", phi), ("as", z)], ds, lam)
    _, dv = O.half_sq_field_diff(v, t)
    _, grads = O.pipeline_backward(dv, [("phase", phi), ("as", z)], ds, lam, saved)
    idx = inputs.splitmix64(57, 16) % np.uint64(n * n)
    _fd_check(loss, phi, [int(i) for i in idx], grads[0])


def test_rs_phase_cascade_pipeline_gradient():
    """RS -> Phase -> RS -> Phase -> RS with sum (|v|^2 - I)^2, both mask gradients."""
    n, ds, lam = 20, (1.0, 1.0), 1.064
    u = O.normalize_power(inputs.gaussian(n, n, 1.0, 1.0, 3.0), ds)
    t = O.normalize_power(inputs.gaussian(n, n, 1.0, 1.0, 3.0, xc=4.0)
                          + inputs.gaussian(n, n, 1.0, 1.0, 3.0, xc=-4.0), ds)
    ti = np.abs(t) ** 2
    p1 = inputs.uniform(61, (n, n), 0, 1.0)
    p2 = inputs.uniform(62, (n, n), 0, 1.0)

    def stages(a, b):
        return [("rs", 15.0), ("phase", a), ("rs", 20.0), ("phase", b), ("rs", 15.0)]

    def loss_pair(a, b):
        v, _ = O.pipeline_forward(u, stages(a, b), ds, lam)
        return O.sq_intensity_diff(v, ti)[0]

    v, saved = O.pipeline_forward(u, stages(p1, p2), ds, lam)
    _, dv = O.sq_intensity_diff(v, ti)
    _, (g1, g2) = O.pipeline_backward(dv, stages(p1, p2), ds, lam, saved)
    idx = [int(i) for i in inputs.splitmix64(63, 12) % np.uint64(n * n)]
    _fd_check(lambda a: loss_pair(a, p2), p1, idx, g1)
    _fd_check(lambda b: loss_pair(p1, b), p2, idx, g2)


def test_bpm_closed_forms():
    n, ds, lam, n0, L, ns = 32, (0.5, 0.5), 1.0, 1.5, 40.0, 8
    u = _cplx((2, n, n), 71)
    v0, _ = O.bpm_forward(u, np.zeros((ns, n, n)), L, n0, lam, ds)
    assert np.max(np.abs(v0 - O.asprop(u, ds, lam, L, n0))) < 1e-12
    c = 0.003
    vc, _ = O.bpm_forward(u, np.full((ns, n, n), c), L, n0, lam, ds)
    ref = O.asprop(u, ds, lam, L, n0) * np.exp(2j * math.pi * L * c / lam)
    assert np.max(np.abs(vc - ref)) < 1e-12


def test_bpm_adjoint_and_dn_gradient():
    n, ds, lams, n0, L, ns = 16, (0.5, 0.5), [1.0, 0.9], 1.5, 12.0, 5
    u = _cplx((2, n, n), 81)
    t = _cplx((2, n, n), 83)
    dn = 0.01 * inputs.normal(85, (ns, n, n))
    v, saved = O.bpm_forward(u, dn, L, n0, lams, ds)
    w = _cplx((2, n, n), 87)
    du, _ = O.bpm_backward(w, dn, L, n0, lams, ds, saved)
    assert abs(np.vdot(v, w) - np.vdot(u, du)) < 1e-11 * abs(np.vdot(v, w))

    def loss(d):
        vv, _ = O.bpm_forward(u, d, L, n0, lams, ds)
        return O.half_sq_field_diff(vv, t)[0]

    _, dv = O.half_sq_field_diff(v, t)
    _, ddn = O.bpm_backward(dv, dn, L, n0, lams, ds, saved)
    idx = [int(i) for i in inputs.splitmix64(89, 12) % np.uint64(ns * n * n)]
    _fd_check(loss, dn, idx, ddn, h=1e-5)


def test_shift_bpm_closed_form():
    """Shift_BPM with a uniform dn = c and whole-pixel shifts: a periodic roll by
    thickness*tan(theta) times cis(2 pi L c / (lam cos(tx) cos(ty))) (bpm.jl:3-6,70,235-242)."""
    n, dx, lam, L, ns, c = 32, 0.5, 1.0, 40.0, 8, 0.002
    u = _cplx((2, n, n), 91)
    tilts = ([math.atan(4 * dx / L), math.atan(-6 * dx / L)], [math.atan(2 * dx / L), 0.0])
    v, _ = O.bpm_forward(u, np.full((ns, n, n), c), L, 1.0, lam, (dx, dx), tilts=tilts, shift=True)
    shifts = [(2, 4), (0, -6)]
    for b in range(2):
        kdz = O.bpm_kdz(L, (tilts[0][b], tilts[1][b]))
        ref = np.roll(u[b], shifts[b], axis=(0, 1)) * np.exp(1j * kdz * c / lam)
        assert np.max(np.abs(v[b] - ref)) < 1e-11


def test_tilted_bpm_reduces_to_asprop():
    """With dn = 0 the tilted (and paraxial) AS_BPM is one ASProp over the thickness with the
    same kernel choice: transfer functions of one tilt compose over z."""
    n, ds, lams, n0, L, ns = 32, (0.5, 0.5), [1.0, 0.8], 1.3, 30.0, 6
    u = _cplx((2, n, n), 93)
    tilts = ([0.05, -0.03], [0.0, 0.02])
    for paraxial in (False, True):
        v, _ = O.bpm_forward(u, np.zeros((ns, n, n)), L, n0, lams, ds, tilts=tilts, paraxial=paraxial)
        ref = O.asprop(u, ds, lams, L, n0, tilts=tilts, paraxial=paraxial)
        assert np.max(np.abs(v - ref)) < 1e-11


@pytest.mark.parametrize("kind", ["tilted", "paraxial", "shift"])
def test_general_bpm_adjoint_and_dn_gradient(kind):
    n, ds, lams, n0, L, ns = 16, (0.5, 0.5), [1.0, 0.9], 1.5, 12.0, 4
    tilts = ([0.08, -0.05], [0.03, 0.1])
    kw = dict(tilts=tilts, paraxial=kind == "paraxial", shift=kind == "shift")
    u = _cplx((2, n, n), 95)
    t = _cplx((2, n, n), 96)
    dn = 0.01 * inputs.normal(97, (ns, n, n))
    v, saved = O.bpm_forward(u, dn, L, n0, lams, ds, **kw)
    w = _cplx((2, n, n), 98)
    du, _ = O.bpm_backward(w, dn, L, n0, lams, ds, saved, **kw)
    assert abs(np.vdot(v, w) - np.vdot(u, du)) < 1e-11 * abs(np.vdot(v, w))

    def loss(d):
        vv, _ = O.bpm_forward(u, d, L, n0, lams, ds, **kw)
        return O.half_sq_field_diff(vv, t)[0]

    _, dv = O.half_sq_field_diff(v, t)
    _, ddn = O.bpm_backward(dv, dn, L, n0, lams, ds, saved, **kw)
    idx = [int(i) for i in inputs.splitmix64(99, 10) % np.uint64(ns * n * n)]
    _fd_check(loss, dn, idx, ddn, h=1e-5)


def test_oracle_sq_field_diff_pins():
    """SquaredFieldDifference: the reference's doctest (4x4 ones against zeros -> 16,
    squared_field_difference.jl), its testset (32x32 ones against 2 ones -> 1024, a field against
    itself -> 0; test/metrics_test.jl:77-99), and the pullback 2 dc (v - t) against central finite
    differences of the loss in the real and imaginary parts of single pixels."""
    c, _ = O.sq_field_diff_values(np.zeros((4, 4)), np.ones((4, 4)))
    assert c.shape == (1,) and c[0] == 16.0
    assert O.sq_field_diff_values(np.ones((32, 32)), 2 * np.ones((32, 32)))[0][0] == 1024.0
    u = inputs.normal(61, (2, 6, 5)) + 1j * inputs.normal(62, (2, 6, 5))
    assert np.all(O.sq_field_diff_values(u, u)[0] == 0.0)
    t = inputs.normal(63, (2, 6, 5)) + 1j * inputs.normal(64, (2, 6, 5))
    dc = np.array([0.7, -1.3])
    _, dv = O.sq_field_diff_values(u, t, dc)
    loss = lambda w: float(np.dot(dc, O.sq_field_diff_values(w, t)[0]))
    h = 1e-6
    for (b, y, x) in [(0, 1, 2), (1, 5, 0), (1, 3, 4)]:
        for unit, part in ((1.0, "re"), (1j, "im")):
            e = np.zeros_like(u)
            e[b, y, x] = unit * h
            fd = (loss(u + e) - loss(u - e)) / (2 * h)
            an = dv[b, y, x].real if part == "re" else dv[b, y, x].imag
            assert abs(fd - an) <= 1e-6 * max(1.0, abs(an))


def test_cpu_beam_splitter_baseline_equals_the_oracle_chain():
    """The timed FP64 CPU restatement of the beam-splitter pipeline (threaded FFTs, cached kernel
    spectra) gives the oracle chain's loss and mask gradients."""
    n, ds, lam = 64, (1.0, 1.0), 1.064
    u0 = inputs.normalize_power(inputs.gaussian(n, n, 1.0, 1.0, 8.0), 1.0, 1.0)
    t = np.abs(inputs.gaussian(n, n, 1.0, 1.0, 8.0, 10.0, 0.0)) ** 2
    phis = [inputs.uniform(1, (n, n), 0, 6.28), inputs.uniform(2, (n, n), 0, 6.28)]
    cpu = O.CpuRSPipelineF64(u0, t, ds, lam, [300.0, 400.0, 300.0], workers=2)
    loss, grads = cpu.iteration(phis)
    st = [("rs", 300.0), ("phase", phis[0]), ("rs", 400.0), ("phase", phis[1]), ("rs", 300.0)]
    v, sv = O.pipeline_forward(u0, st, ds, lam)
    lr, dv = O.sq_intensity_diff(v, t)
    _, gr = O.pipeline_backward(dv, st, ds, lam, sv)
    assert abs(loss - lr) <= 1e-12 * lr
    for a, b in zip(grads, gr):
        assert np.linalg.norm(a - b) <= 1e-12 * np.linalg.norm(b)
