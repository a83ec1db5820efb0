"""GPU parity of ASProp / RSProp / Phase against the FP64 oracle (through the C ABI).

Tolerance: relative L2 error <= 1e-5 for fields and <= 1e-4 for gradients, the
bounds of the complex-FP32 path against complex FP64.  Inputs are rounded to complex64
before the oracle sees them, and both sides use the same wavelength double.
"""

import math

import numpy as np
import pytest
import torch

import fluxoptics as fo
from fluxoptics import inputs
from oracle import fluxoptics_oracle as O
from _util import c64, rel_l2, random_field, speckle_batch, to_dev, to_host

pytestmark = pytest.mark.gpu

FIELD_TOL = 1e-5
GRAD_TOL = 1e-4


@pytest.mark.parametrize("nx,ny,batch,ds,lams,z,n0", [
    (16, 16, 1, (1.0, 1.0), 0.8, 50.0, 1.0),
    (32, 64, 2, (0.5, 0.7), [0.8, 1.064], 120.0, 1.0),
    (128, 128, 3, (1.0, 1.0), [0.64, 0.538, 0.455], 1e4, 1.0),
    (256, 512, 1, (0.5, 0.5), 1.0, -300.0, 1.5),
    (512, 256, 2, (1.0, 0.5), 0.8, 2000.0, 1.0),
    (1024, 1024, 2, (1.0, 1.0), 0.64, 1e4, 1.0),
    # single small fields: the launches use fewer lines per workgroup
    (1024, 1024, 1, (1.0, 1.0), 0.8, 3000.0, 1.0),
    (512, 512, 1, (1.0, 1.0), 1.064, 800.0, 1.0),
    (64, 64, 1, (1.0, 1.0), 0.8, 40.0, 1.0),
    (2048, 1024, 1, (0.5, 1.0), 0.8, 2000.0, 1.0),
    (2048, 2048, 2, (0.5, 0.5), 0.8, 5000.0, 1.0),
    # a table per field (three wavelengths) at 2048: the batched y pass without a shared table
    (2048, 2048, 3, (0.5, 0.5), [0.64, 0.538, 0.455], 5000.0, 1.0),
    (4096, 4096, 1, (0.5, 0.5), 0.8, 5000.0, 1.0),
])
def test_asprop_forward_backward_parity(nx, ny, batch, ds, lams, z, n0):
    print(f"[asprop parity] nx={nx} ny={ny} batch={batch} z={z}", flush=True)
    u0 = c64(speckle_batch(batch, nx, ny, ds[0], ds[1], 0.15 * nx * ds[0], 100 + nx))
    u = fo.ScalarField(to_dev(u0), ds, lams)
    p = fo.ASProp(u, z, n0=n0)
    v = p.propagate(u, fo.Forward)
    ref = O.asprop(u0, ds, lams, z, n0, 1)
    assert rel_l2(to_host(v.electric), ref) <= FIELD_TOL
    w = p.propagate(u, fo.Backward)
    refb = O.asprop(u0, ds, lams, z, n0, -1)
    assert rel_l2(to_host(w.electric), refb) <= FIELD_TOL


def test_asprop_headline_forward_adjoint():
    """8 x 2048^2, ds=0.5, lam=0.8, z=5000: forward, then the adjoint of v - t."""
    print("[asprop headline] 8 x 2048^2", flush=True)
    n, ds, lam, z = 2048, (0.5, 0.5), 0.8, 5000.0
    u0 = c64(speckle_batch(8, n, n, 0.5, 0.5, 150.0, 2000))
    t0 = c64(0.01 * random_field((8, n, n), 77))
    u = fo.ScalarField(to_dev(u0), ds, lam)
    p = fo.ASProp(u, z)
    p.propagate_(u, fo.Forward)
    v = to_host(u.electric)
    vref = O.asprop(u0, ds, lam, z, 1.0, 1)
    err_v = rel_l2(v, vref)
    # adjoint input: dL/dv for L = 1/2 |v - t|^2, formed from the device's own v
    dv0 = c64(v - t0)
    dv = fo.ScalarField(to_dev(dv0), ds, lam)
    p.backpropagate_(dv, fo.Forward)
    du = to_host(dv.electric)
    duref = O.asprop(dv0, ds, lam, z, 1.0, -1)
    err_du = rel_l2(du, duref)
    print(f"rel-L2 field {err_v:.3e}, adjoint {err_du:.3e}")
    assert err_v <= FIELD_TOL and err_du <= FIELD_TOL


def test_asprop_headline_adjoint_central_differences():
    """The 8 x 2048^2 adjoint against central finite differences on 64 random pixels.

    L(u) = 1/2 sum |P u - t|^2 is evaluated with the library's plain complex-FP64 propagator and
    summed in FP64, so the differences see no FP32 rounding.  L is quadratic in each pixel's real
    and imaginary parts, so the central difference (L(u + eps e_j) - L(u - eps e_j)) / (2 eps) is
    exact for any step.  It must equal dL/d(Re u) + i dL/d(Im u) = P^H (P u - t): the fast
    complex-FP32 adjoint applied to v - t, with v from the fast forward."""
    print("[asprop headline FD] 8 x 2048^2, 64 pixels", flush=True)
    n, ds, lam, z, batch = 2048, (0.5, 0.5), 0.8, 5000.0, 8
    u0 = np.asarray(speckle_batch(batch, n, n, 0.5, 0.5, 150.0, 2000), np.complex64)
    t0 = np.asarray(0.01 * random_field((batch, n, n), 77), np.complex64)
    u = fo.ScalarField(to_dev(u0), ds, lam)
    p = fo.ASProp(u, z)
    r = p.propagate(u, fo.Forward)
    r.electric.sub_(to_dev(t0))
    p.backpropagate_(r, fo.Forward)
    g = r.electric
    rng = np.random.default_rng(64)
    pix = [(int(rng.integers(batch)), int(rng.integers(n)), int(rng.integers(n))) for _ in range(64)]
    # the perturbed losses: a one-field plan of the same geometry, complex FP64 propagations
    one = fo.ScalarField(to_dev(u0[:1]), ds, lam)
    p1 = fo.ASProp(one, z)
    eps = 0.25
    got, fd = [], []
    for b, y, x in pix:
        base = torch.from_numpy(u0[b:b + 1].astype(np.complex128)).cuda()
        tb = torch.from_numpy(t0[b:b + 1].astype(np.complex128)).cuda()

        def loss(d):
            w = base.clone()
            w[0, y, x] += d
            p1.propagate_plain_fp64_(w, lam, fo.Forward)
            return 0.5 * float(((w - tb).abs() ** 2).sum().item())

        gre = (loss(eps) - loss(-eps)) / (2 * eps)
        gim = (loss(1j * eps) - loss(-1j * eps)) / (2 * eps)
        fd.append(gre + 1j * gim)
        got.append(complex(g[b, y, x].item()))
    got, fd = np.array(got), np.array(fd)
    err = rel_l2(got, fd)
    worst = float(np.max(np.abs(got - fd) / np.abs(fd)))
    print(f"rel-L2 over 64 pixels {err:.3e}, worst pixel {worst:.3e}")
    assert err <= GRAD_TOL


@pytest.mark.parametrize("nx,ny,batch,ds,lam,z", [
    (16, 16, 1, (1.0, 1.0), 1.064, 30.0),
    (64, 32, 2, (1.0, 1.0), 1.064, 300.0),
    (256, 256, 1, (1.0, 1.0), 1.064, -1500.0),
    (1024, 1024, 1, (1.0, 1.0), 1.064, 2000.0),
    (2048, 2048, 1, (1.0, 1.0), 1.064, 4000.0),
])
def test_rsprop_forward_backward_parity(nx, ny, batch, ds, lam, z):
    print(f"[rsprop parity] nx={nx} ny={ny} batch={batch} z={z}", flush=True)
    u0 = c64(np.stack([inputs.laguerre_gaussian(nx, ny, ds[0], ds[1], 0.1 * nx, b % 3, b - 1)
                       for b in range(batch)]))
    u = fo.ScalarField(to_dev(u0), ds, lam)
    import warnings
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        p = fo.RSProp(u, z)
    v = p.propagate(u, fo.Forward)
    ref = O.rsprop(u0, ds, lam, z, 1, "2n-1")
    assert rel_l2(to_host(v.electric), ref) <= FIELD_TOL
    w = p.propagate(u, fo.Backward)
    refb = O.rsprop(u0, ds, lam, z, -1, "2n-1")
    assert rel_l2(to_host(w.electric), refb) <= FIELD_TOL


def test_phase_apply_and_gradient():
    nx, ny, batch = 96, 64, 3
    u0 = c64(random_field((batch, ny, nx), 5))
    dv0 = c64(random_field((batch, ny, nx), 6))
    phi = inputs.uniform(7, (ny, nx), 0, 2 * math.pi).astype(np.float32)
    u = fo.ScalarField(to_dev(u0), (1.0, 1.0), 0.8)
    ph = fo.Phase(u, phi, trainable=True, buffered=True)
    ph.propagate_and_save_(u, None, fo.Forward)
    assert rel_l2(to_host(u.electric), O.phase_apply(u0, phi.astype(np.float64), 1)) <= FIELD_TOL
    assert np.array_equal(to_host(ph.saved), u0)
    dv = fo.ScalarField(to_dev(dv0), (1.0, 1.0), 0.8)
    du, dphi = ph.backpropagate_with_gradient_(dv, ph.saved, None, fo.Forward)
    du_ref, g_ref = O.phase_backward(dv0, u0, phi.astype(np.float64), 1)
    assert rel_l2(to_host(du.electric), du_ref) <= FIELD_TOL
    torch.cuda.synchronize()
    assert rel_l2(dphi.cpu().numpy(), g_ref) <= GRAD_TOL


@pytest.mark.parametrize("kind,n", [("as", 256), ("as", 1024), ("as", 2048), ("as", 4096), ("rs", 128),
                                    ("rs", 250), ("rs", 512), ("rs", 2048)])
def test_fused_phase_propagate_and_adjoint(kind, n):
    print(f"[fused phase+{kind}] n={n}", flush=True)
    batch, ds, lam, z = 3, (1.0, 1.0), [0.64, 0.538, 0.455], 800.0
    u0 = c64(speckle_batch(batch, n, n, 1.0, 1.0, 0.2 * n, 11))
    dv0 = c64(random_field((batch, n, n), 12))
    phi = inputs.uniform(13, (n, n), 0, 2 * math.pi).astype(np.float32)
    u = fo.ScalarField(to_dev(u0), ds, lam)
    ph = fo.Phase(u, phi, trainable=True, buffered=True)
    import warnings
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        p = fo.ASProp(u, z) if kind == "as" else fo.RSProp(u, z)
    stages = [("phase", phi.astype(np.float64)), (kind, z)]
    fo.phase_propagate_(u, ph, p, ph.saved, fo.Forward)
    vref, saved = O.pipeline_forward(u0, stages, ds, lam)
    assert rel_l2(to_host(u.electric), vref) <= FIELD_TOL
    assert np.array_equal(to_host(ph.saved), u0)
    dv = fo.ScalarField(to_dev(dv0), ds, lam)
    fo.propagate_phase_backward_(dv, ph, p, ph.saved, ph.grad, fo.Forward)
    du_ref, (g_ref,) = O.pipeline_backward(dv0, stages, ds, lam, saved)
    assert rel_l2(to_host(dv.electric), du_ref) <= FIELD_TOL
    torch.cuda.synchronize()
    assert rel_l2(ph.grad.cpu().numpy(), g_ref) <= GRAD_TOL


def test_plain_fp64_comparator():
    n, batch, ds, lam, z = 256, 2, (0.5, 0.5), 0.8, 700.0
    u0 = speckle_batch(batch, n, n, 0.5, 0.5, 40.0, 21)
    u = fo.ScalarField(to_dev(u0), ds, lam)
    p = fo.ASProp(u, z)
    d = torch.from_numpy(u0.copy()).cuda()
    p.propagate_plain_fp64_(d, lam, fo.Forward)
    torch.cuda.synchronize()
    assert rel_l2(d.cpu().numpy(), O.asprop(u0, ds, lam, z, 1.0, 1)) <= 1e-12
    p.propagate_plain_fp64_(d, lam, fo.Backward)
    torch.cuda.synchronize()
    assert rel_l2(d.cpu().numpy(), O.asprop(O.asprop(u0, ds, lam, z, 1.0, 1), ds, lam, z, 1.0, -1)) <= 1e-12


def test_reference_properties_on_device():
    """``test/optical_components_test.jl:85-133`` run through the device path."""
    n = 128
    u0 = inputs.gaussian(n, n, 1.0, 1.0, 20.0)
    u = fo.ScalarField(to_dev(u0), (1.0, 1.0), 1.064)
    p = fo.ASProp(u, 1000.0)
    back = p.propagate(p.propagate(u, fo.Forward), fo.Backward)
    assert float(fo.coupling_efficiency(u, back)[0]) > 0.9999
    r = fo.RSProp(u, 1000.0)
    back_rs = r.propagate(r.propagate(u, fo.Forward), fo.Backward)
    assert float(fo.coupling_efficiency(u, back_rs)[0]) > 0.999
    u1 = fo.ScalarField(to_dev(inputs.gaussian(n, n, 0.5, 0.5, 15.0)), (0.5, 0.5), 1.064)
    u1.normalize_power_()
    p1 = fo.ASProp(u1, 500.0)
    v1 = p1.propagate(u1, fo.Forward)
    assert abs(float(v1.power()[0]) - float(u1.power()[0])) <= 1e-6 * float(u1.power()[0])


def test_error_paths_on_device():
    # 2n - 1 beyond the longest transform: refused (any smaller size runs, padded if needed)
    u = fo.ScalarField(to_dev(np.zeros((16, 3000))), (1.0, 1.0), 0.8)
    with pytest.raises(fo._lib.FluxHipError):
        fo.ASProp(u, 10.0).propagate_(u)
    with pytest.raises(fo._lib.FluxHipError):
        fo.RSProp(u, 10.0).propagate_(u)
    u2 = fo.ScalarField(to_dev(np.zeros((64, 64))), (1.0, 1.0), 0.8)
    with pytest.raises(ValueError):
        fo.ASProp(u, 10.0).propagate_(u2)


@pytest.mark.parametrize("nx,ny,z", [(64, 48, 30.0), (60, 40, -25.0)])
def test_aspropz_distance_gradient(nx, ny, z):
    """ASPropZ (angular_spectrum.jl:236-300): the forward equals ASProp at the current z; the
    distance gradient from the output and a cotangent against the oracle (which finite
    differences of the FP64 propagator pin); a grid with evanescent frequencies (ds 0.4 with
    lambda 0.5, 0.7) and both signs of z; a non-power-of-two grid."""
    print(f"[ASPropZ] {nx}x{ny} z={z}", flush=True)
    ds, lams, n0 = (0.4, 0.4), [0.5, 0.7], 1.2
    u0 = c64(random_field((2, ny, nx), 201))
    c0 = c64(random_field((2, ny, nx), 202))
    u = fo.ScalarField(to_dev(u0), ds, lams)
    pz = fo.ASPropZ(u, z, n0=n0, trainable=True)
    v = pz.propagate(u)
    assert rel_l2(to_host(v.electric), O.asprop(u0, ds, lams, z, n0, 1)) <= FIELD_TOL
    g = pz.z_gradient(v, fo.ScalarField(to_dev(c0), ds, lams))
    ref = O.as_z_gradient(to_host(v.electric), c0, ds, lams, z, n0).sum()
    assert abs(float(g[0]) - ref) <= 1e-4 * max(1.0, abs(ref))


def test_aspropz_in_a_differentiated_system():
    """source | ASPropZ(trainable) with SquaredFieldDifference against the field propagated to
    z = 40: dL/dz against the oracle chain; one Descent step of 0.01 um against the gradient
    sets z to z - eta g and lowers the loss (the field's carrier phase makes L(z) oscillate with
    period lambda, so only small steps are descents)."""
    n, ds, lam = 64, (1.0, 1.0), 0.8
    g0 = c64(inputs.gaussian(n, n, 1.0, 1.0, 4.0))
    t0 = c64(O.asprop(g0, ds, lam, 40.0, 1.0, 1)[0])
    u = fo.ScalarField(to_dev(g0), ds, lam)
    pz = fo.ASPropZ(u, 25.0, trainable=True, buffered=True)
    system = fo.ScalarSource(u) | pz
    metric = fo.SquaredFieldDifference(fo.ScalarField(to_dev(t0), ds, lam))
    value, grads = system.withgradient(metric)
    torch.cuda.synchronize()
    v = O.asprop(g0, ds, lam, 25.0, 1.0, 1)
    c, dv = O.sq_field_diff_values(v, t0)
    ref = O.as_z_gradient(v, dv, ds, lam, 25.0, 1.0).sum()
    assert abs(float(value) - c.sum()) <= 1e-5 * c.sum()
    assert abs(float(grads[pz][0]) - ref) <= 1e-4 * abs(ref) and grads[pz] is pz.grad
    eta = 0.01 / abs(ref)
    opt = fo.setup(fo.Descent(eta), system)
    fo.update_(opt, system, grads)
    z1 = float(pz.z[0])
    assert abs(z1 - (25.0 - eta * float(grads[pz][0]))) <= 1e-5
    value1, _ = system.withgradient(metric)
    assert float(value1) < float(value)
