"""ASProp, paraxial, tilted and ShiftProp on grids that are not powers of two (the padded
convolution form) against the oracle's n-point FFT restatement (tolerance 1e-5), the fused
Phase pair and metric tail on such a grid, and a whole-pixel ShiftProp roll."""

import math

import numpy as np
import pytest
import torch

import fluxoptics as fo
from fluxoptics import inputs
from oracle import fluxoptics_oracle as O
from _util import c64, random_field, rel_l2, speckle_batch, to_dev, to_host

pytestmark = pytest.mark.gpu

FIELD_TOL = 1e-5


def _check(p, u0, ds, lams, tilts, ref):
    u = fo.ScalarField(to_dev(u0), ds, lams, tilts)
    v = p.propagate(u, fo.Forward)
    assert rel_l2(to_host(v.electric), ref(1)) <= FIELD_TOL
    w = p.propagate(u, fo.Backward)
    assert rel_l2(to_host(w.electric), ref(-1)) <= FIELD_TOL


@pytest.mark.parametrize("nx,ny,batch,z,n0", [(250, 250, 1, 3000.0, 1.0), (100, 60, 2, 400.0, 1.33),
                                              (33, 47, 1, -80.0, 1.0), (800, 600, 3, 5000.0, 1.0)])
def test_asprop_any_size(nx, ny, batch, z, n0):
    print(f"[as any size] {nx}x{ny} batch={batch}", flush=True)
    ds, lams = (0.5, 0.5), [0.64, 0.538, 0.455][:batch]
    u0 = c64(speckle_batch(batch, nx, ny, 0.5, 0.5, 0.12 * min(nx, ny) * 0.5, 181))
    u = fo.ScalarField(to_dev(u0), ds, lams)
    _check(fo.ASProp(u, z, n0=n0), u0, ds, lams, (0.0, 0.0), lambda d: O.asprop(u0, ds, lams, z, n0, d))


@pytest.mark.parametrize("kind", ["tilted", "paraxial", "tracked", "shift"])
def test_asprop_any_size_kernels(kind):
    nx, ny, ds, lams, z = 90, 70, (0.5, 0.5), [0.8, 0.6], 300.0
    print(f"[as any size] {kind}", flush=True)
    tilts = ([0.02, -0.03], [0.01, 0.0])
    u0 = c64(random_field((2, ny, nx), 183))
    u = fo.ScalarField(to_dev(u0), ds, lams, tilts)
    if kind == "shift":
        p = fo.ShiftProp(u, z)
        ref = lambda d: O.shiftprop(u0, ds, z, tilts, d)
    else:
        p = fo.ASProp(u, z, paraxial=kind == "paraxial", track_tilts=kind == "tracked")
        ref = lambda d: O.asprop(u0, ds, lams, z, 1.0, d, tilts=tilts, paraxial=kind == "paraxial",
                                 track=kind == "tracked")
    _check(p, u0, ds, lams, tilts, ref)


def test_shiftprop_any_size_roll():
    nx, ny, dx, z = 96, 72, 0.5, 40.0
    print("[shift any size] roll", flush=True)
    tilts = ([math.atan(5 * dx / z)], [math.atan(-3 * dx / z)])
    u0 = c64(random_field((1, ny, nx), 185))
    u = fo.ScalarField(to_dev(u0), (dx, dx), 1.0, tilts)
    v = fo.ShiftProp(u, z).propagate(u, fo.Forward)
    assert rel_l2(to_host(v.electric), np.roll(u0, (-3, 5), axis=(1, 2))) <= FIELD_TOL


def test_asprop_any_size_training_tail():
    """Phase -> ASProp on a 250x250 grid, fused tail and component by component, against the
    oracle chain."""
    n, ds, lam, z = 250, (1.0, 1.0), 0.8, 2000.0
    print("[as any size] fused tail", flush=True)
    u0 = c64(speckle_batch(2, n, n, 1.0, 1.0, 30.0, 187))
    phi = inputs.uniform(188, (n, n), 0, 2 * math.pi).astype(np.float32)
    tg = (np.abs(c64(speckle_batch(2, n, n, 1.0, 1.0, 30.0, 189))) ** 2).astype(np.float32)
    u = fo.ScalarField(to_dev(u0), ds, lam)
    ph = fo.Phase(u, phi, trainable=True, buffered=True)
    system = fo.ScalarSource(u) | ph | fo.ASProp(u, z)
    metric = fo.SquaredIntensityDifference((u, tg))
    stages = [("phase", phi.astype(np.float64)), ("as", z)]
    v_ref, saved = O.pipeline_forward(u0, stages, ds, lam)
    vals_ref, dv_ref = O.sq_intensity_diff_values(v_ref, tg.astype(np.float64))
    _, (g_ref,) = O.pipeline_backward(dv_ref, stages, ds, lam, saved)
    for fuse in (True, False):
        value, grads = system.withgradient(metric, fuse=fuse)
        assert abs(float(value) - vals_ref.sum()) <= FIELD_TOL * vals_ref.sum()
        torch.cuda.synchronize()
        assert rel_l2(ph.grad.cpu().numpy(), g_ref) <= 1e-4


@pytest.mark.parametrize("nx,ny", [(250, 250), (90, 70), (33, 20)])
def test_fourier_operator_any_size(nx, ny):
    """fh_fft2 on sizes that are not powers of two (the DFT as a chirp z-transform): FFT, IFFT
    and the FourierOperator adjoint identity."""
    print(f"[fft2 any size] {nx}x{ny}", flush=True)
    x0 = c64(random_field((2, ny, nx), 191))
    u = fo.ScalarField(to_dev(x0), (1.0, 1.0), 0.8)
    F = fo.FourierOperator(u, True)
    v = F.propagate(u, fo.Forward)
    assert rel_l2(to_host(v.electric), np.fft.fft2(x0)) <= 1e-5
    w = F.propagate(v, fo.Backward)
    assert rel_l2(to_host(w.electric), x0) <= 1e-5


def test_collins_and_czt_any_size():
    nx, ny, ds = 90, 70, (0.5, 0.6)
    print("[collins, czt any size]", flush=True)
    u0 = c64(random_field((2, ny, nx), 193))
    u = fo.ScalarField(to_dev(u0), ds, [0.8, 0.6])
    p = fo.CollinsProp(u, (0.7, 0.45), (0.8, 900.0, -0.3))
    v = p.propagate(u, fo.Forward)
    assert rel_l2(to_host(v.electric), O.collinsprop(u0, ds, (0.7, 0.45), (0.8, 900.0, -0.3), [0.8, 0.6], 1)) <= 1e-5
    x = to_dev(u0)
    a = ((1.0005, 0.3), (0.9995, -0.2))
    w = ((1.0, -0.011), (1.0001, 0.017))
    got = fo.plan_czt(x, (1, 2), a, w, center_on_grid=True) * x
    ref = O.czt(u0, (1, 2), {1: a[0], 2: a[1]}, {1: w[0], 2: w[1]}, True)
    assert rel_l2(to_host(got), ref) <= 1e-5
