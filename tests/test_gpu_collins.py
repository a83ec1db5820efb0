"""GPU parity of CollinsProp / FourierLens / ParaxialProp against the FP64 oracle (tolerance
1e-5 on fields, 1e-4 on gradients), through the C ABI."""

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


@pytest.mark.parametrize("nx,ny,track", [(64, 32, False), (256, 256, True), (1024, 512, False)])
def test_collins_forward_backward(nx, ny, track):
    print(f"[collins] {nx}x{ny} track={track}", flush=True)
    ds, ds_out = (0.5, 0.6), (0.75, 0.45)
    abd = (0.7, 1500.0, -0.2)
    lams = [0.8, 0.65]
    tilts = ([0.01, -0.02], [0.0, 0.015])
    u0 = c64(speckle_batch(2, nx, ny, ds[0], ds[1], 0.1 * min(nx * ds[0], ny * ds[1]), 121))
    u = fo.ScalarField(to_dev(u0), ds, lams, tilts)
    p = fo.CollinsProp(u, ds, ds_out, abd, track_tilts=track)
    v = p.propagate(u, fo.Forward)
    assert v.ds == ds_out
    assert rel_l2(to_host(v.electric), O.collinsprop(u0, ds, ds_out, abd, lams, 1, tilts, track)) <= FIELD_TOL
    w = p.propagate(fo.ScalarField(to_dev(u0), ds_out, lams, tilts), fo.Backward)
    assert w.ds == ds
    assert rel_l2(to_host(w.electric), O.collinsprop(u0, ds, ds_out, abd, lams, -1, tilts, track)) <= FIELD_TOL


def test_fourier_lens_and_paraxial_resampling():
    n, ds = 512, (0.25, 0.25)
    print("[fourier lens] 512^2", flush=True)
    xv, yv = O.spatial_vectors((n, n), ds)
    g = np.exp(-(xv[None, :] ** 2 + yv[:, None] ** 2) / 6.0 ** 2)
    u0 = c64(np.stack([g, g * np.exp(1j * 0.3 * xv[None, :])]))
    u = fo.ScalarField(to_dev(u0), ds, 1.0)
    lens = fo.FourierLens(u, (0.5, 0.5), 200.0)
    v = lens.propagate(u, fo.Forward)
    assert rel_l2(to_host(v.electric), O.collinsprop(u0, ds, (0.5, 0.5), (0.0, 200.0, 0.0), 1.0)) <= FIELD_TOL
    pp = fo.ParaxialProp(u, (0.4, 0.3), 800.0)
    assert isinstance(pp, fo.CollinsProp)
    v = pp.propagate(u, fo.Forward)
    assert rel_l2(to_host(v.electric), O.collinsprop(u0, ds, (0.4, 0.3), (1.0, 800.0, 1.0), 1.0)) <= FIELD_TOL
    assert isinstance(fo.ParaxialProp(u, 800.0), fo.ASProp)
    assert isinstance(fo.ParaxialProp(u, ds, ds, 800.0), fo.ASProp)


def test_collins_backward_is_scaled_adjoint_on_device():
    n, ds, ds_out = 128, (0.5, 0.5), (0.8, 0.35)
    print("[collins adjoint] 128^2", flush=True)
    abd = (1.2, 300.0, 0.5)
    u0, w0 = c64(random_field((1, n, n), 123)), c64(random_field((1, n, n), 125))
    u = fo.ScalarField(to_dev(u0), ds, 1.0)
    p = fo.CollinsProp(u, ds_out, abd)
    pu = to_host(p.propagate(u, fo.Forward).electric).astype(np.complex128)
    bw = to_host(p.propagate(fo.ScalarField(to_dev(w0), ds_out, 1.0), fo.Backward).electric).astype(np.complex128)
    lhs, rhs = np.vdot(pu, w0), np.vdot(u0, bw) / (1.6 * 0.7)
    assert abs(lhs - rhs) <= 1e-5 * np.linalg.norm(pu) * np.linalg.norm(w0)


def test_system_gradient_through_collins():
    """Phase (trainable) -> CollinsProp: loss and mask gradient against the oracle chain, whose
    reverse step is CollinsProp's Backward (the reference's pullback)."""
    n, ds, ds_out, lams = 256, (0.5, 0.5), (0.7, 0.7), [0.8, 0.6]
    print("[collins system] 256^2 x 2", flush=True)
    abd = (1.0, 2000.0, 1.0)
    u0 = c64(speckle_batch(2, n, n, 0.5, 0.5, 25.0, 127))
    phi = inputs.uniform(128, (n, n), 0, 2 * math.pi).astype(np.float32)
    tgt = (np.abs(c64(speckle_batch(2, n, n, 0.7, 0.7, 30.0, 129))) ** 2).astype(np.float32)
    u = fo.ScalarField(to_dev(u0), ds, lams)
    ph = fo.Phase(u, phi, trainable=True, buffered=True)
    system = fo.ScalarSource(u) | ph | fo.CollinsProp(u, ds_out, abd)
    metric = fo.SquaredIntensityDifference((u, tgt))
    stages = [("phase", phi.astype(np.float64)), ("collins", ds, ds_out, abd)]
    v_ref, saved = O.pipeline_forward(u0, stages, ds, lams)
    vals_ref, dv_ref = O.sq_intensity_diff_values(v_ref, tgt.astype(np.float64))
    _, (g_ref,) = O.pipeline_backward(dv_ref, stages, ds, lams, saved)
    value, grads = system.withgradient(metric)
    assert abs(float(value) - vals_ref.sum()) <= FIELD_TOL * vals_ref.sum()
    torch.cuda.synchronize()
    assert rel_l2(ph.grad.cpu().numpy(), g_ref) <= 1e-4
