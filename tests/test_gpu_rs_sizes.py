"""RSProp on grids that are not powers of two (padded to the power of two >= 2n - 1) against
the oracle's (2n-1) padding, including the fused Phase pair, tilted fields and the fused
metric tail (tolerances 1e-5 fields, 1e-4 gradients)."""

import math
import warnings

import numpy as np
import pytest
import torch

import fluxoptics as fo
from fluxoptics import inputs
from oracle import fluxoptics_oracle as O
from _util import c64, random_field, rel_l2, speckle_batch, to_dev, to_host

pytestmark = pytest.mark.gpu

FIELD_TOL = 1e-5


def _rs(u, z, **kw):
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        return fo.RSProp(u, z, **kw)


@pytest.mark.parametrize("nx,ny,batch,z", [(250, 250, 1, 2e5), (100, 60, 2, 900.0), (33, 47, 1, 120.0),
                                           (1000, 600, 1, 4000.0), (17, 16, 3, 50.0)])
def test_rsprop_any_size(nx, ny, batch, z):
    print(f"[rs any size] {nx}x{ny} batch={batch}", flush=True)
    ds = (10.0, 10.0) if nx == 250 else (1.0, 1.0)
    lams = [0.7, 0.6, 0.5][:batch]
    u0 = c64(random_field((batch, ny, nx), 171))
    u = fo.ScalarField(to_dev(u0), ds, lams)
    p = _rs(u, z)
    v = p.propagate(u, fo.Forward)
    assert rel_l2(to_host(v.electric), O.rsprop(u0, ds, lams, z, 1, "2n-1")) <= FIELD_TOL
    w = p.propagate(u, fo.Backward)
    assert rel_l2(to_host(w.electric), O.rsprop(u0, ds, lams, z, -1, "2n-1")) <= FIELD_TOL


def test_rsprop_any_size_tilted_and_fused_phase():
    nx, ny, ds, lam, z = 90, 70, (1.0, 1.0), 1.064, 2500.0
    print("[rs any size] tilted, fused phase", flush=True)
    tilts = ([0.02, -0.01], [0.0, 0.03])
    u0 = c64(random_field((2, ny, nx), 173))
    u = fo.ScalarField(to_dev(u0), ds, lam, tilts)
    p = _rs(u, z)
    v = p.propagate(u, fo.Forward)
    assert rel_l2(to_host(v.electric), O.rsprop(u0, ds, lam, z, 1, "2n-1", tilts=tilts)) <= FIELD_TOL
    # fused Phase -> RS and its adjoint with the mask gradient
    phi = inputs.uniform(174, (ny, nx), 0, 2 * math.pi).astype(np.float32)
    flat = fo.ScalarField(to_dev(u0), ds, lam)
    ph = fo.Phase(flat, phi, trainable=True, buffered=True)
    q = _rs(flat, z)
    t0 = (np.abs(c64(random_field((2, ny, nx), 175))) ** 2).astype(np.float32)
    system = fo.ScalarSource(flat) | ph | q
    metric = fo.SquaredIntensityDifference((flat, t0))
    stages = [("phase", phi.astype(np.float64)), ("rs", z)]
    v_ref, saved = O.pipeline_forward(u0, stages, ds, lam)
    vals_ref, dv_ref = O.sq_intensity_diff_values(v_ref, t0.astype(np.float64))
    _, (g_ref,) = O.pipeline_backward(dv_ref, stages, ds, lam, saved)
    for fuse in (True, False):
        value, grads = system.withgradient(metric, fuse=fuse)
        assert abs(float(value) - vals_ref.sum()) <= FIELD_TOL * vals_ref.sum()
        torch.cuda.synchronize()
        assert rel_l2(ph.grad.cpu().numpy(), g_ref) <= 1e-4
