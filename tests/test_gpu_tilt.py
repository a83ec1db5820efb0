"""GPU parity of the paraxial and tilted transfer functions and of per-field distances
against the FP64 oracle (tolerance 1e-5 on fields)."""

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


def _check(prop, u0, ds, lams, tilts, ref_fn, tag):
    u = fo.ScalarField(to_dev(u0), ds, lams, tilts)
    v = prop.propagate(u, fo.Forward)
    assert rel_l2(to_host(v.electric), ref_fn(u0, 1)) <= FIELD_TOL, tag + " forward"
    w = prop.propagate(u, fo.Backward)
    assert rel_l2(to_host(w.electric), ref_fn(u0, -1)) <= FIELD_TOL, tag + " backward"


@pytest.mark.parametrize("n", [256, 1024])
@pytest.mark.parametrize("track", [False, True])
def test_tilted_asprop(n, track):
    print(f"[tilted AS] n={n} track={track}", flush=True)
    ds, lams, z, n0 = (0.5, 0.5), [0.64, 0.538, 0.455], 400.0, 1.2
    tilts = ([0.01, -0.02, 0.0], [0.015, 0.0, -0.005])
    u0 = c64(speckle_batch(3, n, n, 0.5, 0.5, 0.1 * n, 51))
    u = fo.ScalarField(to_dev(u0), ds, lams, tilts)
    p = fo.ASProp(u, z, n0=n0, track_tilts=track)
    _check(p, u0, ds, lams, tilts,
           lambda a, d: O.asprop(a, ds, lams, z, n0, d, tilts=tilts, track=track), "tilted AS")


@pytest.mark.parametrize("tilted,track", [(False, False), (True, False), (True, True)])
def test_paraxial_asprop(tilted, track):
    n, ds, lams, z = 512, (1.0, 1.0), [0.8, 1.064], -700.0
    print(f"[paraxial AS] tilted={tilted} track={track}", flush=True)
    tilts = ([0.004, -0.003], [0.0, 0.002]) if tilted else (0.0, 0.0)
    u0 = c64(speckle_batch(2, n, n, 1.0, 1.0, 60.0, 52))
    u = fo.ScalarField(to_dev(u0), ds, lams, tilts)
    p = fo.ASProp(u, z, paraxial=True, track_tilts=track)
    _check(p, u0, ds, lams, tilts,
           lambda a, d: O.asprop(a, ds, lams, z, 1.0, d, tilts=tilts if tilted else None, paraxial=True,
                                 track=track), "paraxial AS")


@pytest.mark.parametrize("track", [False, True])
def test_tilted_rsprop(track):
    n, ds, lam, z = 128, (1.0, 1.0), 1.064, 2500.0
    print(f"[tilted RS] track={track}", flush=True)
    tilts = ([0.02, -0.01], [0.0, 0.03])
    u0 = c64(random_field((2, n, n), 53))
    u = fo.ScalarField(to_dev(u0), ds, lam, tilts)
    p = fo.RSProp(u, z, track_tilts=track)
    _check(p, u0, ds, lam, tilts,
           lambda a, d: O.rsprop(a, ds, lam, z, d, "2n", tilts=tilts, track=track), "tilted RS")


@pytest.mark.parametrize("kind", ["as", "rs"])
def test_per_field_distances(kind):
    n, ds, lam = 256, (1.0, 1.0), 1.064
    zs = [600.0, 1900.0, -1200.0, 3100.0]
    print(f"[per-field z] {kind}", flush=True)
    u0 = c64(speckle_batch(4, n, n, 1.0, 1.0, 25.0, 54))
    u = fo.ScalarField(to_dev(u0), ds, lam)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        p = fo.ASProp(u, zs) if kind == "as" else fo.RSProp(u, zs)
    if kind == "as":
        ref = lambda a, d: O.asprop(a, ds, lam, zs, 1.0, d)
    else:
        ref = lambda a, d: O.rsprop(a, ds, lam, zs, d, "2n")
    _check(p, u0, ds, lam, (0.0, 0.0), ref, f"per-field z {kind}")


@pytest.mark.parametrize("n", [256, 1024])
def test_shiftprop(n):
    """``ShiftProp``: per-field geometric shifts by z tan(theta), both directions."""
    ds, lams, z = (0.5, 0.5), [0.64, 0.538, 0.455], 300.0
    print(f"[ShiftProp] n={n}", flush=True)
    tilts = ([0.01, -0.02, 0.0], [0.015, 0.0, -0.005])
    u0 = c64(speckle_batch(3, n, n, 0.5, 0.5, 0.1 * n, 55))
    u = fo.ScalarField(to_dev(u0), ds, lams, tilts)
    p = fo.ShiftProp(u, z)
    _check(p, u0, ds, lams, tilts, lambda a, d: O.shiftprop(a, ds, z, tilts, d), "ShiftProp")


def test_shiftprop_whole_pixels_is_a_roll():
    """A shift of whole pixels on the device equals a periodic roll of the input."""
    n, dx, z = 128, 0.5, 40.0
    print("[ShiftProp roll]", flush=True)
    tilts = ([math.atan(5 * dx / z)], [math.atan(-3 * dx / z)])
    u0 = c64(random_field((1, n, n), 56))
    u = fo.ScalarField(to_dev(u0), (dx, dx), 1.0, tilts)
    v = fo.ShiftProp(u, z).propagate(u, fo.Forward)
    assert rel_l2(to_host(v.electric), np.roll(u0, (-3, 5), axis=(1, 2))) <= FIELD_TOL


def test_offset_tilts_on_device():
    """``offset_tilts!`` on the device: 40 fields (two coefficient chunks), per-field
    wavelengths and tilts, non-square grid; then back to the original tilts."""
    nx, ny, b = 128, 64, 40
    print("[offset_tilts] 40 x 64x128", flush=True)
    ds = (0.5, 0.7)
    lams = [0.5 + 0.01 * k for k in range(b)]
    tx = [0.002 * (k - 20) for k in range(b)]
    ty = [0.03 - 0.001 * k for k in range(b)]
    u0 = c64(random_field((b, ny, nx), 57))
    u = fo.ScalarField(to_dev(u0), ds, lams, (tx, ty))
    u.offset_tilts_((0.01, 0.0))
    assert u.tilts == (0.01, 0.0)
    ref = O.offset_tilts(u0, ds, lams, (tx, ty), (0.01, 0.0))
    assert rel_l2(to_host(u.electric), ref) <= FIELD_TOL
    u.offset_tilts_((tx, ty))
    assert rel_l2(to_host(u.electric), u0) <= FIELD_TOL


def test_system_with_tilt_anchor_gradient():
    """Tilted source -> ASProp (tilted kernel) -> TiltAnchor(0) -> Phase -> ASProp (on-axis):
    loss and Phase gradient against the oracle chain; the cotangent's tilts are restored."""
    n, ds, lam = 256, (0.5, 0.5), 0.8
    print("[tilt anchor system] 256^2 x 2", flush=True)
    th = ([0.03, -0.02], [0.01, 0.02])
    u0 = c64(speckle_batch(2, n, n, 0.5, 0.5, 25.0, 58))
    phi = inputs.uniform(59, (n, n), 0, 2 * math.pi).astype(np.float32)
    tgt = (np.abs(c64(speckle_batch(2, n, n, 0.5, 0.5, 25.0, 60))) ** 2).astype(np.float32)
    u = fo.ScalarField(to_dev(u0), ds, lam, th)
    flat = fo.ScalarField(to_dev(u0), ds, lam)
    ph = fo.Phase(u, phi, trainable=True, buffered=True)
    system = fo.ScalarSource(u) | fo.ASProp(u, 300.0) | fo.TiltAnchor(flat) | ph | fo.ASProp(flat, 200.0)
    metric = fo.SquaredIntensityDifference((u, tgt))
    stages = [("as", 300.0, 1.0, th), ("anchor", th, (0.0, 0.0)), ("phase", phi.astype(np.float64)),
              ("as", 200.0)]
    v_ref, saved = O.pipeline_forward(u0, stages, ds, lam)
    vals_ref, dv_ref = O.sq_intensity_diff_values(v_ref, tgt.astype(np.float64))
    _, (g_ref,) = O.pipeline_backward(dv_ref, stages, ds, lam, saved)
    value, grads = system.withgradient(metric)
    assert abs(float(value) - vals_ref.sum()) <= FIELD_TOL * vals_ref.sum()
    torch.cuda.synchronize()
    assert rel_l2(ph.grad.cpu().numpy(), g_ref) <= 1e-4
