"""GPU parity of the paraxial and tilted transfer functions and of per-field distances
against the FP64 oracle (tolerance 1e-5 on fields)."""

import math
import warnings

import numpy as np
import pytest

import fluxoptics as fo
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
