"""GPU parity of ASProp's frequency filter (angular_spectrum.jl:1-76,187-188) against the FP64
oracle: the reference's filter_lp example (a scalar function, sampled element by element), a
complex asymmetric filter on a non-square multi-wavelength batch for both signs of z and both
directions, the paraxial kernel, tilted fields (one filter grid per field), a grid that is not a
power of two, and a filtered propagator inside a differentiated Phase system."""

import math

import numpy as np
import pytest
import torch

import fluxoptics as fo
from fluxoptics import inputs
from oracle import fluxoptics_oracle as O

from _util import c64, rel_l2, speckle_batch, to_dev, to_host

pytestmark = pytest.mark.gpu

FIELD_TOL = 1e-5
GRAD_TOL = 1e-4


def filter_lp(fx, fy):
    """The reference's docstring example: sqrt(fx^2 + fy^2) < 0.3 ? 1.0 : 0.0."""
    return 1.0 if math.sqrt(fx * fx + fy * fy) < 0.3 else 0.0


def asym(fx, fy):
    """Complex, not even in fx or fy (works on scalars and arrays)."""
    return (1.0 + 0.5j * np.asarray(fx)) * np.exp(7j * np.asarray(fy)) + 0.25 * (np.asarray(fx) > 0)


@pytest.mark.parametrize("z", [800.0, -800.0])
def test_filter_lp_example(z):
    n, ds, lam = 256, (1.0, 1.0), 1.064
    print(f"[filter] filter_lp 256^2 z={z}", flush=True)
    u0 = c64(speckle_batch(2, n, n, ds[0], ds[1], 40.0, 11))
    u = fo.ScalarField(to_dev(u0), ds, lam)
    p = fo.ASProp(u, z, filter=filter_lp)
    for d, s in ((fo.Forward, 1), (fo.Backward, -1)):
        v = to_host(p.propagate(u, d).electric)
        assert rel_l2(v, O.asprop(u0, ds, lam, z, 1.0, s, filt=filter_lp)) <= FIELD_TOL, (z, s)


@pytest.mark.parametrize("z", [350.0, -350.0])
@pytest.mark.parametrize("paraxial", [False, True])
def test_complex_filter_nonsquare_multi_lambda(z, paraxial):
    nx, ny, ds, lams, n0 = 128, 64, (0.5, 0.8), [0.64, 0.532], 1.3
    print(f"[filter] asym {nx}x{ny} paraxial={paraxial} z={z}", flush=True)
    u0 = c64(inputs.normal(21, (2, ny, nx)) + 1j * inputs.normal(22, (2, ny, nx)))
    u = fo.ScalarField(to_dev(u0), ds, lams)
    p = fo.ASProp(u, z, n0=n0, paraxial=paraxial, filter=asym)
    for d, s in ((fo.Forward, 1), (fo.Backward, -1)):
        v = to_host(p.propagate(u, d).electric)
        ref = O.asprop(u0, ds, lams, z, n0, s, paraxial=paraxial, filt=asym)
        assert rel_l2(v, ref) <= FIELD_TOL, (s, rel_l2(v, ref))


def test_filter_with_tilted_fields_and_odd_grid():
    """Tilted fields: per-field filter grids at fx + f0x; a 100 x 60 grid runs the padded form."""
    n, ds, lam = 128, (1.0, 1.0), 0.8
    print("[filter] tilted fields 128^2, then 100x60", flush=True)
    u0 = c64(inputs.normal(31, (2, n, n)) + 1j * inputs.normal(32, (2, n, n)))
    tilts = ([0.02, -0.05], [0.03, 0.0])
    u = fo.ScalarField(to_dev(u0), ds, lam, tilts=tilts)
    p = fo.ASProp(u, 500.0, filter=asym)
    v = to_host(p.propagate(u).electric)
    ref = O.asprop(u0, ds, lam, 500.0, 1.0, 1, tilts=tilts, filt=asym)
    assert rel_l2(v, ref) <= FIELD_TOL
    nx, ny = 100, 60
    w0 = c64(inputs.normal(33, (1, ny, nx)) + 1j * inputs.normal(34, (1, ny, nx)))
    w = fo.ScalarField(to_dev(w0), ds, lam)
    q = fo.ASProp(w, 300.0, filter=filter_lp)
    for d, s in ((fo.Forward, 1), (fo.Backward, -1)):
        got = to_host(q.propagate(w, d).electric)
        assert rel_l2(got, O.asprop(w0, ds, lam, 300.0, 1.0, s, filt=filter_lp)) <= FIELD_TOL, s


def test_filtered_propagator_in_differentiated_system():
    """Phase -> filtered ASProp -> intensity loss: loss and mask gradient against the oracle
    chain with the filtered kernel (the fused block uses the same tables)."""
    n, ds, lam, z = 128, (1.0, 1.0), 0.8, 600.0
    print("[filter] Phase -> filtered ASProp, withgradient", flush=True)
    u0 = c64(inputs.normalize_power(inputs.gaussian(n, n, ds[0], ds[1], 15.0), *ds))
    u = fo.ScalarField(to_dev(u0), ds, lam)
    phi = inputs.uniform(41, (n, n), 0, 2 * math.pi).astype(np.float32)
    tgt = (np.abs(inputs.gaussian(n, n, ds[0], ds[1], 10.0)) ** 2).astype(np.float32)
    doe = fo.Phase(u, phi, trainable=True, buffered=True)
    p = fo.ASProp(u, z, filter=filter_lp)
    system = fo.ScalarSource(u) | doe | p
    metric = fo.SquaredIntensityDifference((u, tgt))
    value, grads = system.withgradient(metric)
    # oracle: v = P_F(u e^{i phi}), L = sum(|v|^2 - t)^2, dL/dphi = sum Im(conj(u e^{i phi}) P_F^H(4 d v))
    a = u0 * np.exp(1j * phi.astype(np.float64))  # u0 is one (ny, nx) field
    v = O.asprop(a, ds, lam, z, 1.0, 1, filt=filter_lp)[0]
    d = np.abs(v) ** 2 - tgt
    loss = float((d * d).sum())
    g = O.asprop(4 * d * v, ds, lam, z, 1.0, -1, filt=filter_lp)[0]
    dphi = np.imag(g * np.conj(a))
    assert abs(float(value) - loss) <= FIELD_TOL * abs(loss)
    assert rel_l2(grads[doe].cpu().numpy(), dphi) <= GRAD_TOL
