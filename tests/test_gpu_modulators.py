"""GPU parity of Mask, TeaDOE/TeaReflector and FourierPhase/FourierMask against the oracle,
alone and inside differentiated systems (1e-5 fields, 1e-4 gradients)."""

import math

import numpy as np
import pytest
import torch

import fluxoptics as fo
from fluxoptics import inputs
from oracle import fluxoptics_oracle as O

from _util import c64, random_field, rel_l2, to_dev, to_host

pytestmark = pytest.mark.gpu

FIELD_TOL, GRAD_TOL = 1e-5, 1e-4


def _c64(a):
    return np.asarray(a, np.complex64).astype(np.complex128)


@pytest.mark.parametrize("direction", [1, -1])
def test_mask_apply_and_gradient(direction):
    b, ny, nx = 3, 96, 128
    print(f"[mask] direction={direction}", flush=True)
    u0, dv0 = c64(random_field((b, ny, nx), 101)), c64(random_field((b, ny, nx), 102))
    m0 = _c64(inputs.normal(103, (ny, nx)) + 1j * inputs.normal(104, (ny, nx)))
    u = fo.ScalarField(to_dev(u0), (1.0, 1.0), 0.8)
    mk = fo.Mask(u, m0.astype(np.complex64), trainable=True, buffered=True)
    d = fo.Forward if direction > 0 else fo.Backward
    mk.propagate_and_save_(u, None, d)
    assert rel_l2(to_host(u.electric), O.mask_apply(u0, m0, direction)) <= FIELD_TOL
    assert np.array_equal(to_host(mk.saved), u0)
    dv = fo.ScalarField(to_dev(dv0), (1.0, 1.0), 0.8)
    du, dm = mk.backpropagate_with_gradient_(dv, mk.saved, None, d)
    du_ref, dm_ref = O.mask_backward(dv0, u0, m0, direction)
    assert rel_l2(to_host(du.electric), du_ref) <= FIELD_TOL
    assert rel_l2(dm.cpu().numpy().astype(np.complex128), dm_ref) <= GRAD_TOL


@pytest.mark.parametrize("direction", [1, -1])
def test_tea_doe_apply_and_gradient(direction):
    b, n, lams, dn = 3, 128, [0.64, 0.538, 0.455], lambda l: 0.5 + 0.01 / l
    r = lambda l: 0.95 * np.exp(1j * 0.2 * l)
    print(f"[tea] direction={direction}", flush=True)
    u0, dv0 = c64(random_field((b, n, n), 111)), c64(random_field((b, n, n), 112))
    h0 = inputs.uniform(113, (n, n), 0.0, 1.5).astype(np.float32)
    u = fo.ScalarField(to_dev(u0), (1.0, 1.0), lams)
    tea = fo.TeaDOE(u, dn, h0, r=r, trainable=True, buffered=True)
    d = fo.Forward if direction > 0 else fo.Backward
    tea.propagate_and_save_(u, None, d)
    ref = np.stack([O.tea_apply(u0[k], h0.astype(np.float64), lams[k], dn(lams[k]), r(lams[k]), direction)[0]
                    for k in range(b)])
    assert rel_l2(to_host(u.electric), ref) <= FIELD_TOL
    dv = fo.ScalarField(to_dev(dv0), (1.0, 1.0), lams)
    du, dh = tea.backpropagate_with_gradient_(dv, tea.saved, None, d)
    du_ref, dh_ref = np.zeros_like(dv0), np.zeros((n, n))
    for k in range(b):
        a, g = O.tea_backward(dv0[k], u0[k], h0.astype(np.float64), lams[k], dn(lams[k]), r(lams[k]), direction)
        du_ref[k], dh_ref = a[0], dh_ref + g
    assert rel_l2(to_host(du.electric), du_ref) <= FIELD_TOL
    assert rel_l2(dh.cpu().numpy(), dh_ref) <= GRAD_TOL


def test_fourier_phase_and_mask():
    n, b, ds = 128, 2, (0.5, 0.5)
    u0 = c64(random_field((b, n, n), 121))
    u = fo.ScalarField(to_dev(u0), ds, 0.8)
    fp = fo.FourierPhase(u, lambda fx, fy: 0.3 * (fx * fx + fy * fy) / (0.1 ** 2))
    fx = np.fft.fftfreq(n, ds[0])[None, :]
    fy = np.fft.fftfreq(n, ds[1])[:, None]
    phi = (0.3 * (fx * fx + fy * fy) / 0.01).astype(np.float32).astype(np.float64)
    ref = O.ift2(O.phase_apply(O.ft2(u0), phi, 1))
    assert rel_l2(to_host(fp.propagate(u, fo.Forward).electric), ref) <= FIELD_TOL
    fm = fo.FourierMask(u, lambda fx, fy: np.exp(-(fx * fx + fy * fy) / 0.2))
    m = _c64(np.exp(-(fx * fx + fy * fy) / 0.2))
    ref = O.ift2(O.mask_apply(O.ft2(u0), m, 1))
    assert rel_l2(to_host(fm.propagate(u, fo.Forward).electric), ref) <= FIELD_TOL


def test_system_with_mask_and_tea_gradients():
    """Mask -> ASProp -> TeaDOE -> ASProp with the intensity loss: both gradients on device."""
    n, ds, lams, z, dn = 256, 1.0, [0.8, 1.064], 900.0, 0.6
    print("[system mask+tea]", flush=True)
    g = inputs.gaussian(n, n, ds, ds, 25.0)
    u0 = c64(inputs.normalize_power(np.stack([g, g]), ds, ds))
    m0 = _c64(np.exp(1j * inputs.uniform(131, (n, n), 0, 1)) * (0.8 + 0.2 * inputs.uniform(132, (n, n))))
    h0 = inputs.uniform(133, (n, n), 0.0, 1.0).astype(np.float32)
    tgt = np.stack([np.abs(inputs.normalize_power(inputs.gaussian(n, n, ds, ds, 20.0, c, 0.0), ds, ds)) ** 2
                    for c in (-30.0, 30.0)]).astype(np.float32)
    u = fo.ScalarField(to_dev(u0), (ds, ds), lams)
    mk = fo.Mask(u, m0.astype(np.complex64), trainable=True, buffered=True)
    tea = fo.TeaDOE(u, dn, h0, trainable=True, buffered=True)
    p = fo.ASProp(u, z)
    system = fo.ScalarSource(u) | mk | p | tea | p
    value, grads = system.withgradient(fo.SquaredIntensityDifference((u, tgt)))
    stages = [("mask", m0), ("as", z), ("tea", h0.astype(np.float64), dn), ("as", z)]
    v_ref, saved = O.pipeline_forward(u0, stages, (ds, ds), lams)
    vals, dv = O.sq_intensity_diff_values(v_ref, tgt.astype(np.float64))
    _, (gm_ref, gh_ref) = O.pipeline_backward(dv, stages, (ds, ds), lams, saved)
    assert abs(float(value) - vals.sum()) <= FIELD_TOL * vals.sum()
    torch.cuda.synchronize()
    assert rel_l2(grads[mk].cpu().numpy().astype(np.complex128), gm_ref) <= GRAD_TOL
    assert rel_l2(grads[tea].cpu().numpy(), gh_ref) <= GRAD_TOL
