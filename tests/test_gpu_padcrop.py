"""Pad / crop (pad_crop_operators.jl) and FieldProbe (field_probe.jl) on the device: the
reference's own testsets (test/optical_components_test.jl:171-236) restated, pad/crop against
numpy slicing, and a padded system differentiated against the oracle chain."""

import math

import numpy as np
import pytest
import torch

import fluxoptics as fo
from fluxoptics import inputs
from oracle import fluxoptics_oracle as O
from _util import c64, rel_l2, random_field, to_dev, to_host

pytestmark = pytest.mark.gpu


def test_reference_pad_and_crop_testset():
    """optical_components_test.jl:171-189: ones(32, 32) padded to 64 x 64 and cropped back."""
    u = to_dev(np.ones((32, 32)))
    up = fo.pad(u, (64, 64))
    assert tuple(up.shape) == (64, 64)
    uc = fo.crop(up, (32, 32))
    assert tuple(uc.shape) == (32, 32)
    off = (64 - 32) // 2
    assert to_host(up)[off, off] == 1.0
    assert np.array_equal(to_host(uc), np.ones((32, 32), np.complex64))


@pytest.mark.parametrize("off", [(0, 0), (5, 3), (20, 7)])
def test_pad_crop_match_slicing(off):
    """pad places u at (x, y) = offset with pad_val around it; crop cuts the region back out; a
    round trip returns u bit for bit.  Non-square fields, a batch of three."""
    print(f"[pad/crop] offset={off}", flush=True)
    u0 = c64(random_field((3, 24, 40), 191))
    u = to_dev(u0)
    ox, oy = off
    up = fo.pad(u, (64, 48), offset=off, pad_val=0.5 - 0.25j)
    ref = np.full((3, 48, 64), 0.5 - 0.25j, np.complex64)
    ref[:, oy:oy + 24, ox:ox + 40] = u0
    assert np.array_equal(to_host(up), ref)
    assert np.array_equal(to_host(fo.crop(up, (40, 24), offset=off)), u0)
    part = fo.crop(u, (16, 10), offset=(ox // 2, oy // 2))
    assert np.array_equal(to_host(part), u0[:, oy // 2:oy // 2 + 10, ox // 2:ox // 2 + 16])
    with pytest.raises(fo._lib.FluxHipError, match="overlap"):
        fo.crop_(u[1:3], u[0:2], offset=(0, 0))
    with pytest.raises(ValueError, match="contiguous"):
        fo.crop(u[:, :, 1:], (8, 8))


def test_reference_field_probe_testset():
    """optical_components_test.jl:212-236: probes before and after a pi/2 Phase."""
    n = 64
    u0 = fo.ScalarField(to_dev(np.ones((n, n))), (1.0, 1.0), 1.0)
    p1, p2 = fo.FieldProbe(), fo.FieldProbe()
    ph = fo.Phase(u0, np.full((n, n), math.pi / 2, np.float32))
    out = (fo.ScalarSource(u0) | p1 | ph | p2)()
    assert p1 in out.probes and p2 in out.probes
    assert np.array_equal(to_host(out.probes[p1].electric), np.ones((n, n), np.complex64))
    assert np.allclose(np.angle(to_host(out.probes[p2].electric)), math.pi / 2, atol=1e-6)
    assert np.allclose(to_host(out.out.electric), to_host(out.probes[p2].electric))


@pytest.mark.parametrize("store_ref", [False, True])
def test_padded_system_gradient_against_oracle(store_ref):
    """source | PadCropOperator | Phase | ASProp | adjoint(PadCropOperator) with the intensity
    loss on the unpadded grid (the reference's pad_crop_operators.jl:281-287 pattern): loss and
    mask gradient against the oracle chain on the zero-padded field, cropped back at offset 0."""
    n, m, ds, lam, z = 96, 256, (1.0, 1.0), [0.8, 0.6], 700.0
    print(f"[padded system] {n} -> {m}, store_ref={store_ref}", flush=True)
    u0 = c64(inputs.gaussian(n, n, 1.0, 1.0, 20.0))[None].repeat(2, axis=0)
    tgt = (np.abs(c64(inputs.gaussian(n, n, 1.0, 1.0, 15.0, 10.0, -5.0))) ** 2).astype(np.float32)
    tgt = np.stack([tgt, 0.5 * tgt])
    phi = inputs.uniform(193, (m, m), 0, 2 * math.pi).astype(np.float32)
    u = fo.ScalarField(to_dev(u0), ds, lam)
    u_tmp = torch.zeros(2, m, m, dtype=torch.complex64, device="cuda")
    pad_op = fo.PadCropOperator(u, u_tmp, store_ref=store_ref)
    up = u.set_field_data(u_tmp)
    ph = fo.Phase(up, phi, trainable=True, buffered=True)
    system = fo.ScalarSource(u) | pad_op | ph | fo.ASProp(up, z) | pad_op.adjoint()
    metric = fo.SquaredIntensityDifference((u, tgt))
    value, grads = system.withgradient(metric)
    torch.cuda.synchronize()
    up0 = np.zeros((2, m, m), np.complex128)
    up0[:, :n, :n] = u0
    stages = [("phase", phi.astype(np.float64)), ("as", z)]
    vp, saved = O.pipeline_forward(up0, stages, ds, lam)
    vals, dv = O.sq_intensity_diff_values(vp[:, :n, :n], tgt.astype(np.float64))
    dvp = np.zeros_like(vp)
    dvp[:, :n, :n] = dv
    _, (g_ref,) = O.pipeline_backward(dvp, stages, ds, lam, saved)
    assert abs(float(value) - vals.sum()) <= 1e-5 * vals.sum()
    assert rel_l2(grads[ph].cpu().numpy(), g_ref) <= 1e-4
    out = system()
    assert tuple(out.out.electric.shape) == (2, n, n)
    assert rel_l2(to_host(out.out.electric), vp[:, :n, :n]) <= 1e-5
