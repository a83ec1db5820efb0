"""GPU parity of AS_BPM (forward, adjoint, dn gradient, checkpointing) against the oracle."""

import math

import numpy as np
import pytest
import torch

import fluxoptics as fo
from fluxoptics import inputs
from oracle import fluxoptics_oracle as O
from _util import c64, rel_l2, random_field, to_dev, to_host

pytestmark = pytest.mark.gpu


def smooth_dn(ns, n, sigma, peak, seed, nx=None):
    """Gaussian-filtered white noise (periodic, FFT filter) scaled to max |dn| = peak; the
    grid is n x n, or n rows of nx points."""
    return inputs.smooth_noise(ns, n, n if nx is None else nx, sigma, peak, seed)


def tilted_gaussian(n, ds, lam, w0, deg):
    x = inputs.grid(n, ds)[None, :]
    return inputs.gaussian(n, n, ds, ds, w0) * np.exp(2j * math.pi * math.sin(math.radians(deg)) * x / lam)


def test_bpm_zero_dn_is_asprop():
    n, ds, lam, n0, L, ns = 128, (0.5, 0.5), 1.0, 1.5, 40.0, 8
    u0 = c64(random_field((2, n, n), 1))
    u = fo.ScalarField(to_dev(u0), ds, lam)
    bpm = fo.AS_BPM(u, L, n0, np.zeros((ns, n, n), np.float32))
    v = bpm.propagate(u, fo.Forward)
    assert rel_l2(to_host(v.electric), O.asprop(u0, ds, lam, L, n0)) <= 1e-5


@pytest.mark.parametrize("n,ns,batch,every", [(64, 6, 2, 1), (64, 6, 2, 4), (256, 12, 1, 5)])
def test_bpm_forward_adjoint_gradient(n, ns, batch, every):
    print(f"[bpm parity] n={n} slices={ns} batch={batch} every={every}", flush=True)
    ds, n0, L = (0.5, 0.5), 1.5, 3.0 * ns
    lams = [1.0, 0.9][:batch]
    dn = smooth_dn(ns, n, 2.0, 0.01, 4)
    u0 = c64(np.stack([tilted_gaussian(n, 0.5, 1.0, 0.15 * n * 0.5, 2.0)] * batch))
    t0 = c64(0.01 * random_field((batch, n, n), 9))
    u = fo.ScalarField(to_dev(u0), ds, lams)
    bpm = fo.AS_BPM(u, L, n0, dn, trainable=True, buffered=True, checkpoint_every=every)
    bpm.propagate_and_save_(u, fo.Forward)
    v = to_host(u.electric)
    vref, saved = O.bpm_forward(u0, dn.astype(np.float64), L, n0, lams, ds)
    assert rel_l2(v, vref) <= 1e-5
    dv0 = c64(v - t0)
    du = fo.ScalarField(to_dev(dv0), ds, lams)
    bpm.backpropagate_with_gradient_(du, None, fo.Forward)
    duref, ddnref = O.bpm_backward(dv0, dn.astype(np.float64), L, n0, lams, ds, saved)
    assert rel_l2(to_host(du.electric), duref) <= 1e-5
    torch.cuda.synchronize()
    assert rel_l2(bpm.grad.cpu().numpy(), ddnref) <= 1e-4


def test_bpm_checkpointing_matches_full_save():
    n, ns, ds, lam, n0, L = 128, 10, (0.5, 0.5), 1.0, 1.5, 10.0
    dn = smooth_dn(ns, n, 2.0, 0.01, 5)
    u0 = c64(tilted_gaussian(n, 0.5, lam, 15.0, 2.0))
    dv0 = c64(random_field((n, n), 3))
    grads = []
    for every in (1, 3, 10):
        u = fo.ScalarField(to_dev(u0), ds, lam)
        bpm = fo.AS_BPM(u, L, n0, dn, trainable=True, buffered=True, checkpoint_every=every)
        bpm.propagate_and_save_(u, fo.Forward)
        du = fo.ScalarField(to_dev(dv0), ds, lam)
        bpm.backpropagate_with_gradient_(du, None, fo.Forward)
        torch.cuda.synchronize()
        grads.append((bpm.grad.cpu().numpy(), to_host(du.electric)))
    for g, d in grads[1:]:
        assert np.array_equal(g, grads[0][0])
        assert np.array_equal(d, grads[0][1])


def test_bpm_graded_index_volume_512x400():
    """512^2, ds=0.5, lam=1, n_bulk=1.5, 400 slices of dz=1, checkpoint every 20."""
    print("[bpm graded-index volume] 512^2 x 400 slices", flush=True)
    n, ns, ds, lam, n0 = 512, 400, (0.5, 0.5), 1.0, 1.5
    dn = smooth_dn(ns, n, 3.0, 0.01, 4)
    u0 = c64(tilted_gaussian(n, 0.5, lam, 80.0, 2.0))
    t0 = c64(0.01 * random_field((n, n), 11))
    u = fo.ScalarField(to_dev(u0), ds, lam)
    bpm = fo.AS_BPM(u, float(ns), n0, dn, trainable=True, buffered=True, checkpoint_every=20)
    torch.cuda.reset_peak_memory_stats()
    bpm.propagate_and_save_(u, fo.Forward)
    v = to_host(u.electric)
    vref, saved = O.bpm_forward(u0, dn.astype(np.float64), float(ns), n0, lam, ds)
    ev = rel_l2(v, vref)
    dv0 = c64(v - t0)
    du = fo.ScalarField(to_dev(dv0), ds, lam)
    bpm.backpropagate_with_gradient_(du, None, fo.Forward)
    duref, ddnref = O.bpm_backward(dv0, dn.astype(np.float64), float(ns), n0, lam, ds, saved)
    edu = rel_l2(to_host(du.electric), duref)
    torch.cuda.synchronize()
    eg = rel_l2(bpm.grad.cpu().numpy(), ddnref)
    ws = bpm.workspace_bytes(u)
    print(f"512^2 x 400 volume: field {ev:.2e} adjoint {edu:.2e} ddn {eg:.2e}; library workspace {ws / 2**20:.1f} MiB "
          f"(a full save of every slice input is {ns * n * n * 8 / 2**20:.0f} MiB)")
    assert ev <= 1e-5 and edu <= 1e-5 and eg <= 1e-4


@pytest.mark.parametrize("kind,every", [("tilted", 3), ("paraxial", 1), ("shift", 2)])
def test_general_bpm_forward_adjoint_gradient(kind, every):
    """Tilted and paraxial AS_BPM and Shift_BPM, per-field tilts and wavelengths (per-field
    kdz = 2 pi dz / (cos tx cos ty)), against the oracle."""
    n, ns, batch = 128, 6, 2
    print(f"[bpm {kind}] n={n} slices={ns} every={every}", flush=True)
    ds, n0, L = (0.5, 0.5), 1.5, 18.0
    lams = [1.0, 0.9]
    tilts = ([0.03, -0.02], [0.01, 0.04])
    kw = dict(tilts=tilts, paraxial=kind == "paraxial", shift=kind == "shift")
    dn = smooth_dn(ns, n, 2.0, 0.01, 6)
    u0 = c64(np.stack([tilted_gaussian(n, 0.5, 1.0, 12.0, 1.0)] * batch))
    t0 = c64(0.01 * random_field((batch, n, n), 12))
    u = fo.ScalarField(to_dev(u0), ds, lams, tilts)
    if kind == "shift":
        bpm = fo.Shift_BPM(u, L, dn, trainable=True, buffered=True, checkpoint_every=every)
    else:
        bpm = fo.AS_BPM(u, L, n0, dn, paraxial=kind == "paraxial", trainable=True, buffered=True,
                        checkpoint_every=every)
    bpm.propagate_and_save_(u, fo.Forward)
    v = to_host(u.electric)
    vref, saved = O.bpm_forward(u0, dn.astype(np.float64), L, n0, lams, ds, **kw)
    assert rel_l2(v, vref) <= 1e-5
    dv0 = c64(v - t0)
    du = fo.ScalarField(to_dev(dv0), ds, lams, tilts)
    bpm.backpropagate_with_gradient_(du, None, fo.Forward)
    duref, ddnref = O.bpm_backward(dv0, dn.astype(np.float64), L, n0, lams, ds, saved, **kw)
    assert rel_l2(to_host(du.electric), duref) <= 1e-5
    torch.cuda.synchronize()
    assert rel_l2(bpm.grad.cpu().numpy(), ddnref) <= 1e-4
    w = bpm.propagate(fo.ScalarField(to_dev(u0), ds, lams, tilts), fo.Backward)
    wref, _ = O.bpm_forward(u0, dn.astype(np.float64), L, n0, lams, ds, direction=-1, **kw)
    assert rel_l2(to_host(w.electric), wref) <= 1e-5


def test_bpm_rejects_other_tilts():
    n = 64
    u = fo.ScalarField(to_dev(c64(random_field((1, n, n), 13))), (0.5, 0.5), 1.0, (0.01, 0.0))
    bpm = fo.Shift_BPM(u, 10.0, np.zeros((3, n, n), np.float32))
    with pytest.raises(ValueError):
        bpm.propagate(fo.ScalarField(u.electric.clone(), (0.5, 0.5), 1.0, (0.02, 0.0)))


@pytest.mark.parametrize("kind,every", [("onaxis", 4), ("tilted", 1), ("shift", 3)])
def test_bpm_any_grid_size(kind, every):
    """BPM on a grid that is not a power of two (general slice propagators, elementwise slice
    phase): forward, adjoint and dn gradient against the oracle."""
    nx, ny, ns, batch = 90, 70, 8, 2
    print(f"[bpm any size] {kind} every={every}", flush=True)
    ds, n0, L = (0.5, 0.5), 1.5, 16.0
    lams = [1.0, 0.9]
    tilts = ([0.03, -0.02], [0.01, 0.04]) if kind != "onaxis" else None
    kw = dict(tilts=tilts, shift=kind == "shift")
    dn = (0.01 * inputs.normal(7, (ns, ny, nx))).astype(np.float32)
    u0 = c64(random_field((batch, ny, nx), 14))
    t0 = c64(0.01 * random_field((batch, ny, nx), 15))
    u = fo.ScalarField(to_dev(u0), ds, lams, tilts if tilts else (0.0, 0.0))
    if kind == "shift":
        bpm = fo.Shift_BPM(u, L, dn, trainable=True, buffered=True, checkpoint_every=every)
    else:
        bpm = fo.AS_BPM(u, L, n0, dn, trainable=True, buffered=True, checkpoint_every=every)
    bpm.propagate_and_save_(u, fo.Forward)
    v = to_host(u.electric)
    vref, saved = O.bpm_forward(u0, dn.astype(np.float64), L, n0, lams, ds, **kw)
    assert rel_l2(v, vref) <= 1e-5
    dv0 = c64(v - t0)
    du = fo.ScalarField(to_dev(dv0), ds, lams, tilts if tilts else (0.0, 0.0))
    bpm.backpropagate_with_gradient_(du, None, fo.Forward)
    duref, ddnref = O.bpm_backward(dv0, dn.astype(np.float64), L, n0, lams, ds, saved, **kw)
    assert rel_l2(to_host(du.electric), duref) <= 1e-5
    torch.cuda.synchronize()
    assert rel_l2(bpm.grad.cpu().numpy(), ddnref) <= 1e-4


def test_bpm_any_grid_size_long_stack():
    """A 200-slice graded-index stack on a 240 x 180 grid (padded slice propagators computed in
    FP64, c64 between launches): forward, adjoint and dn gradient against the FP64 oracle, with
    the bars of the power-of-two volume (1e-5 field and adjoint, 1e-4 gradient).  FP32 slice
    transforms repeat one rounding on every slice and end at 1.6e-5 here."""
    nx, ny, ns, every = 240, 180, 200, 10
    print(f"[bpm any size, long stack] {nx}x{ny} x {ns} slices every={every}", flush=True)
    ds, lam, n0 = (0.5, 0.5), 1.0, 1.5
    dn = smooth_dn(ns, ny, 3.0, 0.01, 8, nx=nx)
    x = inputs.grid(nx, ds[0])[None, :]
    y = inputs.grid(ny, ds[1])[:, None]
    g = np.exp(-(x ** 2 + y ** 2) / 30.0 ** 2) * np.exp(2j * math.pi * math.sin(math.radians(2.0)) * x / lam)
    u0 = c64(g[None])
    t0 = c64(0.01 * random_field((1, ny, nx), 16))
    u = fo.ScalarField(to_dev(u0), ds, lam)
    bpm = fo.AS_BPM(u, float(ns), n0, dn, trainable=True, buffered=True, checkpoint_every=every)
    bpm.propagate_and_save_(u, fo.Forward)
    v = to_host(u.electric)
    vref, saved = O.bpm_forward(u0, dn.astype(np.float64), float(ns), n0, lam, ds)
    ev = rel_l2(v, vref)
    dv0 = c64(v - t0)
    du = fo.ScalarField(to_dev(dv0), ds, lam)
    bpm.backpropagate_with_gradient_(du, None, fo.Forward)
    duref, ddnref = O.bpm_backward(dv0, dn.astype(np.float64), float(ns), n0, lam, ds, saved)
    edu = rel_l2(to_host(du.electric), duref)
    torch.cuda.synchronize()
    eg = rel_l2(bpm.grad.cpu().numpy(), ddnref)
    print(f"{nx}x{ny} x {ns} slices: field {ev:.2e} adjoint {edu:.2e} ddn {eg:.2e}", flush=True)
    assert ev <= 1e-5 and edu <= 1e-5 and eg <= 1e-4


def test_bpm_phase_tables_match_evaluated_phases():
    """The slice phases come from per-segment FP64 tables when two segments' tables fit the
    library's 512 MiB bound (512^2, 8 or 16 slices a segment) and are evaluated inside the slice
    steps otherwise (70 or 80 slices a segment: 2 x 280 MiB and more).  Both paths meet the parity
    bars against the FP64 oracle and each is bitwise reproducible across checkpoint spacings; the
    two differ from each other only at their own rounding level (measured: fields 2.6e-7 apart,
    3.5e-7 and 3.3e-7 from the oracle; the adjoint fields are bitwise equal)."""
    n, ns, ds, lam, n0 = 512, 80, (0.5, 0.5), 1.0, 1.5
    print(f"[bpm phase tables] {n}^2 x {ns} slices, every 8, 16 (tables) and 70, 80 (evaluated)", flush=True)
    dn = smooth_dn(ns, n, 3.0, 0.01, 21)
    u0 = c64(tilted_gaussian(n, 0.5, lam, 60.0, 2.0))
    dv0 = c64(0.01 * random_field((n, n), 22))
    vref, saved = O.bpm_forward(u0, dn.astype(np.float64), float(ns), n0, lam, ds)
    duref, ddnref = O.bpm_backward(dv0, dn.astype(np.float64), float(ns), n0, lam, ds, saved)
    out = {}
    for every in (8, 16, 70, 80):
        u = fo.ScalarField(to_dev(u0), ds, lam)
        bpm = fo.AS_BPM(u, float(ns), n0, dn, trainable=True, buffered=True, checkpoint_every=every)
        bpm.propagate_and_save_(u, fo.Forward)
        v = to_host(u.electric)
        du = fo.ScalarField(to_dev(dv0), ds, lam)
        bpm.backpropagate_with_gradient_(du, None, fo.Forward)
        torch.cuda.synchronize()
        out[every] = (v, to_host(du.electric), bpm.grad.cpu().numpy())
        ev, ed, eg = rel_l2(v, vref), rel_l2(out[every][1], duref), rel_l2(out[every][2], ddnref)
        print(f"every {every}: field {ev:.2e} adjoint {ed:.2e} ddn {eg:.2e}", flush=True)
        assert ev <= 1e-5 and ed <= 1e-5 and eg <= 1e-4
    for a, b in ((8, 16), (70, 80)):
        for x, y in zip(out[a], out[b]):
            assert np.array_equal(x, y)
    v1, d1, g1 = out[8]
    v2, d2, g2 = out[80]
    print(f"tables vs evaluated phases: field {rel_l2(v1, v2):.2e} adjoint {rel_l2(d1, d2):.2e} "
          f"ddn {rel_l2(g1, g2):.2e}")
