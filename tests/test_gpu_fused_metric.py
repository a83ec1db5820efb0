"""The fused [Phase ->] propagator -> SquaredIntensityDifference block (fh_sqint_block) against
the same system run component by component (fuse=False) and against the oracle chain."""

import math

import numpy as np
import pytest
import torch

import fluxoptics as fo
from fluxoptics import inputs
from oracle import fluxoptics_oracle as O
from _util import c64, rel_l2, speckle_batch, to_dev

pytestmark = pytest.mark.gpu


def _compare(system, metric, comps, weights=None):
    v1, g1 = system.withgradient(metric, weights=weights, fuse=True)
    g1 = {c: g.clone() for c, g in g1.items()}
    v0, g0 = system.withgradient(metric, weights=weights, fuse=False)
    torch.cuda.synchronize()
    assert abs(float(v1) - float(v0)) <= 1e-6 * abs(float(v0))
    for c in comps:
        assert rel_l2(g1[c].cpu().numpy(), g0[c].cpu().numpy()) <= 1e-5


@pytest.mark.parametrize("kind,n,tail_phase", [("as", 1024, True), ("as", 256, False), ("rs", 256, True),
                                                ("rs", 128, False), ("as", 4096, True), ("rs", 1024, True)])
def test_fused_tail_matches_unfused(kind, n, tail_phase):
    print(f"[fused metric] {kind} n={n} phase={tail_phase}", flush=True)
    ds, lams = (1.0, 1.0), [0.8, 0.6, 0.5]
    u0 = c64(speckle_batch(3, n, n, 1.0, 1.0, 0.12 * n, 151))
    tgt = (np.abs(c64(speckle_batch(3, n, n, 1.0, 1.0, 0.12 * n, 155))) ** 2).astype(np.float32)
    u = fo.ScalarField(to_dev(u0), ds, lams)
    prop = (lambda z: fo.ASProp(u, z)) if kind == "as" else (lambda z: fo.RSProp(u, z))
    p1 = fo.Phase(u, inputs.uniform(152, (n, n), 0, 2 * math.pi).astype(np.float32), trainable=True)
    p2 = fo.Phase(u, inputs.uniform(153, (n, n), 0, 2 * math.pi).astype(np.float32), trainable=tail_phase)
    system = fo.ScalarSource(u) | p1 | prop(400.0 * n / 256) | p2 | prop(300.0 * n / 256)
    if not tail_phase:
        system = fo.ScalarSource(u) | p1 | prop(400.0 * n / 256) | prop(300.0 * n / 256)
    metric = fo.SquaredIntensityDifference((u, tgt))
    _compare(system, metric, [p1, p2] if tail_phase else [p1], weights=[1.0, -0.5, 2.0])


def test_fused_tail_single_field_spatial_target_and_oracle():
    """Batch 1 with a 2-D target (the spatial-only form, fusable for one field); the fused
    loss and gradient against the oracle chain."""
    n, ds, lam, z = 512, (0.5, 0.5), 0.8, 600.0
    print("[fused metric] single field vs oracle", flush=True)
    u0 = c64(speckle_batch(1, n, n, 0.5, 0.5, 40.0, 157))[0]
    phi = inputs.uniform(158, (n, n), 0, 2 * math.pi).astype(np.float32)
    tgt = (np.abs(c64(speckle_batch(1, n, n, 0.5, 0.5, 40.0, 159))[0]) ** 2).astype(np.float32)
    u = fo.ScalarField(to_dev(u0), ds, lam)
    ph = fo.Phase(u, phi, trainable=True, buffered=True)
    system = fo.ScalarSource(u) | ph | fo.ASProp(u, z)
    metric = fo.SquaredIntensityDifference((u, tgt))
    stages = [("phase", phi.astype(np.float64)), ("as", z)]
    v_ref, saved = O.pipeline_forward(u0, stages, ds, lam)
    vals_ref, dv_ref = O.sq_intensity_diff_values(v_ref, tgt.astype(np.float64))
    _, (g_ref,) = O.pipeline_backward(dv_ref, stages, ds, lam, saved)
    value, grads = system.withgradient(metric)
    assert abs(float(value) - vals_ref.sum()) <= 1e-5 * vals_ref.sum()
    torch.cuda.synchronize()
    assert rel_l2(ph.grad.cpu().numpy(), g_ref) <= 1e-4


def test_captured_fused_tail_replays():
    n = 256
    print("[fused metric] captured replays", flush=True)
    u0 = c64(speckle_batch(2, n, n, 1.0, 1.0, 30.0, 161))
    tgt = (np.abs(c64(speckle_batch(2, n, n, 1.0, 1.0, 30.0, 163))) ** 2).astype(np.float32)
    u = fo.ScalarField(to_dev(u0), (1.0, 1.0), [0.8, 0.7])
    ph = fo.Phase(u, inputs.uniform(162, (n, n), 0, 2 * math.pi).astype(np.float32), trainable=True,
                  buffered=True)
    system = fo.ScalarSource(u) | ph | fo.RSProp(u, 500.0)
    metric = fo.SquaredIntensityDifference((u, tgt))
    v0, g0 = system.withgradient(metric)
    v0, g0 = float(v0), g0[ph].clone()
    cap = system.capture_withgradient(metric)
    for _ in range(3):
        v, g = cap.replay()
        torch.cuda.synchronize()
        assert float(v) == v0
        assert torch.equal(g[ph], g0)
        torch.zeros(1, device="cuda").sum()  # eager work between replays


@pytest.mark.parametrize("kind", ["as", "rs"])
def test_block_reading_the_source_equals_block_on_a_copy(kind):
    """fh_sqint_block_src (the first pass reads the source field; no copy of it, no saved mask
    input) against fh_sqint_block on a copy with a saved input: loss, cotangent and mask gradient
    bit for bit; with the cotangent switched off, the same loss and mask gradient.  The source stays
    unchanged, and a field overlapping it is refused."""
    from fluxoptics import _lib
    from fluxoptics.components import _ptr
    n, b = (256, 3) if kind == "as" else (128, 2)
    print(f"[fused metric from source] {kind} n={n} batch={b}", flush=True)
    u0 = c64(speckle_batch(b, n, n, 1.0, 1.0, 0.1 * n, 171))
    tgt = torch.from_numpy((np.abs(c64(speckle_batch(b, n, n, 1.0, 1.0, 0.1 * n, 173))) ** 2).astype(np.float32)).cuda()
    src = fo.ScalarField(to_dev(u0), (1.0, 1.0), [0.8, 0.6, 0.5][:b])
    prop = fo.ASProp(src, 300.0) if kind == "as" else fo.RSProp(src, 300.0)
    plan = prop.plan(src)
    phi = torch.from_numpy(inputs.uniform(172, (n, n), 0, 2 * math.pi).astype(np.float32)).cuda()
    dc = torch.tensor([1.0, -0.5, 2.0][:b], dtype=torch.float64, device="cuda")

    def run(from_src, cotangent=1):
        u = torch.full_like(src.electric, complex(float("nan"), 0.0)) if from_src else src.electric.clone()
        loss = torch.empty(b, dtype=torch.float64, device="cuda")
        dphi = torch.empty_like(phi)
        if from_src:
            _lib.call("fh_sqint_block_src", plan, _ptr(src.electric), _ptr(u), _ptr(tgt), b, _ptr(dc), _ptr(loss),
                      1, _ptr(phi), _ptr(dphi), 0, cotangent)
        else:
            saved = torch.empty_like(u)
            _lib.call("fh_sqint_block", plan, _ptr(u), _ptr(tgt), b, _ptr(dc), _ptr(loss), 1, _ptr(phi),
                      _ptr(saved), _ptr(dphi), 0)
        return u, loss, dphi

    u1, l1, g1 = run(True)
    u0d, l0, g0 = run(False)
    _, l2, g2 = run(True, cotangent=0)
    torch.cuda.synchronize()
    assert torch.equal(l1, l0) and torch.equal(u1, u0d) and torch.equal(g1, g0)
    assert torch.equal(l2, l0) and torch.equal(g2, g0)
    assert torch.equal(src.electric.cpu(), torch.from_numpy(u0))
    with pytest.raises(_lib.FluxHipError, match="overlap"):
        _lib.call("fh_sqint_block_src", plan, _ptr(src.electric), _ptr(src.electric[1:]), _ptr(tgt), b, _ptr(dc),
                  _ptr(torch.empty(b, dtype=torch.float64, device="cuda")), 1, _ptr(phi), None, 0, 1)


def test_block_reading_the_source_2048_against_oracle():
    """The block right after the source on 2 x 2048^2 (its mask-gradient pass takes the blocked
    spectrum as whole 128-byte lines through LDS): per-field losses, mask gradient and cotangent
    against the FP64 oracle chain; with the cotangent switched off (the pass then writes no field),
    the same losses and mask gradient bit for bit."""
    from fluxoptics import _lib
    from fluxoptics.components import _ptr
    n, b, ds, lam, z = 2048, 2, (0.5, 0.5), 0.8, 3000.0
    print(f"[fused metric from source] as n={n} batch={b} vs oracle", flush=True)
    u0 = c64(speckle_batch(b, n, n, 0.5, 0.5, 150.0, 191))
    tgt = (np.abs(c64(speckle_batch(b, n, n, 0.5, 0.5, 150.0, 193))) ** 2).astype(np.float32)
    phi = inputs.uniform(192, (n, n), 0, 2 * math.pi).astype(np.float32)
    src = fo.ScalarField(to_dev(u0), ds, lam)
    prop = fo.ASProp(src, z)  # owns the plan
    plan = prop.plan(src)
    tg, ph_d = torch.from_numpy(tgt).cuda(), torch.from_numpy(phi).cuda()

    def run(cotangent):
        u = torch.empty_like(src.electric)
        loss = torch.empty(b, dtype=torch.float64, device="cuda")
        dphi = torch.empty_like(ph_d)
        _lib.call("fh_sqint_block_src", plan, _ptr(src.electric), _ptr(u), _ptr(tg), b, None, _ptr(loss), 1,
                  _ptr(ph_d), _ptr(dphi), 0, cotangent)
        return u, loss, dphi

    u1, l1, g1 = run(1)
    _, l0, g0 = run(0)
    torch.cuda.synchronize()
    assert torch.equal(l1, l0) and torch.equal(g1, g0)
    stages = [("phase", phi.astype(np.float64)), ("as", z)]
    v_ref, saved = O.pipeline_forward(u0, stages, ds, lam)
    vals_ref, dv_ref = O.sq_intensity_diff_values(v_ref, tgt.astype(np.float64))
    du_ref, (g_ref,) = O.pipeline_backward(dv_ref, stages, ds, lam, saved)
    errs = (np.abs(l1.cpu().numpy() - vals_ref).max() / vals_ref.max(), rel_l2(g1.cpu().numpy(), g_ref),
            rel_l2(u1.cpu().numpy(), du_ref))
    print(f"[fused metric from source] loss {errs[0]:.2e} mask gradient {errs[1]:.2e} cotangent {errs[2]:.2e}",
          flush=True)
    assert errs[0] <= 1e-5 and errs[1] <= 1e-4 and errs[2] <= 1e-4


def test_trainable_source_through_the_block_reading_it():
    """A trainable ScalarSource right before the fused block (which reads the source field in
    place of a copy): the source's gradient and the mask gradient equal the component-by-component
    path, the source field is unchanged, and an input field given in place of a source (inplace off)
    takes the same path."""
    n, b = 256, 2
    print("[fused metric] trainable source", flush=True)
    u0 = c64(speckle_batch(b, n, n, 1.0, 1.0, 30.0, 181))
    tgt = (np.abs(c64(speckle_batch(b, n, n, 1.0, 1.0, 30.0, 183))) ** 2).astype(np.float32)
    u = fo.ScalarField(to_dev(u0), (1.0, 1.0), [0.8, 0.7])
    src = fo.ScalarSource(u, trainable=True, buffered=True)
    ph = fo.Phase(u, inputs.uniform(182, (n, n), 0, 2 * math.pi).astype(np.float32), trainable=True,
                  buffered=True)
    prop = fo.ASProp(u, 500.0)
    metric = fo.SquaredIntensityDifference((u, tgt))
    system = src | ph | prop
    v1, g1 = system.withgradient(metric)
    g1 = {c: g.clone() for c, g in g1.items()}
    v0, g0 = system.withgradient(metric, fuse=False)
    torch.cuda.synchronize()
    assert abs(float(v1) - float(v0)) <= 1e-6 * abs(float(v0))
    assert rel_l2(g1[src].cpu().numpy(), g0[src].cpu().numpy()) <= 1e-5
    assert rel_l2(g1[ph].cpu().numpy(), g0[ph].cpu().numpy()) <= 1e-5
    assert torch.equal(src.u0.electric.cpu(), torch.from_numpy(u0))
    free = ph | prop
    x = fo.ScalarField(to_dev(u0), (1.0, 1.0), [0.8, 0.7])
    v2, g2 = free.withgradient(metric, x)
    torch.cuda.synchronize()
    assert abs(float(v2) - float(v1)) <= 1e-6 * abs(float(v1))
    assert rel_l2(g2[ph].cpu().numpy(), g1[ph].cpu().numpy()) <= 1e-6
    assert torch.equal(x.electric.cpu(), torch.from_numpy(u0))


def test_trailing_phase_is_skipped_by_the_fused_block():
    """A system ending in a Phase (the RGB cascade's shape, Phi -> AS -> Phi -> AS -> Phi): the
    trailing Phase leaves |u|^2 unchanged, so the fused run skips it; loss and the other mask
    gradients equal the component-by-component run, and the trailing mask's gradient is exactly
    zero (rounding-level in the unfused run)."""
    n = 256
    print("[fused metric] trailing phase", flush=True)
    ds, lams = (1.0, 1.0), [0.64, 0.538, 0.455]
    u0 = c64(speckle_batch(3, n, n, 1.0, 1.0, 30.0, 211))
    tgt = (np.abs(c64(speckle_batch(3, n, n, 1.0, 1.0, 30.0, 212))) ** 2).astype(np.float32)
    u = fo.ScalarField(to_dev(u0), ds, lams)
    ph = [fo.Phase(u, inputs.uniform(213 + k, (n, n), 0, 2 * math.pi).astype(np.float32), trainable=True)
          for k in range(3)]
    system = fo.ScalarSource(u) | ph[0] | fo.ASProp(u, 600.0) | ph[1] | fo.ASProp(u, 600.0) | ph[2]
    metric = fo.SquaredIntensityDifference((u, tgt))
    v1, g1 = system.withgradient(metric)
    g1 = {c: g.clone() for c, g in g1.items()}
    v0, g0 = system.withgradient(metric, fuse=False)
    torch.cuda.synchronize()
    assert abs(float(v1) - float(v0)) <= 1e-6 * abs(float(v0))
    for c in ph[:2]:
        assert rel_l2(g1[c].cpu().numpy(), g0[c].cpu().numpy()) <= 1e-5
    scale = float(g0[ph[0]].abs().max())
    assert float(g1[ph[2]].abs().max()) == 0.0 and float(g0[ph[2]].abs().max()) <= 1e-4 * scale
