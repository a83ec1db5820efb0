"""GPU parity: the SquaredIntensityDifference kernel, and whole systems differentiated on the
device (the beam-splitter pipeline of RSProps and Phases, and the three-wavelength Phase/ASProp
cascade) against the FP64 oracle.  Tolerances: 1e-5 fields and losses, 1e-4 gradients."""

import math
import warnings

import numpy as np
import pytest
import torch

import fluxoptics as fo
from fluxoptics import inputs
from oracle import fluxoptics_oracle as O

from _util import c64, rel_l2, to_dev, to_host

pytestmark = pytest.mark.gpu

FIELD_TOL, GRAD_TOL = 1e-5, 1e-4


def test_sqint_reference_doctest_values_on_device():
    """``squared_intensity_difference.jl:19-45`` doctests through the device kernel."""
    u = fo.ScalarField(to_dev(np.ones((2, 4, 4))), (1.0, 1.0), 1.064)
    m = fo.SquaredIntensityDifference((u, np.zeros((2, 4, 4))))
    (vals,) = m(u)
    assert vals.cpu().tolist() == [16.0, 16.0]
    u3 = fo.ScalarField(to_dev(np.ones((3, 4, 4))), (1.0, 1.0), 1.064)
    m3 = fo.SquaredIntensityDifference((u3, 3 * np.ones((4, 4))))
    (vals,) = m3(u3)
    assert vals.cpu().tolist() == [0.0]


@pytest.mark.parametrize("spatial", [False, True])
def test_sqint_values_and_pullback(spatial):
    b, ny, nx = 3, 192, 256
    print(f"[sqint] spatial={spatial}", flush=True)
    v0 = c64(inputs.normal(5, (b, ny, nx)) + 1j * inputs.normal(6, (b, ny, nx)))
    t = inputs.uniform(7, (ny, nx) if spatial else (b, ny, nx), 0.0, 3.0).astype(np.float32)
    dc = np.array([0.7]) if spatial else np.array([0.5, 1.0, 2.0])
    u = fo.ScalarField(to_dev(v0), (1.0, 1.0), 0.8)
    m = fo.SquaredIntensityDifference((u, t))
    vals_ref, dv_ref = O.sq_intensity_diff_values(v0, t.astype(np.float64), dc)
    (vals,) = m(u)
    assert np.allclose(vals.cpu().numpy(), vals_ref, rtol=1e-12, atol=0)
    (du,) = m.backpropagate_metric([u], [torch.from_numpy(dc)])
    assert rel_l2(to_host(du.electric), dv_ref) <= FIELD_TOL
    # repeated calls give identical sums (fixed reduction order)
    (vals2,) = m(u)
    assert torch.equal(vals, vals2)


def _two_gaussians(n, ds, w0, shift):
    t = inputs.gaussian(n, n, ds, ds, w0, -shift, 0.0) + inputs.gaussian(n, n, ds, ds, w0, shift, 0.0)
    return inputs.normalize_power(t, ds, ds)


@pytest.mark.parametrize("masks", ["zeros", "random"])
def test_beam_splitter_pipeline_gradient(masks):
    """Gaussian source (waist 30 um) on 512x512, ds 1 um, lambda 1.064 um, power-normalised;
    RSProp(1500) -> Phase -> RSProp(2000) -> Phase -> RSProp(1500); loss sum of squared
    intensity differences against two Gaussians at +-60 um; gradients of both masks.
    The oracle pads to 2n-1 as the reference does; the device pads to 2n."""
    n, ds, lam = 512, 1.0, 1.064
    print(f"[beam splitter] masks={masks}", flush=True)
    u0 = c64(inputs.normalize_power(inputs.gaussian(n, n, ds, ds, 30.0), ds, ds))
    target_i = (np.abs(_two_gaussians(n, ds, 30.0, 60.0)) ** 2).astype(np.float32)
    if masks == "zeros":
        phi1 = phi2 = np.zeros((n, n), np.float32)
    else:
        phi1 = inputs.uniform(41, (n, n), 0, 2 * math.pi).astype(np.float32)
        phi2 = inputs.uniform(42, (n, n), 0, 2 * math.pi).astype(np.float32)
    u = fo.ScalarField(to_dev(u0), (ds, ds), lam)
    doe1 = fo.Phase(u, phi1, trainable=True, buffered=True)
    doe2 = fo.Phase(u, phi2, trainable=True, buffered=True)
    prop1, prop2 = fo.RSProp(u, 1500.0), fo.RSProp(u, 2000.0)
    system = fo.ScalarSource(u) | prop1 | doe1 | prop2 | doe2 | prop1
    metric = fo.SquaredIntensityDifference((u, target_i))

    stages = [("rs", 1500.0), ("phase", phi1.astype(np.float64)), ("rs", 2000.0),
              ("phase", phi2.astype(np.float64)), ("rs", 1500.0)]
    v_ref, saved = O.pipeline_forward(u0, stages, (ds, ds), lam)
    l_ref, dv_ref = O.sq_intensity_diff(v_ref, target_i.astype(np.float64))
    _, (g1_ref, g2_ref) = O.pipeline_backward(dv_ref, stages, (ds, ds), lam, saved)

    out = system().out
    assert rel_l2(to_host(out.electric), v_ref) <= FIELD_TOL
    value, grads = system.withgradient(metric)
    assert abs(float(value) - l_ref) <= FIELD_TOL * abs(l_ref)
    torch.cuda.synchronize()
    assert set(grads) == {doe1, doe2}
    assert grads[doe1] is doe1.grad and grads[doe2] is doe2.grad
    assert rel_l2(doe1.grad.cpu().numpy(), g1_ref) <= GRAD_TOL
    assert rel_l2(doe2.grad.cpu().numpy(), g2_ref) <= GRAD_TOL


def test_rgb_cascade_gradients():
    """1024x1024, ds 1 um, lambda in {0.640, 0.538, 0.455} um batched as one field of three
    slices, Gaussian source (waist 100 um); Phase -> ASProp(1e4) -> Phase -> ASProp(1e4) ->
    Phase with masks U[0, 2pi) (seeds 31..33); full-dims intensity loss against three
    Gaussians at (-200, 0), (0, 0), (200, 0); all three mask gradients (summed over lambda)."""
    n, ds, lams, z = 1024, 1.0, [0.640, 0.538, 0.455], 1e4
    print("[rgb cascade] 1024^2 x 3", flush=True)
    g = inputs.gaussian(n, n, ds, ds, 100.0)
    u0 = c64(inputs.normalize_power(np.stack([g, g, g]), ds, ds))
    tgt = np.stack([np.abs(inputs.normalize_power(inputs.gaussian(n, n, ds, ds, 100.0, c, 0.0), ds, ds)) ** 2
                    for c in (-200.0, 0.0, 200.0)]).astype(np.float32)
    phis = [inputs.uniform(31 + k, (n, n), 0, 2 * math.pi).astype(np.float32) for k in range(3)]
    u = fo.ScalarField(to_dev(u0), (ds, ds), lams)
    does = [fo.Phase(u, p, trainable=True, buffered=True) for p in phis]
    prop = fo.ASProp(u, z)
    system = fo.ScalarSource(u) | does[0] | prop | does[1] | prop | does[2]
    metric = fo.SquaredIntensityDifference((u, tgt))

    stages = [("phase", phis[0].astype(np.float64)), ("as", z), ("phase", phis[1].astype(np.float64)),
              ("as", z), ("phase", phis[2].astype(np.float64))]
    v_ref, saved = O.pipeline_forward(u0, stages, (ds, ds), lams)
    vals_ref, dv_ref = O.sq_intensity_diff_values(v_ref, tgt.astype(np.float64))
    _, g_ref = O.pipeline_backward(dv_ref, stages, (ds, ds), lams, saved)

    value, grads = system.withgradient(metric)
    assert abs(float(value) - vals_ref.sum()) <= FIELD_TOL * vals_ref.sum()
    torch.cuda.synchronize()
    for k in range(2):
        assert rel_l2(does[k].grad.cpu().numpy(), g_ref[k]) <= GRAD_TOL, k
    # the last mask does not change any intensity: its gradient is zero up to rounding
    # (the oracle's is ~1e-27), so it is checked against the scale of the others
    scale = np.linalg.norm(g_ref[0])
    assert np.linalg.norm(g_ref[2]) <= 1e-12 * scale
    assert np.linalg.norm(does[2].grad.cpu().numpy()) <= GRAD_TOL * scale


def test_merged_system_matches_unmerged():
    """``merge_components=true``: Phase+Phase and ASProp+ASProp merge without changing the
    output (beyond FP32 rounding)."""
    n, ds, lam = 256, 1.0, 0.8
    u0 = c64(inputs.gaussian(n, n, ds, ds, 30.0) * np.exp(1j * inputs.uniform(9, (n, n), 0, 1)))
    u = fo.ScalarField(to_dev(u0), (ds, ds), lam)
    a = fo.Phase(u, inputs.uniform(10, (n, n), 0, 1).astype(np.float32))
    b = fo.Phase(u, inputs.uniform(11, (n, n), 0, 1).astype(np.float32))
    p, q = fo.ASProp(u, 300.0), fo.ASProp(u, 700.0)
    s = a | b | p | q
    sm = s | {"merge_components": True}
    assert len(sm.components) == 2
    out = to_host(s(u).out.electric)
    outm = to_host(sm(u).out.electric)
    assert rel_l2(outm, out) <= 2 * FIELD_TOL
    ref = O.asprop(O.phase_apply(u0, (a.phi + b.phi).cpu().numpy().astype(np.float64), 1), (ds, ds), lam,
                   1000.0, 1.0, 1)
    assert rel_l2(outm, ref) <= FIELD_TOL


def test_captured_withgradient_replays_match_eager():
    """A recorded withgradient replays bit for bit like the eager calls, with eager work and
    synchronisations between replays and the masks changed in place between them."""
    n, ds, lam = 256, 1.0, 1.064
    print("[captured withgradient] 256^2 RS/Phase system", flush=True)
    u0 = c64(inputs.normalize_power(inputs.gaussian(n, n, ds, ds, 20.0), ds, ds))
    tgt = (np.abs(_two_gaussians(n, ds, 20.0, 30.0)) ** 2).astype(np.float32)
    u = fo.ScalarField(to_dev(u0), (ds, ds), lam)
    doe1 = fo.Phase(u, inputs.uniform(61, (n, n), 0, 1).astype(np.float32), trainable=True, buffered=True)
    doe2 = fo.Phase(u, inputs.uniform(62, (n, n), 0, 1).astype(np.float32), trainable=True, buffered=True)
    p1, p2 = fo.RSProp(u, 900.0), fo.ASProp(u, 700.0)
    system = fo.ScalarSource(u) | p1 | doe1 | p2 | doe2 | p1
    metric = fo.SquaredIntensityDifference((u, tgt))
    step = system.capture_withgradient(metric)
    for rep in range(3):
        with torch.no_grad():
            doe1.phi.add_(0.05 * rep)
        _ = torch.randn(4096, device="cuda").sum()
        torch.cuda.synchronize()
        value, grads = step.replay()
        torch.cuda.synchronize()
        g1, g2, v = grads[doe1].clone(), grads[doe2].clone(), float(value)
        ev, eg = system.withgradient(metric)
        torch.cuda.synchronize()
        assert v == float(ev), rep
        assert torch.equal(g1, eg[doe1]) and torch.equal(g2, eg[doe2]), rep


def test_dot_product_and_power_coupling_on_device():
    """Values against the oracle (complex FP64 sums of FP32 data) and the pullbacks; the
    reference's doctest values (1 per normalised mode)."""
    from test_system_host import _doc_modes
    u = fo.ScalarField(to_dev(_doc_modes()), (1.0, 1.0), 1.064)
    (p,) = fo.PowerCoupling(u)(u)
    assert np.allclose(p.cpu().numpy(), [1.0, 1.0], atol=1e-7)
    (c,) = fo.DotProduct(u)(u)
    assert np.allclose(c.cpu().numpy(), [1.0, 1.0], atol=1e-7)
    b, ny, nx, ds = 3, 128, 256, (0.5, 0.5)
    u0 = c64(inputs.normal(21, (b, ny, nx)) + 1j * inputs.normal(22, (b, ny, nx)))
    v0 = c64(inputs.normal(23, (b, ny, nx)) + 1j * inputs.normal(24, (b, ny, nx)))
    u = fo.ScalarField(to_dev(u0), ds, 0.8)
    v = fo.ScalarField(to_dev(v0), ds, 0.8)
    dm = fo.DotProduct(v)
    (c,) = dm(u)
    assert np.allclose(c.cpu().numpy(), O.dot_product(u0, v0, ds), rtol=1e-10, atol=0)
    dc = np.array([0.3 - 0.7j, -1.1 + 0.2j, 2.0 + 0.0j])
    (du,) = dm.backpropagate_metric([u], [torch.from_numpy(dc)])
    assert rel_l2(to_host(du.electric), O.dot_product_pullback(v0, dc)) <= FIELD_TOL
    pm = fo.PowerCoupling(v)
    (p,) = pm(u)
    p_ref, c_ref, vh = O.power_coupling(u0, v0, ds)
    assert np.allclose(p.cpu().numpy(), p_ref, rtol=1e-6, atol=0)  # targets normalised in FP32
    dp = np.array([1.0, -0.5, 2.0])
    (du,) = pm.backpropagate_metric([u], [torch.from_numpy(dp)])
    assert rel_l2(to_host(du.electric), O.power_coupling_pullback(vh, c_ref, dp)) <= FIELD_TOL


def test_mode_coupling_system_gradient():
    """Maximise the power coupled into a shifted Gaussian through Phase -> ASProp: the mask
    gradient of -PowerCoupling against the oracle chain."""
    n, ds, lam, z = 256, 1.0, 0.8, 600.0
    u0 = c64(inputs.normalize_power(inputs.gaussian(n, n, ds, ds, 20.0), ds, ds))
    t0 = inputs.gaussian(n, n, ds, ds, 15.0, 25.0, -10.0)
    phi = inputs.uniform(91, (n, n), 0, 1).astype(np.float32)
    u = fo.ScalarField(to_dev(u0), (ds, ds), lam)
    t = fo.ScalarField(to_dev(t0), (ds, ds), lam)
    doe = fo.Phase(u, phi, trainable=True, buffered=True)
    system = fo.ScalarSource(u) | doe | fo.ASProp(u, z)
    value, grads = system.withgradient(fo.PowerCoupling(t), weights=-1.0)
    stages = [("phase", phi.astype(np.float64)), ("as", z)]
    v_ref, saved = O.pipeline_forward(u0, stages, (ds, ds), lam)
    p_ref, c_ref, vh = O.power_coupling(v_ref, c64(t0), (ds, ds))
    dv = O.power_coupling_pullback(vh, c_ref, [-1.0])
    _, (g_ref,) = O.pipeline_backward(dv, stages, (ds, ds), lam, saved)
    assert abs(float(value) + p_ref.sum()) <= FIELD_TOL * p_ref.sum()
    torch.cuda.synchronize()
    assert rel_l2(doe.grad.cpu().numpy(), g_ref) <= GRAD_TOL


def test_sq_field_difference_on_device():
    """SquaredFieldDifference: the reference's testset values (test/metrics_test.jl:77-99) and its
    "Metric Backpropagation" system (source | trainable Phase, :101-131), loss and mask gradient
    against the oracle chain; values and weighted pullback of a batch of three against the oracle."""
    print("[sq field difference]", flush=True)
    n = 32
    ones = fo.ScalarField(to_dev(np.ones((n, n))), (1.0, 1.0), 1.0)
    twos = fo.ScalarField(to_dev(2 * np.ones((n, n))), (1.0, 1.0), 1.0)
    assert float(fo.SquaredFieldDifference(twos)(ones)[0][0]) == 1024.0
    assert float(fo.SquaredFieldDifference(ones)(ones)[0][0]) == 0.0
    u0 = c64(np.stack([inputs.gaussian(48, 40, 1.0, 1.0, 8.0 + 2 * k) for k in range(3)]) * (1 + 0.5j))
    t0 = c64(np.stack([inputs.gaussian(48, 40, 1.0, 1.0, 9.0, 3.0 - k, 1.0) for k in range(3)]))
    u = fo.ScalarField(to_dev(u0), (1.0, 1.0), [0.8, 0.7, 0.6])
    t = fo.ScalarField(to_dev(t0), (1.0, 1.0), [0.8, 0.7, 0.6])
    m = fo.SquaredFieldDifference(t)
    dc = [0.5, -1.0, 2.0]
    (vals,), (du,) = m.value_and_pullback((u,), [dc])
    c_ref, dv_ref = O.sq_field_diff_values(u0, t0, dc)
    assert rel_l2(vals.cpu().numpy(), c_ref) <= 1e-6
    assert rel_l2(to_host(du.electric), dv_ref) <= 1e-6
    # the reference's backpropagation testset: source | Phase(zeros, trainable) against a shifted target
    x = inputs.grid(n, 1.0)
    g0 = c64(inputs.gaussian(n, n, 1.0, 1.0, 10.0))
    tg = c64(inputs.gaussian(n, n, 1.0, 1.0, 10.0, 5.0, 0.0))
    src = fo.ScalarField(to_dev(g0), (1.0, 1.0), 1.0)
    phi = (0.3 * np.cos(2 * math.pi * x[None, :] / 16.0) * np.ones((n, 1))).astype(np.float32)
    ph = fo.Phase(src, phi, trainable=True, buffered=True)
    metric = fo.SquaredFieldDifference(fo.ScalarField(to_dev(tg), (1.0, 1.0), 1.0))
    value, grads = (fo.ScalarSource(src) | ph).withgradient(metric)
    torch.cuda.synchronize()
    stages = [("phase", phi.astype(np.float64))]
    v, saved = O.pipeline_forward(g0, stages, (1.0, 1.0), 1.0)
    c, dv = O.sq_field_diff_values(v, tg)
    _, (g_ref,) = O.pipeline_backward(dv, stages, (1.0, 1.0), 1.0, saved)
    assert abs(float(value) - c.sum()) <= 1e-6 * c.sum()
    assert rel_l2(grads[ph].cpu().numpy(), g_ref) <= 1e-5


def test_reference_field_testsets_on_device():
    """test/fields_test.jl:49-130 on device fields: power and normalize_power! (a waist-20 Gaussian
    on 128^2, ds 0.5), arithmetic, copy independence, intensity and phase of exp(i(x + y))."""
    x = inputs.grid(128, 0.5)
    u = fo.ScalarField(to_dev(inputs.gaussian(128, 128, 0.5, 0.5, 20.0)), (0.5, 0.5), 1.064)
    assert float(fo.power(u)[0]) > 0
    fo.normalize_power_(u, 1.0)
    assert abs(float(fo.power(u)[0]) - 1.0) <= 1e-6
    fo.normalize_power_(u, 2.5)
    assert abs(float(fo.power(u)[0]) - 2.5) <= 2.5e-6
    u1 = fo.ScalarField(to_dev(np.ones((32, 32))), (1.0, 1.0), 1.0)
    u2 = fo.ScalarField(to_dev(2 * np.ones((32, 32))), (1.0, 1.0), 1.0)
    assert np.all(to_host((u1 + u2).electric) == 3.0)
    assert np.all(to_host((u2 - u1).electric) == 1.0)
    assert np.all(to_host((3 * u1).electric) == 3.0)
    u3 = u1.copy()
    u3.electric[0, 0] = 2.0
    assert to_host(u1.electric)[0, 0] == 1.0 and to_host(u3.electric)[0, 0] == 2.0
    y = inputs.grid(64, 1.0)
    xx = inputs.grid(64, 1.0)
    f = fo.ScalarField(to_dev(np.exp(1j * (xx[None, :] + y[:, None]))), (1.0, 1.0), 1.064)
    assert np.allclose(fo.intensity(f).cpu().numpy(), 1.0, atol=1e-6)
    ph = fo.phase(f).cpu().numpy()
    assert ph.shape == (64, 64) and np.all(np.abs(ph) <= math.pi)
    assert np.allclose(to_host(f.conj().electric), np.conj(to_host(f.electric)))
    with pytest.raises(ValueError):
        u1 + fo.ScalarField(to_dev(np.ones((32, 32))), (1.0, 1.0), 0.8)


def test_reference_multiple_fields_metric():
    """test/metrics_test.jl:133-152: a SquaredIntensityDifference of two (field, target) pairs
    gives one value per pair, both zero for matching intensities."""
    u1 = fo.ScalarField(to_dev(np.ones((32, 32))), (1.0, 1.0), 1.0)
    u2 = fo.ScalarField(to_dev(2 * np.ones((32, 32))), (1.0, 1.0), 1.0)
    metric = fo.SquaredIntensityDifference((u1, np.ones((32, 32))), (u2, 4 * np.ones((32, 32))))
    result = metric(u1, u2)
    assert len(result) == 2 and all(float(r.sum()) == 0.0 for r in result)


def test_function_forms_of_field_and_fft_helpers():
    """make_fft_plans / compute_ft! / compute_ift! (FFTutils.jl:13-28) against numpy's FFT and as a
    round trip; dot (Fields.jl:810-814); field accessors and setters."""
    u0 = c64(np.random.default_rng(3).normal(size=(2, 24, 40)) + 1j * np.random.default_rng(4).normal(size=(2, 24, 40)))
    u = fo.ScalarField(to_dev(u0), (1.0, 0.5), [0.8, 0.7], tilts=([0.0, 0.01], 0.0))
    plans = fo.make_fft_plans(u)
    fo.compute_ft_(plans, u)
    assert rel_l2(to_host(u.electric), np.fft.fft2(u0.astype(np.complex128))) <= 1e-6
    fo.compute_ift_(plans, u)
    assert rel_l2(to_host(u.electric), u0) <= 1e-6
    d = fo.dot(u, u).cpu().numpy()
    assert np.allclose(d, (np.abs(to_host(u.electric).astype(np.complex128)) ** 2).sum(axis=(1, 2)), rtol=1e-12)
    assert fo.get_lambdas_collection(u) == [0.8, 0.7] and not fo.is_on_axis(u)
    assert fo.get_tilts_collection(u)[0] == [0.0, 0.01]
    fo.set_field_ds_(u, (2.0, 2.0))
    assert u.ds == (2.0, 2.0) and fo.set_field_data(u, u.electric * 0).ds == (2.0, 2.0)


def test_recorded_step_survives_metric_partials_growth():
    """A metric call with a large batch grows the context's partials buffer after a training
    step was recorded; the recorded step keeps its own (retired, still allocated) buffer, so its
    replays still equal the eager calls."""
    n, ds, lam = 128, 1.0, 1.064
    print("[captured withgradient] then DotProduct over 1000 fields", flush=True)
    u0 = c64(inputs.normalize_power(inputs.gaussian(n, n, ds, ds, 12.0), ds, ds))
    tgt = (np.abs(_two_gaussians(n, ds, 12.0, 20.0)) ** 2).astype(np.float32)
    u = fo.ScalarField(to_dev(u0), (ds, ds), lam)
    doe = fo.Phase(u, inputs.uniform(63, (n, n), 0, 1).astype(np.float32), trainable=True, buffered=True)
    p = fo.RSProp(u, 900.0)
    system = fo.ScalarSource(u) | doe | p
    metric = fo.SquaredIntensityDifference((u, tgt))
    step = system.capture_withgradient(metric)
    value, grads = step.replay()
    torch.cuda.synchronize()
    v_before, g_before = float(value), grads[doe].clone()
    # 1000 fields: 1000 x 256 complex partials, more than any earlier call asked for
    big0 = c64(inputs.normal(64, (1000, 16, 16)) + 1j * inputs.normal(65, (1000, 16, 16)))
    big = fo.ScalarField(to_dev(big0), (ds, ds), lam)
    (c,) = fo.DotProduct(big)(big)
    assert np.allclose(c.cpu().numpy(), O.dot_product(big0, big0, (ds, ds)), rtol=1e-10, atol=0)
    for rep in range(3):
        _ = torch.randn(4096, device="cuda").sum()
        value, grads = step.replay()
        torch.cuda.synchronize()
        assert float(value) == v_before and torch.equal(grads[doe], g_before), rep
        ev, eg = system.withgradient(metric)
        torch.cuda.synchronize()
        assert float(ev) == v_before and torch.equal(eg[doe], g_before), rep
