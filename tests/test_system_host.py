"""CPU tests: the oracle's SquaredIntensityDifference against the reference's own doctest
values and finite differences, and the host-side system logic (pipe, merge rules)."""

import math

import numpy as np
import pytest
import torch

import fluxoptics as fo
from fluxoptics import inputs
from oracle import fluxoptics_oracle as O

TWO_PI = 2.0 * math.pi


def test_sqint_reference_doctest_values():
    """``src/metrics/squared_intensity_difference.jl:19-45``: ones(4,4,2) against zeros gives
    16 per field; ones(4,4,3) against 3*ones(4,4) (spatial only) gives 0."""
    vals, _ = O.sq_intensity_diff_values(np.ones((2, 4, 4), complex), np.zeros((2, 4, 4)))
    assert np.array_equal(vals, [16.0, 16.0])
    vals, dv = O.sq_intensity_diff_values(np.ones((3, 4, 4), complex), 3 * np.ones((4, 4)))
    assert np.array_equal(vals, [0.0]) and np.all(dv == 0)


@pytest.mark.parametrize("spatial", [False, True])
def test_sqint_cotangent_by_finite_differences(spatial):
    """The cotangent dL/dRe + i dL/dIm of the summed metric, by central differences."""
    b, n = 3, 6
    v = inputs.normal(1, (b, n, n)) + 1j * inputs.normal(2, (b, n, n))
    t = inputs.uniform(3, (n, n) if spatial else (b, n, n), 0.0, 2.0)
    dc = np.array([0.7]) if spatial else np.array([0.5, 1.0, 2.0])
    _, dv = O.sq_intensity_diff_values(v, t, dc)
    f = lambda w: float((dc * O.sq_intensity_diff_values(w, t)[0]).sum())
    h = 1e-6
    for k, (bb, i, j) in enumerate([(0, 1, 2), (1, 4, 4), (2, 0, 5), (1, 3, 0)]):
        e = np.zeros_like(v)
        e[bb, i, j] = h
        g_re = (f(v + e) - f(v - e)) / (2 * h)
        g_im = (f(v + 1j * e) - f(v - 1j * e)) / (2 * h)
        assert abs(g_re - dv[bb, i, j].real) <= 1e-6 * (1 + abs(g_re))
        assert abs(g_im - dv[bb, i, j].imag) <= 1e-6 * (1 + abs(g_im))


def _cpu_field(n=16, batch=1):
    return fo.ScalarField(torch.zeros((batch, n, n), dtype=torch.complex64), (1.0, 1.0), 0.8)


def test_merge_rules():
    """``merge_rules.jl``: static Phase + Phase adds masks, trainable ones stay, two
    on-axis ASProps in one medium become one over the summed distance, RSProps stay."""
    u = _cpu_field()
    a = np.arange(256, dtype=np.float32).reshape(16, 16)
    p1, p2 = fo.Phase(u, a), fo.Phase(u, 2 * a)
    m = fo.merge([p1, p2])
    assert len(m) == 1 and torch.equal(m[0].phi, torch.from_numpy(3 * a))
    pt = fo.Phase(u, a, trainable=True)
    assert len(fo.merge([p1, pt])) == 2
    m = fo.merge([fo.ASProp(u, 100.0), fo.ASProp(u, 250.0), fo.ASProp(u, 1.0, n0=1.5)])
    assert len(m) == 2 and m[0].z == 350.0 and m[1].n0 == 1.5
    with pytest.warns(UserWarning):
        r = fo.RSProp(u, 1.0)
    assert len(fo.merge([r, r])) == 2


def test_pipe_builds_systems():
    u = _cpu_field()
    src = fo.ScalarSource.__new__(fo.ScalarSource)  # a source without device data
    src.trainable = False
    p, q = fo.ASProp(u, 10.0), fo.ASProp(u, 20.0)
    ph = fo.Phase(u, None)
    s = src | p | ph | q
    assert s.source is src and s.components == (p, ph, q)
    s2 = s | {"merge_components": True}
    assert s2.merge_components and len(s2.components) == 3
    s3 = src | p | q | {"merge_components": True}
    assert len(s3.components) == 1 and s3.components[0].z == 30.0
    chain = p | ph
    assert chain.source is None and chain.components == (p, ph)
    with pytest.raises(ValueError):
        s(u)  # a system with a source takes no input field
    with pytest.raises(ValueError):
        chain()  # a system without a source needs one


def test_fourier_operator_merge_and_wrapper():
    """``merge_rules.jl``: FourierOperators of opposite direction cancel; FourierWrapper is
    the sequence FFT -> components -> IFFT and a system runs it flattened."""
    u = _cpu_field()
    ft, ift = fo.FourierOperator(u, True), fo.FourierOperator(u, False)
    assert fo.merge([ft, ift]) == [] and fo.merge([ift, ft]) == []
    assert fo.merge([ft, ft]) == [ft, ft]
    ph = fo.Phase(u, None)
    w = fo.FourierWrapper(u, ph)
    assert isinstance(w, fo.OpticalSequence) and len(w) == 3 and w.optical_components[1] is ph
    s = fo.OpticalSystem(None, fo.ASProp(u, 1.0), w)
    assert len(s.components) == 4
    p2 = fo.ASProp(u, 2.0)
    sm = fo.OpticalSystem(None, w, fo.FourierWrapper(u, p2), merge_components=True)
    assert [type(c).__name__ for c in sm.components] == ["FourierOperator", "Phase", "ASProp", "FourierOperator"]


def _doc_modes():
    """``dot_product.jl`` / ``power_coupling.jl`` doctest field: ones(4, 4, 2) with the second
    mode's rows 3:4 (x) set to -1, power-normalised; as (B, ny, nx): u[1, :, 2:] = -1."""
    u = np.ones((2, 4, 4), complex)
    u[1, :, 2:] = -1
    return O.normalize_power(u, (1.0, 1.0))


def test_dot_and_power_coupling_reference_doctests():
    u = _doc_modes()
    assert np.allclose(O.dot_product(u, u, (1.0, 1.0)), [1.0, 1.0], atol=1e-15)
    p, c, vh = O.power_coupling(u, u, (1.0, 1.0))
    assert np.allclose(p, [1.0, 1.0], atol=1e-15)
    # the two modes are orthogonal (the non-selective doctest's off-diagonal zeros)
    assert abs(O.dot_product(u[0], u[1], (1.0, 1.0))[0]) < 1e-15


@pytest.mark.parametrize("which", ["dot", "power"])
def test_dot_and_power_pullbacks_by_finite_differences(which):
    """At ds = 1 the pullbacks equal the gradient of Re(sum conj(dc) c) (DotProduct) and of
    sum dp |c|^2 (PowerCoupling), checked by central differences."""
    b, n = 2, 5
    u = inputs.normal(11, (b, n, n)) + 1j * inputs.normal(12, (b, n, n))
    v = inputs.normal(13, (b, n, n)) + 1j * inputs.normal(14, (b, n, n))
    ds = (1.0, 1.0)
    if which == "dot":
        dc = np.array([0.3 - 0.7j, -1.1 + 0.2j])
        f = lambda w: float(np.real(np.sum(np.conj(dc) * O.dot_product(w, v, ds))))
        g = O.dot_product_pullback(v, dc)
    else:
        dp = np.array([0.5, -2.0])
        f = lambda w: float(np.sum(dp * O.power_coupling(w, v, ds)[0]))
        _, c, vh = O.power_coupling(u, v, ds)
        g = O.power_coupling_pullback(vh, c, dp)
    h = 1e-6
    for (bb, i, j) in [(0, 1, 2), (1, 4, 0), (1, 2, 3)]:
        e = np.zeros_like(u)
        e[bb, i, j] = h
        gr = (f(u + e) - f(u - e)) / (2 * h)
        gi = (f(u + 1j * e) - f(u - 1j * e)) / (2 * h)
        assert abs(gr - g[bb, i, j].real) <= 1e-6 * (1 + abs(gr))
        assert abs(gi - g[bb, i, j].imag) <= 1e-6 * (1 + abs(gi))


@pytest.mark.parametrize("direction", [1, -1])
def test_mask_gradient_by_finite_differences(direction):
    """The complex mask gradient dL/dRe(m) + i dL/dIm(m) of L = 1/2 |v - t|^2, v = mask(u)."""
    b, n = 2, 5
    u = inputs.normal(31, (b, n, n)) + 1j * inputs.normal(32, (b, n, n))
    t = inputs.normal(33, (b, n, n)) + 1j * inputs.normal(34, (b, n, n))
    m = inputs.normal(35, (n, n)) + 1j * inputs.normal(36, (n, n))
    f = lambda mm: O.half_sq_field_diff(O.mask_apply(u, mm, direction), t)[0]
    dv = O.mask_apply(u, m, direction) - t
    du, dm = O.mask_backward(dv, u, m, direction)
    h = 1e-6
    for (i, j) in [(0, 1), (3, 4), (2, 2)]:
        e = np.zeros((n, n), complex)
        e[i, j] = h
        gr = (f(m + e) - f(m - e)) / (2 * h)
        gi = (f(m + 1j * e) - f(m - 1j * e)) / (2 * h)
        assert abs(gr - dm[i, j].real) <= 1e-6 * (1 + abs(gr))
        assert abs(gi - dm[i, j].imag) <= 1e-6 * (1 + abs(gi))
    # du is the exact adjoint: <mask(x), y> = <x, mask^H(y)>
    x, y = inputs.normal(37, (b, n, n)) + 0j, inputs.normal(38, (b, n, n)) + 1j
    assert np.isclose(np.vdot(y, O.mask_apply(x, m, direction)), np.vdot(O.mask_backward(y, x, m, direction)[0], x))


def test_tea_doe_reduces_to_phase_and_gradient_by_fd():
    """With r = 1 a TeaDOE is the Phase 2 pi dn h / lambda per wavelength; its height gradient
    matches central differences (with a complex r as well)."""
    b, n, lams, dn = 2, 6, [0.8, 1.064], 0.45
    u = inputs.normal(41, (b, n, n)) + 1j * inputs.normal(42, (b, n, n))
    h = inputs.uniform(43, (n, n), 0.0, 2.0)
    for k in range(b):
        ref = np.reshape(O.phase_apply(u[k], TWO_PI * dn * h / lams[k], 1), (n, n))
        assert np.allclose(O.tea_apply(u[k], h, lams[k], dn)[0], ref, rtol=1e-14, atol=0)
    r = 0.9 * np.exp(0.3j)
    t = inputs.normal(44, (b, n, n)) + 1j * inputs.normal(45, (b, n, n))
    for direction in (1, -1):
        f = lambda hh: O.half_sq_field_diff(O.tea_apply(u, hh, lams, dn, r, direction), t)[0]
        dv = O.tea_apply(u, h, lams, dn, r, direction) - t
        _, dh = O.tea_backward(dv, u, h, lams, dn, r, direction)
        eps = 1e-6
        for (i, j) in [(0, 0), (2, 5), (4, 1)]:
            e = np.zeros((n, n))
            e[i, j] = eps
            g = (f(h + e) - f(h - e)) / (2 * eps)
            assert abs(g - dh[i, j]) <= 1e-6 * (1 + abs(g))


def test_prox_reference_doctests():
    """``proximal_operators/clamp_prox.jl``, ``ista_prox.jl``, ``positive_prox.jl`` doctests."""
    assert np.allclose(O.prox_clamp([-0.5, 0.3, 1.2, 0.8], 0.0, 1.0), [0.0, 0.3, 1.0, 0.8])
    assert np.allclose(O.prox_ista([-0.5, -0.005, 0.0, 0.003, 0.2], 0.01, 0.0), [-0.49, 0.0, 0.0, 0.0, 0.19])
    assert np.allclose(O.prox_positive([-0.1, 0.0, 0.5, -0.3, 1.2]), [0.0, 0.0, 0.5, 0.0, 1.2])


def test_rule_invariants():
    """Momentum and Nesterov with rho = 0 are Descent; Fista's first step is Descent with the
    stored (squared) rate and t_2 = golden ratio; on f(x) = |x|^2 / 2 Fista meets the
    O(1/k^2) bound f(x_k) <= 2 |x_0|^2 / (eta (k+1)^2)."""
    x = inputs.normal(51, (7,))
    g = inputs.normal(52, (7,))
    d = O.optim_step("descent", x, g, eta=0.3)[0]
    assert np.allclose(O.optim_step("momentum", x, g, np.zeros(7), eta=0.3, rho=0.0)[0], d, rtol=1e-15)
    assert np.allclose(O.optim_step("nesterov", x, g, np.zeros(7), eta=0.3, rho=0.0)[0], d, rtol=1e-15)
    xf, xk, t = O.optim_step("fista", x, g, x.copy(), 1.0, eta=0.3)
    assert np.allclose(xf, d, rtol=1e-15) and abs(t - (1 + 5 ** 0.5) / 2) < 1e-15
    eta = 0.5
    xx, st, t = x.copy(), x.copy(), 1.0
    for k in range(1, 30):
        xx, st, t = O.optim_step("fista", xx, xx, st, t, eta=eta)
        assert 0.5 * np.dot(xx, xx) <= 2 * np.dot(x, x) / (eta * (k + 1) ** 2) + 1e-15


def test_momentum_nesterov_constant_gradient_closed_forms():
    """Momentum and Nesterov at rho != 0 on a constant gradient g, from the textbook definitions
    rather than the oracle's update formulas.  Heavy-ball momentum (v_k = rho v_(k-1) + eta g,
    x_k = x_(k-1) - v_k) sums two geometric series: x_k = x_0 - eta g / (1 - rho) * (k - rho (1 -
    rho^k) / (1 - rho)).  Sutskever's Nesterov step takes the gradient at x + rho v, which for a
    constant gradient gives the heavy-ball velocities; Optimisers.jl tracks the look-ahead point
    x_k - rho v_k, so its iterate is the heavy-ball one minus rho eta g (1 - rho^k) / (1 - rho).  Both
    step sizes tend to eta g / (1 - rho)."""
    x0 = inputs.normal(53, (5,))
    g = inputs.normal(54, (5,))
    for rho, eta in ((0.9, 0.1), (0.5, 0.3)):
        xm, sm = x0.copy(), np.zeros(5)
        xn, sn = x0.copy(), np.zeros(5)
        for k in range(1, 40):
            xm, sm, _ = O.optim_step("momentum", xm, g, sm, eta=eta, rho=rho)
            xn, sn, _ = O.optim_step("nesterov", xn, g, sn, eta=eta, rho=rho)
            geo = (1 - rho ** k) / (1 - rho)
            heavy = x0 - eta * g / (1 - rho) * (k - rho * geo)
            assert np.allclose(xm, heavy, rtol=1e-12, atol=1e-12)
            assert np.allclose(xn, heavy - rho * eta * g * geo, rtol=1e-12, atol=1e-12)


def test_oracle_adam_closed_forms():
    """Adam's first step is eta * g / (|g| + eps) (bias corrections undo (1 - b)); with b1 = b2 = 0
    every step is that sign-like step; the state stays FP32."""
    from oracle import fluxoptics_oracle as O
    x = np.linspace(-1, 1, 11).astype(np.float32)
    g = np.linspace(-3, 2, 11).astype(np.float32)
    z = np.zeros(11, np.float32)
    xn, m, v, bt = O.adam_step(x, g, z, z, (0.9, 0.999), eta=0.1)
    ref = x.astype(np.float64) - 0.1 * g / (np.abs(g.astype(np.float64)) + 1e-8)
    assert np.max(np.abs(xn - ref)) <= 1e-6 and m.dtype == np.float32 and bt == (0.81, 0.999 ** 2)
    xk = x
    for _ in range(3):
        xk2, m0, v0, _ = O.adam_step(xk, g, z, z, (0.0, 0.0), eta=0.05, beta=(0.0, 0.0))
        assert np.max(np.abs(xk2 - (xk.astype(np.float64) - 0.05 * np.sign(g) * (np.abs(g) / (np.abs(g) + 1e-8)))))\
            <= 1e-6
        xk = xk2


def test_host_dispatch_and_argument_parsing():
    """Host-only logic of the newer components (no device work): ParaxialProp's choice between
    ASProp and CollinsProp, FourierLens/CollinsProp argument forms, CZT argument conversion,
    setup(rule, model), TiltAnchor tilts, ShiftProp options."""
    import cmath
    import fluxoptics as fo
    from fluxoptics import czt as czt_mod
    u = fo.ScalarField(torch.zeros((2, 32, 32), dtype=torch.complex64), (0.5, 0.5), [0.8, 0.6],
                       ([0.01, 0.02], 0.0))
    assert isinstance(fo.ParaxialProp(u, 100.0), fo.ASProp) and fo.ParaxialProp(u, 100.0).paraxial
    assert isinstance(fo.ParaxialProp(u, (0.5, 0.5), 100.0), fo.ASProp)
    c = fo.ParaxialProp(u, (1.0, 0.25), 100.0)
    assert isinstance(c, fo.CollinsProp) and c.abd == (1.0, 100.0, 1.0) and c.ds_out == (1.0, 0.25)
    lens = fo.FourierLens(u, (0.3, 0.3), 500.0)
    assert lens.abd == (0.0, 500.0, 0.0) and lens.ds == (0.5, 0.5)
    lens2 = fo.FourierLens(u, (0.25, 0.25), (0.3, 0.3), 500.0)
    assert lens2.ds == (0.25, 0.25)
    with pytest.raises(ValueError):
        fo.CollinsProp(u, (1.0, 1.0), (1.0, 0.0, 1.0))
    with pytest.raises(TypeError):
        fo.CollinsProp(u, (1.0, 1.0))
    with pytest.raises(NotImplementedError):
        fo.ShiftProp(u, 10.0, double_precision_kernel=False)
    # CZT argument forms (convert_czt_arg)
    assert czt_mod._convert_czt_arg(2j) == [(2.0, cmath.phase(2j))] * 2
    assert czt_mod._convert_czt_arg((1.0, -1.0)) == [(1.0, 0.0), (1.0, math.pi)]
    assert czt_mod._convert_czt_arg(((1.5, 0.2), (0.5, -0.1))) == [(1.5, 0.2), (0.5, -0.1)]
    with pytest.raises(ValueError):
        czt_mod._convert_czt_arg((1.0, 2.0, 3.0))
    # setup(rule, model) gives every trainable parameter the rule
    p1 = fo.Phase(u, np.zeros((32, 32), np.float32), trainable=True)
    p2 = fo.Phase(u, np.zeros((32, 32), np.float32), trainable=True)
    st = fo.setup(fo.Descent(0.5), [p1, p2])
    assert len(st) == 2 and all(isinstance(l.base, fo.Descent) for l in st)
    st = fo.setup(fo.Adam(0.1), [p1])
    assert st[0].state.shape == (2, 32, 32)
    # TiltAnchor keeps the template's tilts
    a = fo.TiltAnchor(u)
    assert a.tilts == ([0.01, 0.02], 0.0) and not a.trainable


def test_pad_crop_host_checks():
    """pad / crop (pad_crop_operators.jl:1-16,61-79) check offsets and sizes before any device
    work, and take device tensors only; PadCropOperator's adjoint swaps pad and crop."""
    import torch
    u = torch.zeros(2, 8, 6, dtype=torch.complex64)
    with pytest.raises(ValueError):
        fo.pad(u, (10, 10), offset=(-1, 0))
    with pytest.raises(ValueError):
        fo.pad(u, (10, 10), offset=(5, 0))  # 6 + 5 > 10
    with pytest.raises(ValueError):
        fo.crop(u, (4, 4), offset=(3, 0))   # 4 + 3 > 6
    with pytest.raises(ValueError):
        fo.pad(u, (10,))
    with pytest.raises(TypeError):
        fo.pad(u, (10, 10))                 # no CPU path
    op = fo.PadCropOperator(u, torch.zeros(2, 16, 12, dtype=torch.complex64))
    assert op.ispad and not op.adjoint().ispad and op.adjoint().adjoint().ispad
    assert op.p_pad.size_pad == (12, 16) and op.p_crop.size_crop == (6, 8)


def test_make_rules_keys_like_the_reference():
    """test/optimisers_test.jl:87-106: make_rules(phase1 => rule1, phase2 => rule2) is keyed by
    the parameter arrays (rules[phase1.phi] is rule1); setup takes each parameter's rule."""
    import torch
    u = fo.ScalarField(torch.ones(32, 32, dtype=torch.complex64), (1.0, 1.0), 1.064)
    p1 = fo.Phase.__new__(fo.Phase)
    p2 = fo.Phase.__new__(fo.Phase)
    for p in (p1, p2):
        p.phi, p.trainable = torch.zeros(32, 32), True
    r1, r2 = fo.Fista(0.1), fo.Fista(0.05)
    rules = fo.make_rules((p1, r1), (p2, r2))
    assert p1.phi in rules and rules[p1.phi] is r1 and rules[p2.phi] is r2
    leaves = fo.setup(rules, [p1, p2])
    assert [leaf.rule for leaf in leaves] == [r1, r2]


def test_reference_spatial_vector_testsets():
    """test/gridutils_test.jl:2-33: the 1-D, 2-D and tuple forms of spatial_vectors, with offsets."""
    x, = fo.spatial_vectors(64, 2.0)
    assert len(x) == 64 and np.allclose(np.diff(x), 2.0) and abs(x[31] + x[32]) < 1e-10
    assert abs(fo.spatial_vectors(64, 2.0, xc=5.0)[0].mean() + 5.0) < 1e-10
    xv, yv = fo.spatial_vectors(128, 128, 2.0, 2.0)
    assert len(xv) == len(yv) == 128 and np.allclose(np.diff(yv), 2.0)
    xo, yo = fo.spatial_vectors(128, 128, 2.0, 2.0, xc=-5.0, yc=30.5)
    assert abs(xo.mean() - 5.0) < 1e-10 and abs(yo.mean() + 30.5) < 1e-10
    x2, y2 = fo.spatial_vectors((128, 128), (2.0, 2.0), offset=(-5.0, 30.5))
    assert np.array_equal(x2, xo) and np.array_equal(y2, yo)


def test_recording_refuses_systems_with_aspropz():
    """ASPropZ chooses its transfer function from z on the host, so capture_withgradient and
    capture_train_step refuse a system holding one before any capture begins."""
    from types import SimpleNamespace
    from fluxoptics.components import ASPropZ
    from fluxoptics.system import OpticalSystem, _capture_train_step, _check_recordable
    sysz = SimpleNamespace(components=(object(), ASPropZ.__new__(ASPropZ)))
    with pytest.raises(NotImplementedError, match="ASPropZ"):
        _check_recordable(sysz)
    with pytest.raises(NotImplementedError, match="ASPropZ"):
        _capture_train_step(sysz, None, None)
    with pytest.raises(NotImplementedError, match="ASPropZ"):
        OpticalSystem.capture_withgradient(sysz, None)
    _check_recordable(SimpleNamespace(components=(object(),)))


def test_overlap_matrix_reference_doctest_and_diagonal():
    """The non-selective DotProduct doctest (dot_product.jl:35-39): the two normalised modes give the
    2x2 identity; the diagonal equals the mode-selective values for any fields; PowerCoupling's
    matrix is |c|^2 entrywise against normalised targets."""
    u = _doc_modes()
    assert np.allclose(O.dot_product_matrix(u, u, (1.0, 1.0)), np.eye(2), atol=1e-15)
    a = inputs.normal(1, (3, 5, 6)) + 1j * inputs.normal(2, (3, 5, 6))
    b = inputs.normal(3, (3, 5, 6)) + 1j * inputs.normal(4, (3, 5, 6))
    c = O.dot_product_matrix(a, b, (0.5, 0.7))
    assert np.allclose(np.diag(c), O.dot_product(a, b, (0.5, 0.7)), rtol=1e-13)
    # entry (i, j) pairs target i with field j
    assert np.isclose(c[0, 2], O.dot_product(a[2], b[0], (0.5, 0.7))[0], rtol=1e-13)
    p, cm, vh = O.power_coupling_matrix(a, b, (0.5, 0.7))
    assert np.allclose(p, np.abs(O.dot_product_matrix(a, vh, (0.5, 0.7))) ** 2, rtol=1e-13)


def test_overlap_matrix_pullbacks_by_finite_differences():
    """At ds = 1 the reference's pullbacks are the exact cotangents dL/dRe + i dL/dIm of
    L = Re sum conj(dc) o c (DotProduct) and L = sum dp o |c|^2 (PowerCoupling)."""
    n, ny, nx = 3, 4, 5
    u = inputs.normal(5, (n, ny, nx)) + 1j * inputs.normal(6, (n, ny, nx))
    v = inputs.normal(7, (n, ny, nx)) + 1j * inputs.normal(8, (n, ny, nx))
    dc = inputs.normal(9, (n, n)) + 1j * inputs.normal(10, (n, n))
    dp = inputs.uniform(11, (n, n), 0.5, 2.0)
    ds = (1.0, 1.0)
    g_dot = O.dot_product_matrix_pullback(v, dc)
    p, c, vh = O.power_coupling_matrix(u, v, ds)
    g_pow = O.power_coupling_matrix_pullback(vh, c, dp)
    f_dot = lambda w: float(np.real(np.sum(np.conj(dc) * O.dot_product_matrix(w, v, ds))))
    f_pow = lambda w: float(np.sum(dp * np.abs(O.dot_product_matrix(w, vh, ds)) ** 2))
    h = 1e-6
    for (b, i, j) in [(0, 1, 2), (2, 3, 4), (1, 0, 0)]:
        e = np.zeros_like(u)
        e[b, i, j] = h
        for f, g in ((f_dot, g_dot), (f_pow, g_pow)):
            fd = (f(u + e) - f(u - e)) / (2 * h) + 1j * (f(u + 1j * e) - f(u - 1j * e)) / (2 * h)
            assert abs(fd - g[b, i, j]) <= 1e-6 * max(1.0, abs(fd)), (b, i, j, fd, g[b, i, j])
