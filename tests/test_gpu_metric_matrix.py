"""GPU parity of the non-selective DotProduct / PowerCoupling (metrics/dot_product.jl:53-97,
power_coupling.jl:55-69): the overlap matrix against the oracle, the reference's doctest identity,
both pullbacks, a Phase -> ASProp system differentiated through the PowerCoupling matrix, and the
refusal of more than 32 modes."""

import math

import numpy as np
import pytest
import torch

import fluxoptics as fo
from fluxoptics import inputs
from oracle import fluxoptics_oracle as O

from _util import c64, rel_l2, to_dev, to_host

pytestmark = pytest.mark.gpu


def test_overlap_matrix_values_and_pullbacks():
    n, ny, nx, ds = 5, 96, 128, (0.5, 0.7)
    print("[matrix] 5 modes 96x128", flush=True)
    u0 = c64(inputs.normal(1, (n, ny, nx)) + 1j * inputs.normal(2, (n, ny, nx)))
    v0 = c64(inputs.normal(3, (n, ny, nx)) + 1j * inputs.normal(4, (n, ny, nx)))
    u = fo.ScalarField(to_dev(u0), ds, 0.8)
    v = fo.ScalarField(to_dev(v0), ds, 0.8)
    dm = fo.DotProduct(v, mode_selective=False)
    (c,) = dm(u)
    cr = O.dot_product_matrix(u0, v0, ds)
    assert np.allclose(c.cpu().numpy(), cr, rtol=1e-10, atol=1e-10 * np.abs(cr).max())
    dc = inputs.normal(5, (n, n)) + 1j * inputs.normal(6, (n, n))
    (du,) = dm.backpropagate_metric([u], [torch.from_numpy(dc)])
    assert rel_l2(to_host(du.electric), O.dot_product_matrix_pullback(v0, dc)) <= 1e-6
    pm = fo.PowerCoupling(v, mode_selective=False)
    (p,) = pm(u)
    pr, cm, vh = O.power_coupling_matrix(u0, v0, ds)
    # the targets are normalised in the field's precision (FP32), as in the reference
    assert np.allclose(p.cpu().numpy(), pr, rtol=2e-6, atol=1e-9 * pr.max())
    vdev = to_host(pm.v[0].electric)
    assert np.allclose(p.cpu().numpy(), np.abs(O.dot_product_matrix(u0, vdev, ds)) ** 2, rtol=1e-9,
                       atol=1e-12 * pr.max())
    dp = inputs.uniform(7, (n, n), 0.5, 2.0)
    (dup,) = pm.backpropagate_metric([u], [torch.from_numpy(dp)])
    assert rel_l2(to_host(dup.electric), O.power_coupling_matrix_pullback(vh, cm, dp)) <= 1e-6
    # the reference's doctest: two normalised orthogonal modes give the identity
    from test_system_host import _doc_modes
    w = fo.ScalarField(to_dev(_doc_modes()), (1.0, 1.0), 1.064)
    (ci,) = fo.DotProduct(w, mode_selective=False)(w)
    assert np.allclose(ci.cpu().numpy(), np.eye(2), atol=1e-7)
    big = fo.ScalarField(torch.zeros((33, 16, 16), dtype=torch.complex64, device="cuda"), (1.0, 1.0), 1.0)
    with pytest.raises(RuntimeError, match="32 modes"):
        fo.DotProduct(big, mode_selective=False)(big)


def test_system_through_power_coupling_matrix():
    """ScalarSource (4 modes) -> Phase -> ASProp -> PowerCoupling(mode_selective=false): the loss
    sum p[i, j] and the mask gradient against the oracle chain."""
    n, m, ds, lam, z = 128, 4, (1.0, 1.0), 1.064, 700.0
    print("[matrix] Phase -> ASProp -> PowerCoupling matrix, 4 modes 128^2", flush=True)
    modes = np.stack([inputs.laguerre_gaussian(n, n, 1.0, 1.0, 12.0, p, l) for p, l in
                      [(0, 0), (0, 1), (1, 0), (0, -2)]])
    u0 = c64(modes)
    tg = c64(np.stack([inputs.gaussian(n, n, 1.0, 1.0, 10.0, 12.0 * k - 18.0, 0.0) for k in range(m)]))
    u = fo.ScalarField(to_dev(u0), ds, lam)
    t = fo.ScalarField(to_dev(tg), ds, lam)
    phi = inputs.uniform(8, (n, n), 0, 2 * math.pi).astype(np.float32)
    doe = fo.Phase(u, phi, trainable=True, buffered=True)
    p = fo.ASProp(u, z)
    system = fo.ScalarSource(u) | doe | p
    metric = fo.PowerCoupling(t, mode_selective=False)
    value, grads = system.withgradient(metric)
    a = u0 * np.exp(1j * phi.astype(np.float64))
    v = O.asprop(a, ds, lam, z)
    pr, cm, vh = O.power_coupling_matrix(v, tg, ds)
    dv = O.power_coupling_matrix_pullback(vh, cm)
    g = O.asprop(dv, ds, lam, z, 1.0, -1)
    dphi = np.imag(g * np.conj(a)).sum(axis=0)
    assert abs(float(value) - pr.sum()) <= 1e-5 * pr.sum()
    assert rel_l2(grads[doe].cpu().numpy(), dphi) <= 1e-4


@pytest.mark.parametrize("kind,n", [("as", 128), ("rs", 128), ("as", 4096)])
def test_fused_power_coupling_tail_matches_unfused(kind, n):
    """Phase -> ASProp/RSProp -> PowerCoupling: the fused tail (the pullback read by the adjoint's
    first pass, fh_dot_pullback_propagate) against the component-by-component run and the oracle,
    with weights; and the C entry point alone against pullback + adjoint propagation."""
    m, ds, lam, z = 3, (1.0, 1.0), 1.064, 900.0 * n / 128
    w0 = 12.0 * n / 128
    print(f"[fused PowerCoupling] {kind} n={n}", flush=True)
    u0 = c64(np.stack([inputs.laguerre_gaussian(n, n, 1.0, 1.0, w0, p, l) for p, l in [(0, 0), (0, 1), (1, -1)]]))
    tg = c64(np.stack([inputs.gaussian(n, n, 1.0, 1.0, 0.75 * w0, (10.0 * k - 10.0) * n / 128, 5.0) for k in range(m)]))
    u = fo.ScalarField(to_dev(u0), ds, lam)
    t = fo.ScalarField(to_dev(tg), ds, lam)
    phi = inputs.uniform(12, (n, n), 0, 2 * math.pi).astype(np.float32)
    doe = fo.Phase(u, phi, trainable=True, buffered=True)
    import warnings
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        p = fo.RSProp(u, z) if kind == "rs" else fo.ASProp(u, z)
    system = fo.ScalarSource(u) | doe | p
    metric = fo.PowerCoupling(t)
    w = [-1.0, 0.5, 2.0]
    vf, gf = system.withgradient(metric, weights=w)
    gf = gf[doe].clone()
    vu, gu = system.withgradient(metric, weights=w, fuse=False)
    assert abs(float(vf) - float(vu)) <= 1e-12 * abs(float(vu))
    assert rel_l2(gf.cpu().numpy(), gu[doe].cpu().numpy()) <= 1e-6
    if n > 1024:  # the large case checks the fused path against the component-by-component one
        return
    a = u0 * np.exp(1j * phi.astype(np.float64))
    prop = (lambda x, d: O.rsprop(x, ds, lam, z, d)) if kind == "rs" else (lambda x, d: O.asprop(x, ds, lam, z, 1.0, d))
    v = prop(a, 1)
    pr, cr, vh = O.power_coupling(v, tg, ds)
    dv = O.power_coupling_pullback(vh, cr, np.array(w))
    dphi = np.imag(prop(dv, -1) * np.conj(a)).sum(axis=0)
    assert abs(float(vf) - float((np.array(w) * pr).sum())) <= 1e-5 * abs(float((np.array(w) * pr).sum()))
    assert rel_l2(gf.cpu().numpy(), dphi) <= 1e-4
    # the entry point alone: P^H(v g) equals the pullback kernel followed by the adjoint propagation
    from fluxoptics import _lib
    from fluxoptics.components import _ptr, context
    plan = p.plan(u)
    (pv,) = metric(u)
    c = metric._c[0]
    out = torch.empty_like(u.electric)
    _lib.call("fh_dot_pullback_propagate", plan, _ptr(metric.v[0].electric), None, _ptr(c), None, _ptr(out), 1,
              None, None, None, 0)
    ref = torch.empty_like(u.electric)
    _lib.call("fh_dot_product_pullback", context(u.electric.device), _ptr(metric.v[0].electric), None, _ptr(c), None,
              n, n, m, _ptr(ref))
    _lib.call("fh_propagate", plan, _ptr(ref), -1)
    assert rel_l2(to_host(out), to_host(ref)) <= 1e-6
