"""GPU parity of the fused optimiser steps (rules + proximal operators) against the oracle,
and a short training loop (withgradient + update_) against the same loop on the oracle."""

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


class _P:
    """A stand-in trainable component holding one real parameter."""

    def __init__(self, x):
        self.phi = x
        self.trainable = True


RULES = [("descent", lambda: fo.Descent(0.3), {}), ("momentum", lambda: fo.Momentum(0.2, 0.8), {"rho": 0.8}),
         ("nesterov", lambda: fo.Nesterov(0.1, 0.7), {"rho": 0.7}), ("fista", lambda: fo.Fista(0.5), {})]
PROX = [("none", None, lambda x: x), ("clamp", fo.ClampProx(-0.5, 0.8), lambda x: O.prox_clamp(x, -0.5, 0.8)),
        ("ista", fo.IstaProx(0.05, 0.1), lambda x: O.prox_ista(x, 0.05, 0.1)),
        ("positive", fo.PositiveProx(), O.prox_positive)]


@pytest.mark.parametrize("rname,make,kw", RULES, ids=[r[0] for r in RULES])
@pytest.mark.parametrize("pname,prox,pfn", PROX, ids=[p[0] for p in PROX])
def test_rule_and_prox_steps(rname, make, kw, pname, prox, pfn):
    n = 4096
    x0 = inputs.normal(201, (n,)).astype(np.float32)
    x = torch.from_numpy(x0.copy()).cuda()
    comp = fo.Phase.__new__(fo.Phase)
    comp.phi, comp.trainable = x, True
    rule = make()
    state = fo.setup({comp: fo.ProxRule(rule, prox) if prox is not None else rule}, [comp])
    xr, st, t = x0.astype(np.float64), None, 1.0
    if rname in ("momentum", "nesterov"):
        st = np.zeros(n)
    if rname == "fista":
        st = xr.copy()
    for k in range(4):
        g = inputs.normal(210 + k, (n,)).astype(np.float32)
        fo.update_(state, [comp], {comp: torch.from_numpy(g).cuda()})
        xr, st, t = O.optim_step(rname, xr, g.astype(np.float64), st, t, eta=rule.eta, **kw)
        xr = pfn(xr)
    torch.cuda.synchronize()
    assert np.allclose(x.cpu().numpy(), xr, rtol=2e-5, atol=2e-6)


def test_complex_parameter_with_ista():
    """A trainable Mask (complex) under ProxRule(Fista, IstaProx): real and imaginary parts are
    thresholded separately."""
    n = 64
    u = fo.ScalarField(torch.zeros((n, n), dtype=torch.complex64, device="cuda"), (1.0, 1.0), 0.8)
    m0 = (inputs.normal(220, (n, n)) + 1j * inputs.normal(221, (n, n))).astype(np.complex64)
    mk = fo.Mask(u, m0, trainable=True)
    state = fo.setup(fo.make_rules((mk, fo.ProxRule(fo.Fista(0.4), fo.IstaProx(0.02)))), [mk])
    xr, st, t = m0.astype(np.complex128), m0.astype(np.complex128), 1.0
    for k in range(3):
        g = (inputs.normal(230 + k, (n, n)) + 1j * inputs.normal(240 + k, (n, n))).astype(np.complex64)
        fo.update_(state, [mk], {mk: torch.from_numpy(g).cuda()})
        xr, st, t = O.optim_step("fista", xr, g.astype(np.complex128), st, t, eta=0.16)
        xr = O.prox_ista(xr, 0.02)
    torch.cuda.synchronize()
    assert np.allclose(mk.m.cpu().numpy(), xr, rtol=2e-5, atol=2e-6)


def test_training_loop_matches_oracle():
    """Five iterations of withgradient + update_ (Fista with a clamp) on a 128^2
    RS -> Phase -> RS system with the intensity loss, against the same loop on the oracle."""
    n, ds, lam = 128, 1.0, 1.064
    u0 = c64(inputs.normalize_power(inputs.gaussian(n, n, ds, ds, 12.0), ds, ds))
    t = inputs.normalize_power(inputs.gaussian(n, n, ds, ds, 10.0, -15.0, 0.0)
                               + inputs.gaussian(n, n, ds, ds, 10.0, 15.0, 0.0), ds, ds)
    tgt = (np.abs(t) ** 2).astype(np.float32)
    u = fo.ScalarField(to_dev(u0), (ds, ds), lam)
    doe = fo.Phase(u, np.zeros((n, n), np.float32), trainable=True, buffered=True)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        p1, p2 = fo.RSProp(u, 400.0), fo.RSProp(u, 500.0)
    system = fo.ScalarSource(u) | p1 | doe | p2
    metric = fo.SquaredIntensityDifference((u, tgt))
    eta = 300.0
    opt = fo.setup(fo.make_rules((doe, fo.ProxRule(fo.Fista(eta), fo.ClampProx(-math.pi, math.pi)))), system)
    phi, st, tt = np.zeros((n, n)), np.zeros((n, n)), 1.0
    for it in range(5):
        _, grads = system.withgradient(metric)
        fo.update_(opt, system, grads)
        stages = [("rs", 400.0), ("phase", phi), ("rs", 500.0)]
        v, saved = O.pipeline_forward(u0, stages, (ds, ds), lam)
        _, dv = O.sq_intensity_diff(v, tgt.astype(np.float64))
        _, (g,) = O.pipeline_backward(dv, stages, (ds, ds), lam, saved)
        phi, st, tt = O.optim_step("fista", phi, g, st, tt, eta=eta * eta)
        phi = O.prox_clamp(phi, -math.pi, math.pi)
    torch.cuda.synchronize()
    assert rel_l2(doe.phi.cpu().numpy(), phi) <= 1e-3
    assert np.abs(phi).max() > 1e-2  # the loop moved the mask


@pytest.mark.parametrize("pname,prox,pfn", PROX, ids=[p[0] for p in PROX])
def test_adam_steps(pname, prox, pfn):
    """Adam (Optimisers.jl rule) fused with each proximal operator, 5 steps, against the
    oracle's step (FP32 state, FP64 step formation)."""
    n = 4096
    x0 = inputs.normal(251, (n,)).astype(np.float32)
    x = torch.from_numpy(x0.copy()).cuda()
    comp = fo.Phase.__new__(fo.Phase)
    comp.phi, comp.trainable = x, True
    rule = fo.Adam(0.1)
    state = fo.setup({comp: fo.ProxRule(rule, prox) if prox is not None else rule}, [comp])
    xr = x0.copy()
    m = np.zeros(n, np.float32)
    v = np.zeros(n, np.float32)
    bt = rule.beta
    for k in range(5):
        g = inputs.normal(260 + k, (n,)).astype(np.float32)
        fo.update_(state, [comp], {comp: torch.from_numpy(g).cuda()})
        xr, m, v, bt = O.adam_step(xr, g, m, v, bt, eta=rule.eta, beta=rule.beta, eps=rule.epsilon)
        xr = pfn(xr.astype(np.float64)).astype(np.float32)
    torch.cuda.synchronize()
    assert np.allclose(x.cpu().numpy(), xr, rtol=1e-6, atol=1e-6)
