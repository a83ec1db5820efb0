"""Optimisation rules and proximal operators for trainable components, stepped on the device.

Mirrors ``src/OptimisersExt.jl`` (``Fista``, ``NoDescent``, ``make_rules``, ``setup``,
``update!``, ``ProxRule``) and ``src/proximal_operators`` (``ClampProx``, ``IstaProx``,
``PositiveProx``, composition).  ``Descent``, ``Momentum`` and ``Nesterov`` restate the rules of
Optimisers.jl 0.4.6, the reference's dependency (``Manifest.toml``).  Each step is one fused
device kernel (rule + proximal operator) over the parameter (``fh_optim_step``).
"""

from __future__ import annotations

from typing import Dict, Iterable, Optional

import torch

from . import _lib
from .components import ASPropZ, BPM, _Modulator, _ptr, context, Phase
from .system import OpticalSystem, ScalarSource

_RULE_IDS = {"Descent": 0, "Momentum": 1, "Nesterov": 2, "Fista": 3}


class Descent:
    """x -= eta dx (Optimisers.jl ``Descent``; default eta 0.1)."""

    def __init__(self, eta: float = 0.1):
        self.eta = float(eta)


class Momentum:
    """m = rho m + eta dx; x -= m (Optimisers.jl ``Momentum``; defaults 0.01, 0.9)."""

    def __init__(self, eta: float = 0.01, rho: float = 0.9):
        self.eta, self.rho = float(eta), float(rho)


class Nesterov:
    """d = -rho^2 v + (1 + rho) eta dx; v = rho v - eta dx; x -= d (Optimisers.jl ``Nesterov``;
    defaults 0.001, 0.9)."""

    def __init__(self, eta: float = 0.001, rho: float = 0.9):
        self.eta, self.rho = float(eta), float(rho)


class Adam:
    """m = b1 m + (1 - b1) g, v = b2 v + (1 - b2) g^2, x -= eta m/(1 - b1^t)/(sqrt(v/(1 - b2^t)) + eps)
    (Optimisers.jl ``Adam``; defaults 0.001, (0.9, 0.999), 1e-8)."""

    def __init__(self, eta: float = 0.001, beta=(0.9, 0.999), epsilon: float = 1e-8):
        self.eta, self.beta, self.epsilon = float(eta), (float(beta[0]), float(beta[1])), float(epsilon)


class Fista:
    """``Fista(eta)`` (``OptimisersExt.jl:12-40``): stores eta^2 as the reference does; with
    t_1 = 1, t_{k+1} = (1 + sqrt(1 + 4 t_k^2)) / 2 and beta = (t_k - 1) / t_{k+1}:
    d = eta^2 dx - beta (x - x_prev), x_prev = x, x -= d."""

    def __init__(self, eta: float):
        self.eta = float(eta) ** 2


class NoDescent:
    """No update (``OptimisersExt.jl:52-56``)."""


class ClampProx:
    """``ClampProx(lo, hi)`` (``proximal_operators/clamp_prox.jl``)."""

    code = 1

    def __init__(self, lo: float, hi: float):
        self.p0, self.p1 = float(lo), float(hi)


class IstaProx:
    """``IstaProx(s, c=0)`` (``proximal_operators/ista_prox.jl``): soft threshold by s around c
    (real and imaginary parts separately for complex parameters)."""

    code = 2

    def __init__(self, s: float, c: float = 0.0):
        self.p0, self.p1 = float(s), float(c)


class PositiveProx:
    """``PositiveProx()`` (``proximal_operators/positive_prox.jl``): negatives to zero."""

    code = 3
    p0 = p1 = 0.0


class CompositeProx:
    """``a ∘ b``: applies the operators right to left (``ProximalOperators.jl:111-133``)."""

    def __init__(self, *ops):
        self.ops = tuple(ops)


def compose(*ops) -> CompositeProx:
    return CompositeProx(*ops)


class ProxRule:
    """``ProxRule(rule, prox)``: the rule's step, then the proximal operator."""

    def __init__(self, rule, prox):
        self.rule, self.prox = rule, prox


def _params(component):
    """(name, tensor) of a trainable component's parameters (``trainable(p)``)."""
    if isinstance(component, Phase):
        return [("phi", component.phi)]
    if isinstance(component, _Modulator):
        return [(type(component).__name__, component._param)]
    if isinstance(component, BPM):
        return [("dn", component.dn)]
    if isinstance(component, ScalarSource):
        return [("u0", component.u0.electric)]
    if isinstance(component, ASPropZ):
        return [("z", component.z)]
    return []


def _trainable_components(model) -> list:
    if isinstance(model, OpticalSystem):
        comps = ([model.source] if model.source is not None else []) + list(model.components)
    elif isinstance(model, (list, tuple)):
        comps = list(model)
    else:
        comps = [model]
    out = []
    for c in comps:
        if getattr(c, "trainable", False) and c not in out:
            out.append(c)
    return out


def make_rules(*pairs) -> Dict:
    """``make_rules(component => rule, ...)`` (``OptimisersExt.jl:160-190``): a dict keyed, as
    the reference's, by each component's trainable parameter arrays (``rules[phase.phi] is
    rule``), and by the component itself."""
    out = {}
    for c, r in pairs:
        out[c] = r
        for _, x in _params(c):
            out[x] = r
    return out


class _Leaf:
    def __init__(self, component, tensor, rule):
        self.component, self.x, self.rule = component, tensor, rule
        base = rule.rule if isinstance(rule, ProxRule) else rule
        self.base = base
        self.state = None
        if isinstance(base, (Momentum, Nesterov)):
            self.state = torch.zeros_like(tensor)
        elif isinstance(base, Adam):
            self.state = torch.zeros((2,) + tuple(tensor.shape), dtype=tensor.dtype, device=tensor.device)
            # b1^t, b2^t of the next step, on the device (advanced by each step)
            self.bt = torch.tensor(base.beta, dtype=torch.float64, device=tensor.device)
        elif isinstance(base, Fista):
            self.state = tensor.detach().clone()
            # {t_k, beta} on the device (each step advances them), t_1 = 1
            self.tb = torch.tensor([1.0, 0.0], dtype=torch.float64, device=tensor.device)


def _rule_of(rules: Dict, x: torch.Tensor, c, default_rule):
    # parameter keys first (tensors hash by identity), then the component
    for k, r in rules.items():
        if k is x:
            return r
    return rules.get(c, default_rule)


def setup(rules, model, default_rule=None):
    """``setup(rules, [default_rule,] model)`` (``OptimisersExt.jl:95-140``): one optimiser state
    per trainable parameter; parameters without a rule get ``default_rule`` (NoDescent).
    ``setup(rule, model)`` with a single rule applies it to every parameter, as
    ``FluxOptics.setup(Optimisers.Adam(0.1), system)`` does."""
    if not isinstance(rules, dict):
        rules, default_rule = {}, rules
    default_rule = NoDescent() if default_rule is None else default_rule
    leaves = []
    for c in _trainable_components(model):
        for _, x in _params(c):
            rule = _rule_of(rules, x, c, default_rule)
            leaves.append(_Leaf(c, x, rule))
    if not leaves:
        import warnings
        warnings.warn("setup found no trainable parameters in this model")
    return leaves


def _prox_list(prox) -> list:
    if prox is None:
        return []
    if isinstance(prox, CompositeProx):
        out = []
        for p in reversed(prox.ops):
            out.extend(_prox_list(p))
        return out
    return [prox]


def _step(leaf: _Leaf, grad: torch.Tensor) -> None:
    base = leaf.base
    if isinstance(base, NoDescent):
        return
    x = leaf.x
    if grad.shape != x.shape or grad.dtype != x.dtype:
        raise ValueError("gradient does not match its parameter")
    proxes = _prox_list(leaf.rule.prox if isinstance(leaf.rule, ProxRule) else None)
    if x.is_complex() and any(p.code in (1, 3) for p in proxes):
        raise NotImplementedError("clamp and positive projections are defined for real parameters only")
    xr = torch.view_as_real(x) if x.is_complex() else x
    gr = torch.view_as_real(grad) if grad.is_complex() else grad
    sr = None
    if leaf.state is not None:
        sr = torch.view_as_real(leaf.state) if leaf.state.is_complex() else leaf.state
    n = xr.numel()
    ctx = context(x.device)
    if isinstance(base, Adam):
        if x.is_complex():
            raise NotImplementedError("Adam is applied to real parameters here")
        first = proxes[0] if proxes else None
        _lib.call("fh_adam_step", ctx, _ptr(xr), _ptr(gr), _ptr(leaf.state), n, base.eta, base.beta[0], base.beta[1],
                  base.epsilon, _ptr(leaf.bt), first.code if first else 0, first.p0 if first else 0.0,
                  first.p1 if first else 0.0)
        zero = torch.zeros_like(xr)
        for p in proxes[1:]:
            _lib.call("fh_optim_step", ctx, 0, _ptr(xr), _ptr(zero), None, n, 0.0, 0.0, 0.0, p.code, p.p0, p.p1)
        return
    eta = base.eta
    rho = getattr(base, "rho", 0.0)
    first = proxes[0] if proxes else None
    if isinstance(base, Fista):
        _lib.call("fh_fista_step", ctx, _ptr(xr), _ptr(gr), _ptr(sr), n, eta, _ptr(leaf.tb),
                  first.code if first else 0, first.p0 if first else 0.0, first.p1 if first else 0.0)
    else:
        _lib.call("fh_optim_step", ctx, _RULE_IDS[type(base).__name__], _ptr(xr), _ptr(gr), _ptr(sr), n, eta, rho,
                  0.0, first.code if first else 0, first.p0 if first else 0.0, first.p1 if first else 0.0)
    zero = torch.zeros_like(xr)
    for p in proxes[1:]:  # further operators of a composition: a prox-only pass each
        _lib.call("fh_optim_step", ctx, 0, _ptr(xr), _ptr(zero), None, n, 0.0, 0.0, 0.0, p.code, p.p0, p.p1)


def update_(opt_state, model, grads: Dict) -> None:
    """``update!(opt_state, model, grads)``: one step for every parameter with a gradient
    (``grads`` as returned by ``OpticalSystem.withgradient``)."""
    for leaf in opt_state:
        g = grads.get(leaf.component)
        if g is not None:
            _step(leaf, g)
