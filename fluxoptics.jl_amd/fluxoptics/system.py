"""Sources, sequences, systems and the gradient driver.

Mirrors ``src/optical_components/sources/scalar_source.jl`` (``ScalarSource``),
``src/optical_components/system/optical_sequence.jl`` (``OpticalSequence``: forward loop,
reverse loop with the reversed direction, :54-66), ``optical_system.jl`` (``OpticalSystem``,
the pipe ``|>`` written ``|`` here, ``system()`` returning ``(out, probes)``, :126-214) and
``merge_rules.jl`` (static Phase + Phase adds the masks; consecutive Fourier operators of
opposite directions cancel, which for two on-axis angular-spectrum propagators in one medium
leaves one propagator over the summed distance).

The reference differentiates a system with Zygote through the component rrules
(``ChainRulesCoreExt.jl:21-111``): the forward runs ``propagate_and_save!`` on trainable
components and the pullback runs ``backpropagate_with_gradient!`` in reverse order.
``withgradient`` does exactly that sequence with the device kernels, and fuses each
``Phase`` that precedes a propagator into the propagator's first pass (forward) and last
pass (adjoint + mask gradient).
"""

from __future__ import annotations

from typing import Callable, Dict, List, Optional, Sequence

import torch

from .components import (ASPropZ, BPM, CollinsProp, TiltAnchor, ASProp, FourierOperator, Phase, RSProp, _KernelPropagator,
                         _Modulator, _ptr, check_field)
from .components import phase_propagate_, propagate_phase_backward_
from .fields import Backward, Forward, ScalarField, dir_sign
from .metrics import PowerCoupling, SquaredIntensityDifference
from .padcrop import FieldProbe
from . import _lib


class ScalarSource:
    """``ScalarSource(u; trainable=false, buffered=false)``: reproduces a stored copy of ``u``."""

    def __init__(self, u: ScalarField, *, trainable: bool = False, buffered: bool = False):
        check_field(u.electric)
        self.u0 = u.copy()
        self.trainable = trainable
        self.grad = torch.empty_like(u.electric) if (trainable and buffered) else None

    def propagate(self) -> ScalarField:
        return self.u0.copy()

    def backpropagate_with_gradient(self, dv: ScalarField) -> torch.Tensor:
        g = self.grad if self.grad is not None else torch.empty_like(dv.electric)
        g.copy_(dv.electric)
        return g

    def __or__(self, other):
        return OpticalSystem(self, *_components_of(other))


def _components_of(x) -> list:
    if isinstance(x, OpticalSystem):
        if x.source is not None:
            raise ValueError("cannot pipe into a system that has a source")
        return list(x.components)
    if isinstance(x, OpticalSequence):
        return list(x.optical_components)
    return [x]


def _is_static_phase(c) -> bool:
    return isinstance(c, Phase) and not c.trainable


def _merge_pair(a, b):
    """``merge_rules.jl``: returns the merged component, ``_CANCEL`` when the pair is the
    identity (FFT followed by IFFT), or None."""
    if (isinstance(a, FourierOperator) and isinstance(b, FourierOperator) and a.direct != b.direct
            and (a.nx, a.ny) == (b.nx, b.ny)):
        return _CANCEL
    if _is_static_phase(a) and _is_static_phase(b) and a.phi.shape == b.phi.shape:
        return Phase._from_tensor(a.phi + b.phi)
    if (type(a) is ASProp and type(b) is ASProp and a.n0 == b.n0 and a.ds == b.ds
            and (a.nx, a.ny) == (b.nx, b.ny) and not a.paraxial and not b.paraxial
            and not a.track_tilts and not b.track_tilts
            and getattr(a, "filter", None) is None and getattr(b, "filter", None) is None
            and not isinstance(a.z, list) and not isinstance(b.z, list)
            and (a.z >= 0) == (b.z >= 0)):
        # H(z1) H(z2) = H(z1 + z2) for distances of one sign (evanescent terms decay with |z|)
        return ASProp._from_params(a.nx, a.ny, a.ds, a.z + b.z, a.n0)
    return None


_CANCEL = object()


def merge(components: Sequence) -> list:
    out: list = []
    for c in components:
        while out:
            m = _merge_pair(out[-1], c)
            if m is None:
                break
            out.pop()
            c = m
            if c is _CANCEL:
                break
        if c is not _CANCEL:
            out.append(c)
    return out


class OpticalSequence:
    """``OpticalSequence(components...)``: forward loop / reverse loop with the reversed
    direction (``optical_sequence.jl:54-66``)."""

    def __init__(self, *components):
        self.optical_components = tuple(components)

    def __len__(self):
        return len(self.optical_components)

    @property
    def trainable(self) -> bool:
        return any(getattr(c, "trainable", False) for c in self.optical_components)

    def propagate_(self, u: ScalarField, direction=Forward) -> ScalarField:
        comps = self.optical_components if dir_sign(direction) > 0 else self.optical_components[::-1]
        for c in comps:
            u = c.propagate_(u, direction)
        return u

    def propagate(self, u: ScalarField, direction=Forward) -> ScalarField:
        return self.propagate_(u.copy(), direction)

    def backpropagate_(self, dv: ScalarField, direction=Forward) -> ScalarField:
        return self.propagate_(dv, Backward if dir_sign(direction) > 0 else Forward)


class SystemOutput:
    def __init__(self, out: ScalarField, probes: Optional[Dict] = None):
        self.out = out
        self.probes = probes or {}


class OpticalSystem:
    """``OpticalSystem(source, components...; merge_components=false, inplace=false)``.

    ``source | c1 | c2`` builds it; ``system | {"merge_components": True}`` rebuilds it with
    keywords; ``system()`` (or ``system(x)`` without a source) returns ``.out``.
    """

    def __init__(self, source, *components, inplace: bool = False, direction=Forward,
                 merge_components: bool = False):
        if source is not None and not isinstance(source, ScalarSource):
            components = (source,) + components
            source = None
        self.source = source
        flat = []
        for c in components:  # sequences (e.g. FourierWrapper) run as their components
            flat.extend(c.optical_components if isinstance(c, OpticalSequence) else [c])
        components = tuple(flat)
        self.input_components = tuple(components)
        self.components = tuple(merge(components)) if merge_components else tuple(components)
        self.inplace = inplace
        self.direction = direction
        self.merge_components = merge_components

    def __len__(self):
        return len(self.components)

    def __or__(self, other):
        if isinstance(other, dict):
            kw = dict(inplace=self.inplace, direction=self.direction,
                      merge_components=self.merge_components)
            kw.update(other)
            return OpticalSystem(self.source, *self.input_components, **kw)
        return OpticalSystem(self.source, *self.input_components, *_components_of(other),
                             inplace=self.inplace, direction=self.direction,
                             merge_components=self.merge_components)

    def get_source(self):
        return self.source

    def get_components(self):
        return self.components

    # -- forward
    def _start(self, x: Optional[ScalarField]) -> ScalarField:
        if (x is None) == (self.source is None):
            raise ValueError("give an input field exactly when the system has no source")
        if x is None:
            return self.source.propagate()
        return x if self.inplace else x.copy()

    def __call__(self, x: Optional[ScalarField] = None) -> SystemOutput:
        """``system(x)`` (``optical_system.jl:189-214``): ``out`` and, for every ``FieldProbe``,
        a copy of the field where it sits (``probes``)."""
        u = self._start(x)
        probes: Dict = {}
        _run_forward(list(self.components), u, self.direction, save=False, probes=probes)
        return SystemOutput(u, probes)

    # -- gradient
    def withgradient(self, loss, x: Optional[ScalarField] = None, weights=None, fuse: bool = True):
        """Loss value and gradients of every trainable component.

        ``loss`` is a metric with ``value_and_pullback`` (``SquaredIntensityDifference``,
        ``PowerCoupling``); the objective is the sum of its values, each times its weight
        (``weights``: one per value or a scalar, e.g. -1 to maximise a coupling).  Returns
        ``(value, grads)`` where ``value`` is the weighted sum as a float64 device scalar
        and ``grads`` maps each trainable component (and a trainable source) to its gradient.
        A system ending in [Phase ->] ASProp/RSProp with a one-pair SquaredIntensityDifference
        (one target per field) runs its tail as one fused block (``fh_sqint_block``: two HBM
        passes fewer); ``fuse=False`` runs the components and the metric one by one.
        """
        losses, grads = self._withgradient_launches(loss, x, weights, fuse)
        return _total(losses, weights), grads

    def _withgradient_launches(self, loss, x, weights=None, fuse: bool = True):
        """The device work of ``withgradient``: per-pair loss tensors and the gradients."""
        if (x is None) == (self.source is None):
            raise ValueError("give an input field exactly when the system has no source")
        start = self.source.u0 if x is None else x
        comps = list(self.components)
        tail = _fusable_tail(comps, loss, start, self.direction) if fuse else None
        grads: Dict = {}
        if tail is not None and not tail[0] and (x is None or not self.inplace):
            # the whole system is the fused block: it reads the start field itself (no copy of
            # the source field, and the Phase gradient takes it as its saved input)
            _, ph, prop, trail = tail
            u = start.set_field_data(torch.empty_like(start.electric))
            # the cotangent at the source is needed only for a trainable source's gradient
            need_du = self.source is not None and self.source.trainable
            losses = (_sqint_block(loss, u, ph, prop, [], grads, weights, src=start.electric,
                                   cotangent=need_du),)
            _zero_grads(trail, grads)
            du = u
        elif fuse and (dtail := _dot_tail(comps, loss, start, self.direction)) is not None:
            # [Phase ->] propagator -> PowerCoupling: the pullback rides on the adjoint's first pass
            u = self._start(x)
            head, ph, prop = dtail
            tape = _run_forward(head, u, self.direction, save=True)
            losses = (_dot_block(loss, u, ph, prop, tape, grads, weights),)
            du = _run_backward(head, tape, u, self.direction, grads)
        elif tail is not None:
            u = self._start(x)
            head, ph, prop, trail = tail
            tape = _run_forward(head, u, self.direction, save=True)
            losses = (_sqint_block(loss, u, ph, prop, tape, grads, weights),)
            _zero_grads(trail, grads)
            du = _run_backward(head, tape, u, self.direction, grads)
        else:
            u = self._start(x)
            tape = _run_forward(comps, u, self.direction, save=True)
            dcs = None if weights is None else [weights]
            losses, (du,) = loss.value_and_pullback((u,), dcs, inplace=True)
            du = _run_backward(comps, tape, du, self.direction, grads)
        if self.source is not None and self.source.trainable:
            grads[self.source] = self.source.backpropagate_with_gradient(du)
        return losses, grads

    def capture_train_step(self, loss, opt_state, x: Optional[ScalarField] = None, weights=None) -> "CapturedGradient":
        """Record one optimisation iteration -- ``withgradient`` and ``update_(opt_state, ...)``
        -- as one HIP graph; each ``replay()`` runs an iteration (parameters and optimiser state
        move on) and returns ``(value, grads)`` of that iteration.  The recording itself changes
        nothing; one eager ``withgradient`` runs first so that every plan and workspace exists."""
        return _capture_train_step(self, loss, opt_state, x, weights)

    def capture_withgradient(self, loss, x: Optional[ScalarField] = None, weights=None) -> "CapturedGradient":
        """Record the whole ``withgradient`` launch sequence (forward with saves, metric and
        pullback, adjoint with gradients) once as a HIP graph; ``replay()`` then reruns it with
        one graph launch and returns ``(value, grads)`` in the same buffers every time.

        The inputs are those of the capture: the source's field (or ``x``'s buffer), the
        masks and volumes of the components and the metric's targets; change them in place
        between replays.  One eager call runs first so that every plan, table and workspace
        exists before capture.
        """
        _check_recordable(self)
        if weights is not None:  # a device copy that lives as long as the graph
            weights = torch.as_tensor(weights, dtype=torch.float64).cuda()
        self._withgradient_launches(loss, x, weights)
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            losses, grads = self._withgradient_launches(loss, x, weights)
        return CapturedGradient(graph, losses, grads, weights)


def _check_recordable(system):
    """An ``ASPropZ`` reads its distance on the host to pick its transfer function, which a
    recorded graph can neither do (a host read during capture) nor follow (the tables of the
    recorded z would replay after z moved on), so systems holding one are not recorded."""
    for c in system.components:
        if isinstance(c, ASPropZ):
            raise NotImplementedError("ASPropZ picks its transfer function from z on the host; a system "
                                      "holding one runs eagerly (withgradient), not as a recorded graph")


def _capture_train_step(system, loss, opt_state, x=None, weights=None):
    from .optim import update_
    _check_recordable(system)
    if weights is not None:
        weights = torch.as_tensor(weights, dtype=torch.float64).cuda()
    system._withgradient_launches(loss, x, weights)  # plans, tables, workspaces; parameters unchanged
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        losses, grads = system._withgradient_launches(loss, x, weights)
        update_(opt_state, system, grads)
    return CapturedGradient(graph, losses, grads, weights)


def _total(losses, weights=None):
    if weights is None:
        return sum(l.sum() for l in losses)
    w = torch.as_tensor(weights, dtype=torch.float64, device=losses[0].device).reshape(-1)
    return sum((l * w).sum() for l in losses)


class CapturedGradient:
    """A recorded ``withgradient`` (see ``OpticalSystem.capture_withgradient``)."""

    def __init__(self, graph, losses, grads, weights=None):
        self.graph, self.losses, self.grads, self.weights = graph, losses, grads, weights

    def replay(self):
        self.graph.replay()
        return _total(self.losses, self.weights), self.grads


# --------------------------------------------------------------------------- executor

def _is_prop(c) -> bool:
    """A propagator whose plan a preceding Phase can ride on (its first pass has no chirp)."""
    return isinstance(c, _KernelPropagator) and not isinstance(c, CollinsProp)


def _fusable_tail(comps: List, loss, u: ScalarField, direction):
    """(head, Phase or None, propagator, trailing Phases) when the system ends in [Phase ->]
    ASProp/RSProp, possibly followed by Phases, and the loss is a one-pair
    SquaredIntensityDifference with one target per field; else None.  A Phase leaves |u|^2 as it
    is, so trailing Phases change neither the loss nor the pullback, and their mask gradients are
    exactly zero."""
    if dir_sign(direction) < 0 or not comps or not isinstance(loss, SquaredIntensityDifference):
        return None
    if getattr(u, "nd", 2) == 1:  # 1-D propagations are one pass already
        return None
    k = len(comps)
    while k > 0 and type(comps[k - 1]) is Phase:
        k -= 1
    trail, comps = comps[k:], comps[:k]
    if len(loss.targets) != 1 or not comps or not _is_prop(comps[-1]):
        return None
    t = loss.targets[0]
    tb = 1 if t.dim() == 2 else t.shape[0]
    if not (tb == u.batch or (tb == 1 and u.batch == 1)):
        return None
    if len(comps) >= 2 and isinstance(comps[-2], Phase):
        return comps[:-2], comps[-2], comps[-1], trail
    return comps[:-1], None, comps[-1], trail


def _dot_tail(comps: List, loss, u: ScalarField, direction):
    """(head, Phase or None, propagator) when the system ends in [Phase ->] ASProp/RSProp and the
    loss is a mode-selective PowerCoupling with one target field; else None."""
    if dir_sign(direction) < 0 or not comps or not isinstance(loss, PowerCoupling):
        return None
    if not loss.mode_selective or len(loss.v) != 1 or not _is_prop(comps[-1]):
        return None
    if tuple(loss.v[0].electric.shape) != tuple(u.electric.shape):
        return None
    if len(comps) >= 2 and isinstance(comps[-2], Phase):
        return comps[:-2], comps[-2], comps[-1]
    return comps[:-1], None, comps[-1]


def _dot_block(loss, u: ScalarField, ph, prop, tape, grads: Dict, weights):
    """[Phase ->] propagation, the PowerCoupling values, then the pullback fused with the adjoint
    propagation and [the Phase adjoint with its gradient] (``fh_dot_pullback_propagate``): u ends as
    the cotangent at the block's input."""
    plan = prop.plan(u)
    phi = saved = g = None
    acc = False
    if ph is not None:
        phi = ph.phi
        if ph.trainable:
            saved = ph.saved if (ph.saved is not None and not any(e[2] is ph.saved for e in tape)) \
                else torch.empty_like(u.electric)
            g, acc = _grad_buffer(grads, ph, ph.phi)
        _lib.call("fh_phase_propagate", plan, _ptr(u.electric), _ptr(phi), _ptr(saved), dir_sign(Forward))
    else:
        prop.propagate_(u, Forward)
    (vals,) = loss.compute_metric(u)
    dp = None
    if weights is not None:
        dp = torch.as_tensor(weights, dtype=torch.float64, device=u.electric.device).reshape(-1)
        dp = (dp.expand(u.batch) if dp.numel() == 1 else dp).contiguous()
    _lib.call("fh_dot_pullback_propagate", plan, _ptr(loss.v[0].electric), None, _ptr(loss._c[0]), _ptr(dp),
              _ptr(u.electric), dir_sign(Forward), _ptr(phi), _ptr(saved), _ptr(g), int(acc))
    return vals


def _zero_grads(phases, grads: Dict) -> None:
    """Gradients of Phases that the loss does not see (their sum is unchanged where a Phase
    already has one)."""
    for ph in phases:
        if ph.trainable:
            g, acc = _grad_buffer(grads, ph, ph.phi)
            if not acc:
                g.zero_()


def _sqint_block(loss, u: ScalarField, ph, prop, tape, grads: Dict, weights, src=None, cotangent=True):
    """The fused tail (``fh_sqint_block``); u ends as the cotangent at the tail's input.  With
    ``src`` (the tail's input, left unchanged) the first pass reads it and u is only written
    (``fh_sqint_block_src``); there ``cotangent=False`` leaves u unspecified and skips writing it."""
    t = loss.targets[0]
    tb = 1 if t.dim() == 2 else t.shape[0]
    if t.shape[-2:] != (u.ny, u.nx):
        raise ValueError("field does not match the metric's target")
    nout = u.batch if tb == u.batch else 1
    dev = u.electric.device
    out = torch.empty(nout, dtype=torch.float64, device=dev)
    dc = None
    if weights is not None:
        dc = torch.as_tensor(weights, dtype=torch.float64, device=dev).reshape(-1)
        dc = (dc.expand(nout) if dc.numel() == 1 else dc).contiguous()
        if dc.numel() != nout:
            raise ValueError(f"dc must hold {nout} values")
    phi = saved = g = None
    acc = False
    if ph is not None:
        phi = ph.phi
        if ph.trainable:
            if src is None:
                saved = ph.saved if (ph.saved is not None and not any(e[2] is ph.saved for e in tape)) \
                    else torch.empty_like(u.electric)
            g, acc = _grad_buffer(grads, ph, ph.phi)
    plan = prop.plan(u)
    if src is None:
        _lib.call("fh_sqint_block", plan, _ptr(u.electric), _ptr(t), tb, _ptr(dc), _ptr(out), dir_sign(Forward),
                  _ptr(phi), _ptr(saved), _ptr(g), int(acc))
    else:
        _lib.call("fh_sqint_block_src", plan, _ptr(src), _ptr(u.electric), _ptr(t), tb, _ptr(dc), _ptr(out),
                  dir_sign(Forward), _ptr(phi), _ptr(g), int(acc), int(cotangent))
    return out


def _run_forward(comps: List, u: ScalarField, direction, save: bool, probes: Optional[Dict] = None) -> list:
    """Run the components in order; returns the tape of (kind, components, saved buffer).
    A Phase followed by a propagator runs fused with it; a FieldProbe stores a copy of the field
    in ``probes`` when that is given."""
    tape = []
    seq = comps if dir_sign(direction) > 0 else comps[::-1]
    i = 0
    while i < len(seq):
        c = seq[i]
        if isinstance(c, FieldProbe):
            if probes is not None:
                probes[c] = u.copy()
            tape.append(("plain", (c,), None))
            i += 1
            continue
        if isinstance(c, Phase) and i + 1 < len(seq) and _is_prop(seq[i + 1]):
            p = seq[i + 1]
            saved = None
            if save and c.trainable:
                saved = c.saved if (c.saved is not None and not any(t[2] is c.saved for t in tape)) \
                    else torch.empty_like(u.electric)
            phase_propagate_(u, c, p, saved, direction)
            tape.append(("phase+prop", (c, p), saved))
            i += 2
            continue
        if save and isinstance(c, Phase) and c.trainable:
            saved = c.saved if (c.saved is not None and not any(t[2] is c.saved for t in tape)) \
                else torch.empty_like(u.electric)
            c.propagate_and_save_(u, saved, direction)
            tape.append(("phase", (c,), saved))
        elif save and isinstance(c, _Modulator) and c.trainable:
            saved = c.saved if (c.saved is not None and not any(t[2] is c.saved for t in tape)) \
                else torch.empty_like(u.electric)
            c.propagate_and_save_(u, saved, direction)
            tape.append(("mod", (c,), saved))
        elif save and isinstance(c, ASPropZ) and c.trainable:
            if dir_sign(direction) < 0:
                raise NotImplementedError("ASPropZ's distance gradient is taken in the Forward direction")
            c.propagate_(u, direction)
            tape.append(("asz", (c,), u.copy()))
        elif save and isinstance(c, TiltAnchor):
            tape.append(("anchor", (c,), u.tilts))
            c.propagate_(u, direction)
        elif save and isinstance(c, BPM) and c.trainable:
            if any(t[0] == "bpm" and t[1][0] is c for t in tape):
                raise NotImplementedError("a trainable BPM may appear once in a differentiated system "
                                          "(it keeps one set of checkpoints)")
            c.propagate_and_save_(u, direction)
            tape.append(("bpm", (c,), None))
        else:
            c.propagate_(u, direction)
            tape.append(("plain", (c,), None))
        i += 1
    return tape


def _grad_buffer(grads: Dict, c, like: torch.Tensor):
    """(buffer, accumulate) for a trainable component's gradient."""
    if c in grads:
        return grads[c], True
    g = c.grad if getattr(c, "grad", None) is not None else torch.empty_like(like)
    grads[c] = g
    return g, False


def _run_backward(comps: List, tape: list, du: ScalarField, direction, grads: Dict) -> ScalarField:
    for kind, cs, saved in reversed(tape):
        if kind == "asz":
            (c,) = cs
            gz = c.z_gradient(saved, du)
            if c in grads:
                grads[c] += gz
            else:
                grads[c] = c.grad.copy_(gz) if c.grad is not None else gz
            c.backpropagate_(du, direction)
            continue
        if kind == "phase+prop":
            ph, p = cs
            if ph.trainable:
                g, acc = _grad_buffer(grads, ph, ph.phi)
                propagate_phase_backward_(du, ph, p, saved, g, direction, acc)
            else:
                propagate_phase_backward_(du, ph, p, None, None, direction)
        elif kind == "phase":
            (ph,) = cs
            g, acc = _grad_buffer(grads, ph, ph.phi)
            ph.backpropagate_with_gradient_(du, saved, g, direction, accumulate=acc)
        elif kind == "mod":
            (c,) = cs
            g, acc = _grad_buffer(grads, c, c._param)
            c.backpropagate_with_gradient_(du, saved, g, direction, accumulate=acc)
        elif kind == "anchor":
            du.offset_tilts_(saved)
        elif kind == "bpm":
            (m,) = cs
            if m in grads:
                tmp = torch.empty_like(m.dn)
                m.backpropagate_with_gradient_(du, tmp, direction)
                grads[m] += tmp
            else:
                g = m.grad if m.grad is not None else torch.empty_like(m.dn)
                m.backpropagate_with_gradient_(du, g, direction)
                grads[m] = g
        else:
            (c,) = cs
            c.backpropagate_(du, direction)
    return du
