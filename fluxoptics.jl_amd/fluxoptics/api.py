"""Function forms of the reference's field, FFT and component helpers, for code written against
its names (``Fields.jl``, ``GridUtils.jl``, ``FFTutils.jl``, ``OpticalComponents.jl``).  Each
maps onto the methods and kernels of this package; nothing here computes on the host except the
grid vectors, which the reference builds on the host too.
"""
from __future__ import annotations

from typing import Sequence

import numpy as np
import torch

from . import _lib
from .components import _ptr, check_field, context
from .fields import Forward, ScalarField


# ------------------------------------------------------------------ grids (GridUtils.jl:56-68)

def spatial_vectors(*args, offset=None, xc: float = 0.0, yc: float = 0.0):
    """``spatial_vectors(ns, ds; offset)``, ``spatial_vectors(nx, dx; xc)`` and
    ``spatial_vectors(nx, ny, dx, dy; xc, yc)``: centred grids ((0:n-1) - (n-1)/2) dx - c, one
    float64 vector per dimension."""
    if len(args) == 2 and isinstance(args[0], (tuple, list)):
        ns, ds = args
        off = offset if offset is not None else (0.0,) * len(ns)
    elif len(args) == 2:
        ns, ds, off = (args[0],), (args[1],), (xc,)
    elif len(args) == 4:
        ns, ds, off = args[:2], args[2:], (xc, yc)
    else:
        raise TypeError("spatial_vectors(ns, ds), (nx, dx) or (nx, ny, dx, dy)")
    return tuple((np.arange(n, dtype=np.float64) - (n - 1) / 2.0) * d - c for n, d, c in zip(ns, ds, off))


# ------------------------------------------------------------------ fields (Fields.jl)

def get_lambdas(u: ScalarField):
    """``get_lambdas`` (``Fields.jl:153-155``)."""
    return u.lambdas


def get_lambdas_collection(u: ScalarField):
    """``get_lambdas_collection`` (``Fields.jl:157-159``): one wavelength per field."""
    return u.lambda_list()


def get_tilts(u: ScalarField):
    """``get_tilts`` (``Fields.jl:168-170``)."""
    return u.tilts


def get_tilts_collection(u: ScalarField):
    """``get_tilts_collection`` (``Fields.jl:172-174``): per-field angles for x and y."""
    return u.tilt_lists()


def is_on_axis(u: ScalarField) -> bool:
    """``is_on_axis`` (``Fields.jl:254-256``)."""
    return u.is_on_axis()


def set_field_data(u: ScalarField, data: torch.Tensor) -> ScalarField:
    """``set_field_data`` (``Fields.jl:213-215``)."""
    return u.set_field_data(data)


def set_field_tilts(u: ScalarField, tilts) -> ScalarField:
    """``set_field_tilts`` (``Fields.jl:243-245``)."""
    return u.set_field_tilts(tilts)


def set_field_ds_(u: ScalarField, ds) -> ScalarField:
    """``set_field_ds!`` (``Fields.jl:182-185``)."""
    if len(ds) != 2:
        raise ValueError("ds must be (dx, dy)")
    u.ds = (float(ds[0]), float(ds[1]))
    return u


def offset_tilts_(u: ScalarField, tilts) -> ScalarField:
    """``offset_tilts!`` (``Fields.jl:280-293``), on the device."""
    return u.offset_tilts_(tilts)


def dot(u: ScalarField, v: ScalarField) -> torch.Tensor:
    """``dot(u, v)`` (``Fields.jl:810-814``): per field sum conj(u) v (no area factor), complex128."""
    check_field(u.electric)
    check_field(v.electric)
    if tuple(u.electric.shape) != tuple(v.electric.shape):
        raise ValueError("fields must have the same shape")
    a = u.electric.reshape(u.batch, -1).to(torch.complex128)
    b = v.electric.reshape(v.batch, -1).to(torch.complex128)
    return (a.conj() * b).sum(dim=1)


# ------------------------------------------------------------------ FFT plans (FFTutils.jl:13-28)

class FFTPlans:
    """``make_fft_plans(u, dims)``: the in-place 2-D FFT (sign -1, scale 1) and IFFT (sign +1,
    scale 1/N) over the field's spatial dimensions, run by ``fh_fft2``."""

    def __init__(self, u, dims=(1, 2)):
        if tuple(dims) != (1, 2):
            raise NotImplementedError("FFT plans over both spatial dimensions only (dims = (1, 2))")
        e = u.electric if isinstance(u, ScalarField) else u
        self.shape = tuple(e.shape)

    def _run(self, e: torch.Tensor, sign: int, scale: float) -> torch.Tensor:
        check_field(e)
        if tuple(e.shape) != self.shape:
            raise ValueError("array does not match the plan")
        nx, ny = e.shape[-1], e.shape[-2]
        batch = 1 if e.dim() == 2 else e.shape[0]
        _lib.call("fh_fft2", context(e.device), _ptr(e), nx, ny, batch, sign, float(scale))
        return e

    def ft(self, e: torch.Tensor) -> torch.Tensor:
        return self._run(e, -1, 1.0)

    def ift(self, e: torch.Tensor) -> torch.Tensor:
        return self._run(e, +1, 1.0 / (e.shape[-1] * e.shape[-2]))


def make_fft_plans(u, dims=(1, 2)) -> FFTPlans:
    """``make_fft_plans(u, dims)`` (``FFTutils.jl:13-18``)."""
    return FFTPlans(u, dims)


def compute_ft_(plans: FFTPlans, u: ScalarField) -> ScalarField:
    """``compute_ft!(p, u)`` (``FFTutils.jl:20-23``): in place."""
    plans.ft(u.electric)
    return u


def compute_ift_(plans: FFTPlans, u: ScalarField) -> ScalarField:
    """``compute_ift!(p, u)`` (``FFTutils.jl:25-28``): in place."""
    plans.ift(u.electric)
    return u


# ------------------------------------------------------------------ components (OpticalComponents.jl)

def istrainable(c) -> bool:
    """``istrainable`` (``OpticalComponents.jl``): the component has trainable parameters."""
    return bool(getattr(c, "trainable", False))


def isbuffered(c) -> bool:
    """``isbuffered``: trainable with pre-allocated saved input and gradient buffers."""
    return bool(getattr(c, "buffered", False)) or getattr(c, "grad", None) is not None


def trainable(c) -> dict:
    """``trainable(p)``: the component's trainable parameters by name (empty when static)."""
    from .optim import _params
    return dict(_params(c)) if istrainable(c) else {}


def get_data(c) -> torch.Tensor:
    """``get_data``: the component's parameter array (a source's field data)."""
    from .optim import _params
    ps = _params(c)
    if not ps:
        raise ValueError(f"{type(c).__name__} has no parameter array")
    return ps[0][1]


def get_source(src) -> ScalarField:
    """``get_source`` (``scalar_source.jl``): the source's stored field."""
    return src.u0


def compute_metric(m, *fields):
    """``compute_metric(m, fields...)``: one value tensor per (field, target) pair."""
    return m.compute_metric(*fields)


def backpropagate_metric(m, fields: Sequence[ScalarField], dcs=None):
    """``backpropagate_metric(m, fields, ∂c)``: the cotangent fields."""
    return m.backpropagate_metric(fields, dcs)


def backpropagate(dv: ScalarField, c, direction=Forward) -> ScalarField:
    """``backpropagate(∂v, p, direction)``: the adjoint of ``propagate`` on a copy of ∂v."""
    return c.backpropagate_(dv.copy(), direction)
