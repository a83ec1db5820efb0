"""Pad and crop of fields (``src/optical_components/utilities/pad_crop_operators.jl``) and the
field probe of systems (``src/optical_components/system/field_probe.jl``).

Sizes and offsets are given in the reference's order, (x, y) = (dim 1, dim 2) of its
column-major ``[nx, ny, batch]`` arrays, which are the last and second-to-last dimensions of
the tensors here.  The data movement runs on the device (``fh_window``); CPU tensors raise.
"""
from __future__ import annotations

from typing import Sequence, Tuple

import torch

from . import _lib
from .components import _PipeComponent, _ptr, check_field, context
from .fields import ScalarField, dir_sign, Forward


def _xy(ns: Sequence[int], what: str) -> Tuple[int, int]:
    if len(ns) != 2:
        raise ValueError(f"{what} must have two entries (x, y)")
    return int(ns[0]), int(ns[1])


def _window_(dst: torch.Tensor, src: torch.Tensor, x0: int, y0: int, fill: complex = 0.0) -> torch.Tensor:
    check_field(src, "source")
    check_field(dst, "destination")
    sb = 1 if src.dim() == 2 else src.shape[0]
    db = 1 if dst.dim() == 2 else dst.shape[0]
    if sb != db:
        raise ValueError("source and destination must hold the same number of fields")
    fill = complex(fill)
    _lib.call("fh_window", context(src.device), _ptr(src), src.shape[-1], src.shape[-2], _ptr(dst), dst.shape[-1],
              dst.shape[-2], sb, int(x0), int(y0), fill.real, fill.imag)
    return dst


def _resized(u: torch.Tensor, nx: int, ny: int) -> torch.Tensor:
    shape = (ny, nx) if u.dim() == 2 else (u.shape[0], ny, nx)
    return torch.empty(shape, dtype=u.dtype, device=u.device)


def pad_(u_pad: torch.Tensor, u: torch.Tensor, offset=(0, 0), pad_val: complex = 0.0) -> torch.Tensor:
    """``pad!(u_pad, u; offset, pad_val)`` (``pad_crop_operators.jl:1-16``): u at ``offset``
    inside u_pad, ``pad_val`` elsewhere."""
    ox, oy = _xy(offset, "offset")
    if ox < 0 or oy < 0:
        raise ValueError("offsets must be non-negative")
    if u.shape[-1] + ox > u_pad.shape[-1] or u.shape[-2] + oy > u_pad.shape[-2]:
        raise ValueError("the field does not fit the padded array at this offset")
    return _window_(u_pad, u, ox, oy, pad_val)


def pad(u: torch.Tensor, ns, offset=(0, 0), pad_val: complex = 0.0) -> torch.Tensor:
    """``pad(u, ns; offset, pad_val)`` (``pad_crop_operators.jl:53-59``): a new array of (x, y)
    size ``ns`` holding u at ``offset``."""
    nx, ny = _xy(ns, "ns")
    return pad_(_resized(u, nx, ny), u, offset, pad_val)


def crop_(u_crop: torch.Tensor, u: torch.Tensor, offset=(0, 0)) -> torch.Tensor:
    """``crop!(u_crop, u; offset)`` (``pad_crop_operators.jl:61-79``): the region of u of u_crop's
    size starting at ``offset``."""
    ox, oy = _xy(offset, "offset")
    if ox < 0 or oy < 0:
        raise ValueError("offsets must be non-negative")
    if u_crop.shape[-1] + ox > u.shape[-1] or u_crop.shape[-2] + oy > u.shape[-2]:
        raise ValueError("the cropped region lies outside the array")
    return _window_(u_crop, u, -ox, -oy)


def crop(u: torch.Tensor, ns, offset=(0, 0)) -> torch.Tensor:
    """``crop(u, ns; offset)`` (``pad_crop_operators.jl:114-119``): a new array of (x, y) size
    ``ns`` cut from u at ``offset``."""
    nx, ny = _xy(ns, "ns")
    return crop_(_resized(u, nx, ny), u, offset)


def _electric(u):
    return u.electric if isinstance(u, ScalarField) else u


class PadOperator:
    """``PadOperator(u, ns or u_tmp; offset, pad_val)`` (``pad_crop_operators.jl:122-167``):
    called on a field, returns it with its data padded (into ``u_tmp`` when given)."""

    def __init__(self, u, ns_or_tmp, offset=(0, 0), pad_val: complex = 0.0):
        self.u_tmp = ns_or_tmp if isinstance(ns_or_tmp, torch.Tensor) else None
        self.size_pad = ((ns_or_tmp.shape[-1], ns_or_tmp.shape[-2]) if self.u_tmp is not None
                         else _xy(ns_or_tmp, "ns"))
        self.offset = _xy(offset, "offset")
        self.pad_value = complex(pad_val)

    def __call__(self, u: ScalarField) -> ScalarField:
        e = (pad(u.electric, self.size_pad, self.offset, self.pad_value) if self.u_tmp is None
             else pad_(self.u_tmp, u.electric, self.offset, self.pad_value))
        return u.set_field_data(e)


class CropOperator:
    """``CropOperator(u, ns or u_tmp; offset)`` (``pad_crop_operators.jl:170-211``)."""

    def __init__(self, u, ns_or_tmp, offset=(0, 0)):
        self.u_tmp = ns_or_tmp if isinstance(ns_or_tmp, torch.Tensor) else None
        self.size_crop = ((ns_or_tmp.shape[-1], ns_or_tmp.shape[-2]) if self.u_tmp is not None
                          else _xy(ns_or_tmp, "ns"))
        self.offset = _xy(offset, "offset")

    def __call__(self, u: ScalarField, u_tmp: torch.Tensor = None) -> ScalarField:
        dst = u_tmp if u_tmp is not None else self.u_tmp
        e = (crop(u.electric, self.size_crop, self.offset) if dst is None
             else crop_(dst, u.electric, self.offset))
        return u.set_field_data(e)


class PadCropOperator(_PipeComponent):
    """``PadCropOperator(u, u_tmp; offset, pad_val, store_ref)`` (``pad_crop_operators.jl:214-356``):
    pads a field to ``u_tmp``'s (x, y) size in the Forward direction and crops it back in the
    Backward direction; ``adjoint()`` swaps the two.  The crop is placed at offset zero of the
    padded array, as the reference builds it (``CropOperator(u_tmp, size(u))``).  With
    ``store_ref`` the operator keeps the array the pad received and the matching crop writes
    into it (no allocation).  Its pullback is the reverse direction (``backpropagate!``)."""

    trainable = False

    def __init__(self, u, u_tmp: torch.Tensor = None, offset=(0, 0), pad_val: complex = 0.0,
                 store_ref: bool = False, *, _parts=None):
        if _parts is not None:
            self.p_pad, self.p_crop, self._ref, self.ispad, self.store_ref = _parts
            return
        e = _electric(u)
        self.p_pad = PadOperator(e, u_tmp, offset, pad_val)
        self.p_crop = CropOperator(u_tmp, (e.shape[-1], e.shape[-2]))
        self._ref = [None]  # shared with the adjoint: the reference's Ref{U}
        self.ispad = True
        self.store_ref = bool(store_ref)

    def adjoint(self) -> "PadCropOperator":
        """``adjoint(p)``: the same operator with pad and crop swapped."""
        return PadCropOperator(None, _parts=(self.p_pad, self.p_crop, self._ref, not self.ispad, self.store_ref))

    def _pad(self, u: ScalarField) -> ScalarField:
        if self.store_ref:
            self._ref[0] = u.electric
        v = self.p_pad(u)
        u.electric = v.electric
        return u

    def _crop(self, u: ScalarField) -> ScalarField:
        v = self.p_crop(u, self._ref[0]) if self.store_ref else self.p_crop(u)
        u.electric = v.electric
        return u

    def propagate_(self, u: ScalarField, direction=Forward) -> ScalarField:
        """``propagate!`` (``pad_crop_operators.jl:318-350``): the field's data is replaced by
        the padded (or cropped) array."""
        if (dir_sign(direction) > 0) == self.ispad:
            return self._pad(u)
        return self._crop(u)

    def propagate(self, u: ScalarField, direction=Forward) -> ScalarField:
        return self.propagate_(u.copy(), direction)

    def backpropagate_(self, du: ScalarField, direction=Forward) -> ScalarField:
        """``backpropagate!`` (``pad_crop_operators.jl:352-355``): the reverse direction (the
        crop is the adjoint of the zero pad)."""
        return self.propagate_(du, -dir_sign(direction))


class FieldProbe(_PipeComponent):
    """``FieldProbe()`` (``field_probe.jl``): in a system, ``system().probes[probe]`` is a copy of
    the field where the probe sits; the field passes unchanged (and so does a cotangent)."""

    trainable = False

    def propagate_(self, u: ScalarField, direction=Forward) -> ScalarField:
        return u

    def propagate(self, u: ScalarField, direction=Forward) -> ScalarField:
        return u.copy()

    def backpropagate_(self, du: ScalarField, direction=Forward) -> ScalarField:
        return du
