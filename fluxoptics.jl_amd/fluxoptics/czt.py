"""Chirp z-transform plans on the device (``src/FFTutils.jl:30-177``).

``plan_czt(x, dims[, a, w]; center_on_grid)`` returns a plan applying
X[m] = sum_k x[k] a^-k w^(m k) along ``dims`` (1 = x, the fastest axis; 2 = y) of a complex64
HIP tensor of shape (ny, nx) or (B, ny, nx).  ``plan * x`` returns a new tensor (in place for
``plan_czt_``); ``plan.adjoint() * y`` applies the adjoint plan.  The transform runs in
libfluxhip as one chirp / zero-padded convolution / chirp plan (``fh_czt_create``).
"""

from __future__ import annotations

import cmath
import ctypes
import math
from numbers import Number

import torch

from . import _lib
from .components import _ptr, check_field, context


def _convert_czt_arg(v):
    """``convert_czt_arg`` (``FFTutils.jl:105-115``): a number, one number per dimension, or one
    (rho, theta) pair per dimension -> [(rho, theta) for x, (rho, theta) for y]."""
    if isinstance(v, Number):
        return [(abs(v), cmath.phase(v))] * 2
    v = list(v)
    if len(v) != 2:
        raise ValueError("need one value per spatial dimension (x, y)")
    out = []
    for e in v:
        if isinstance(e, Number):
            out.append((abs(e), cmath.phase(e)))
        elif len(e) == 2:
            out.append((float(e[0]), float(e[1])))
        else:
            raise ValueError("wrong CZT argument")
    return out


class CZTPlan:
    """``CZTPlan`` (``FFTutils.jl:52-89``)."""

    def __init__(self, x: torch.Tensor, dims, a=None, w=None, inplace: bool = False,
                 center_on_grid: bool = False, _adjoint: bool = False):
        check_field(x, "x")
        dims = tuple(int(d) for d in (dims if not isinstance(dims, Number) else (dims,)))
        if len(set(dims)) != len(dims) or not all(d in (1, 2) for d in dims):
            raise ValueError("dims must be distinct values among 1 (x) and 2 (y)")
        self.ny, self.nx = x.shape[-2], x.shape[-1]
        self.dims, self.inplace, self.center_on_grid = dims, inplace, bool(center_on_grid)
        self.a = _convert_czt_arg(a) if a is not None else [(1.0, 0.0)] * 2
        self.w = (_convert_czt_arg(w) if w is not None
                  else [(1.0, -2 * math.pi / self.nx), (1.0, -2 * math.pi / self.ny)])
        self._plans = {}
        self._adjoint = _adjoint
        self._parent = None

    def _plan(self, x: torch.Tensor) -> ctypes.c_void_p:
        if x.shape[-2:] != (self.ny, self.nx):
            raise ValueError("array shape does not match the plan")
        check_field(x, "x")
        batch = 1 if x.dim() == 2 else x.shape[0]
        key = (x.device.index, batch)
        p = self._plans.get(key)
        ctx = context(x.device)
        if p is None:
            p = ctypes.c_void_p()
            arr = lambda v: (ctypes.c_double * 4)(v[0][0], v[0][1], v[1][0], v[1][1])
            mask = sum(1 << (d - 1) for d in self.dims)
            _lib.call("fh_czt_create", ctx, self.nx, self.ny, batch, mask, arr(self.a), arr(self.w),
                      int(self.center_on_grid), ctypes.byref(p))
            self._plans[key] = p
        return p

    def czt_(self, x: torch.Tensor) -> torch.Tensor:
        """``czt!(p, x)``: in place."""
        _lib.call("fh_propagate", self._plan(x), _ptr(x), -1 if self._adjoint else 1)
        return x

    def __mul__(self, x: torch.Tensor) -> torch.Tensor:
        return self.czt_(x if self.inplace else x.clone())

    def adjoint(self) -> "CZTPlan":
        """``adjoint(p)`` (``FFTutils.jl:91-98``): the same tables, every factor conjugated and
        applied in reverse order."""
        q = CZTPlan.__new__(CZTPlan)
        q.__dict__.update(self.__dict__)
        q._adjoint = not self._adjoint
        q._parent = self  # shares the plans
        return q

    def __del__(self):
        try:
            if self._parent is None:
                for p in self._plans.values():
                    _lib.load().fh_plan_destroy(p)
        except Exception:
            pass


def plan_czt(x: torch.Tensor, dims, a=None, w=None, *, center_on_grid: bool = False, inplace: bool = False):
    """``plan_czt(x, dims[, a, w]; center_on_grid)`` (``FFTutils.jl:117-125``)."""
    return CZTPlan(x, dims, a, w, inplace=inplace, center_on_grid=center_on_grid)


def plan_czt_(x: torch.Tensor, dims, a=None, w=None, *, center_on_grid: bool = False):
    """``plan_czt!`` (``FFTutils.jl:127-129``): a plan whose ``*`` works in place."""
    return CZTPlan(x, dims, a, w, inplace=True, center_on_grid=center_on_grid)
