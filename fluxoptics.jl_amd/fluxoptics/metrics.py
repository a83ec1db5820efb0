"""Loss metrics evaluated on the device, with their pullbacks.

``SquaredIntensityDifference`` mirrors ``src/metrics/squared_intensity_difference.jl``
(constructor from ``(field, target_intensity)`` pairs, ``compute_metric`` at :78-83 and
``backpropagate_metric`` at :85-90); ``SquaredFieldDifference`` mirrors
``src/metrics/squared_field_difference.jl``; ``DotProduct`` and ``PowerCoupling`` mirror
``src/metrics/dot_product.jl`` and ``power_coupling.jl`` (mode-selective).  The sums run in
libfluxhip (FP64 accumulation in a fixed order); nothing is computed on the host.
"""

from __future__ import annotations

from typing import Optional, Sequence, Tuple

import torch

from . import _lib
from .components import _ptr, check_field, context
from .fields import ScalarField


def _target(t, field: ScalarField) -> torch.Tensor:
    t = torch.as_tensor(t, dtype=torch.float32, device=field.electric.device).contiguous()
    if field.nd == 1:  # a 1-D field's targets (nx,) or (B, nx) as rows of the library's layout
        if t.dim() == 1 and t.shape[0] == field.nx:
            t = t.reshape(1, field.nx)
        elif t.dim() == 2 and tuple(t.shape) == (field.batch, field.nx) and field.batch > 1:
            t = t.reshape(field.batch, 1, field.nx)
        else:
            raise ValueError(f"target intensity shape {tuple(t.shape)} matches neither the 1-D field's "
                             f"spatial dims ({field.nx},) nor its full dims {(field.batch, field.nx)}")
    spatial = (field.ny, field.nx)
    if tuple(t.shape) == spatial:
        return t
    if t.dim() == 3 and tuple(t.shape) == (field.batch, field.ny, field.nx):
        return t
    if t.dim() == 2 and field.batch == 1 and tuple(t.shape) == spatial:
        return t
    raise ValueError(f"target intensity shape {tuple(t.shape)} matches neither the field's spatial "
                     f"dims {spatial} nor its full dims {(field.batch,) + spatial}")


class SquaredIntensityDifference:
    """``SquaredIntensityDifference(field_intensity_pairs...)``.

    Per pair: with a target of the field's full dims, one value per field,
    ``sum (|u_k|^2 - I_k)^2``; with a target of the spatial dims only, one value,
    ``sum (sum_k |u_k|^2 - I)^2``.  ``compute_metric`` returns one float64 device tensor
    per pair; ``backpropagate_metric`` returns the cotangent fields ``4 dc (..) u``.
    """

    def __init__(self, *pairs: Tuple[ScalarField, object]):
        if not pairs:
            raise ValueError("SquaredIntensityDifference needs at least one (field, target) pair")
        self.targets = []
        for field, t in pairs:
            check_field(field.electric)
            self.targets.append(_target(t, field))

    def _run(self, k: int, u: ScalarField, dc: Optional[torch.Tensor], du: Optional[torch.Tensor]):
        check_field(u.electric)
        t = self.targets[k]
        tb = 1 if t.dim() == 2 else t.shape[0]
        if t.shape[-2:] != (u.ny, u.nx) or tb not in (1, u.batch):
            raise ValueError("field does not match the metric's target")
        nout = u.batch if tb == u.batch else 1
        loss = torch.empty(nout, dtype=torch.float64, device=u.electric.device)
        if dc is not None:
            dc = torch.as_tensor(dc, dtype=torch.float64, device=u.electric.device).reshape(-1).contiguous()
            if dc.numel() == 1 and nout > 1:
                dc = dc.expand(nout).contiguous()
            if dc.numel() != nout:
                raise ValueError(f"dc must hold {nout} values")
        _lib.call("fh_sq_intensity_diff", context(u.electric.device), _ptr(u.electric), _ptr(t),
                  u.nx, u.ny, u.batch, tb, _ptr(dc), _ptr(loss), _ptr(du))
        return loss

    def compute_metric(self, *fields: ScalarField):
        if len(fields) != len(self.targets):
            raise ValueError(f"expected {len(self.targets)} fields")
        return tuple(self._run(k, u, None, None) for k, u in enumerate(fields))

    __call__ = compute_metric

    def backpropagate_metric(self, fields: Sequence[ScalarField], dcs=None):
        """Cotangent fields 4 dc (|u|^2 - I) u, one per pair (new buffers)."""
        out = []
        for k, u in enumerate(fields):
            du = torch.empty_like(u.electric)
            self._run(k, u, None if dcs is None else dcs[k], du)
            out.append(u.set_field_data(du))
        return tuple(out)

    def value_and_pullback(self, fields: Sequence[ScalarField], dcs=None, inplace: bool = False):
        """Loss values and cotangents in one pass over each field (``inplace`` overwrites the
        fields with their cotangents)."""
        losses, cots = [], []
        for k, u in enumerate(fields):
            du = u.electric if inplace else torch.empty_like(u.electric)
            losses.append(self._run(k, u, None if dcs is None else dcs[k], du))
            cots.append(u if inplace else u.set_field_data(du))
        return tuple(losses), tuple(cots)


class SquaredFieldDifference:
    """``SquaredFieldDifference(fields...)`` (``metrics/squared_field_difference.jl``): per target
    field v, one value per field of the batch, ``sum |u_b - v_b|^2`` (no area factor, as the
    reference's ``sum!(abs2, c, u - v)``), and the pullback ``2 dc (u - v)``."""

    def __init__(self, *fields: ScalarField):
        if not fields:
            raise ValueError("SquaredFieldDifference needs at least one target field")
        for v in fields:
            check_field(v.electric)
        self.targets = [v.electric for v in fields]

    def _run(self, k: int, u: ScalarField, dc, du: Optional[torch.Tensor]):
        check_field(u.electric)
        v = self.targets[k]
        if tuple(v.shape) != tuple(u.electric.shape):
            raise ValueError("field does not match the metric's target field")
        loss = torch.empty(u.batch, dtype=torch.float64, device=u.electric.device)
        if dc is not None:
            dc = torch.as_tensor(dc, dtype=torch.float64, device=u.electric.device).reshape(-1).contiguous()
            if dc.numel() == 1 and u.batch > 1:
                dc = dc.expand(u.batch).contiguous()
            if dc.numel() != u.batch:
                raise ValueError(f"dc must hold {u.batch} values")
        _lib.call("fh_sq_field_diff", context(u.electric.device), _ptr(u.electric), _ptr(v), u.nx, u.ny, u.batch,
                  _ptr(dc), _ptr(loss), _ptr(du))
        return loss

    def compute_metric(self, *fields: ScalarField):
        if len(fields) != len(self.targets):
            raise ValueError(f"expected {len(self.targets)} fields")
        return tuple(self._run(k, u, None, None) for k, u in enumerate(fields))

    __call__ = compute_metric

    def backpropagate_metric(self, fields: Sequence[ScalarField], dcs=None):
        """Cotangent fields 2 dc (u - v), one per target (new buffers)."""
        out = []
        for k, u in enumerate(fields):
            du = torch.empty_like(u.electric)
            self._run(k, u, None if dcs is None else dcs[k], du)
            out.append(u.set_field_data(du))
        return tuple(out)

    def value_and_pullback(self, fields: Sequence[ScalarField], dcs=None, inplace: bool = False):
        losses, cots = [], []
        for k, u in enumerate(fields):
            du = u.electric if inplace else torch.empty_like(u.electric)
            losses.append(self._run(k, u, None if dcs is None else dcs[k], du))
            cots.append(u if inplace else u.set_field_data(du))
        return tuple(losses), tuple(cots)


def _matrix(c: torch.Tensor, n: int) -> torch.Tensor:
    """The library's column-major n x n buffer as the matrix M[i, j] (a view)."""
    return c.view(n, n).t()


def _colmajor(m, n: int, dtype, device) -> torch.Tensor:
    """An (n, n) matrix M[i, j] as the library's column-major buffer (contiguous copy)."""
    m = torch.as_tensor(m, dtype=dtype, device=device).reshape(n, n)
    return m.t().contiguous().reshape(-1)


class DotProduct:
    """``DotProduct(fields...; mode_selective=true)`` (``metrics/dot_product.jl``): one complex
    value per mode, c_b = ds * sum u_b conj(v_b), with the pullback du_b = v_b dc_b; with
    ``mode_selective=False`` the overlap matrix c[i, j] = ds * sum conj(v_i) u_j of the n <= 32
    modes (an (n, n) complex128 device tensor) with the pullback du_j = sum_i v_i dc[i, j]."""

    def __init__(self, *fields: ScalarField, mode_selective: bool = True):
        if not fields:
            raise ValueError("DotProduct needs at least one reference field")
        for v in fields:
            check_field(v.electric)
        self.v = [v.copy() for v in fields]
        self.mode_selective = bool(mode_selective)

    def _dot(self, k, u, p=None):
        v = self.v[k]
        check_field(u.electric)
        if u.electric.shape != v.electric.shape:
            raise ValueError("field does not match the metric's reference field")
        dev = u.electric.device
        if self.mode_selective:
            c = torch.empty(u.batch, dtype=torch.complex128, device=dev)
            _lib.call("fh_dot_product", context(dev), _ptr(u.electric), _ptr(v.electric), u.nx, u.ny, u.batch,
                      u.cell(), _ptr(c), _ptr(p))
            return c
        c = torch.empty(u.batch * u.batch, dtype=torch.complex128, device=dev)
        _lib.call("fh_dot_matrix", context(dev), _ptr(u.electric), _ptr(v.electric), u.nx, u.ny, u.batch, u.cell(),
                  _ptr(c), _ptr(p))
        return c

    def compute_metric(self, *fields: ScalarField):
        out = tuple(self._dot(k, u) for k, u in enumerate(fields))
        return out if self.mode_selective else tuple(_matrix(c, u.batch) for c, u in zip(out, fields))

    __call__ = compute_metric

    def backpropagate_metric(self, fields, dcs):
        out = []
        for k, (u, dc) in enumerate(zip(fields, dcs)):
            dev = u.electric.device
            du = torch.empty_like(u.electric)
            v = self.v[k]
            if self.mode_selective:
                dc = torch.as_tensor(dc, dtype=torch.complex128, device=dev).reshape(-1)
                dc = (dc.expand(u.batch) if dc.numel() == 1 else dc).contiguous()
                _lib.call("fh_dot_product_pullback", context(dev), _ptr(v.electric), _ptr(dc), None, None,
                          u.nx, u.ny, u.batch, _ptr(du))
            else:
                dcm = _colmajor(dc, u.batch, torch.complex128, dev)
                _lib.call("fh_dot_matrix_pullback", context(dev), _ptr(v.electric), _ptr(dcm), None, None,
                          u.nx, u.ny, u.batch, _ptr(du))
            out.append(u.set_field_data(du))
        return tuple(out)


class PowerCoupling:
    """``PowerCoupling(fields...; mode_selective=true)`` (``metrics/power_coupling.jl``): the
    targets are copied and normalised to unit power; one value per mode, p_b = |<u_b, v_b>|^2,
    with the pullback 2 c_b dp_b v_b; with ``mode_selective=False`` the (n, n) matrix
    p[i, j] = |c[i, j]|^2 of the overlap matrix, with the pullback du_j = sum_i v_i 2 c[i, j] dp[i, j]."""

    def __init__(self, *fields: ScalarField, mode_selective: bool = True):
        if not fields:
            raise ValueError("PowerCoupling needs at least one target field")
        self.v = []
        for v in fields:
            check_field(v.electric)
            self.v.append(v.copy().normalize_power_())
        self.mode_selective = bool(mode_selective)
        self._c = [None] * len(self.v)

    def _run(self, k, u):
        v = self.v[k]
        check_field(u.electric)
        if u.electric.shape != v.electric.shape:
            raise ValueError("field does not match the metric's target field")
        dev = u.electric.device
        n = u.batch if self.mode_selective else u.batch * u.batch
        c = torch.empty(n, dtype=torch.complex128, device=dev)
        p = torch.empty(n, dtype=torch.float64, device=dev)
        _lib.call("fh_dot_product" if self.mode_selective else "fh_dot_matrix", context(dev), _ptr(u.electric),
                  _ptr(v.electric), u.nx, u.ny, u.batch, u.cell(), _ptr(c), _ptr(p))
        self._c[k] = c
        return p if self.mode_selective else _matrix(p, u.batch)

    def compute_metric(self, *fields: ScalarField):
        return tuple(self._run(k, u) for k, u in enumerate(fields))

    __call__ = compute_metric

    def backpropagate_metric(self, fields, dps=None, inplace: bool = False):
        """Cotangent fields 2 c dp v (needs a preceding compute_metric on the same fields)."""
        out = []
        for k, u in enumerate(fields):
            c = self._c[k]
            if c is None:
                raise RuntimeError("backpropagate_metric needs compute_metric first")
            dev = u.electric.device
            dp = None
            if self.mode_selective:
                if dps is not None:
                    dp = torch.as_tensor(dps[k], dtype=torch.float64, device=dev).reshape(-1)
                    dp = (dp.expand(u.batch) if dp.numel() == 1 else dp).contiguous()
                du = u.electric if inplace else torch.empty_like(u.electric)
                _lib.call("fh_dot_product_pullback", context(dev), _ptr(self.v[k].electric), None, _ptr(c),
                          _ptr(dp), u.nx, u.ny, u.batch, _ptr(du))
            else:
                if dps is not None:
                    dpk = torch.as_tensor(dps[k], dtype=torch.float64, device=dev)
                    dp = (dpk.expand(u.batch, u.batch) if dpk.numel() == 1 else dpk)
                    dp = _colmajor(dp, u.batch, torch.float64, dev)
                du = torch.empty_like(u.electric)  # the matrix pullback reads every target per pixel
                _lib.call("fh_dot_matrix_pullback", context(dev), _ptr(self.v[k].electric), None, _ptr(c), _ptr(dp),
                          u.nx, u.ny, u.batch, _ptr(du))
                if inplace:
                    u.electric.copy_(du)
                    du = u.electric
            out.append(u if inplace else u.set_field_data(du))
        return tuple(out)

    def value_and_pullback(self, fields, dcs=None, inplace: bool = False):
        vals = self.compute_metric(*fields)
        return vals, self.backpropagate_metric(fields, dcs, inplace)
