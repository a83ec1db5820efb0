"""Loss metrics evaluated on the device, with their pullbacks.

``SquaredIntensityDifference`` mirrors ``src/metrics/squared_intensity_difference.jl``
(constructor from ``(field, target_intensity)`` pairs, ``compute_metric`` at :78-83 and
``backpropagate_metric`` at :85-90).  The sums run in libfluxhip (FP64 accumulation in a
fixed order); nothing is computed on the host.
"""

from __future__ import annotations

from typing import Optional, Sequence, Tuple

import torch

from . import _lib
from .components import _ptr, check_field, context
from .fields import ScalarField


def _target(t, field: ScalarField) -> torch.Tensor:
    t = torch.as_tensor(t, dtype=torch.float32, device=field.electric.device).contiguous()
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
