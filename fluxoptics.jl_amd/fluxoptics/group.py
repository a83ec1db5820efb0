"""Uneven batches: many (propagator, field) pairs run as one hipGraph.

The reference propagates each field with its own launches; ``PropagationGroup`` is this
library's batching of fields of different sizes, wavelengths and distances (C ABI
``fh_group_*``): the plans' launch chains are captured once on parallel streams and
replayed as one graph launch.
"""

from __future__ import annotations

import ctypes
from typing import Optional, Sequence, Tuple

from . import _lib
from .components import _KernelPropagator, check_field, context
from .fields import Forward, ScalarField, dir_sign


class PropagationGroup:
    """``PropagationGroup([(prop, field), ...], autotune_reps=0)``.

    ``propagate_(direction)`` propagates every registered field in place with its
    propagator; ``propagate_(direction, fields)`` does so for other fields of the same
    shapes.  ``autotune_reps > 0`` tunes each plan's y pass on a scratch field first.
    """

    def __init__(self, pairs: Sequence[Tuple[_KernelPropagator, ScalarField]], autotune_reps: int = 0):
        self.pairs = list(pairs)
        if not self.pairs:
            raise ValueError("a group needs at least one (propagator, field) pair")
        dev = self.pairs[0][1].electric.device
        for _, u in self.pairs:
            check_field(u.electric)
            if u.electric.device != dev:
                raise ValueError("every field of a group must be on one device")
        self.device = dev
        self._plans = [p.plan(u) for p, u in self.pairs]
        n = len(self._plans)
        arr = (ctypes.c_void_p * n)(*[p.value for p in self._plans])
        self._g = ctypes.c_void_p()
        _lib.call("fh_group_create", context(dev), n, arr, int(autotune_reps), ctypes.byref(self._g))

    def __len__(self):
        return len(self.pairs)

    def propagate_(self, direction=Forward, fields: Optional[Sequence[ScalarField]] = None):
        fields = [u for _, u in self.pairs] if fields is None else list(fields)
        if len(fields) != len(self.pairs):
            raise ValueError(f"expected {len(self.pairs)} fields")
        for (p, u0), u in zip(self.pairs, fields):
            check_field(u.electric)
            if u.electric.shape != u0.electric.shape or u.electric.device != self.device:
                raise ValueError("fields must match the shapes and device the group was built for")
        ptrs = (ctypes.c_void_p * len(fields))(*[u.electric.data_ptr() for u in fields])
        context(self.device)  # bind the library to torch's current stream
        _lib.call("fh_group_propagate", self._g, ptrs, dir_sign(direction))
        return fields

    def graph_count(self) -> int:
        c = ctypes.c_int()
        _lib.call("fh_group_graph_count", self._g, ctypes.byref(c))
        return int(c.value)

    def __del__(self):
        try:
            if self._g:
                _lib.load().fh_group_destroy(self._g)
        except Exception:
            pass
