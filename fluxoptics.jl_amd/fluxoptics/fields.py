"""ScalarField and propagation directions (host-side mirror of the reference's data model).

Memory layout: ``electric`` is a torch tensor of shape ``(B, ny, nx)`` (or ``(ny, nx)``)
in C order, byte-identical to the reference's column-major ``electric[nx, ny, B]``
(``src/Fields.jl:118-149``).  Device fields are complex64 on a HIP device.
"""

from __future__ import annotations

import ctypes
import math
from typing import Sequence, Union

import numpy as np
import torch


class Direction:
    """``Forward``/``Backward`` markers with ``sign`` and ``reverse``
    (``src/optical_components/OpticalComponents.jl:34-99``)."""

    sign = 0

    @classmethod
    def reverse(cls):
        return Backward if cls is Forward else Forward


class Forward(Direction):
    sign = 1


class Backward(Direction):
    sign = -1


def dir_sign(direction) -> int:
    if direction in (Forward, Backward):
        return direction.sign
    if direction in (1, -1):
        return int(direction)
    raise ValueError(f"direction must be Forward or Backward, got {direction!r}")


class ScalarField:
    """Scalar optical field: complex data + sampling ``ds`` + wavelength(s).

    ``lambdas`` is a scalar or one value per field (``src/Fields.jl:28-32``).  A
    ComplexF32 field in the reference stores lambda as Float32; pass
    ``float(np.float32(lam))`` to reproduce that exactly.

    The number of spatial dimensions is ``len(ds)`` (``Fields.jl:135-136``: Nd in (1, 2)):
    2-D fields are ``(ny, nx)`` or ``(B, ny, nx)`` with ``ds = (dx, dy)`` and
    ``tilts = (theta_x, theta_y)``; 1-D fields are ``(nx,)`` or ``(B, nx)`` (the reference's
    ``electric[nx, B]``) with ``ds = (dx,)`` and ``tilts = (theta_x,)``, and reach the library as
    fields of one row (``ny = 1``).
    """

    def __init__(self, electric, ds: Sequence[float], lambdas: Union[float, Sequence[float]],
                 tilts=None):
        if not isinstance(electric, torch.Tensor):
            electric = torch.as_tensor(np.asarray(electric))
        self.nd = len(ds)
        if self.nd not in (1, 2):
            raise ValueError("ds must hold one (1-D field) or two (2-D field) samplings")
        if electric.dim() not in ((2, 3) if self.nd == 2 else (1, 2)):
            raise ValueError("electric must be (ny, nx) or (B, ny, nx)" if self.nd == 2 else
                             "a 1-D field's electric must be (nx,) or (B, nx)")
        self.electric = electric
        self.ds = tuple(float(d) for d in ds)
        lam = np.atleast_1d(np.asarray(lambdas, dtype=np.float64))
        if lam.size not in (1, self.batch):
            raise ValueError("lambdas must be a scalar or one value per field")
        if not np.all(lam > 0):
            raise ValueError("wavelengths must be positive")
        self.lambdas = float(lam[0]) if lam.size == 1 and np.ndim(lambdas) == 0 else [float(v) for v in lam]
        if tilts is None:
            tilts = (0.0,) * self.nd
        if len(tilts) != self.nd:
            raise ValueError("tilts must be (theta_x, theta_y)" if self.nd == 2 else "a 1-D field's tilts are (theta_x,)")
        tl = []
        for t in tilts:
            a = np.atleast_1d(np.asarray(t, dtype=np.float64))
            if a.size not in (1, self.batch):
                raise ValueError("each tilt must be a scalar or one value per field")
            tl.append(float(a[0]) if np.ndim(t) == 0 else [float(v) for v in a])
        self.tilts = tuple(tl)

    # -- shape helpers
    @property
    def nx(self) -> int:
        return self.electric.shape[-1]

    @property
    def ny(self) -> int:
        return 1 if self.nd == 1 else self.electric.shape[-2]

    @property
    def batch(self) -> int:
        return 1 if self.electric.dim() == self.nd else self.electric.shape[0]

    @property
    def ds2(self):
        """(dx, dy) as the library takes them (dy = 1 for a 1-D field: its fy is 0)."""
        return self.ds if self.nd == 2 else (self.ds[0], 1.0)

    def cell(self) -> float:
        """The sampling cell dx dy (dx for a 1-D field)."""
        return float(np.prod(self.ds))

    def lambda_list(self):
        lam = self.lambdas if isinstance(self.lambdas, list) else [self.lambdas]
        return lam if len(lam) == self.batch else lam * self.batch

    def lambda_min(self) -> float:
        return min(self.lambda_list())

    def tilt_lists(self):
        """(theta_x per field, theta_y per field) in radians (``Fields.jl:168-180``); theta_y is
        0 for a 1-D field."""
        out = []
        for t in self.tilts:
            lst = t if isinstance(t, list) else [t]
            out.append(lst if len(lst) == self.batch else lst * self.batch)
        if self.nd == 1:
            out.append([0.0] * self.batch)
        return tuple(out)

    def is_on_axis(self) -> bool:
        return all(v == 0.0 for lst in self.tilt_lists() for v in lst)

    def copy(self) -> "ScalarField":
        return ScalarField(self.electric.clone(), self.ds, self.lambdas, self.tilts)

    def set_field_data(self, data) -> "ScalarField":
        """New field over ``data`` with the same grid, wavelengths and tilts
        (``Fields.jl:213-215``)."""
        return ScalarField(data, self.ds, self.lambdas, self.tilts)

    def set_field_tilts(self, tilts) -> "ScalarField":
        """New field over the same data with other tilts (``Fields.jl:220-230``)."""
        return ScalarField(self.electric, self.ds, self.lambdas, tilts)

    def offset_tilts_(self, tilts) -> "ScalarField":
        """``offset_tilts!(u, tilts)`` (``Fields.jl:280-293``): multiplies each field by
        cis(2 pi / lambda ((sin(theta_x) - sin(theta_x')) x + (sin(theta_y) - sin(theta_y')) y))
        on the device and takes the new tilts; nothing happens when they equal the current ones."""
        from . import _lib
        from .components import check_field, context
        head = self.electric[..., :1] if self.nd == 1 else self.electric[..., :1, :1]
        new = ScalarField(head, self.ds, self.lambdas, tilts)  # parses the tilts
        if tuple(new.tilts) == tuple(self.tilts):
            return self
        check_field(self.electric)
        arr = lambda v: (ctypes.c_double * len(v))(*v)
        (fx, fy), (tx, ty) = self.tilt_lists(), new.tilt_lists()
        lams = self.lambda_list()
        _lib.call("fh_offset_tilts", context(self.electric.device), ctypes.c_void_p(self.electric.data_ptr()),
                  self.nx, self.ny, self.batch, self.ds2[0], self.ds2[1], arr(lams), len(lams), arr(fx), arr(fy),
                  arr(tx), arr(ty), self.batch)
        self.tilts = new.tilts
        return self

    def power(self) -> torch.Tensor:
        """Per-field power sum|u|^2 dx dy, shape (B,) (``Fields.jl:843-846``)."""
        e = self.electric.reshape(self.batch, self.ny, self.nx)
        return (e.abs() ** 2).sum(dim=(-2, -1)) * self.cell()

    def normalize_power_(self, target: float = 1.0) -> "ScalarField":
        """``normalize_power!`` (``Fields.jl:889-892``)."""
        scale = torch.sqrt(target / self.power()).to(self.electric.dtype)
        per_field = self.electric.dim() > self.nd
        self.electric *= scale.reshape((-1,) + (1,) * self.nd if per_field else (1,) * self.nd)
        return self

    # -- arithmetic (the reference's test/fields_test.jl:74-93): elementwise on the data, the
    # result keeps this field's grid, wavelengths and tilts
    def _same_grid(self, other: "ScalarField") -> None:
        if tuple(other.electric.shape) != tuple(self.electric.shape) or other.ds != self.ds:
            raise ValueError("fields must share shape and sampling")
        if other.lambda_list() != self.lambda_list():
            raise ValueError("fields must share wavelengths")

    def __add__(self, other):
        if isinstance(other, ScalarField):
            self._same_grid(other)
            return self.set_field_data(self.electric + other.electric)
        return NotImplemented

    def __sub__(self, other):
        if isinstance(other, ScalarField):
            self._same_grid(other)
            return self.set_field_data(self.electric - other.electric)
        return NotImplemented

    def __mul__(self, a):
        if isinstance(a, (int, float, complex)):
            return self.set_field_data(self.electric * a)
        return NotImplemented

    __rmul__ = __mul__

    def __neg__(self):
        return self.set_field_data(-self.electric)

    def conj(self) -> "ScalarField":
        """``conj(u)`` (``Fields.jl:899``): the complex conjugate field."""
        return self.set_field_data(self.electric.conj().resolve_conj())

    def __repr__(self):
        return (f"ScalarField(shape={tuple(self.electric.shape)}, dtype={self.electric.dtype}, "
                f"device={self.electric.device}, ds={self.ds}, lambdas={self.lambdas}, "
                f"tilts={self.tilts})")


def intensity(u) -> torch.Tensor:
    """``intensity(u)`` (``Fields.jl:608-616``): sum over the fields of |u|^2, shape (ny, nx)
    ((nx,) for a 1-D field)."""
    e = u.electric if isinstance(u, ScalarField) else u
    nd = u.nd if isinstance(u, ScalarField) else 2
    i = e.abs() ** 2
    return i if i.dim() == nd else i.sum(dim=0)


def phase(u) -> torch.Tensor:
    """``phase(u)`` (``Fields.jl:667-673``): arg(u) in [-pi, pi], the data's shape."""
    e = u.electric if isinstance(u, ScalarField) else u
    return torch.angle(e)


def power(u: ScalarField) -> torch.Tensor:
    """``power(u)`` (``Fields.jl:843-846``)."""
    return u.power()


def normalize_power_(u: ScalarField, target: float = 1.0) -> ScalarField:
    """``normalize_power!(u, P)`` (``Fields.jl:889-892``)."""
    return u.normalize_power_(target)


def coupling_efficiency(u: ScalarField, v: ScalarField) -> torch.Tensor:
    """|<u,v>|^2 / (<u,u><v,v>) per field, shape (B,) (``Fields.jl:748-756``)."""
    a = u.electric.reshape(u.batch, u.ny, u.nx)
    b = v.electric.reshape(v.batch, v.ny, v.nx)
    num = (a.conj() * b).sum(dim=(-2, -1)).abs() ** 2
    den = (a.abs() ** 2).sum(dim=(-2, -1)) * (b.abs() ** 2).sum(dim=(-2, -1))
    return num / den


def rs_valid_distance(nx: int, ny: int, dx: float, dy: float, lam: float) -> float:
    """Critical RS distance z_c (``rayleigh_sommerfeld.jl:29-36``)."""
    if dx < lam / 2 and dy < lam / 2:
        return 0.0
    zx = 0.0 if dx < lam / 2 else (nx * dx / 2) * math.sqrt(4 * dx * dx / lam ** 2 - 1)
    zy = 0.0 if dy < lam / 2 else (ny * dy / 2) * math.sqrt(4 * dy * dy / lam ** 2 - 1)
    return max(zx, zy)
