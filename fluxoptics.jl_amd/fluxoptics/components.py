"""Propagators and modulators on the accelerated path: ASProp, RSProp, Phase.

Each class mirrors the reference component's constructor keywords and its
``propagate!`` / ``backpropagate!`` / ``propagate_and_save!`` /
``backpropagate_with_gradient!`` contract
(``src/optical_components/OpticalComponents.jl:726-816``), with Python names
``propagate_`` (in place), ``propagate`` (copy), ``backpropagate_``,
``propagate_and_save_`` and ``backpropagate_with_gradient_``.  All arithmetic runs in
libfluxhip on the field's HIP device.
"""

from __future__ import annotations

import ctypes
import math
import warnings
from typing import Callable, Optional, Union

import numpy as np
import torch

from . import _lib
from .fields import Backward, Forward, ScalarField, dir_sign, rs_valid_distance

# --------------------------------------------------------------------------- contexts

_contexts = {}


def context(device: torch.device) -> ctypes.c_void_p:
    """The library context of a HIP device, bound to torch's current stream."""
    idx = device.index if device.index is not None else torch.cuda.current_device()
    ctx = _contexts.get(idx)
    if ctx is None:
        ctx = ctypes.c_void_p()
        _lib.call("fh_ctx_create", idx, None, ctypes.byref(ctx))
        _contexts[idx] = ctx
    stream = torch.cuda.current_stream(idx).cuda_stream
    _lib.call("fh_ctx_set_stream", ctx, ctypes.c_void_p(stream))
    return ctx


def _ptr(t: Optional[torch.Tensor]):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def check_field(e: torch.Tensor, what: str = "field") -> None:
    if not isinstance(e, torch.Tensor) or e.device.type != "cuda":
        raise TypeError(f"{what} must be a torch tensor on a HIP device (libfluxhip has no CPU path)")
    if e.dtype != torch.complex64:
        raise TypeError(f"{what} must be complex64, got {e.dtype}")
    if not e.is_contiguous():
        raise ValueError(f"{what} must be contiguous")


def _check_mask(phi: torch.Tensor, ny: int, nx: int, device, nd: int = 2) -> None:
    if phi.dtype != torch.float32 or phi.device != device or not phi.is_contiguous():
        raise TypeError("mask must be a contiguous float32 tensor on the field's device")
    want = (ny, nx) if nd == 2 else (nx,)
    if tuple(phi.shape) != want:
        raise ValueError(f"mask shape {tuple(phi.shape)} does not match the field {want}")


# --------------------------------------------------------------------------- pipe

class _PipeComponent:
    """``c1 | c2`` builds an ``OpticalSystem`` (the reference's ``c1 |> c2``,
    ``optical_system.jl:150-180``); ``c | {"merge_components": True}`` passes keywords."""

    trainable = False

    def __or__(self, other):
        from .system import OpticalSystem, _components_of
        if isinstance(other, dict):
            return OpticalSystem(None, self, **other)
        return OpticalSystem(None, self, *_components_of(other))


# --------------------------------------------------------------------------- propagators

class _KernelPropagator(_PipeComponent):
    """Shared plan cache of ASProp / RSProp (one plan per device, batch and wavelengths)."""

    def __init__(self, u: ScalarField, ds, z):
        self.ds = tuple(float(d) for d in (ds if ds is not None else u.ds))
        if len(self.ds) != getattr(u, "nd", 2):
            raise ValueError("ds must have one sampling per spatial dimension of the field")
        # one distance, or one per field of the batch (an extension for uneven batches)
        self.z = float(z) if np.ndim(z) == 0 else [float(v) for v in np.asarray(z).ravel()]
        self.nx, self.ny = u.nx, u.ny
        self._plans = {}

    track_tilts = False
    _create_ex = ""

    def _desc(self, batch, lams, tilts, flags) -> "_lib.PropDesc":
        arr = lambda v: (ctypes.c_double * len(v))(*v)
        d = _lib.PropDesc()
        d.nx, d.ny, d.batch = self.nx, self.ny, batch
        d.dx, d.dy = self.ds if len(self.ds) == 2 else (self.ds[0], 1.0)  # a 1-D field's fy is 0
        zs = self.z if isinstance(self.z, list) else [self.z]
        if len(zs) not in (1, batch):
            raise ValueError(f"{len(zs)} distances for a batch of {batch} fields")
        d._keep = (arr(lams), arr(zs), arr(tilts[0]), arr(tilts[1]))
        d.lambdas = ctypes.cast(d._keep[0], _lib.c_double_p)
        d.nlambdas = len(lams)
        d.z = ctypes.cast(d._keep[1], _lib.c_double_p)
        d.nz = len(zs)
        d.tilts_x = ctypes.cast(d._keep[2], _lib.c_double_p)
        d.tilts_y = ctypes.cast(d._keep[3], _lib.c_double_p)
        d.ntilts = batch
        d.n0 = getattr(self, "n0", 1.0)
        d.flags = flags
        return d

    def _flags(self) -> int:
        return _lib.FH_PROP_TRACK_TILTS if self.track_tilts else 0

    def plan(self, u: ScalarField) -> ctypes.c_void_p:
        if (u.nx, u.ny) != (self.nx, self.ny):
            raise ValueError(f"field grid {(u.nx, u.ny)} does not match the propagator {(self.nx, self.ny)}")
        check_field(u.electric)
        lams = tuple(u.lambda_list())
        tilts = tuple(tuple(t) for t in u.tilt_lists())
        key = (u.electric.device.index, u.batch, lams, tilts)
        p = self._plans.get(key)
        ctx = context(u.electric.device)
        if p is None:
            p = ctypes.c_void_p()
            desc = self._desc(u.batch, lams, tilts, self._flags())
            self._create(ctx, desc, p)
            self._plans[key] = p
        return p

    def _create(self, ctx, desc, p) -> None:
        _lib.call(self._create_ex, ctx, ctypes.byref(desc), ctypes.byref(p))

    def __del__(self):
        try:
            for p in self._plans.values():
                _lib.load().fh_plan_destroy(p)
        except Exception:
            pass

    # -- the component contract
    def propagate_(self, u: ScalarField, direction=Forward) -> ScalarField:
        _lib.call("fh_propagate", self.plan(u), _ptr(u.electric), dir_sign(direction))
        return u

    def propagate(self, u: ScalarField, direction=Forward) -> ScalarField:
        return self.propagate_(u.copy(), direction)

    def backpropagate_(self, dv: ScalarField, direction=Forward) -> ScalarField:
        """Adjoint = propagation in the reverse direction (``freespace.jl:61-63``)."""
        return self.propagate_(dv, Direction_reverse(direction))

    def propagate_plain_fp64_(self, u_c128: torch.Tensor, lambdas, direction=Forward) -> torch.Tensor:
        """The plain comparator: complex128 per-field, unfused propagation."""
        if u_c128.dtype != torch.complex128 or u_c128.device.type != "cuda" or not u_c128.is_contiguous():
            raise TypeError("the plain FP64 path takes a contiguous complex128 HIP tensor")
        batch = 1 if u_c128.dim() == 2 else u_c128.shape[0]
        tmp = ScalarField(torch.empty((batch, self.ny, self.nx), dtype=torch.complex64,
                                      device=u_c128.device), self.ds, lambdas)
        _lib.call("fh_propagate_plain_fp64", self.plan(tmp), _ptr(u_c128), dir_sign(direction))
        return u_c128


def Direction_reverse(direction):
    return Backward if dir_sign(direction) > 0 else Forward


class ASProp(_KernelPropagator):
    """Angular-spectrum propagation in a medium of index ``n0``.

    ``ASProp(u, z; use_cache=true, n0=1)`` (``angular_spectrum.jl:193-229``): the
    transfer function is computed once in FP64 and cached as complex FP32, as the
    reference does with ``use_cache=true`` (``double_precision_kernel=use_cache``).
    The kernel follows the reference's choice (``angular_spectrum.jl:126-139``): paraxial
    when ``paraxial=true``, else on-axis for untilted fields, else tilted (with the field's
    tilts, ``track_tilts`` adding the walk-off phase).  ``filter``, a function of the
    frequencies (fx, fy) returning a (complex) transmission, multiplies the kernel as in the
    reference (``angular_spectrum.jl:5,34,57``): it is sampled on the host in FP64 on the
    fftfreq grid of ``ds`` -- at (fx, fy) for on-axis and paraxial kernels, at
    (fx + f0x, fy + f0y) per field for the tilted one -- and folded into the FP64 transfer
    function before rounding (``fh_asprop_create_filtered``).  It is called once on whole
    frequency arrays; a filter written for scalars (``if`` on its arguments) is evaluated
    element by element instead.
    """

    _create_ex = "fh_asprop_create_ex"

    def __init__(self, u: ScalarField, z: float, ds=None, *, n0: float = 1.0,
                 use_cache: bool = True, paraxial: bool = False, filter: Optional[Callable] = None,
                 track_tilts: bool = False, double_precision_kernel: Optional[bool] = None):
        if double_precision_kernel is False:
            raise NotImplementedError("transfer functions are always computed in FP64")
        if filter is not None and not callable(filter):
            raise TypeError("filter must be a function of (fx, fy)")
        super().__init__(u, ds, z)
        self.n0 = float(n0)
        self.paraxial = bool(paraxial)
        self.track_tilts = bool(track_tilts)
        self.filter = filter
        if not self.n0 > 0:
            raise ValueError("n0 must be positive")

    def _flags(self) -> int:
        return super()._flags() | (_lib.FH_PROP_PARAXIAL if self.paraxial else 0)

    def _create(self, ctx, desc, p) -> None:
        if self.filter is None:
            return super()._create(ctx, desc, p)
        grids = self.filter_grids(desc)
        buf = np.ascontiguousarray(grids, dtype=np.complex128)
        _lib.call("fh_asprop_create_filtered", ctx, ctypes.byref(desc), buf.ctypes.data_as(ctypes.c_void_p),
                  len(grids), ctypes.byref(p))

    def filter_grids(self, desc) -> np.ndarray:
        """The filter's samples for a plan: (1 or batch, ny, nx) complex128 on the fftfreq grid
        (FFT order), at the arguments the reference's kernel passes it."""
        fx = np.fft.fftfreq(self.nx, self.ds[0])[None, :]
        fy = np.fft.fftfreq(self.ny, self.ds[1])[:, None] if len(self.ds) == 2 else np.zeros((1, 1))
        filt = self.filter if len(self.ds) == 2 else (lambda a, b: self.filter(a))  # 1-D: filter(fx)
        tx = [desc.tilts_x[b] for b in range(desc.batch)]
        ty = [desc.tilts_y[b] for b in range(desc.batch)]
        if self.paraxial or all(a == 0.0 for a in tx + ty):
            return _sample_filter(filt, fx, fy)[None]
        lams = [desc.lambdas[0 if desc.nlambdas == 1 else b] for b in range(desc.batch)]
        out = []
        for b in range(desc.batch):
            le = lams[b] / self.n0
            out.append(_sample_filter(filt, fx + math.sin(tx[b]) / le, fy + math.sin(ty[b]) / le))
        return np.stack(out)

    @classmethod
    def _from_params(cls, nx: int, ny: int, ds, z: float, n0: float) -> "ASProp":
        """An ASProp for an (ny, nx) grid without a template field (used by merging)."""
        self = cls.__new__(cls)
        _KernelPropagator.__init__(self, _GridOnly(nx, ny, ds), ds, z)
        self.n0 = float(n0)
        self.paraxial = False
        self.track_tilts = False
        self.filter = None
        return self


def _sample_filter(filt, fx, fy) -> np.ndarray:
    """filt over the broadcast frequency grids as complex128: one call on whole arrays when the
    filter accepts them, else one call per sample (as the reference broadcasts a scalar
    function)."""
    fx, fy = np.broadcast_arrays(np.asarray(fx, np.float64), np.asarray(fy, np.float64))
    try:
        with np.errstate(all="ignore"):
            v = np.asarray(filt(fx, fy), dtype=np.complex128)
        if v.shape == fx.shape:
            return v
        if v.ndim == 0:  # a constant filter
            return np.full(fx.shape, complex(v))
    except (ValueError, TypeError):
        pass
    return np.vectorize(lambda a, b: complex(filt(float(a), float(b))), otypes=[np.complex128])(fx, fy)


class ASPropZ(_PipeComponent):
    """``ASPropZ(u, z; n0, trainable)`` (``angular_spectrum.jl:236-300``): an on-axis ASProp whose
    distance z is a trainable parameter (a one-element FP32 device tensor, ``self.z``).  The
    propagation is ASProp's (its plan is rebuilt when z changes, as the reference recomputes its
    kernel each call); inside ``withgradient`` the distance gradient comes from the output and its
    cotangent (``fh_as_z_gradient``).  Paraxial kernels, frequency filters and tilted fields are not
    provided here; the gradient is taken in the Forward direction."""

    def __init__(self, u: ScalarField, z: float, *, n0: float = 1.0, trainable: bool = False,
                 buffered: bool = False, paraxial: bool = False, track_tilts: bool = False,
                 filter: Optional[Callable] = None, double_precision_kernel: bool = False):
        check_field(u.electric)
        if paraxial or filter is not None:
            raise NotImplementedError("ASPropZ: paraxial kernels and frequency filters are not provided")
        if not u.is_on_axis():
            raise NotImplementedError("ASPropZ: tilted fields are not provided")
        self.n0 = float(n0)
        self.z = torch.tensor([float(z)], dtype=torch.float32, device=u.electric.device)
        self.trainable = bool(trainable)
        self.buffered = self.trainable and buffered
        self.grad = torch.zeros_like(self.z) if self.buffered else None
        self._prop, self._zp = None, None

    def _plan(self, u: ScalarField) -> ASProp:
        zv = float(self.z.item())
        if self._prop is None or zv != self._zp:
            self._prop, self._zp = ASProp(u, zv, n0=self.n0), zv
        return self._prop

    def propagate_(self, u: ScalarField, direction=Forward) -> ScalarField:
        return self._plan(u).propagate_(u, direction)

    def propagate(self, u: ScalarField, direction=Forward) -> ScalarField:
        return self.propagate_(u.copy(), direction)

    def backpropagate_(self, du: ScalarField, direction=Forward) -> ScalarField:
        return self._plan(du).backpropagate_(du, direction)

    def z_gradient(self, v: ScalarField, dv: ScalarField) -> torch.Tensor:
        """dL/dz (summed over the fields, shape (1,), FP32) from the Forward output v and its
        cotangent dv."""
        V = v.electric.clone()
        DV = dv.electric.clone()
        ctx = context(v.electric.device)
        for a in (V, DV):
            _lib.call("fh_fft2", ctx, _ptr(a), v.nx, v.ny, v.batch, -1, 1.0)
        lams = v.lambda_list()
        arr = (ctypes.c_double * len(lams))(*lams)
        out = torch.empty(v.batch, dtype=torch.float64, device=v.electric.device)
        _lib.call("fh_as_z_gradient", ctx, _ptr(V), _ptr(DV), v.nx, v.ny, v.batch, v.ds[0], v.ds[1], arr, len(lams),
                  self.n0, float(self._zp if self._zp is not None else self.z.item()), _ptr(out))
        return out.sum().reshape(1).to(torch.float32)


class ShiftProp(_KernelPropagator):
    """``ShiftProp(u, z)`` (``freespace_propagators/shift_kernel.jl:47-117``): the geometric
    shift of each field by z tan(theta) according to its tilts, no diffraction, applied as the
    Fourier-domain phase cis(-2 pi z (tan(theta_x) fx + tan(theta_y) fy))."""

    _create_ex = "fh_asprop_create_ex"

    def __init__(self, u: ScalarField, z, ds=None, *, use_cache: bool = True,
                 double_precision_kernel: Optional[bool] = None):
        if double_precision_kernel is False:
            raise NotImplementedError("transfer functions are always computed in FP64")
        super().__init__(u, ds, z)
        self.n0 = 1.0

    def _flags(self) -> int:
        return _lib.FH_PROP_SHIFT


class RSProp(_KernelPropagator):
    """Rayleigh-Sommerfeld propagation (``rayleigh_sommerfeld.jl:128-164``).

    Warns like the reference when |z| is below the critical distance.  Pads to 2n
    (the reference pads to 2n-1; the cropped result is the same linear convolution).
    Tilted fields use the tilted kernel (``rayleigh_sommerfeld.jl:14-27,90-96``).
    """

    _create_ex = "fh_rsprop_create_ex"

    def __init__(self, u: ScalarField, z: float, ds=None, *, use_cache: bool = True,
                 track_tilts: bool = False, double_precision_kernel: Optional[bool] = None):
        if getattr(u, "nd", 2) == 1:
            raise NotImplementedError("the reference's RSProp has no 1-D kernel (rs_kernel takes x and y)")
        super().__init__(u, ds, z)
        self.track_tilts = bool(track_tilts)
        zc = rs_valid_distance(self.nx, self.ny, self.ds[0], self.ds[1], u.lambda_min())
        zmin = min(abs(v) for v in (self.z if isinstance(self.z, list) else [self.z]))
        if zmin < zc:
            warnings.warn(f"RSProp: propagation distance z={zmin} is below critical distance "
                          f"zc={zc}. Numerical artifacts expected. Consider using ASProp or finer "
                          "sampling (dx < lambda/2).")


class CollinsProp(_KernelPropagator):
    """``CollinsProp(u, ds', (A, B, D))`` or ``CollinsProp(u, ds, ds', (A, B, D))``
    (``collins_integral.jl:200-301``): the Collins integral of the ABCD system with
    AD - BC = 1, resampled from ``ds`` (default ``u.ds``) to ``ds'``.  One fused plan: the A
    chirp rides on the first x FFT, the zero-padded convolution on the y pass, the D chirp
    on the last x IFFT.  Forward relabels the field's sampling to ``ds'``, Backward to
    ``ds``.  Backward (and so the pullback, ``freespace.jl:61-63``) conjugates every factor
    and normalises by prod(ds' / sqrt(-i B)) as the reference does, which is alpha_x alpha_y
    times the plain adjoint when ds' differs from ds."""

    _create_ex = "fh_collins_create"

    def __init__(self, u: ScalarField, *args, use_cache: bool = True, track_tilts: bool = False,
                 double_precision_kernel: Optional[bool] = None):
        if len(args) == 2:
            ds, ds_out, abd = None, args[0], args[1]
        elif len(args) == 3:
            ds, ds_out, abd = args
        else:
            raise TypeError("CollinsProp(u, ds_out, (A, B, D)) or CollinsProp(u, ds, ds_out, (A, B, D))")
        if double_precision_kernel is False:
            raise NotImplementedError("chirps and kernels are always computed in FP64")
        super().__init__(u, ds, 0.0)
        self.ds_out = (float(ds_out[0]), float(ds_out[1]))
        a, b, d = (float(v) for v in abd)
        if b == 0.0:
            raise ValueError("CollinsProp cannot be used with B = 0")
        self.abd = (a, b, d)
        self.track_tilts = bool(track_tilts)

    def _desc(self, batch, lams, tilts, flags) -> "_lib.PropDesc":
        d = super()._desc(batch, lams, tilts, flags)
        d.z, d.nz = None, 0
        return d

    def _create(self, ctx, desc, p) -> None:
        _lib.call("fh_collins_create", ctx, ctypes.byref(desc), self.ds_out[0], self.ds_out[1], *self.abd,
                  ctypes.byref(p))

    def propagate_(self, u: ScalarField, direction=Forward) -> ScalarField:
        super().propagate_(u, direction)
        u.ds = self.ds_out if dir_sign(direction) > 0 else self.ds
        return u


def FourierLens(u: ScalarField, *args, use_cache: bool = True,
                double_precision_kernel: Optional[bool] = None) -> CollinsProp:
    """``FourierLens(u, ds', fl)`` or ``FourierLens(u, ds, ds', fl)`` (``fourier_lens.jl:38-64``):
    the ideal Fourier lens, ``CollinsProp`` with ABCD = (0, fl, 0)."""
    if len(args) == 2:
        return CollinsProp(u, args[0], (0.0, args[1], 0.0), use_cache=use_cache,
                           double_precision_kernel=double_precision_kernel)
    if len(args) == 3:
        return CollinsProp(u, args[0], args[1], (0.0, args[2], 0.0), use_cache=use_cache,
                           double_precision_kernel=double_precision_kernel)
    raise TypeError("FourierLens(u, ds_out, fl) or FourierLens(u, ds, ds_out, fl)")


def ParaxialProp(u: ScalarField, *args, use_cache: bool = True, filter: Optional[Callable] = None,
                 double_precision_kernel: Optional[bool] = None):
    """``ParaxialProp(u, z)``, ``ParaxialProp(u, ds', z)`` or ``ParaxialProp(u, ds, ds', z)``
    (``paraxial_propagation.jl:33-82``): the paraxial ASProp without resampling, else
    ``CollinsProp`` with ABCD = (1, z, 1)."""
    if len(args) == 1:
        return ASProp(u, args[0], paraxial=True, filter=filter, double_precision_kernel=double_precision_kernel)
    if len(args) == 2:
        ds, ds_out, z = tuple(u.ds), args[0], args[1]
    elif len(args) == 3:
        ds, ds_out, z = args
    else:
        raise TypeError("ParaxialProp(u, z), ParaxialProp(u, ds_out, z) or ParaxialProp(u, ds, ds_out, z)")
    ds, ds_out = tuple(float(v) for v in ds), tuple(float(v) for v in ds_out)
    if ds == ds_out:
        return ASProp(u, z, ds, paraxial=True, filter=filter, double_precision_kernel=double_precision_kernel)
    return CollinsProp(u, ds, ds_out, (1.0, z, 1.0), use_cache=use_cache,
                       double_precision_kernel=double_precision_kernel)


# --------------------------------------------------------------------------- Fourier operators

class FourierOperator(_PipeComponent):
    """``FourierOperator(u, direct)`` (``fourier_operator.jl:37-79``): FFT2 of the field when
    ``direct`` (IFFT2 otherwise); the Backward direction swaps them.  Unnormalised forward,
    1/(nx*ny) inverse (``FFTutils.jl:13-28``)."""

    def __init__(self, u: ScalarField, direct: bool):
        self.nx, self.ny = u.nx, u.ny
        self.direct = bool(direct)

    def _transform(self, direction):
        """(sign, scale) of the transform applied in ``direction``."""
        fft = self.direct if dir_sign(direction) > 0 else not self.direct
        return (-1, 1.0) if fft else (1, 1.0 / (self.nx * self.ny))

    def _run(self, u: ScalarField, sign: int, scale: float) -> ScalarField:
        check_field(u.electric)
        if (u.nx, u.ny) != (self.nx, self.ny):
            raise ValueError("field grid does not match the Fourier operator")
        _lib.call("fh_fft2", context(u.electric.device), _ptr(u.electric), u.nx, u.ny, u.batch, sign,
                  float(scale))
        return u

    def propagate_(self, u: ScalarField, direction=Forward) -> ScalarField:
        return self._run(u, *self._transform(direction))

    def propagate(self, u: ScalarField, direction=Forward) -> ScalarField:
        return self.propagate_(u.copy(), direction)

    def backpropagate_(self, dv: ScalarField, direction=Forward) -> ScalarField:
        """The exact adjoint, as Zygote's pullback through the FFT plan gives it
        (``OpticalComponents.jl:690-696``): the FFT's adjoint is the unnormalised inverse and
        the normalised IFFT's is FFT/(nx*ny)."""
        sign, scale = self._transform(direction)
        return self._run(dv, -sign, scale)


def FourierWrapper(u: ScalarField, *components):
    """``FourierWrapper(u, components...)`` (``fourier_wrapper.jl:1-5,34-39``): the sequence
    FFT -> components -> IFFT."""
    from .system import OpticalSequence
    return OpticalSequence(FourierOperator(u, True), *components, FourierOperator(u, False))


# --------------------------------------------------------------------------- Phase

class _GridOnly:
    """Grid description standing in for a template field."""

    def __init__(self, nx, ny, ds):
        self.nx, self.ny, self.ds = int(nx), int(ny), tuple(float(d) for d in ds)
        self.nd = len(self.ds)


def _spatial(n: int, d: float) -> np.ndarray:
    return (np.arange(n, dtype=np.float64) - (n - 1) / 2.0) * d


class Phase(_PipeComponent):
    """Phase mask u -> u * cis(s*phi) (``phasemask.jl:38-150``).

    ``phi`` is a function ``(x, y) -> value`` evaluated on the centred grid
    (``function_to_array``) or a (ny, nx) array.  ``trainable=True`` enables the
    gradient; ``buffered=True`` pre-allocates the saved input and the gradient.
    """

    def __init__(self, u: ScalarField, phi: Union[Callable, np.ndarray, torch.Tensor, None] = None,
                 ds=None, *, trainable: bool = False, buffered: bool = False):
        ds = tuple(ds if ds is not None else u.ds)
        device = u.electric.device
        nd = getattr(u, "nd", 2)
        shape = (u.ny, u.nx) if nd == 2 else (u.nx,)
        if phi is None:
            arr = np.zeros(shape)
        elif callable(phi):
            x = _spatial(u.nx, ds[0])[None, :]
            if nd == 2:
                y = _spatial(u.ny, ds[1])[:, None]
                arr = np.broadcast_to(np.vectorize(phi)(x, y), shape)
            else:
                arr = np.broadcast_to(np.vectorize(phi)(x[0]), shape)
        else:
            arr = phi
        self.phi = torch.as_tensor(np.asarray(arr) if not isinstance(arr, torch.Tensor) else arr,
                                   dtype=torch.float32, device=device).contiguous()
        _check_mask(self.phi, u.ny, u.nx, device, nd)
        self.trainable = trainable
        self.buffered = trainable and buffered
        self.grad = torch.zeros_like(self.phi) if self.buffered else None
        self.saved = torch.empty_like(u.electric) if self.buffered else None

    @classmethod
    def _from_tensor(cls, phi: torch.Tensor, trainable: bool = False) -> "Phase":
        """A static Phase over an existing (ny, nx) float32 device mask (used by merging)."""
        self = cls.__new__(cls)
        self.phi = phi.to(torch.float32).contiguous()
        self.trainable = trainable
        self.buffered = False
        self.grad = None
        self.saved = None
        return self

    def propagate_(self, u: ScalarField, direction=Forward) -> ScalarField:
        check_field(u.electric)
        _check_mask(self.phi, u.ny, u.nx, u.electric.device, u.nd)
        _lib.call("fh_phase_apply", context(u.electric.device), _ptr(u.electric), _ptr(self.phi),
                  u.nx, u.ny, u.batch, dir_sign(direction), None)
        return u

    def propagate(self, u: ScalarField, direction=Forward) -> ScalarField:
        return self.propagate_(u.copy(), direction)

    def backpropagate_(self, dv: ScalarField, direction=Forward) -> ScalarField:
        return self.propagate_(dv, Direction_reverse(direction))

    def alloc_saved_buffer(self, u: ScalarField) -> torch.Tensor:
        return torch.empty_like(u.electric)

    def propagate_and_save_(self, u: ScalarField, saved: Optional[torch.Tensor] = None,
                            direction=Forward) -> ScalarField:
        saved = self.saved if saved is None else saved
        check_field(u.electric)
        check_field(saved, "saved buffer")
        _lib.call("fh_phase_apply", context(u.electric.device), _ptr(u.electric), _ptr(self.phi),
                  u.nx, u.ny, u.batch, dir_sign(direction), _ptr(saved))
        return u

    def backpropagate_with_gradient_(self, dv: ScalarField, saved: torch.Tensor,
                                     dphi: Optional[torch.Tensor] = None, direction=Forward,
                                     accumulate: bool = False):
        """Returns (du, dphi) with du written over dv (``phasemask.jl:142-150``);
        ``accumulate`` adds to dphi instead of overwriting it."""
        dphi = self.grad if dphi is None else dphi
        if dphi is None:
            dphi = torch.empty_like(self.phi)
        check_field(dv.electric)
        check_field(saved, "saved buffer")
        _lib.call("fh_phase_backward", context(dv.electric.device), _ptr(dv.electric), _ptr(self.phi),
                  _ptr(saved), _ptr(dphi), dv.nx, dv.ny, dv.batch, dir_sign(direction), int(accumulate))
        return dv, dphi


# --------------------------------------------------------------------------- Mask, TeaDOE

def _grid_array(u: ScalarField, ds, f, dtype):
    """A function (x, y) -> value on the centred grid (``function_to_array``,
    ``OpticalComponents.jl:927-935``) or an array, as a (ny, nx) tensor on the field's device."""
    ds = tuple(ds if ds is not None else u.ds)
    if callable(f):
        x = _spatial(u.nx, ds[0])[None, :]
        y = _spatial(u.ny, ds[1])[:, None]
        arr = np.broadcast_to(np.vectorize(f, otypes=[complex])(x, y), (u.ny, u.nx))
        if dtype == torch.float32:
            arr = arr.real
    else:
        arr = f
    t = torch.as_tensor(np.asarray(arr) if not isinstance(arr, torch.Tensor) else arr, device=u.electric.device)
    t = t.to(dtype).contiguous()
    if tuple(t.shape) != (u.ny, u.nx):
        raise ValueError(f"array shape {tuple(t.shape)} does not match the field ({u.ny}, {u.nx})")
    return t


class _Modulator(_PipeComponent):
    """Trainable elementwise modulators: the component contract of
    ``OpticalComponents.jl:726-816`` (``propagate_and_save!``,
    ``backpropagate_with_gradient!`` with a pre-allocated gradient when buffered)."""

    def _init_training(self, u: ScalarField, trainable: bool, buffered: bool):
        self.trainable = trainable
        self.buffered = trainable and buffered
        self.grad = torch.zeros_like(self._param) if self.buffered else None
        self.saved = torch.empty_like(u.electric) if self.buffered else None

    def propagate(self, u: ScalarField, direction=Forward) -> ScalarField:
        return self.propagate_(u.copy(), direction)

    def backpropagate_(self, dv: ScalarField, direction=Forward) -> ScalarField:
        return self._backward(dv, None, None, direction, False)

    def alloc_saved_buffer(self, u: ScalarField) -> torch.Tensor:
        return torch.empty_like(u.electric)

    def propagate_and_save_(self, u: ScalarField, saved: Optional[torch.Tensor] = None,
                            direction=Forward) -> ScalarField:
        saved = self.saved if saved is None else saved
        check_field(saved, "saved buffer")
        return self._apply(u, saved, direction)

    def backpropagate_with_gradient_(self, dv: ScalarField, saved: torch.Tensor, grad=None,
                                     direction=Forward, accumulate: bool = False):
        grad = self.grad if grad is None else grad
        if grad is None:
            grad = torch.empty_like(self._param)
        check_field(saved, "saved buffer")
        self._backward(dv, saved, grad, direction, accumulate)
        return dv, grad

    def propagate_(self, u: ScalarField, direction=Forward) -> ScalarField:
        return self._apply(u, None, direction)


class Mask(_Modulator):
    """``Mask(u, f; trainable, buffered)`` (``modulators/mask.jl``): u -> u * m(x, y), complex
    transmission (conjugated in the Backward direction); the gradient is complex."""

    def __init__(self, u: ScalarField, f=None, ds=None, *, trainable: bool = False, buffered: bool = False):
        self.m = _grid_array(u, ds, (lambda x, y: 1.0) if f is None else f, torch.complex64)
        self._init_training(u, trainable, buffered)

    @property
    def _param(self):
        return self.m

    def _apply(self, u, saved, direction):
        check_field(u.electric)
        _lib.call("fh_mask_apply", context(u.electric.device), _ptr(u.electric), _ptr(self.m), u.nx, u.ny, u.batch,
                  dir_sign(direction), _ptr(saved))
        return u

    def _backward(self, dv, saved, grad, direction, accumulate):
        check_field(dv.electric)
        _lib.call("fh_mask_backward", context(dv.electric.device), _ptr(dv.electric), _ptr(self.m), _ptr(saved),
                  _ptr(grad), dv.nx, dv.ny, dv.batch, dir_sign(direction), int(accumulate))
        return dv


class TeaDOE(_Modulator):
    """``TeaDOE(u, dn, f; r=1, trainable, buffered)`` (``modulators/tea_doe.jl``): a height map
    h(x, y) with transmission r(lambda) * cis(2 pi dn(lambda) h / lambda) per field
    wavelength; ``dn`` and ``r`` are numbers or functions of lambda (evaluated on the host
    once per wavelength set)."""

    def __init__(self, u: ScalarField, dn, f=None, ds=None, *, r=1.0, trainable: bool = False,
                 buffered: bool = False):
        self.h = _grid_array(u, ds, (lambda x, y: 0.0) if f is None else f, torch.float32)
        self.dn = dn
        self.r = r
        self._coef = {}
        self._init_training(u, trainable, buffered)

    @property
    def _param(self):
        return self.h

    def _coefficients(self, u: ScalarField):
        lams = tuple(u.lambda_list())
        key = (u.electric.device.index, lams)
        c = self._coef.get(key)
        if c is None:
            dn = [float(self.dn(l)) if callable(self.dn) else float(self.dn) for l in lams]
            kl = torch.tensor([2.0 * np.pi * d / l for d, l in zip(dn, lams)], dtype=torch.float64,
                              device=u.electric.device)
            rs = [complex(self.r(l)) if callable(self.r) else complex(self.r) for l in lams]
            r = None
            if any(x != 1 for x in rs):
                r = torch.tensor([[x.real, x.imag] for x in rs], dtype=torch.float64, device=u.electric.device)
            c = (kl, r)
            self._coef[key] = c
        return c

    def _apply(self, u, saved, direction):
        check_field(u.electric)
        kl, r = self._coefficients(u)
        _lib.call("fh_tea_apply", context(u.electric.device), _ptr(u.electric), _ptr(self.h), u.nx, u.ny, u.batch,
                  _ptr(kl), _ptr(r), dir_sign(direction), _ptr(saved))
        return u

    def _backward(self, dv, saved, grad, direction, accumulate):
        check_field(dv.electric)
        kl, r = self._coefficients(dv)
        _lib.call("fh_tea_backward", context(dv.electric.device), _ptr(dv.electric), _ptr(self.h), _ptr(saved),
                  _ptr(grad), dv.nx, dv.ny, dv.batch, _ptr(kl), _ptr(r), dir_sign(direction), int(accumulate))
        return dv


def TeaReflector(u: ScalarField, f=None, ds=None, *, r=1.0, trainable: bool = False, buffered: bool = False):
    """``TeaReflector(u, f; r)`` (``tea_doe.jl:110-131``): a TeaDOE with dn = 2 (phase 4 pi h / lambda)."""
    return TeaDOE(u, 2.0, f, ds, r=r, trainable=trainable, buffered=buffered)


def _fourier_array(u: ScalarField, ds, f, dtype):
    """f(fx, fy) on the FFT-order frequency grid (``function_to_array(..., true)``)."""
    ds = tuple(ds if ds is not None else u.ds)
    if not callable(f):
        return f
    fx = np.fft.fftfreq(u.nx, ds[0])[None, :]
    fy = np.fft.fftfreq(u.ny, ds[1])[:, None]
    return np.broadcast_to(np.vectorize(f, otypes=[complex if dtype == torch.complex64 else float])(fx, fy),
                           (u.ny, u.nx))


def FourierPhase(u: ScalarField, f=None, ds=None, *, trainable: bool = False, buffered: bool = False):
    """``FourierPhase(u, f)`` (``fourier/fourier_phase.jl``): a Phase on the spectrum,
    FFT -> Phase(f(fx, fy)) -> IFFT."""
    arr = _fourier_array(u, ds, (lambda a, b: 0.0) if f is None else f, torch.float32)
    return FourierWrapper(u, Phase(u, np.asarray(arr, dtype=np.float64), ds, trainable=trainable,
                                   buffered=buffered))


def FourierMask(u: ScalarField, f=None, ds=None, *, trainable: bool = False, buffered: bool = False):
    """``FourierMask(u, f)`` (``fourier/fourier_mask.jl``): a Mask on the spectrum."""
    arr = _fourier_array(u, ds, (lambda a, b: 1.0) if f is None else f, torch.complex64)
    return FourierWrapper(u, Mask(u, np.asarray(arr), ds, trainable=trainable, buffered=buffered))


# --------------------------------------------------------------------------- fused pairs

def phase_propagate_(u: ScalarField, phase: Phase, prop: _KernelPropagator,
                     saved: Optional[torch.Tensor] = None, direction=Forward) -> ScalarField:
    """Phase then propagator in one pass over the field (mask fused into the first FFT pass)."""
    _lib.call("fh_phase_propagate", prop.plan(u), _ptr(u.electric), _ptr(phase.phi), _ptr(saved),
              dir_sign(direction))
    return u


def propagate_phase_backward_(dv: ScalarField, phase: Phase, prop: _KernelPropagator,
                              saved: Optional[torch.Tensor], dphi: Optional[torch.Tensor],
                              direction=Forward, accumulate: bool = False) -> ScalarField:
    """Adjoint of ``phase_propagate_``: propagator adjoint, mask adjoint and mask gradient
    (summed over the batch in field order) in one pass."""
    _lib.call("fh_propagate_phase_backward", prop.plan(dv), _ptr(dv.electric), _ptr(phase.phi),
              _ptr(saved), _ptr(dphi), dir_sign(direction), int(accumulate))
    return dv


# --------------------------------------------------------------------------- AS_BPM

class TiltAnchor(_PipeComponent):
    """``TiltAnchor(u; trainable=false, buffered=false)`` (``tilt_anchor.jl:28-77``): resets
    the tilts of the fields passing through to the template's (``offset_tilts!``), in either
    direction.  It has no parameters.  Inside a differentiated system the reverse pass
    restores the tilts the field carried on entry (``backpropagate_with_gradient!`` with the
    saved tilts), which is the exact adjoint of the forward carrier multiplication."""

    def __init__(self, u: ScalarField, trainable: bool = False, buffered: bool = False):
        self.tilts = tuple(list(t) if isinstance(t, list) else t for t in u.tilts)
        self.trainable = False

    def propagate_(self, u: ScalarField, direction=Forward) -> ScalarField:
        return u.offset_tilts_(self.tilts)

    def propagate(self, u: ScalarField, direction=Forward) -> ScalarField:
        return self.propagate_(u.copy(), direction)

    def backpropagate_(self, du: ScalarField, direction=Forward) -> ScalarField:
        """``backpropagate!`` (``tilt_anchor.jl:70-72``): offsets to the anchor's tilts."""
        return du.offset_tilts_(self.tilts)


class BPM(_PipeComponent):
    """Split-step beam propagation through a graded-index volume (``BPM``, ``bpm.jl:8-74``);
    built as ``AS_BPM`` or ``Shift_BPM``.

    ``dn`` has shape (nslices, ny, nx) (the reference's ``dn[nx, ny, nslices]``);
    ``thickness`` is the total length.  kdz = 2 pi dz / (cos(theta_x) cos(theta_y)) is taken
    from the template field's tilts, as the reference does at construction (``bpm.jl:70``), so
    a propagated field must carry the template's tilts.  For the gradient the library keeps
    one checkpoint every ``checkpoint_every`` slices instead of every slice input and
    recomputes one segment at a time in the adjoint.  ``aperture`` is accepted and, as in the
    reference (whose ``propagate!`` never applies ``aperture_mask``), has no effect.
    """

    _flags = 0

    def __init__(self, u: ScalarField, thickness: float, n0: float, dn, *, trainable: bool = False,
                 buffered: bool = False, checkpoint_every: int = 20, use_cache: bool = True,
                 aperture: Optional[Callable] = None, double_precision_kernel: Optional[bool] = None):
        if double_precision_kernel is False:
            raise NotImplementedError("slice transfer functions are always FP64")
        check_field(u.electric)
        device = u.electric.device
        self.dn = torch.as_tensor(np.asarray(dn) if not isinstance(dn, torch.Tensor) else dn,
                                  dtype=torch.float32, device=device).contiguous()
        if u.nd == 2 and (self.dn.dim() != 3 or tuple(self.dn.shape[1:]) != (u.ny, u.nx)):
            raise ValueError("dn must have shape (nslices, ny, nx)")
        if u.nd == 1 and (self.dn.dim() != 2 or self.dn.shape[1] != u.nx):
            raise ValueError("a 1-D field's dn must have shape (nslices, nx)")
        if self.dn.shape[0] < 2:
            raise ValueError("BPM needs at least 2 slices (bpm.jl:40)")
        self.nslices = int(self.dn.shape[0])
        self.thickness, self.n0 = float(thickness), float(n0)
        self.nx, self.ny, self.ds = u.nx, u.ny, u.ds2
        self.tilts = tuple(tuple(t) for t in u.tilt_lists())
        self.checkpoint_every = int(checkpoint_every)
        self.trainable = trainable
        self.grad = torch.zeros_like(self.dn) if (trainable and buffered) else None
        self._objs = {}

    def _tilts_for(self, u: ScalarField):
        tilts = tuple(tuple(t) for t in u.tilt_lists())
        same = tilts == self.tilts
        if not same and len(set(self.tilts[0])) == 1 and len(set(self.tilts[1])) == 1:
            same = all(v == self.tilts[0][0] for v in tilts[0]) and all(v == self.tilts[1][0] for v in tilts[1])
        if not same:
            raise ValueError("the field's tilts differ from the template's, for which kdz was computed")
        return tilts

    def _obj(self, u: ScalarField) -> ctypes.c_void_p:
        if (u.nx, u.ny) != (self.nx, self.ny):
            raise ValueError("field grid does not match the BPM volume")
        check_field(u.electric)
        lams = tuple(u.lambda_list())
        tilts = self._tilts_for(u)
        key = (u.electric.device.index, u.batch, lams, tilts)
        ctx = context(u.electric.device)
        m = self._objs.get(key)
        if m is None:
            m = ctypes.c_void_p()
            arr = lambda v: (ctypes.c_double * len(v))(*v)
            d = _lib.PropDesc()
            d.nx, d.ny, d.batch = self.nx, self.ny, u.batch
            d.dx, d.dy = self.ds
            keep = (arr(lams), arr(tilts[0]), arr(tilts[1]))
            d.lambdas = ctypes.cast(keep[0], _lib.c_double_p)
            d.nlambdas = len(lams)
            d.tilts_x = ctypes.cast(keep[1], _lib.c_double_p)
            d.tilts_y = ctypes.cast(keep[2], _lib.c_double_p)
            d.ntilts = u.batch
            d.n0 = self.n0
            d.flags = self._flags
            _lib.call("fh_bpm_create_ex", ctx, ctypes.byref(d), self.thickness, self.nslices,
                      self.checkpoint_every, ctypes.byref(m))
            self._objs[key] = m
        return m

    def __del__(self):
        try:
            for m in self._objs.values():
                _lib.load().fh_bpm_destroy(m)
        except Exception:
            pass

    def workspace_bytes(self, u: ScalarField) -> int:
        b = ctypes.c_size_t()
        _lib.call("fh_bpm_workspace_bytes", self._obj(u), ctypes.byref(b))
        return int(b.value)

    def propagate_(self, u: ScalarField, direction=Forward) -> ScalarField:
        _lib.call("fh_bpm_forward", self._obj(u), _ptr(u.electric), _ptr(self.dn), dir_sign(direction), 0)
        return u

    def propagate(self, u: ScalarField, direction=Forward) -> ScalarField:
        return self.propagate_(u.copy(), direction)

    def propagate_and_save_(self, u: ScalarField, direction=Forward) -> ScalarField:
        _lib.call("fh_bpm_forward", self._obj(u), _ptr(u.electric), _ptr(self.dn), dir_sign(direction), 1)
        return u

    def backpropagate_(self, du: ScalarField, direction=Forward) -> ScalarField:
        _lib.call("fh_bpm_backward", self._obj(du), _ptr(du.electric), _ptr(self.dn), None, dir_sign(direction))
        return du

    def backpropagate_with_gradient_(self, du: ScalarField, ddn: Optional[torch.Tensor] = None,
                                     direction=Forward):
        """Returns (du, ddn); needs a preceding ``propagate_and_save_`` in the same direction."""
        ddn = self.grad if ddn is None else ddn
        if ddn is None:
            ddn = torch.empty_like(self.dn)
        _lib.call("fh_bpm_backward", self._obj(du), _ptr(du.electric), _ptr(self.dn), _ptr(ddn),
                  dir_sign(direction))
        return du, ddn


class AS_BPM(BPM):
    """``AS_BPM(u, thickness, n0, dn; paraxial=false, trainable, buffered, aperture)``
    (``bpm.jl:130-142``): the slice propagator is ASProp(dz; n0) with the reference's kernel
    choice (paraxial, on-axis or tilted by the field's tilts; tilts untracked)."""

    def __init__(self, u: ScalarField, thickness: float, n0: float, dn, *, paraxial: bool = False, **kw):
        super().__init__(u, thickness, n0, dn, **kw)
        self.paraxial = bool(paraxial)
        self._flags = _lib.FH_PROP_PARAXIAL if paraxial else 0


class Shift_BPM(BPM):
    """``Shift_BPM(u, thickness, dn; trainable, buffered, aperture)`` (``bpm.jl:207-217``): the
    slice propagator is ShiftProp(dz), a geometric shift by dz tan(theta) per field."""

    _flags = _lib.FH_PROP_SHIFT

    def __init__(self, u: ScalarField, thickness: float, dn, **kw):
        super().__init__(u, thickness, 1.0, dn, **kw)
