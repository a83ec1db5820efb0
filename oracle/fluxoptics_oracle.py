"""FP64 CPU oracle for the FluxOptics.jl propagation path (ASProp, RSProp, Phase, AS_BPM).

TEST INFRASTRUCTURE, NOT PRODUCT CODE.  Only the tests under ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may import
this module, and only as the checker (or, for ``cpu_baseline``, as the timed CPU
restatement).  The product path (``fluxoptics.jl_amd/``) never imports it and fails
loudly when its HIP library is missing.

This is a plain numpy restatement of the reference algorithm in complex FP64.  The
reference itself (Julia + FFTW + Zygote) cannot run in this container (no Julia, no
network), so this restatement is the oracle; every function cites the reference lines
it follows and is pinned by the checks in ``tests/test_oracle_*.py`` (closed forms,
brute-force DFTs, mpmath, invariants, finite differences).

Array convention.  A field batch is a numpy array of shape ``(B, ny, nx)`` in C order.
That is byte-for-byte the reference's column-major ``electric[nx, ny, B]`` (x fastest,
``src/Fields.jl:118-123``), so element ``(i, j, b)`` sits at offset ``i + nx*(j + ny*b)``.
A 2-D array ``(ny, nx)`` is a batch of one.

Direction convention.  ``direction=+1`` is ``Forward`` and ``-1`` is ``Backward``
(``src/optical_components/OpticalComponents.jl:34-99``).
"""

from __future__ import annotations

import math
from typing import Sequence

import numpy as np

TWO_PI = 2.0 * math.pi


# ----------------------------------------------------------------------------- grids

def spatial_vectors(ns: Sequence[int], ds: Sequence[float], offset=None):
    """Centred grid vectors x_i = (i - (n-1)/2)*dx - xc.

    Follows ``src/GridUtils.jl:56-60``.  ``ns``/``ds`` are ``(nx, ny)``/``(dx, dy)``.
    """
    if offset is None:
        offset = (0.0,) * len(ns)
    return tuple((np.arange(n, dtype=np.float64) - (n - 1) / 2.0) * d - c
                 for n, d, c in zip(ns, ds, offset))


def fftfreq(n: int, dx: float) -> np.ndarray:
    """FFT-order frequencies k/(n*dx), k in [0, n/2) then negative.

    AbstractFFTs ``fftfreq(n, 1/dx)`` as called at
    ``src/optical_components/freespace_propagators/fourier_kernel.jl:17``.
    """
    k = np.arange(n, dtype=np.float64)
    k[k >= (n + 1) // 2] -= n
    return k / (n * dx)


# ----------------------------------------------------------------------------- FFT

def ft2(u: np.ndarray) -> np.ndarray:
    """Unnormalised forward 2-D DFT over the two transverse axes (e^{-2*pi*i*jk/n}).

    ``compute_ft!`` with an in-place ``plan_fft!`` over dims (1, 2): ``src/FFTutils.jl:13-23``.
    """
    return np.fft.fft2(u, axes=(-2, -1))


def ift2(u: np.ndarray) -> np.ndarray:
    """Inverse 2-D DFT normalised by 1/(nx*ny): ``compute_ift!``, ``src/FFTutils.jl:25-28``."""
    return np.fft.ifft2(u, axes=(-2, -1))


# ----------------------------------------------------------------------------- helpers

def _as_batch(u: np.ndarray) -> np.ndarray:
    u = np.asarray(u)
    return u[None] if u.ndim == 2 else u


def _lambdas(lams, batch: int) -> np.ndarray:
    """Per-slice wavelengths.  A scalar applies to every slice (``Fields.jl:28-32``)."""
    lams = np.atleast_1d(np.asarray(lams, dtype=np.float64))
    if lams.size == 1:
        lams = np.repeat(lams, batch)
    if lams.size != batch:
        raise ValueError("one wavelength per batch slice expected")
    return lams


def conj_direction(h: np.ndarray, direction: int) -> np.ndarray:
    """``OpticalComponents.jl:919-925``: the kernel for Forward, its conjugate for Backward."""
    return h if direction > 0 else np.conj(h)


# ----------------------------------------------------------------------------- ASProp

def as_kernel(fx, fy, lam: float, n0: float, z: float):
    """Angular-spectrum transfer function H(fx, fy), complex FP64.

    ``src/optical_components/freespace_propagators/angular_spectrum.jl:1-12``:
    with lam' = lam/n0, H = cis(2*pi*z*sqrt(1/lam'^2 - fx^2 - fy^2)) using the principal
    complex square root, so evanescent components decay as exp(-2*pi*|z|*kappa);
    for z < 0, H(z) = conj(H(-z)).
    """
    if z < 0:
        return np.conj(as_kernel(fx, fy, lam, n0, -z))
    lam_eff = lam / n0
    f2 = complex(1.0 / lam_eff ** 2)
    arg = np.asarray(f2 - np.asarray(fx) ** 2 - np.asarray(fy) ** 2, dtype=np.complex128)
    return np.exp(1j * (TWO_PI * z) * np.sqrt(arg))


def as_paraxial_kernel(fx, fy, lam: float, tx: float, ty: float, track: bool, n0: float, z: float):
    """Paraxial (Fresnel) transfer function, complex FP64.

    ``angular_spectrum.jl:27-35``: lam' = lam/n0, f0 = sin(theta)/lam',
    H = cis(-pi*lam'*z*((fx+f0x)^2 + (fy+f0y)^2) - offset + 2*pi*z/lam') with
    offset = 2*(f0x*fx + f0y*fy) when tilts are tracked, else 0 (for either sign of z).
    """
    le = lam / n0
    f0x, f0y = math.sin(tx) / le, math.sin(ty) / le
    fx, fy = np.asarray(fx, dtype=np.float64), np.asarray(fy, dtype=np.float64)
    off = 2.0 * (f0x * fx + f0y * fy) if track else 0.0
    return np.exp(1j * (-math.pi * le * z * ((fx + f0x) ** 2 + (fy + f0y) ** 2) - off + TWO_PI * z / le))


def tilted_as_kernel(fx, fy, lam: float, tx: float, ty: float, track: bool, n0: float, z: float):
    """Angular-spectrum transfer function of a tilted field, complex FP64.

    ``angular_spectrum.jl:45-62``: f0 = sin(theta)/lam', H evaluated at (fx+f0x, fy+f0y),
    plus the phase 2*pi*z*(f0x*fx + f0y*fy)*lam' when tilts are tracked; z < 0 gives
    conj(H(-z)).
    """
    if z < 0:
        return np.conj(tilted_as_kernel(fx, fy, lam, tx, ty, track, n0, -z))
    le = lam / n0
    f0x, f0y = math.sin(tx) / le, math.sin(ty) / le
    fx, fy = np.asarray(fx, dtype=np.float64), np.asarray(fy, dtype=np.float64)
    off = (f0x * fx + f0y * fy) * le if track else 0.0
    arg = np.asarray(complex(1.0 / le ** 2) - (fx + f0x) ** 2 - (fy + f0y) ** 2, dtype=np.complex128)
    return np.exp(1j * (TWO_PI * z) * (np.sqrt(arg) + off))


def offset_tilts(u, ds, lams, from_tilts, to_tilts):
    """``offset_tilts!`` (``Fields.jl:280-293``): u_b * cis(2 pi / lam_b ((sin(fx_b) - sin(tx_b)) x
    + (sin(fy_b) - sin(ty_b)) y)) on the centred grid ``spatial_vectors``."""
    u = _as_batch(np.asarray(u, dtype=np.complex128))
    b, ny, nx = u.shape
    lam_b = _lambdas(lams, b)
    xv, yv = spatial_vectors((nx, ny), ds)
    out = np.empty_like(u)
    for k in range(b):
        f0 = [_per_field(from_tilts[i], b)[k] for i in (0, 1)]
        t0 = [_per_field(to_tilts[i], b)[k] for i in (0, 1)]
        ph = TWO_PI / lam_b[k] * ((math.sin(f0[0]) - math.sin(t0[0])) * xv[None, :]
                                  + (math.sin(f0[1]) - math.sin(t0[1])) * yv[:, None])
        out[k] = u[k] * np.exp(1j * ph)
    return out


def shift_kernel(fx, fy, tx: float, ty: float, z: float):
    """``shift_kernel`` (``freespace_propagators/shift_kernel.jl:1-6``):
    cis(-2 pi z (tan(tx) fx + tan(ty) fy))."""
    return np.exp(-1j * TWO_PI * z * (math.tan(tx) * np.asarray(fx) + math.tan(ty) * np.asarray(fy)))


def shiftprop(u, ds, z, tilts, direction: int = 1):
    """``ShiftProp``: FFT2 -> shift kernel (conjugate for Backward) -> IFFT2, per field tilts."""
    u = _as_batch(np.asarray(u, dtype=np.complex128))
    b, ny, nx = u.shape
    z_b = _per_field(z, b)
    tx_b, ty_b = _per_field(tilts[0], b), _per_field(tilts[1], b)
    fx = fftfreq(nx, ds[0])[None, :]
    fy = fftfreq(ny, ds[1])[:, None]
    v = ft2(u)
    for k in range(b):
        v[k] *= conj_direction(shift_kernel(fx, fy, tx_b[k], ty_b[k], z_b[k]), direction)
    return ift2(v)


def filter_samples(filt, fx, fy):
    """``Complex(filter(fx, fy))`` evaluated element by element over broadcast (fx, fy), as the
    reference's kernel broadcast calls it (``angular_spectrum.jl:5,34,57``), complex FP64."""
    fx, fy = np.broadcast_arrays(np.asarray(fx, np.float64), np.asarray(fy, np.float64))
    out = np.empty(fx.shape, np.complex128)
    for idx in np.ndindex(fx.shape):
        out[idx] = complex(filt(float(fx[idx]), float(fy[idx])))
    return out


def as_transfer(nx: int, ny: int, dx: float, dy: float, lam: float, n0: float, z: float,
                tilts=(0.0, 0.0), paraxial: bool = False, track: bool = False, filt=None):
    """H sampled on the FFT-order frequency grid, shape (ny, nx).

    ``fill_kernel_cache`` broadcasting ``as_kernel`` over ``f_vec = (x=fx, y=fy')``:
    ``abstract_kernel.jl:15-25``, ``fourier_kernel.jl:17-18``; the kernel choice of
    ``angular_spectrum.jl:126-139`` (paraxial, else on-axis when both tilts are zero, else
    tilted).  ``filt``: the optional frequency filter, a function of (fx, fy) multiplying the
    kernel (``angular_spectrum.jl:5,34,57``): at (fx, fy) for the on-axis and paraxial kernels,
    at (fx + f0x, fy + f0y) for the tilted one; the on-axis and tilted kernels for z <= 0 are
    conj(kernel(-z)) with the filter inside (``:9-12,64-68``, chosen by ``Val(sign(z) > 0)``,
    ``:119``), so there the filter enters conjugated; the paraxial kernel never conjugates it.
    """
    fx = fftfreq(nx, dx)[None, :]
    fy = fftfreq(ny, dy)[:, None]
    if paraxial:
        h = as_paraxial_kernel(fx, fy, lam, tilts[0], tilts[1], track, n0, z)
        return h if filt is None else h * filter_samples(filt, fx, fy)
    if tilts[0] == 0 and tilts[1] == 0:
        h = as_kernel(fx, fy, lam, n0, z)
        gx, gy = fx, fy
    else:
        h = tilted_as_kernel(fx, fy, lam, tilts[0], tilts[1], track, n0, z)
        le = lam / n0
        gx, gy = fx + math.sin(tilts[0]) / le, fy + math.sin(tilts[1]) / le
    if filt is None:
        return h
    f = filter_samples(filt, gx, gy)
    return h * (f if z > 0 else np.conj(f))


def as_z_gradient(v, dv, ds, lams, z: float, n0: float = 1.0):
    """dL/dz of an on-axis ASProp with a trainable distance (``ASPropZ``,
    ``angular_spectrum.jl:236-300``, kernel of ``:1-12``), per field, given its output v and the
    cotangent dv at v: with V = FFT(v) = H FFT(u), DV = FFT(dv) and dH/dz = 2 pi i w~ H, where
    w = sqrt((n0/lam)^2 - fx^2 - fy^2) (principal root) and w~ = w for z >= 0, conj(w) for z < 0
    (H(z) = conj(H(-z))), Parseval gives Re (1/N) sum conj(DV) 2 pi i w~ V."""
    v = _as_batch(np.asarray(v, np.complex128))
    dv = _as_batch(np.asarray(dv, np.complex128))
    b, ny, nx = v.shape
    lam_b = _lambdas(lams, b)
    fx = fftfreq(nx, ds[0])[None, :]
    fy = fftfreq(ny, ds[1])[:, None]
    V, DV = ft2(v), ft2(dv)
    out = np.empty(b)
    for k in range(b):
        w = np.sqrt(np.asarray((n0 / lam_b[k]) ** 2 - fx ** 2 - fy ** 2, np.complex128))
        wt = w if z >= 0 else np.conj(w)
        out[k] = np.real(np.sum(np.conj(DV[k]) * (TWO_PI * 1j * wt) * V[k])) / (nx * ny)
    return out


def _per_field(v, b: int):
    a = np.atleast_1d(np.asarray(v, dtype=np.float64))
    return np.broadcast_to(a, (b,)) if a.size == 1 else a


def asprop(u, ds, lams, z, n0: float = 1.0, direction: int = 1, tilts=None, paraxial: bool = False,
           track: bool = False, filt=None):
    """``ASProp`` propagation: FFT2 -> multiply by H (or conj H) per slice -> IFFT2.

    Sequence (ft, ASKernelProp, ift): ``angular_spectrum.jl:193-229``; per-slice kernel
    multiply ``abstract_kernel.jl:64-83``; Backward = reverse direction
    ``freespace.jl:61-63`` + ``fourier_operator.jl:65-79``.  ``z`` and each tilt angle
    may be one value or one per field; ``tilts`` = (theta_x, theta_y) in radians; ``filt`` the
    optional frequency filter (see ``as_transfer``).
    """
    u = _as_batch(np.asarray(u, dtype=np.complex128))
    b, ny, nx = u.shape
    dx, dy = ds
    lam_b = _lambdas(lams, b)
    z_b = _per_field(z, b)
    tx_b = _per_field(0.0 if tilts is None else tilts[0], b)
    ty_b = _per_field(0.0 if tilts is None else tilts[1], b)
    v = ft2(u)
    for k in range(b):
        h = as_transfer(nx, ny, dx, dy, lam_b[k], n0, z_b[k], (tx_b[k], ty_b[k]), paraxial, track, filt)
        v[k] *= conj_direction(h, direction)
    return ift2(v)


# ----------------------------------------------------------------------------- RSProp

def rs_kernel(x, y, lam: float, z: float, nrm_f: float):
    """Rayleigh-Sommerfeld impulse response h(x, y), complex FP64.

    ``rayleigh_sommerfeld.jl:1-12``: k = 2*pi/lam, r = sqrt(x^2+y^2+z^2),
    h = nrm_f * (e^{ikr}/r) * (z/r) * (1/r - i*k); for z < 0, h(z) = conj(h(-z)).
    """
    if z < 0:
        return np.conj(rs_kernel(x, y, lam, -z, nrm_f))
    k = TWO_PI / lam
    r = np.sqrt(np.asarray(x) ** 2 + np.asarray(y) ** 2 + z * z)
    return nrm_f * (np.exp(1j * k * r) / r) * (z / r) * (1.0 / r - 1j * k)


def rs_tilted_kernel(x, y, lam: float, tx: float, ty: float, track: bool, z: float, nrm_f: float):
    """RS impulse response of a tilted field (``rayleigh_sommerfeld.jl:14-27``): h times
    cis(-2*pi*(x*f0x + y*f0y)), f0 = sin(theta)/lam, unless tilts are tracked; z < 0 gives
    conj(h(-z))."""
    if z < 0:
        return np.conj(rs_tilted_kernel(x, y, lam, tx, ty, track, -z, nrm_f))
    h = rs_kernel(x, y, lam, z, nrm_f)
    if track:
        return h
    f0x, f0y = math.sin(tx) / lam, math.sin(ty) / lam
    return h * np.exp(-1j * TWO_PI * (np.asarray(x) * f0x + np.asarray(y) * f0y))


def rs_offsets(n: int, dx: float, pad: str = "2n-1") -> np.ndarray:
    """Sample offsets of the padded convolution kernel along one axis.

    ``pad="2n-1"``: ``circshift((1-n):(n-1), n) .* dx`` = [0..n-1, -(n-1)..-1]*dx,
    exactly as ``convolution_kernel.jl:16``.  ``pad="2n"``: the power-of-two padding
    [0..n-1, n, -(n-1)..-1]*dx (offset n is never read by a cropped output, and
    sampling it at +n makes the kernel even, hence its DFT even).
    """
    if pad == "2n-1":
        j = np.concatenate([np.arange(n), np.arange(1 - n, 0)])
    elif pad == "2n":
        j = np.concatenate([np.arange(n + 1), np.arange(1 - n, 0)])
    else:
        raise ValueError(pad)
    return j.astype(np.float64) * dx


def rs_transfer(nx: int, ny: int, dx: float, dy: float, lam: float, z: float,
                pad: str = "2n-1", tilts=(0.0, 0.0), track: bool = False):
    """FFT2 of the sampled RS kernel on the padded grid (``convolution_kernel.jl:47-55``).

    nrm_f = dx*dy/(2*pi) (``rayleigh_sommerfeld.jl:55``).  The reference rounds h to the
    field precision before an FP32 FFT; the oracle keeps everything in FP64.  Nonzero
    tilts select the tilted kernel (``rayleigh_sommerfeld.jl:90-96``).
    """
    sx = rs_offsets(nx, dx, pad)[None, :]
    sy = rs_offsets(ny, dy, pad)[:, None]
    if tilts[0] == 0 and tilts[1] == 0:
        h = rs_kernel(sx, sy, lam, z, dx * dy / TWO_PI)
    else:
        h = rs_tilted_kernel(sx, sy, lam, tilts[0], tilts[1], track, z, dx * dy / TWO_PI)
    return ft2(h)


def pad(u: np.ndarray, shape) -> np.ndarray:
    """Zero-pad at offset 0 (``pad_crop_operators.jl:1-16``)."""
    u = _as_batch(u)
    out = np.zeros(u.shape[:-2] + tuple(shape), dtype=u.dtype)
    out[..., : u.shape[-2], : u.shape[-1]] = u
    return out


def crop(u: np.ndarray, ny: int, nx: int) -> np.ndarray:
    """Crop [0, n) at offset 0 (``pad_crop_operators.jl:61-79``)."""
    return np.ascontiguousarray(u[..., :ny, :nx])


def rsprop(u, ds, lams, z, direction: int = 1, pad_mode: str = "2n-1", tilts=None,
           track: bool = False):
    """``RSProp``: pad -> FFT2 -> multiply by FFT2(h) (or its conjugate) -> IFFT2 -> crop.

    Sequence (pad, ft, RSKernelProp, ift, crop): ``rayleigh_sommerfeld.jl:135-156``;
    Backward reverses the sequence and conjugates the kernel (``freespace.jl:61-63``,
    ``pad_crop_operators.jl:338-356``).  ``z`` and the tilts may be per field.
    """
    u = _as_batch(np.asarray(u, dtype=np.complex128))
    b, ny, nx = u.shape
    dx, dy = ds
    lam_b = _lambdas(lams, b)
    z_b = _per_field(z, b)
    tx_b = _per_field(0.0 if tilts is None else tilts[0], b)
    ty_b = _per_field(0.0 if tilts is None else tilts[1], b)
    py, px = (2 * ny - 1, 2 * nx - 1) if pad_mode == "2n-1" else (2 * ny, 2 * nx)
    v = ft2(pad(u, (py, px)))
    for k in range(b):
        h = rs_transfer(nx, ny, dx, dy, lam_b[k], z_b[k], pad_mode, (tx_b[k], ty_b[k]), track)
        v[k] *= conj_direction(h, direction)
    return crop(ift2(v), ny, nx)


def rs_valid_distance(nx: int, ny: int, dx: float, dy: float, lam: float) -> float:
    """Critical distance z_c below which RSProp warns (``rayleigh_sommerfeld.jl:29-36``)."""
    if dx < lam / 2 and dy < lam / 2:
        return 0.0
    zx = 0.0 if dx < lam / 2 else (nx * dx / 2) * math.sqrt(4 * dx * dx / lam ** 2 - 1)
    zy = 0.0 if dy < lam / 2 else (ny * dy / 2) * math.sqrt(4 * dy * dy / lam ** 2 - 1)
    return max(zx, zy)


# ----------------------------------------------------------------------------- Phase

def phase_apply(u, phi, direction: int = 1):
    """u * cis(s*phi) with s = sign(direction) (``phasemask.jl:90-98``).

    ``phi`` broadcasts over the batch (2-D mask shared by every slice).
    """
    return np.asarray(u, dtype=np.complex128) * np.exp(1j * direction * np.asarray(phi, np.float64))


def phase_grad(du, u_saved, direction: int = 1):
    """Mask gradient s * sum_b Im(du * conj(u_saved)) (``phasemask.jl:115-140``).

    ``du`` is the cotangent after un-applying the mask (``backpropagate_with_gradient!``,
    ``phasemask.jl:142-150``); ``u_saved`` is the mask input saved by
    ``propagate_and_save!`` (``phasemask.jl:100-113``).
    """
    g = direction * np.imag(_as_batch(du) * np.conj(_as_batch(u_saved)))
    return g.sum(axis=0)


def phase_backward(dv, u_saved, phi, direction: int = 1):
    """``backpropagate_with_gradient!`` for ``Phase``: returns (du, dphi)."""
    du = phase_apply(dv, phi, -direction)
    return du, phase_grad(du, u_saved, direction)


def mask_apply(u, m, direction: int = 1):
    """``Mask`` ``propagate!`` (``modulators/mask.jl:95-101``): u * m, or u * conj(m) Backward."""
    u = _as_batch(np.asarray(u, np.complex128))
    m = np.asarray(m, np.complex128)
    return u * (m if direction > 0 else np.conj(m))


def mask_backward(dv, u_saved, m, direction: int = 1):
    """``Mask`` ``backpropagate!`` with gradient (``mask.jl:112-141``): dm = sum_b
    conj_direction(dv conj(u_saved)) from the cotangent at the mask output, then
    du = dv * conj_direction(m, reverse(direction)).  Returns (du, dm)."""
    dv = _as_batch(np.asarray(dv, np.complex128))
    g = (dv * np.conj(_as_batch(u_saved))).sum(axis=0)
    dm = g if direction > 0 else np.conj(g)
    m = np.asarray(m, np.complex128)
    return dv * (np.conj(m) if direction > 0 else m), dm


def tea_transmission(h, lam: float, dn: float, r: complex = 1.0, direction: int = 1):
    """``TeaDOE`` transmission (``modulators/tea_doe.jl:146-152``): r cis(2 pi dn h / lam)
    Forward, conj(r) cis(-2 pi dn h / lam) Backward."""
    ph = TWO_PI * dn * np.asarray(h, np.float64) / lam
    return (r if direction > 0 else np.conj(r)) * np.exp(1j * direction * ph)


def _dn_of(dn, lam: float, b: int):
    # dn(lambda) as TeaDOE takes it: a number, a function of lambda or one value per field
    if callable(dn):
        return float(dn(lam))
    if isinstance(dn, (list, tuple, np.ndarray)):
        return float(dn[b])
    return float(dn)


def tea_apply(u, h, lams, dn, r=1.0, direction: int = 1):
    u = _as_batch(np.asarray(u, np.complex128))
    lam_b = _lambdas(lams, u.shape[0])
    return np.stack([u[b] * tea_transmission(h, lam_b[b], _dn_of(dn, lam_b[b], b), r, direction)
                     for b in range(u.shape[0])])


def tea_backward(dv, u_saved, h, lams, dn, r=1.0, direction: int = 1):
    """``tea_doe.jl:180-199``: du = dv * conj(t); dh = s * sum_b (2 pi dn / lam_b) Im(du conj(u_saved))."""
    dv = _as_batch(np.asarray(dv, np.complex128))
    lam_b = _lambdas(lams, dv.shape[0])
    dns = [_dn_of(dn, lam_b[b], b) for b in range(dv.shape[0])]
    du = np.stack([dv[b] * np.conj(tea_transmission(h, lam_b[b], dns[b], r, direction)) for b in range(dv.shape[0])])
    us = _as_batch(u_saved)
    dh = sum(TWO_PI * dns[b] / lam_b[b] * np.imag(du[b] * np.conj(us[b])) for b in range(dv.shape[0]))
    return du, direction * dh


# ----------------------------------------------------------------------------- CollinsProp

def collins_chirps(nx: int, ny: int, ds, lam: float, alpha, a: float, b: float, d: float, tilts=None,
                   track: bool = False):
    """The A and D chirps on the centred input grid (``collins_integral.jl:5-43``):
    A = cis(pi (x^2 (a - ax) + y^2 (a - ay)) / (b lam)) / lam and
    D = cis(pi (x^2 ax (d ax - 1) + y^2 ay (d ay - 1)) / (b lam)); a tilted field whose tilts
    are not tracked adds cis(2 pi (x f0x + y f0y)) to A and its conjugate to D (f0 = sin/lam)."""
    xv, yv = spatial_vectors((nx, ny), ds)
    x, y = xv[None, :], yv[:, None]
    A = np.exp(1j * math.pi * (x ** 2 * (a - alpha[0]) + y ** 2 * (a - alpha[1])) / (b * lam)) / lam
    D = np.exp(1j * math.pi * (x ** 2 * alpha[0] * (d * alpha[0] - 1) + y ** 2 * alpha[1] * (d * alpha[1] - 1))
               / (b * lam))
    if tilts is not None and not track and (tilts[0] != 0 or tilts[1] != 0):
        lin = np.exp(1j * TWO_PI * (x * math.sin(tilts[0]) / lam + y * math.sin(tilts[1]) / lam))
        A, D = A * lin, D * np.conj(lin)
    return A, D


def collins_conv_spectrum(nx: int, ny: int, ds, lam: float, alpha, b: float):
    """FFT of ``collins_convolution_kernel`` cis(pi (x^2 ax + y^2 ay) / (b lam)) sampled on the
    (2n-1) grid of offsets circshift((1-n):(n-1), n) * ds (``convolution_kernel.jl:13-18``,
    ``abstract_kernel.jl:15-25``)."""
    sx = np.roll(np.arange(1 - nx, nx), nx) * ds[0]
    sy = np.roll(np.arange(1 - ny, ny), ny) * ds[1]
    h = np.exp(1j * math.pi * (sx[None, :] ** 2 * alpha[0] + sy[:, None] ** 2 * alpha[1]) / (b * lam))
    return np.fft.fft2(h)


def collinsprop(u, ds, ds_out, abd, lams, direction: int = 1, tilts=None, track: bool = False):
    """``CollinsProp`` (``collins_integral.jl:262-296``): the sequence A chirp, pad to (2n-1),
    FFT, convolution spectrum times prod(ds ./ sqrt(i b)), IFFT, crop, D chirp; Backward runs it
    reversed with every factor conjugated and prod(ds' ./ sqrt(-i b)) (``:189-198``)."""
    u = _as_batch(np.asarray(u, dtype=np.complex128))
    nb, ny, nx = u.shape
    a, b, d = abd
    alpha = (ds_out[0] / ds[0], ds_out[1] / ds[1])
    lam_b = _lambdas(lams, nb)
    tx = _per_field(0.0 if tilts is None else tilts[0], nb)
    ty = _per_field(0.0 if tilts is None else tilts[1], nb)
    on_axis = all(v == 0 for v in tx) and all(v == 0 for v in ty)
    out = np.empty_like(u)
    for k in range(nb):
        A, D = collins_chirps(nx, ny, ds, lam_b[k], alpha, a, b, d, None if on_axis else (tx[k], ty[k]), track)
        H = collins_conv_spectrum(nx, ny, ds, lam_b[k], alpha, b)
        if direction > 0:
            pre, post, nrm = A, D, np.prod([complex(v) / np.sqrt(1j * b) for v in ds])
        else:
            pre, post, H = np.conj(D), np.conj(A), np.conj(H)
            nrm = np.prod([complex(v) / np.sqrt(-1j * b) for v in ds_out])
        w = np.zeros((2 * ny - 1, 2 * nx - 1), np.complex128)
        w[:ny, :nx] = u[k] * pre
        w = np.fft.ifft2(np.fft.fft2(w) * H) * nrm
        out[k] = w[:ny, :nx] * post
    return out


def collins_direct(u, ds, ds_out, abd, lams, tilts=None):
    """Brute-force Collins integral (Collins, JOSA 60, 1168 (1970)) on the grids:
    U(x'_m) = dx dy / (i lam b) sum_k u_k c(x_k) exp(i pi (a x_k^2 - 2 x_k x'_m + d x'_m^2) / (lam b))
    / c(x~_m), x'_m = alpha x~_m with x~ the centred input grid and c the tilt carrier."""
    u = _as_batch(np.asarray(u, dtype=np.complex128))
    nb, ny, nx = u.shape
    a, b, d = abd
    alpha = (ds_out[0] / ds[0], ds_out[1] / ds[1])
    lam_b = _lambdas(lams, nb)
    xv, yv = spatial_vectors((nx, ny), ds)
    out = np.empty_like(u)
    for k in range(nb):
        lam = lam_b[k]
        f0 = (0.0, 0.0) if tilts is None else (math.sin(_per_field(tilts[0], nb)[k]) / lam,
                                               math.sin(_per_field(tilts[1], nb)[k]) / lam)
        kx = np.exp(1j * math.pi * (a * xv[None, :] ** 2 - 2 * xv[None, :] * (alpha[0] * xv[:, None])
                                    + d * (alpha[0] * xv[:, None]) ** 2) / (lam * b))  # [m, k]
        ky = np.exp(1j * math.pi * (a * yv[None, :] ** 2 - 2 * yv[None, :] * (alpha[1] * yv[:, None])
                                    + d * (alpha[1] * yv[:, None]) ** 2) / (lam * b))
        cx = np.exp(1j * TWO_PI * f0[0] * xv)
        cy = np.exp(1j * TWO_PI * f0[1] * yv)
        w = u[k] * cy[:, None] * cx[None, :]
        v = ky @ w @ kx.T
        out[k] = v / (cy[:, None] * cx[None, :]) * ds[0] * ds[1] / (1j * lam * b)
    return out


# ----------------------------------------------------------------------------- CZT

def _czt_data(n: int, a, w, center: bool):
    """``prepare_czt_data`` (``FFTutils.jl:30-50``) for one dimension, N = 2n - 1: input chirp
    w^{k^2/2} a^{-k}, output chirp w^{k^2/2} (k shifted by (n-1)/2 when centred) and the kernel
    w^{-f^2/2} at f = N fftfreq(N, 1); a, w in polar form (rho, theta)."""
    N = 2 * n - 1
    k = np.arange(N, dtype=np.float64)
    if center:
        k = k - (n - 1) / 2.0
    f = np.round(N * np.fft.fftfreq(N, 1.0))
    (ra, ta), (rw, tw) = a, w
    lw, la = math.log(rw), math.log(ra)
    inp = np.exp(k ** 2 / 2 * lw - k * la) * np.exp(1j * (k ** 2 / 2 * tw - k * ta))
    outc = np.exp(k ** 2 / 2 * lw) * np.exp(1j * (k ** 2 / 2 * tw))
    h = np.exp(-f ** 2 / 2 * lw) * np.exp(-1j * (f ** 2 / 2 * tw))
    return inp, outc, h


def czt(x, dims, a=None, w=None, center_on_grid: bool = False, adjoint: bool = False):
    """``czt!`` of a ``CZTPlan`` (``FFTutils.jl:52-141``) and of its adjoint plan (``:143-155``)
    on an array [..., ny, nx]; ``dims`` holds 1 (x) and/or 2 (y); ``a``, ``w`` are per-dimension
    (rho, theta) pairs, defaulting to (1, 0) and (1, -2 pi / n)."""
    x = np.asarray(x, dtype=np.complex128)
    shape = x.shape
    ns = {1: shape[-1], 2: shape[-2]}
    axis = {1: -1, 2: -2}
    a = a or {1: (1.0, 0.0), 2: (1.0, 0.0)}
    w = w or {1: (1.0, -TWO_PI / ns[1]), 2: (1.0, -TWO_PI / ns[2])}
    tmp_shape = list(shape)
    for d in dims:
        tmp_shape[axis[d]] = 2 * ns[d] - 1
    inp = np.ones(tmp_shape[-2:], np.complex128)
    outc = np.ones(tmp_shape[-2:], np.complex128)
    h = np.ones(tmp_shape[-2:], np.complex128)
    for d in dims:
        i1, o1, h1 = _czt_data(ns[d], a[d], w[d], center_on_grid)
        sh = (1, -1) if d == 1 else (-1, 1)
        inp, outc, h = inp * i1.reshape(sh), outc * o1.reshape(sh), h * h1.reshape(sh)
    ax = tuple(axis[d] for d in dims)
    hf = np.fft.fftn(h, axes=ax)
    t = np.zeros(tmp_shape, np.complex128)
    t[..., :shape[-2], :shape[-1]] = x
    if adjoint:
        t = np.fft.ifftn(np.fft.fftn(t * np.conj(outc), axes=ax) * np.conj(hf), axes=ax) * np.conj(inp)
    else:
        t = np.fft.ifftn(np.fft.fftn(t * inp, axes=ax) * hf, axes=ax) * outc
    return t[..., :shape[-2], :shape[-1]]


# ----------------------------------------------------------------------------- AS_BPM

def bpm_kdz(dz: float, tilts=(0.0, 0.0)) -> float:
    """kdz = 2*pi*dz / prod(cos(theta)) (``bpm.jl:3-6,70``)."""
    return TWO_PI * dz / float(np.prod(np.cos(np.asarray(tilts, np.float64))))


def _dn_phase(dn_s, kdz, lam_b, s):
    # apply_dn_slice!: cis(s*kdz/lambdas*dn) with vacuum lambdas (bpm.jl:235-242)
    return np.exp(1j * (s * kdz / lam_b)[:, None, None] * np.asarray(dn_s, np.float64)[None])


def _bpm_setup(b: int, thickness: float, ns: int, lams, tilts):
    """dz, per-field kdz (``bpm.jl:3-6,70``, with each field's tilts) and vacuum lambdas."""
    if ns < 2:
        raise ValueError("BPM needs at least 2 slices (bpm.jl:40)")
    dz = thickness / ns
    tx = _per_field(0.0 if tilts is None else tilts[0], b)
    ty = _per_field(0.0 if tilts is None else tilts[1], b)
    kdz = np.array([bpm_kdz(dz, (tx[k], ty[k])) for k in range(b)])
    return dz, kdz, _lambdas(lams, b)


def _bpm_prop(u, ds, lam_b, z, n0, direction, tilts, paraxial, track, shift):
    # the slice propagator P: ASProp(z; n0, paraxial) or ShiftProp(z) (bpm.jl:54-73,140-142,207-217)
    if shift:
        return shiftprop(u, ds, z, (0.0, 0.0) if tilts is None else tilts, direction)
    return asprop(u, ds, lam_b, z, n0, direction, tilts=tilts, paraxial=paraxial, track=track)


def bpm_forward(u, dn, thickness: float, n0: float, lams, ds, direction: int = 1, tilts=None,
                paraxial: bool = False, track: bool = False, shift: bool = False):
    """``AS_BPM`` / ``Shift_BPM`` split-step forward; returns (v, saved) with saved[s] =
    slice-s input.

    ``bpm.jl:54-73`` (construction: dz = L/N, P(dz) and P(dz/2) are ASProps with n0, or
    ShiftProps for ``shift``), ``bpm.jl:244-260`` (half step, N-1 x [save, phase, full step],
    save, phase, half step).  ``dn`` has shape (N, ny, nx) (the reference's dn[nx, ny, N]).
    """
    u = _as_batch(np.asarray(u, dtype=np.complex128)).copy()
    dn = np.asarray(dn, np.float64)
    ns = dn.shape[0]
    dz, kdz, lam_b = _bpm_setup(u.shape[0], thickness, ns, lams, tilts)
    prop = lambda a, z: _bpm_prop(a, ds, lam_b, z, n0, direction, tilts, paraxial, track, shift)
    saved = []
    u = prop(u, dz / 2)
    for s in range(ns - 1):
        saved.append(u.copy())
        u = u * _dn_phase(dn[s], kdz, lam_b, direction)
        u = prop(u, dz)
    saved.append(u.copy())
    u = u * _dn_phase(dn[ns - 1], kdz, lam_b, direction)
    u = prop(u, dz / 2)
    return u, saved


def bpm_backward(dv, dn, thickness: float, n0: float, lams, ds, saved, direction: int = 1, tilts=None,
                 paraxial: bool = False, track: bool = False, shift: bool = False):
    """``AS_BPM`` / ``Shift_BPM`` adjoint with the dn gradient (``bpm.jl:275-312``);
    returns (du, ddn).

    ddn[s] = s * kdz_b/lam_b * sum_b Im(du_s * conj(saved[s])), du_s being the cotangent
    after un-applying slice s's phase.
    """
    du = _as_batch(np.asarray(dv, dtype=np.complex128)).copy()
    dn = np.asarray(dn, np.float64)
    ns = dn.shape[0]
    dz, kdz, lam_b = _bpm_setup(du.shape[0], thickness, ns, lams, tilts)
    prop = lambda a, z: _bpm_prop(a, ds, lam_b, z, n0, -direction, tilts, paraxial, track, shift)
    ddn = np.zeros_like(dn)
    coef = (direction * kdz / lam_b)[:, None, None]
    du = prop(du, dz / 2)
    for s in range(ns - 1, -1, -1):
        du = du * _dn_phase(dn[s], kdz, lam_b, -direction)
        ddn[s] = (coef * np.imag(du * np.conj(saved[s]))).sum(axis=0)
        du = prop(du, dz if s > 0 else dz / 2)
    return du, ddn


# ----------------------------------------------------------------------------- fields, losses

def power(u, ds) -> np.ndarray:
    """Per-slice power sum|u|^2 * dx*dy (``Fields.jl:843-846``)."""
    u = _as_batch(u)
    return (np.abs(u) ** 2).sum(axis=(-2, -1)) * ds[0] * ds[1]


def normalize_power(u, ds, target=1.0) -> np.ndarray:
    """``normalize_power!`` (``Fields.jl:889-892``)."""
    u = _as_batch(np.asarray(u, dtype=np.complex128))
    return u * np.sqrt(target / power(u, ds))[:, None, None]


def sq_intensity_diff_values(v, target_intensity, dc=None):
    """Per-slot values and cotangent of ``SquaredIntensityDifference``.

    ``src/metrics/squared_intensity_difference.jl:60-90``: a target with the field's
    full dims gives one value per field, sum (|v_b|^2 - I_b)^2; a target with the spatial
    dims only (and more than one field) gives one value, sum (sum_b |v_b|^2 - I)^2.  The
    cotangent is 4 dc d v_b with d the difference of the field's slot.
    """
    v = _as_batch(np.asarray(v, dtype=np.complex128))
    t = np.asarray(target_intensity, dtype=np.float64)
    spatial = t.ndim == 2 and v.shape[0] > 1
    if spatial:
        d = (np.abs(v) ** 2).sum(axis=0) - t
        vals = np.array([(d * d).sum()])
        g = 4.0 * (1.0 if dc is None else float(np.ravel(dc)[0]))
        return vals, g * d[None] * v
    d = np.abs(v) ** 2 - t.reshape(v.shape)
    vals = (d * d).sum(axis=(-2, -1))
    g = 4.0 * (np.ones(v.shape[0]) if dc is None else np.asarray(dc, dtype=np.float64).reshape(-1))
    return vals, g[:, None, None] * d * v


def sq_intensity_diff(v, target_intensity):
    """Objective sum of the ``SquaredIntensityDifference`` values and its cotangent
    (``backpropagate_metric`` with dc = 1)."""
    vals, dv = sq_intensity_diff_values(v, target_intensity)
    return float(vals.sum()), dv.reshape(np.shape(v))


def dot_product(u, v, ds):
    """Mode-selective ``DotProduct`` (``metrics/dot_product.jl:60-72``): per field
    c_b = sum u_b conj(v_b) * dx * dy (the code's form; its docstring writes <u, v>)."""
    u, v = _as_batch(np.asarray(u, np.complex128)), _as_batch(np.asarray(v, np.complex128))
    return (u * np.conj(v)).sum(axis=(-2, -1)) * (ds[0] * ds[1])


def dot_product_pullback(v, dc):
    """``backpropagate_metric`` of DotProduct (``dot_product.jl:97-106``): du_b = v_b * dc_b
    (the reference applies no ds factor here)."""
    v = _as_batch(np.asarray(v, np.complex128))
    return v * np.asarray(dc, np.complex128).reshape(-1, 1, 1)


def power_coupling(u, v, ds):
    """Mode-selective ``PowerCoupling`` (``power_coupling.jl:47-64``): targets normalised to
    unit power, p_b = |c_b|^2 with c the DotProduct against the normalised targets.  Returns
    (p, c, v_normalised)."""
    vh = normalize_power(v, ds)
    c = dot_product(u, vh, ds)
    return np.abs(c) ** 2, c, vh


def power_coupling_pullback(vh, c, dp=None):
    """``power_coupling.jl:66-69``: dc = 2 c dp, then the DotProduct pullback."""
    dp = np.ones(len(c)) if dp is None else np.asarray(dp, np.float64)
    return dot_product_pullback(vh, 2.0 * c * dp)


def dot_product_matrix(u, v, ds):
    """Non-selective ``DotProduct`` (``metrics/dot_product.jl:73-82``): c = V' U * prod(ds) with
    the fields reshaped to (pixels, modes), i.e. c[i, j] = sum conj(v_i) u_j dx dy; returned as an
    (n, n) array indexed [i, j]."""
    u, v = _as_batch(np.asarray(u, np.complex128)), _as_batch(np.asarray(v, np.complex128))
    U = u.reshape(u.shape[0], -1).T
    V = v.reshape(v.shape[0], -1).T
    return (V.conj().T @ U) * (ds[0] * ds[1])


def dot_product_matrix_pullback(v, dc):
    """``backpropagate_metric`` of the non-selective DotProduct (``dot_product.jl:92-96``):
    dU = V dc with the fields reshaped to (pixels, modes); no ds factor."""
    v = _as_batch(np.asarray(v, np.complex128))
    V = v.reshape(v.shape[0], -1).T
    return (V @ np.asarray(dc, np.complex128)).T.reshape(v.shape)


def power_coupling_matrix(u, v, ds):
    """Non-selective ``PowerCoupling`` (``power_coupling.jl:55-64``): |c|^2 entrywise with c the
    overlap matrix against the power-normalised targets.  Returns (p, c, v_normalised)."""
    vh = normalize_power(v, ds)
    c = dot_product_matrix(u, vh, ds)
    return np.abs(c) ** 2, c, vh


def power_coupling_matrix_pullback(vh, c, dp=None):
    """``power_coupling.jl:66-69``: dc = 2 c o dp entrywise, then the matrix pullback."""
    dp = np.ones(c.shape) if dp is None else np.asarray(dp, np.float64)
    return dot_product_matrix_pullback(vh, 2.0 * c * dp)


def half_sq_field_diff(v, t):
    """L = 1/2 sum|v - t|^2 and its cotangent v - t (harness loss for the fwd+adjoint unit)."""
    d = np.asarray(v) - np.asarray(t)
    return 0.5 * float(np.vdot(d, d).real), d


def sq_field_diff_values(v, t, dc=None):
    """``SquaredFieldDifference`` (``metrics/squared_field_difference.jl``, compute_metric and
    backpropagate_metric): per field b, c_b = sum |v_b - t_b|^2 (no area factor) and the pullback
    2 dc_b (v_b - t_b); returns (c, dv)."""
    v, t = _as_batch(np.asarray(v, np.complex128)), _as_batch(np.asarray(t, np.complex128))
    d = v - t
    c = (np.abs(d) ** 2).sum(axis=(-2, -1))
    g = 2.0 * (np.ones(len(c)) if dc is None else np.asarray(dc, np.float64).reshape(-1))
    return c, g[:, None, None] * d


def coupling_efficiency(u, v) -> np.ndarray:
    """|<u, v>|^2 / (<u,u><v,v>) per slice (``Fields.jl:748-756``)."""
    u, v = _as_batch(u), _as_batch(v)
    num = np.abs((np.conj(u) * v).sum(axis=(-2, -1))) ** 2
    den = (np.abs(u) ** 2).sum(axis=(-2, -1)) * (np.abs(v) ** 2).sum(axis=(-2, -1))
    return num / den


# ----------------------------------------------------------------------------- optimisation

def prox_clamp(x, lo, hi):
    """``ClampProx`` (``proximal_operators/clamp_prox.jl:33-35``)."""
    return np.clip(x, lo, hi)


def prox_ista(x, s, c=0.0):
    """``IstaProx`` (``proximal_operators/ista_prox.jl:1-18``): per real component,
    c if |x - c| <= s, else x - s or x + s."""
    def shrink(r):
        xc = r - c
        return np.where(np.abs(xc) <= s, c, np.where(xc > s, r - s, r + s))
    x = np.asarray(x)
    if np.iscomplexobj(x):
        return shrink(x.real) + 1j * shrink(x.imag)
    return shrink(x)


def prox_positive(x):
    """``PositiveProx`` (``proximal_operators/positive_prox.jl:1-7``)."""
    return np.where(np.asarray(x) < 0, 0.0, x)


def fista_t(t):
    """t_{k+1} of ``Fista`` (``OptimisersExt.jl:34-36``)."""
    return (1.0 + math.sqrt(1.0 + 4.0 * t * t)) / 2.0


def optim_step(rule, x, dx, state=None, t=1.0, eta=0.1, rho=0.9):
    """One step of a rule on (x, dx) returning (x_new, state_new, t_new).  ``rule``: "descent",
    "momentum", "nesterov" (Optimisers.jl 0.4.6 ``apply!``) or "fista" (``OptimisersExt.jl:31-40``;
    ``eta`` is the squared rate the reference stores)."""
    x = np.asarray(x)
    dx = np.asarray(dx)
    if rule == "descent":
        return x - eta * dx, state, t
    if rule == "momentum":
        m = rho * state + eta * dx
        return x - m, m, t
    if rule == "nesterov":
        d = -rho * rho * state + (1.0 + rho) * eta * dx
        return x - d, rho * state - eta * dx, t
    if rule == "fista":
        tn = fista_t(t)
        beta = (t - 1.0) / tn
        d = eta * dx - beta * (x - state)
        return x - d, x.copy(), tn
    raise ValueError(rule)


def adam_step(x, dx, m, v, bt, eta=0.001, beta=(0.9, 0.999), eps=1e-8):
    """Optimisers.jl 0.4.6 ``Adam`` ``apply!`` on FP32 parameters: m, v updated and stored in
    FP32, the step formed in FP64, x rounded back to FP32; ``bt`` = (b1^t, b2^t) of this step.
    Returns (x, m, v, bt_next)."""
    g = np.asarray(dx, np.float64)
    m = (beta[0] * np.asarray(m, np.float64) + (1 - beta[0]) * g).astype(np.float32)
    v = (beta[1] * np.asarray(v, np.float64) + (1 - beta[1]) * g * g).astype(np.float32)
    d = m.astype(np.float64) / (1 - bt[0]) / (np.sqrt(v.astype(np.float64) / (1 - bt[1])) + eps) * eta
    xn = (np.asarray(x, np.float64) - d).astype(np.float32)
    return xn, m, v, (bt[0] * beta[0], bt[1] * beta[1])


# ----------------------------------------------------------------------------- pipelines

def pipeline_forward(u, stages, ds, lams):
    """Run a list of stages forward, saving the inputs of every Phase.

    ``OpticalSequence`` forward loop (``optical_sequence.jl:54-60``) with trainable
    modulators using ``propagate_and_save!`` (``phasemask.jl:100-113``).  A stage is
    ``("as", z, n0[, tilts])``, ``("rs", z, pad_mode)``, ``("phase", phi)``, ``("mask", m)``,
    ``("tea", h, dn, r)``, ``("anchor", from_tilts, to_tilts)`` (``TiltAnchor``, whose
    reverse restores the entry tilts, ``tilt_anchor.jl:74-77``) or ``("collins", ds, ds', abd)``
    (whose reverse is CollinsProp's Backward, as the reference's pullback).
    """
    u = _as_batch(np.asarray(u, dtype=np.complex128))
    saved = []
    for st in stages:
        if st[0] == "as":
            u = asprop(u, ds, lams, st[1], st[2] if len(st) > 2 else 1.0, 1, tilts=st[3] if len(st) > 3 else None)
        elif st[0] == "anchor":
            u = offset_tilts(u, ds, lams, st[1], st[2])
        elif st[0] == "collins":
            u = collinsprop(u, st[1], st[2], st[3], lams, 1)
        elif st[0] == "rs":
            u = rsprop(u, ds, lams, st[1], 1, st[2] if len(st) > 2 else "2n-1")
        elif st[0] == "phase":
            saved.append(u.copy())
            u = phase_apply(u, st[1], 1)
        elif st[0] == "mask":
            saved.append(u.copy())
            u = mask_apply(u, st[1], 1)
        elif st[0] == "tea":
            saved.append(u.copy())
            u = tea_apply(u, st[1], lams, st[2], st[3] if len(st) > 3 else 1.0, 1)
        else:
            raise ValueError(st[0])
    return u, saved


def pipeline_backward(dv, stages, ds, lams, saved):
    """Reverse loop with reversed direction (``optical_sequence.jl:62-66``) and the Phase
    pullbacks of ``ChainRulesCoreExt.jl:35-65,81-111``; returns (du, [dphi per Phase])."""
    du = _as_batch(np.asarray(dv, dtype=np.complex128))
    grads = []
    k = len(saved)
    for st in reversed(stages):
        if st[0] == "as":
            du = asprop(du, ds, lams, st[1], st[2] if len(st) > 2 else 1.0, -1, tilts=st[3] if len(st) > 3 else None)
        elif st[0] == "anchor":
            du = offset_tilts(du, ds, lams, st[2], st[1])
        elif st[0] == "collins":
            du = collinsprop(du, st[1], st[2], st[3], lams, -1)
        elif st[0] == "rs":
            du = rsprop(du, ds, lams, st[1], -1, st[2] if len(st) > 2 else "2n-1")
        elif st[0] == "phase":
            k -= 1
            du, g = phase_backward(du, saved[k], st[1], 1)
            grads.append(g)
        elif st[0] == "mask":
            k -= 1
            du, g = mask_backward(du, saved[k], st[1], 1)
            grads.append(g)
        elif st[0] == "tea":
            k -= 1
            du, g = tea_backward(du, saved[k], st[1], lams, st[2], st[3] if len(st) > 3 else 1.0, 1)
            grads.append(g)
    return du, grads[::-1]


# ----------------------------------------------------------------------------- CPU baseline

class CpuASPropC64:
    """The reference's CPU ASProp at complex-FP32 precision, for the timed CPU baseline.

    Restates the reference's CPU algorithm: an in-place FFT2 (``FFTutils.jl:13-28``,
    FFTW there; pocketfft via ``scipy.fft`` with ``workers`` threads here), a multiply by
    the cached complex-FP32 kernel computed once in FP64 (``abstract_kernel.jl:15-25,
    39-46``; ``angular_spectrum.jl:94-100``) and the inverse FFT2.  ``forward`` and
    ``adjoint`` work in place on a (ny, nx) complex64 array.
    """

    def __init__(self, nx, ny, dx, dy, lam, z, n0=1.0, workers=1):
        import scipy.fft as sfft
        self._fft = sfft
        self.workers = int(workers)
        self.h = as_transfer(nx, ny, dx, dy, lam, n0, z).astype(np.complex64)
        self.hc = np.conj(self.h)

    def _apply(self, u, h):
        f = self._fft.fft2(u, workers=self.workers, overwrite_x=True)
        np.multiply(f, h, out=f)
        u[...] = self._fft.ifft2(f, workers=self.workers, overwrite_x=True)
        return u

    def forward(self, u):
        return self._apply(u, self.h)

    def adjoint(self, u):
        return self._apply(u, self.hc)


class CpuRSPipelineF64:
    """The reference's CPU path for the beam-splitter pipeline in Float64, for the timed CPU
    baseline: RSProp(z1) -> Phase -> RSProp(z2) -> Phase -> RSProp(z3) with the intensity loss and
    the gradients of both masks (the reference's config run through ``Zygote.withgradient``).

    Each RSProp is the reference's (pad to 2n-1, FFT2, multiply by the cached kernel spectrum, IFFT2,
    crop; ``rayleigh_sommerfeld.jl:135-156``, ``convolution_kernel.jl:47-55``) with pocketfft
    (``scipy.fft``, ``workers`` threads) in place of FFTW; the Phases save their inputs and give
    s * Im(du conj(u_saved)) (``phasemask.jl:100-150``); the loss and its pullback are
    ``squared_intensity_difference.jl:78-90``.  ``iteration(phis)`` returns (loss, [dphi1, dphi2]).
    """

    def __init__(self, u0, target_i, ds, lam, zs, workers=1):
        import scipy.fft as sfft
        self._fft = sfft
        self.workers = int(workers)
        self.u0 = np.asarray(u0, np.complex128)
        self.t = np.asarray(target_i, np.float64)
        ny, nx = self.u0.shape
        self.n = (ny, nx)
        self.h = {z: rs_transfer(nx, ny, ds[0], ds[1], lam, z, "2n-1") for z in set(zs)}
        self.zs = list(zs)

    def _rs(self, u, z, direction):
        ny, nx = self.n
        f = self._fft.fft2(u, s=(2 * ny - 1, 2 * nx - 1), workers=self.workers)
        f *= self.h[z] if direction > 0 else np.conj(self.h[z])
        return self._fft.ifft2(f, workers=self.workers, overwrite_x=True)[:ny, :nx]

    def iteration(self, phis):
        saved = []
        u = self._rs(self.u0, self.zs[0], 1)
        for z, phi in zip(self.zs[1:], phis):
            saved.append(u)
            u = self._rs(u * np.exp(1j * phi), z, 1)
        d = np.abs(u) ** 2 - self.t
        loss = float((d * d).sum())
        du = 4.0 * d * u
        grads = []
        for z, phi, s in zip(self.zs[:0:-1], phis[::-1], saved[::-1]):
            du = self._rs(du, z, -1) * np.exp(-1j * phi)
            grads.append(np.imag(du * np.conj(s)))
        return loss, grads[::-1]
