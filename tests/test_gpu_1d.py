"""GPU parity of 1-D fields (Fields.jl:135-136: electric[nx, B], Nd = 1) against the FP64 oracle
(which runs them as one-row 2-D fields, pinned against the reference's 1-D kernels on the CPU):
ASProp on power-of-two and other lengths (one launch per propagation), multi-wavelength, both
signs of z and both directions, paraxial, tilted, a frequency filter (the waveguide-tomography
NA diaphragm, angular_spectrum.jl:14-25 with filter(fx)), ShiftProp, the 1-D FFT, a Phase ->
ASProp system differentiated on the device, and AS_BPM / Shift_BPM with the dn gradient (one
launch per direction: the slice loop of a field in one workgroup)."""

import math

import numpy as np
import pytest
import torch

import fluxoptics as fo
from fluxoptics import inputs
from oracle import fluxoptics_oracle as O

from _util import c64, rel_l2, to_dev, to_host

pytestmark = pytest.mark.gpu

FIELD_TOL = 1e-5
GRAD_TOL = 1e-4


def rows(a):
    """(B, nx) -> the oracle's (B, 1, nx)."""
    return np.asarray(a)[:, None, :]


def field_1d(b, n, seed):
    return c64(inputs.normal(seed, (b, n)) + 1j * inputs.normal(seed + 1, (b, n)))


@pytest.mark.parametrize("n", [1024, 1600])
@pytest.mark.parametrize("z", [40.0, -40.0])
def test_1d_asprop_on_axis(n, z):
    print(f"[1d] ASProp n={n} z={z}", flush=True)
    dx, lams, n0 = 0.1, [0.455, 0.532, 0.64], 1.52
    u0 = field_1d(3, n, 1)
    u = fo.ScalarField(to_dev(u0), (dx,), lams)
    p = fo.ASProp(u, z, n0=n0)
    for d, s in ((fo.Forward, 1), (fo.Backward, -1)):
        v = to_host(p.propagate(u, d).electric)
        ref = O.asprop(rows(u0), (dx, 1.0), lams, z, n0, s)[:, 0, :]
        assert rel_l2(v, ref) <= FIELD_TOL, (n, z, s, rel_l2(v, ref))


@pytest.mark.parametrize("n", [512, 1600])
def test_1d_tilted_paraxial_filtered_and_shift(n):
    print(f"[1d] tilted / paraxial / filter / shift n={n}", flush=True)
    dx, lam, n0, z = 0.1, 0.455, 1.52, -6.0
    b = 5
    th = list(np.linspace(-0.3, 0.3, b))
    u0 = field_1d(b, n, 11)
    u = fo.ScalarField(to_dev(u0), (dx,), lam, tilts=(th,))
    na = 1.4
    diaphragm = lambda fx: 1.0 if abs(fx) <= abs(na / lam) else 0.0  # the notebook's make_diaphragm
    cases = [
        ("tilted", fo.ASProp(u, z, n0=n0), dict(tilts=(th, [0.0] * b))),
        ("tilted tracked", fo.ASProp(u, z, n0=n0, track_tilts=True), dict(tilts=(th, [0.0] * b), track=True)),
        ("paraxial", fo.ASProp(u, z, n0=n0, paraxial=True), dict(tilts=(th, [0.0] * b), paraxial=True)),
        ("tilted + NA filter", fo.ASProp(u, z, n0=n0, filter=diaphragm),
         dict(tilts=(th, [0.0] * b), filt=lambda fx, fy: diaphragm(fx))),
    ]
    for name, p, kw in cases:
        for d, s in ((fo.Forward, 1), (fo.Backward, -1)):
            v = to_host(p.propagate(u, d).electric)
            ref = O.asprop(rows(u0), (dx, 1.0), lam, z, n0, s, **kw)[:, 0, :]
            assert rel_l2(v, ref) <= FIELD_TOL, (name, s, rel_l2(v, ref))
    sp = fo.ShiftProp(u, 25.0)
    v = to_host(sp.propagate(u).electric)
    ref = O.shiftprop(rows(u0), (dx, 1.0), 25.0, (th, [0.0] * b), 1)[:, 0, :]
    assert rel_l2(v, ref) <= FIELD_TOL


def test_1d_fft_and_rsprop_refusal():
    n = 2048
    u0 = field_1d(2, n, 21)
    u = fo.ScalarField(to_dev(u0), (0.5,), 1.0)
    f = fo.FourierOperator(u, True)
    v = to_host(f.propagate(u).electric)
    assert rel_l2(v, np.fft.fft(u0, axis=-1)) <= FIELD_TOL
    w = to_host(f.propagate(u, fo.Backward).electric)  # Backward swaps FFT and IFFT (1/n)
    assert rel_l2(w, np.fft.ifft(u0, axis=-1)) <= FIELD_TOL
    a = to_host(f.backpropagate(u).electric) if hasattr(f, "backpropagate") else \
        to_host(f.backpropagate_(u.copy()).electric)  # the exact adjoint: unnormalised inverse
    assert rel_l2(a, np.fft.ifft(u0, axis=-1) * n) <= FIELD_TOL
    with pytest.raises(NotImplementedError):
        fo.RSProp(u, 100.0)


def test_1d_phase_system_gradient():
    """ScalarSource -> Phase -> ASProp -> intensity loss on 43 tilted 1-D fields of 1600 points."""
    print("[1d] Phase -> ASProp withgradient, 43 x 1600", flush=True)
    n, b, dx, lam, n0, z = 1600, 43, 0.1, 0.455, 1.52, 20.0
    th = list(np.linspace(-0.35, 0.35, b))
    u0 = c64(np.ones((b, n)))
    u = fo.ScalarField(to_dev(u0), (dx,), lam, tilts=(th,))
    phi = (0.3 * inputs.normal(31, (n,))).astype(np.float32)
    tgt = inputs.uniform(32, (b, n), 0.5, 1.5).astype(np.float32)
    doe = fo.Phase(u, phi, trainable=True, buffered=True)
    p = fo.ASProp(u, z, n0=n0)
    system = fo.ScalarSource(u) | doe | p
    metric = fo.SquaredIntensityDifference((u, tgt))
    value, grads = system.withgradient(metric)
    tilts = (th, [0.0] * b)
    a = rows(u0) * np.exp(1j * phi.astype(np.float64))
    v = O.asprop(a, (dx, 1.0), lam, z, n0, 1, tilts=tilts)
    dd = np.abs(v) ** 2 - rows(tgt)
    loss = float((dd * dd).sum())
    g = O.asprop(4 * dd * v, (dx, 1.0), lam, z, n0, -1, tilts=tilts)
    dphi = np.imag(g * np.conj(a)).sum(axis=0)[0]
    assert abs(float(value) - loss) <= FIELD_TOL * loss
    assert rel_l2(grads[doe].cpu().numpy(), dphi) <= GRAD_TOL


@pytest.mark.parametrize("shift", [False, True])
def test_1d_bpm_forward_adjoint_gradient(shift):
    """43 tilted 1-D fields of 1600 points through 300 slices (the waveguide-tomography volume:
    n_bulk 1.52, lambda 0.455, 20 um thick), forward + dn gradient against the oracle."""
    print(f"[1d] BPM 43 x 1600 x 300 shift={shift}", flush=True)
    n, b, ns, dx, lam, n0, L = 1600, 43, 300, 0.1, 0.455, 1.52, 20.0
    th = list(np.linspace(-0.35, 0.35, b))
    u0 = c64(np.exp(1j * 0.1 * inputs.normal(41, (b, n))))
    x = (np.arange(n) - (n - 1) / 2) * dx
    dn = (0.01 * np.exp(-(x / 3.0) ** 2)[None, :] * np.cos(np.linspace(0, 3, ns))[:, None]).astype(np.float32)
    u = fo.ScalarField(to_dev(u0), (dx,), lam, tilts=(th,))
    bpm = (fo.Shift_BPM(u, L, dn, trainable=True, buffered=True) if shift else
           fo.AS_BPM(u, L, n0, dn, trainable=True, buffered=True))
    t0 = c64(inputs.normal(42, (b, n)) + 1j * inputs.normal(43, (b, n)))
    w = u.copy()
    bpm.propagate_and_save_(w)
    v = to_host(w.electric)
    tilts = (th, [0.0] * b)
    vr, saved = O.bpm_forward(rows(u0), dn[:, None, :], L, n0, lam, (dx, 1.0), tilts=tilts, shift=shift)
    assert rel_l2(v, vr[:, 0, :]) <= FIELD_TOL, rel_l2(v, vr[:, 0, :])
    dv = fo.ScalarField(to_dev(v - t0), (dx,), lam, tilts=(th,))
    du, ddn = bpm.backpropagate_with_gradient_(dv)
    dur, ddnr = O.bpm_backward(rows(v - t0), dn[:, None, :], L, n0, lam, (dx, 1.0), saved, tilts=tilts, shift=shift)
    assert rel_l2(to_host(du.electric), dur[:, 0, :]) <= FIELD_TOL
    assert rel_l2(ddn.cpu().numpy(), ddnr[:, 0, :]) <= GRAD_TOL, rel_l2(ddn.cpu().numpy(), ddnr[:, 0, :])
    # zero dn: the volume is ASProp(L) in the bulk index
    if not shift:
        z = fo.AS_BPM(u, L, n0, np.zeros_like(dn))
        got = to_host(z.propagate(u).electric)
        ref = O.asprop(rows(u0), (dx, 1.0), lam, L, n0, 1, tilts=tilts)[:, 0, :]
        assert rel_l2(got, ref) <= FIELD_TOL
