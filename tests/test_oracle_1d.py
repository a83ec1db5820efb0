"""CPU pins of the oracle on 1-D fields (Fields.jl:135-136).  The oracle runs a 1-D field
[nx, B] as a batch of one-row 2-D fields (B, 1, nx) with ds = (dx, 1) and theta_y = 0; these tests
check that reduction against the reference's own 1-D kernels written out here
(angular_spectrum.jl:14-25 as_kernel, :37-43 as_paraxial_kernel, :64-76 tilted_as_kernel):
exact plane-wave eigenvectors for each kernel, both signs of z and both directions; and the 1-D
split-step BPM against ASProp(L) at dn = 0 and its dn gradient against central differences."""

import cmath
import math

import numpy as np
import pytest

from fluxoptics import inputs
from oracle import fluxoptics_oracle as O


def h_1d(f, lam, n0, z, kind, t0=0.0, track=False):
    """The reference's 1-D kernels at frequency f (FP64, principal square root)."""
    le = lam / n0
    if kind == "paraxial":
        f0 = math.sin(t0) / le
        off = 2 * f0 * f if track else 0.0
        return cmath.exp(1j * (-math.pi * le * z * (f + f0) ** 2 - off + 2 * math.pi * z / le))
    if z < 0:
        return h_1d(f, lam, n0, -z, kind, t0, track).conjugate()
    if kind == "axis":
        return cmath.exp(1j * 2 * math.pi * z * cmath.sqrt(complex(1 / le ** 2 - f * f)))
    f0 = math.sin(t0) / le
    off = f0 * f * le if track else 0.0
    return cmath.exp(1j * 2 * math.pi * z * (cmath.sqrt(complex(1 / le ** 2 - (f + f0) ** 2)) + off))


@pytest.mark.parametrize("kind,track", [("axis", False), ("paraxial", False), ("paraxial", True),
                                        ("tilted", False), ("tilted", True)])
@pytest.mark.parametrize("z", [23.0, -23.0])
def test_1d_asprop_plane_wave_eigenvectors(kind, track, z):
    n, dx, lam, n0 = 24, 0.35, 0.8, 1.1   # 1/(2 dx) > n0/lam: the edge of the band is evanescent
    t0 = 0.07 if kind != "axis" else 0.0
    j = np.arange(n)
    for k in (0, 3, n // 2, n - 5, n - 1):
        u = np.exp(2j * math.pi * k * j / n)[None, None, :]
        f = (k if k < n // 2 else k - n) / (n * dx)
        h = h_1d(f, lam, n0, z, kind, t0, track)
        for d in (1, -1):
            v = O.asprop(u, (dx, 1.0), lam, z, n0, d, tilts=(t0, 0.0), paraxial=kind == "paraxial", track=track)
            want = (h if d > 0 else h.conjugate()) * u
            assert np.max(np.abs(v - want)) < 1e-12, (k, d)


def test_1d_bpm_zero_dn_and_gradient():
    n, dx, lam, n0, L, ns = 32, 0.5, 1.0, 1.5, 12.0, 6
    u = (inputs.normal(1, (2, 1, n)) + 1j * inputs.normal(2, (2, 1, n)))
    tilts = ([0.03, -0.02], [0.0, 0.0])
    v, _ = O.bpm_forward(u, np.zeros((ns, 1, n)), L, n0, [1.0, 0.8], (dx, 1.0), tilts=tilts)
    ref = O.asprop(u, (dx, 1.0), [1.0, 0.8], L, n0, 1, tilts=tilts)
    assert np.linalg.norm(v - ref) <= 1e-12 * np.linalg.norm(ref)
    # dn gradient of L = 1/2 |v - t|^2 by central differences on a few voxels
    dn = 0.01 * inputs.normal(3, (ns, 1, n))
    t = inputs.normal(4, (2, 1, n)) + 1j * inputs.normal(5, (2, 1, n))
    lams = [1.0, 0.8]
    loss = lambda d: 0.5 * np.sum(np.abs(O.bpm_forward(u, d, L, n0, lams, (dx, 1.0), tilts=tilts)[0] - t) ** 2)
    v, saved = O.bpm_forward(u, dn, L, n0, lams, (dx, 1.0), tilts=tilts)
    _, ddn = O.bpm_backward(v - t, dn, L, n0, lams, (dx, 1.0), saved, tilts=tilts)
    h = 1e-6
    for (s, x) in [(0, 3), (2, 17), (5, 30)]:
        e = np.zeros_like(dn)
        e[s, 0, x] = h
        fd = (loss(dn + e) - loss(dn - e)) / (2 * h)
        assert abs(fd - ddn[s, 0, x]) <= 1e-6 * max(1.0, abs(fd)), (s, x, fd, ddn[s, 0, x])
