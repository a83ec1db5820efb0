"""GPU parity of the standalone Fourier operators (FourierOperator, FourierWrapper) against
the oracle's FFT2/IFFT2 (tolerance 1e-5), and their exact adjoints (dot-product test)."""

import numpy as np
import pytest
import torch

import fluxoptics as fo
from fluxoptics import inputs
from oracle import fluxoptics_oracle as O

from _util import c64, random_field, rel_l2, to_dev, to_host

pytestmark = pytest.mark.gpu

FIELD_TOL = 1e-5


@pytest.mark.parametrize("ny,nx,batch", [(16, 16, 1), (64, 32, 3), (512, 256, 2), (2048, 2048, 1)])
def test_fft2_matches_oracle(ny, nx, batch):
    print(f"[fft2] {ny}x{nx} b{batch}", flush=True)
    u0 = c64(random_field((batch, ny, nx), 70 + nx))
    u = fo.ScalarField(to_dev(u0), (1.0, 1.0), 0.8)
    ft, ift = fo.FourierOperator(u, True), fo.FourierOperator(u, False)
    v = ft.propagate(u, fo.Forward)
    assert rel_l2(to_host(v.electric), O.ft2(u0)) <= FIELD_TOL
    w = ift.propagate(u, fo.Forward)
    assert rel_l2(to_host(w.electric), O.ift2(u0)) <= FIELD_TOL
    # Backward swaps FFT and IFFT (fourier_operator.jl:73-79)
    assert torch.equal(ft.propagate(u, fo.Backward).electric, w.electric)
    back = ift.propagate(v, fo.Forward)
    assert rel_l2(to_host(back.electric), u0) <= FIELD_TOL


@pytest.mark.parametrize("direct", [True, False])
def test_fourier_adjoint_dot_product(direct):
    """<A u, v> = <u, A^H v> for the operator and its backpropagate_."""
    n, b = 128, 2
    u0, v0 = c64(random_field((b, n, n), 80)), c64(random_field((b, n, n), 81))
    op = fo.FourierOperator(fo.ScalarField(to_dev(u0), (1.0, 1.0), 0.8), direct)
    au = to_host(op.propagate(fo.ScalarField(to_dev(u0), (1.0, 1.0), 0.8), fo.Forward).electric)
    ahv = to_host(op.backpropagate_(fo.ScalarField(to_dev(v0), (1.0, 1.0), 0.8), fo.Forward).electric)
    lhs, rhs = np.vdot(v0, au), np.vdot(ahv, u0)
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)


def test_fourier_wrapper_phase_is_a_frequency_filter():
    """FourierWrapper(u, Phase) = IFFT(cis(phi) * FFT(u)) (fourier_wrapper.jl:34-39)."""
    n, b = 256, 2
    u0 = c64(random_field((b, n, n), 82))
    phi = inputs.uniform(83, (n, n), 0, 2 * np.pi).astype(np.float32)
    u = fo.ScalarField(to_dev(u0), (1.0, 1.0), 0.8)
    w = fo.FourierWrapper(u, fo.Phase(u, phi))
    v = w.propagate(u, fo.Forward)
    ref = O.ift2(O.phase_apply(O.ft2(u0), phi.astype(np.float64), 1))
    assert rel_l2(to_host(v.electric), ref) <= FIELD_TOL
