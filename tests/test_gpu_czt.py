"""GPU parity of the chirp z-transform plans against the oracle (relative L2 1e-5)."""

import math

import numpy as np
import pytest
import torch

import fluxoptics as fo
from oracle import fluxoptics_oracle as O
from _util import c64, random_field, rel_l2, to_dev, to_host

pytestmark = pytest.mark.gpu


def test_default_czt_is_the_dft():
    print("[czt] default = DFT, 3 x 64x128", flush=True)
    x0 = c64(random_field((3, 64, 128), 141))
    x = to_dev(x0)
    v = fo.plan_czt(x, (1, 2)) * x
    assert rel_l2(to_host(v), np.fft.fft2(x0)) <= 1e-5
    assert np.array_equal(to_host(x), x0)  # out of place
    v = fo.plan_czt(x, 1) * x
    assert rel_l2(to_host(v), np.fft.fft(x0, axis=-1)) <= 1e-5


@pytest.mark.parametrize("dims,center", [((1, 2), False), ((1, 2), True), ((2,), True), ((1,), False)])
def test_czt_general_and_adjoint(dims, center):
    print(f"[czt] dims={dims} center={center}", flush=True)
    ny, nx = 128, 256
    a = ((1.0005, 0.3), (0.9995, -0.2))
    w = ((1.0, -0.011), (1.0001, 0.017))
    x0 = c64(random_field((2, ny, nx), 143))
    x = to_dev(x0)
    p = fo.plan_czt(x, dims, a, w, center_on_grid=center)
    ad = {1: a[0], 2: a[1]}
    wd = {1: w[0], 2: w[1]}
    assert rel_l2(to_host(p * x), O.czt(x0, dims, ad, wd, center)) <= 1e-5
    assert rel_l2(to_host(p.adjoint() * x), O.czt(x0, dims, ad, wd, center, adjoint=True)) <= 1e-5
    q = fo.plan_czt_(x, dims, a, w, center_on_grid=center)
    y = q * x
    assert y.data_ptr() == x.data_ptr()
    assert rel_l2(to_host(x), O.czt(x0, dims, ad, wd, center)) <= 1e-5
