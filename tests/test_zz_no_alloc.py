"""GPU test of fh_sq_intensity_diff_noalloc, the metric entry point that allocates nothing
(its partials live in a fixed per-GPU array): through the C ABI, its loss values and pullback
equal fh_sq_intensity_diff's bit for bit on two calls in a row (the second smaller, and
spatial-only), both agree with the FP64 oracle, and sizes beyond its maximum are refused."""

import ctypes

import numpy as np
import pytest
import torch

from fluxoptics import _lib, inputs
from fluxoptics.components import context
from oracle import fluxoptics_oracle as O

pytestmark = pytest.mark.gpu


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _run(name, ctx, u, tgt, target_batch, dc):
    b, ny, nx = u.shape
    nout = 1 if (target_batch == 1 and b > 1) else b
    loss = torch.full((nout,), float("nan"), dtype=torch.float64, device=u.device)
    du = torch.full_like(u, float("nan"))
    _lib.call(name, ctx, _ptr(u), _ptr(tgt), nx, ny, b, target_batch, None if dc is None else _ptr(dc), _ptr(loss),
              _ptr(du))
    torch.cuda.synchronize()
    return loss, du


@pytest.mark.parametrize("order", ["noalloc_first", "library_first"])
def test_noalloc_metric_equals_library_metric(order):
    dev = torch.device("cuda", 0)
    ctx = context(dev)
    cases = [  # (batch, ny, nx, target_batch, dc): a larger call, then a smaller spatial-only one
        (6, 192, 256, 6, np.array([0.5, 1.0, 2.0, -1.0, 0.25, 3.0])),
        (3, 48, 64, 1, None),
    ]
    for k, (b, ny, nx, tb, dc) in enumerate(cases):
        print(f"[noalloc] call {k}: batch {b}, {ny}x{nx}, target batch {tb}", flush=True)
        uh = (inputs.normal(70 + k, (b, ny, nx)) + 1j * inputs.normal(80 + k, (b, ny, nx))).astype(np.complex64)
        th = inputs.uniform(90 + k, (tb, ny, nx), 0.0, 2.0).astype(np.float32)
        u = torch.from_numpy(uh).to(dev)
        t = torch.from_numpy(th).to(dev)
        dcd = None if dc is None else torch.from_numpy(dc).to(dev)
        names = ["fh_sq_intensity_diff_noalloc", "fh_sq_intensity_diff"]
        if order == "library_first":
            names.reverse()
        res = {n: _run(n, ctx, u, t, tb, dcd) for n in names}
        (l0, d0), (l1, d1) = res["fh_sq_intensity_diff_noalloc"], res["fh_sq_intensity_diff"]
        assert torch.equal(l0, l1), (k, l0, l1)
        assert torch.equal(d0, d1), k
        vals, dv = O.sq_intensity_diff_values(uh.astype(np.complex128), th[0] if tb == 1 and b > 1 else th,
                                              dc)
        assert np.allclose(l0.cpu().numpy(), vals, rtol=1e-10, atol=0)
        err = np.linalg.norm(d0.cpu().numpy() - dv) / np.linalg.norm(dv)
        assert err <= 1e-6, err


def test_noalloc_metric_refuses_more_slots_than_its_array_holds():
    dev = torch.device("cuda", 0)
    ctx = context(dev)
    u = torch.zeros((1025, 2, 2), dtype=torch.complex64, device=dev)
    t = torch.zeros((1025, 2, 2), dtype=torch.float32, device=dev)
    loss = torch.zeros(1025, dtype=torch.float64, device=dev)
    with pytest.raises(_lib.FluxHipError, match="no-allocation maximum"):
        _lib.call("fh_sq_intensity_diff_noalloc", ctx, _ptr(u), _ptr(t), 2, 2, 1025, 1025, None, _ptr(loss), None)
    # 1024 fields is the limit and runs
    u = u[:1024].contiguous()
    t = t[:1024].contiguous()
    _lib.call("fh_sq_intensity_diff_noalloc", ctx, _ptr(u), _ptr(t), 2, 2, 1024, 1024, None, _ptr(loss), None)
    torch.cuda.synchronize()
    assert torch.count_nonzero(loss[:1024]) == 0
