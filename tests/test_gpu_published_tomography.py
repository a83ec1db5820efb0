"""GPU check of the waveguide-tomography loop (notebooks/04_waveguide_tomography.ipynb) at its
sizes: 43 tilted 1-D fields of 1600 points, AS_BPM of 300 slices (trainable dn) and the NA-limited
ASProp(-6 um; n0 1.52); the first iteration's loss and dn gradient against the FP64 oracle chain,
then Fista + Ista steps (ProxRule(Fista(0.01), IstaProx(3e-5))) lower the loss, recorded replays
equal to host-driven iterations."""

import numpy as np
import pytest
import torch

import fluxoptics as fo
from fluxoptics import inputs
from oracle import fluxoptics_oracle as O

from _util import rel_l2

pytestmark = pytest.mark.gpu


def _setup():
    nx, dx, lam, nb, na, nz, L = 1600, 0.0405, 0.455, 1.52, 1.4, 300, 20.0
    th = list(np.linspace(-0.45, 0.45, 49)[5:48])
    b = len(th)
    u0 = fo.ScalarField(torch.ones((b, nx), dtype=torch.complex64, device="cuda"), (dx,), lam, tilts=(th,))
    diaphragm = lambda fx: np.abs(fx) <= abs(na / lam)
    target = inputs.uniform(5, (b, nx), 0.5, 1.5).astype(np.float32)
    sample = fo.AS_BPM(u0, L, nb, np.zeros((nz, nx), np.float32), trainable=True, buffered=True)
    p = fo.ASProp(u0, -6.0, n0=nb, filter=diaphragm)  # a 1-D filter takes fx only
    system = fo.ScalarSource(u0) | sample | p
    metric = fo.SquaredIntensityDifference((u0, target))
    geo = dict(nx=nx, dx=dx, lam=lam, nb=nb, nz=nz, L=L, th=th, b=b, diaphragm=diaphragm, target=target)
    return system, sample, metric, geo


def test_tomography_iteration_against_oracle_and_descent():
    print("[tomography] 43 x 1600 x 300, first iteration vs oracle", flush=True)
    system, sample, metric, g = _setup()
    value, grads = system.withgradient(metric)
    torch.cuda.synchronize()
    b, nx, nz = g["b"], g["nx"], g["nz"]
    tilts = (g["th"], [0.0] * b)
    ds = (g["dx"], 1.0)
    filt = lambda fx, fy: float(g["diaphragm"](fx))
    u0 = np.ones((b, 1, nx), np.complex128)
    dn = np.zeros((nz, 1, nx))
    w, saved = O.bpm_forward(u0, dn, g["L"], g["nb"], g["lam"], ds, tilts=tilts)
    v = O.asprop(w, ds, g["lam"], -6.0, g["nb"], 1, tilts=tilts, filt=filt)
    d = np.abs(v) ** 2 - g["target"][:, None, :]
    loss = float((d * d).sum())
    dw = O.asprop(4 * d * v, ds, g["lam"], -6.0, g["nb"], -1, tilts=tilts, filt=filt)
    _, ddn = O.bpm_backward(dw, dn, g["L"], g["nb"], g["lam"], ds, saved, tilts=tilts)
    assert abs(float(value) - loss) <= 1e-5 * loss
    assert rel_l2(grads[sample].cpu().numpy(), ddn[:, 0, :]) <= 1e-4
    # a few steps of the notebook's rule lower the loss; recorded replays follow the same values
    opt = fo.setup(fo.ProxRule(fo.Fista(0.01), fo.IstaProx(3e-5)), system)
    fo.update_(opt, system, grads)
    vals = []
    for _ in range(3):
        val, gr = system.withgradient(metric)
        fo.update_(opt, system, gr)
        vals.append(float(val))
    assert vals[-1] < float(value)
