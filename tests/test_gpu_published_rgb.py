"""A reduced copy of the reference's RGB DOE notebook loop (03_RGB_beam_shaping.ipynb): three
wavelengths as one batch on a non-power-of-two grid, RSProp -> TeaDOE(dn(lambda)) -> RSProp ->
TeaDOE -> RSProp, per-wavelength intensity targets, ProxRule(Fista, ClampProx); the losses of
the first iterations against the same loop on the FP64 oracle."""

import warnings

import numpy as np
import pytest
import torch

import fluxoptics as fo
from fluxoptics import inputs
from oracle import fluxoptics_oracle as O

pytestmark = pytest.mark.gpu


def test_rgb_doe_loop_losses():
    print("[rgb doe loop] 96^2 x 3", flush=True)
    n, ds, lams, z = 96, 0.5, [0.640, 0.538, 0.455], 300.0
    dn_tab = {0.455: 0.518, 0.538: 0.512, 0.640: 0.508}
    dn = lambda lam: dn_tab[round(lam, 3)]
    g = inputs.gaussian(n, n, ds, ds, 6.0)
    u0 = np.stack([2 * g, g, g]).astype(np.complex64)
    tg = np.stack([np.abs(inputs.gaussian(n, n, ds, ds, 4.0, c, 2.0)) ** 2 for c in (-5.0, 0.0, 5.0)])
    tg = (tg / tg.sum(axis=(1, 2), keepdims=True)).astype(np.float32)
    u = fo.ScalarField(torch.from_numpy(u0).cuda(), (ds, ds), lams)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        p = fo.RSProp(u, z)
    does = [fo.TeaDOE(u, dn, trainable=True, buffered=True) for _ in range(2)]
    system = fo.ScalarSource(u) | p | does[0] | p | does[1] | p
    metric = fo.SquaredIntensityDifference((u, tg))
    opt = fo.setup(fo.ProxRule(fo.Fista(2.0), fo.ClampProx(-0.5, 0.5)), system)
    hs = [np.zeros((n, n)) for _ in range(2)]
    prev = [h.copy() for h in hs]
    t = 1.0
    for it in range(3):
        val, grads = system.withgradient(metric)
        fo.update_(opt, system, grads)
        stages = [("rs", z), ("tea", hs[0], dn), ("rs", z), ("tea", hs[1], dn), ("rs", z)]
        v, saved = O.pipeline_forward(u0.astype(np.complex128), stages, (ds, ds), lams)
        vals, dv = O.sq_intensity_diff_values(v, tg.astype(np.float64))
        _, gs = O.pipeline_backward(dv, stages, (ds, ds), lams, saved)
        assert abs(float(val) - vals.sum()) <= 1e-4 * vals.sum(), (it, float(val), vals.sum())
        tn = O.fista_t(t)
        for k in range(2):
            hs[k], prev[k], _ = O.optim_step("fista", hs[k], gs[k], prev[k], t, eta=4.0)
            hs[k] = O.prox_clamp(hs[k], -0.5, 0.5)
        t = tn
    torch.cuda.synchronize()
    for k in range(2):
        assert np.allclose(does[k].h.cpu().numpy(), hs[k], rtol=1e-3, atol=1e-4)
