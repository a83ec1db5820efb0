"""The reference's published beam-splitter optimisation (docs/src/api/index.md:79-119) on the
device: 250x250 RS fields (padded to 512^2), three trainable Phases, PowerCoupling, Adam(0.1);
the loss of the first iterations against the same loop on the FP64 oracle."""

import math
import warnings

import numpy as np
import pytest
import torch

import fluxoptics as fo
from fluxoptics import inputs
from oracle import fluxoptics_oracle as O

pytestmark = pytest.mark.gpu


def test_published_splitter_loss_trajectory():
    print("[published splitter] 250^2, 3 iterations", flush=True)
    n, ds, lam, w0, z = 250, 10.0, 0.7, 150.0, 2e5
    u0 = inputs.normalize_power(inputs.gaussian(n, n, ds, ds, w0), ds, ds).astype(np.complex64)
    off = 3.8 * w0
    vf = sum(inputs.gaussian(n, n, ds, ds, w0, sx * off, sy * off) for sx in (-1, 1) for sy in (-1, 1))
    vf = inputs.normalize_power(vf, ds, ds).astype(np.complex64)
    u = fo.ScalarField(torch.from_numpy(u0).cuda(), (ds, ds), lam)
    v = fo.ScalarField(torch.from_numpy(vf).cuda(), (ds, ds), lam)
    does = [fo.Phase(u, np.zeros((n, n), np.float32), trainable=True, buffered=True) for _ in range(3)]
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        prop = fo.RSProp(u, z)
    system = fo.ScalarSource(u) | does[0] | prop | does[1] | prop | does[2] | prop
    metric = fo.PowerCoupling(v)
    opt = fo.setup(fo.Adam(0.1), system)

    phis = [np.zeros((n, n), np.float32) for _ in range(3)]
    ms = [np.zeros((n, n), np.float32) for _ in range(3)]
    vs = [np.zeros((n, n), np.float32) for _ in range(3)]
    bt = (0.9, 0.999)
    for it in range(3):
        val, grads = system.withgradient(metric, weights=-1.0)
        fo.update_(opt, system, grads)
        stages = []
        for p in phis:
            stages += [("phase", p.astype(np.float64)), ("rs", z)]
        out, saved = O.pipeline_forward(u0.astype(np.complex128), stages, (ds, ds), lam)
        pc, c, vh = O.power_coupling(out, vf.astype(np.complex128), (ds, ds))
        dv = O.power_coupling_pullback(vh, c, [-1.0])
        _, gs = O.pipeline_backward(dv, stages, (ds, ds), lam, saved)
        loss_ref = 1.0 - pc.sum()
        assert abs((1.0 + float(val)) - loss_ref) <= 1e-4 * max(loss_ref, 1e-3), (it, float(val), loss_ref)
        for k in range(3):
            phis[k], ms[k], vs[k], _ = O.adam_step(phis[k], gs[k].astype(np.float32), ms[k], vs[k], bt, eta=0.1)
        bt = (bt[0] * 0.9, bt[1] * 0.999)
    torch.cuda.synchronize()


@pytest.mark.parametrize("rule", ["adam", "fista"])
def test_recorded_train_step_matches_eager_loop(rule):
    """capture_train_step: replays of one recorded iteration (gradient + Adam step, or Fista with a
    clamp, whose coefficients live in device memory) follow the eager loop's losses and masks bit
    for bit (eager work between the replays)."""
    n, ds, lam = 128, 10.0, 0.7
    print(f"[recorded train step] 128^2 {rule}", flush=True)
    g = inputs.normalize_power(inputs.gaussian(n, n, ds, ds, 80.0), ds, ds).astype(np.complex64)
    t = inputs.normalize_power(inputs.gaussian(n, n, ds, ds, 80.0, 200.0, -150.0), ds, ds).astype(np.complex64)
    losses = {}
    masks = {}
    for mode in ("eager", "graph"):
        u = fo.ScalarField(torch.from_numpy(g).cuda(), (ds, ds), lam)
        v = fo.ScalarField(torch.from_numpy(t).cuda(), (ds, ds), lam)
        does = [fo.Phase(u, np.zeros((n, n), np.float32), trainable=True, buffered=True) for _ in range(2)]
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            prop = fo.RSProp(u, 5e4)
        system = fo.ScalarSource(u) | does[0] | prop | does[1] | prop
        metric = fo.PowerCoupling(v)
        opt = fo.setup(fo.Adam(0.05) if rule == "adam" else fo.ProxRule(fo.Fista(3.0), fo.ClampProx(-1.0, 1.0)),
                       system)
        vals = []
        if mode == "eager":
            for _ in range(5):
                val, grads = system.withgradient(metric, weights=-1.0)
                fo.update_(opt, system, grads)
                vals.append(float(val))
        else:
            step = system.capture_train_step(metric, opt, weights=-1.0)
            for _ in range(5):
                val, _ = step.replay()
                vals.append(float(val))
                torch.zeros(3, device="cuda").sum()
        torch.cuda.synchronize()
        losses[mode] = vals
        masks[mode] = [d.phi.cpu().numpy() for d in does]
    assert losses["eager"] == losses["graph"]
    for a, b in zip(masks["eager"], masks["graph"]):
        assert np.array_equal(a, b)
