"""GPU tests of PropagationGroup (uneven batches as one hipGraph): every graph replay equals
the eager launches bit for bit, with eager work and synchronisations between replays, and
the results match the FP64 oracle."""

import math
import warnings

import numpy as np
import pytest
import torch

import fluxoptics as fo
from fluxoptics import inputs
from oracle import fluxoptics_oracle as O

from _util import c64, rel_l2, to_dev, to_host

pytestmark = pytest.mark.gpu

FIELD_TOL = 1e-5


def lg_batch(n, count, ds, seed):
    """Random Laguerre-Gaussian modes, p in [0, 3], l in [-4, 4], waist 0.1 * n * ds."""
    ps = (inputs.uniform(seed, (count,), 0, 4)).astype(int)
    ls = (inputs.uniform(seed + 1, (count,), -4, 5)).astype(int)
    return np.stack([inputs.laguerre_gaussian(n, n, ds, ds, 0.1 * n * ds, int(p), int(l))
                     for p, l in zip(ps, ls)])


def build(sizes, per, seed, kinds=("rs",)):
    pairs, host = [], []
    ds, lam = 1.0, 1.064
    for k, n in enumerate(sizes):
        u0 = c64(lg_batch(n, per, ds, seed + 10 * k))
        zs = list(inputs.uniform(seed + 10 * k + 5, (per,), 500.0, 5000.0))
        u = fo.ScalarField(to_dev(u0), (ds, ds), lam)
        kind = kinds[k % len(kinds)]
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            p = fo.RSProp(u, zs) if kind == "rs" else fo.ASProp(u, zs)
        pairs.append((p, u))
        host.append((kind, u0, zs))
    return pairs, host


def test_group_replays_equal_eager_and_oracle():
    print("[group] sizes 32..256, rs+as", flush=True)
    pairs, host = build([32, 64, 128, 256], 3, 500, kinds=("rs", "as"))
    g = fo.PropagationGroup(pairs)
    inputs_dev = [u.electric.clone() for _, u in pairs]
    eager_f, eager_b = [], []
    for p, u in pairs:
        v = p.propagate(u, fo.Forward)
        eager_f.append(v.electric.clone())
        eager_b.append(p.propagate(u, fo.Backward).electric.clone())
    for rep in range(3):
        for (_, u), x in zip(pairs, inputs_dev):
            u.electric.copy_(x)
        _ = torch.randn(1000, device="cuda").sum()  # eager work between replays
        g.propagate_(fo.Forward)
        torch.cuda.synchronize()
        for (_, u), e in zip(pairs, eager_f):
            assert torch.equal(u.electric, e), f"forward replay {rep}"
        for (_, u), x in zip(pairs, inputs_dev):
            u.electric.copy_(x)
        g.propagate_(fo.Backward)
        torch.cuda.synchronize()
        for (_, u), e in zip(pairs, eager_b):
            assert torch.equal(u.electric, e), f"backward replay {rep}"
    assert g.graph_count() == 2
    # parity of the group's results with the oracle
    for (kind, u0, zs), e in zip(host, eager_f):
        ref = O.rsprop(u0, (1.0, 1.0), 1.064, zs, 1, "2n") if kind == "rs" else O.asprop(u0, (1.0, 1.0), 1.064, zs)
        assert rel_l2(to_host(e), ref) <= FIELD_TOL
    # other buffers of the same shapes: a new graph, same results
    others = [u.set_field_data(x.clone()) for (_, u), x in zip(pairs, inputs_dev)]
    g.propagate_(fo.Forward, others)
    torch.cuda.synchronize()
    for o, e in zip(others, eager_f):
        assert torch.equal(o.electric, e)
    assert g.graph_count() == 3


def test_group_autotune_and_shape_checks():
    pairs, host = build([64, 128], 2, 600)
    g = fo.PropagationGroup(pairs, autotune_reps=2)
    with pytest.raises(ValueError):
        g.propagate_(fo.Forward, [pairs[1][1], pairs[0][1]])
    g.propagate_(fo.Forward)
    torch.cuda.synchronize()
    for (_, u), (kind, u0, zs) in zip(pairs, host):
        assert rel_l2(to_host(u.electric), O.rsprop(u0, (1.0, 1.0), 1.064, zs, 1, "2n")) <= FIELD_TOL
