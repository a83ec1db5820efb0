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


def _config5_pairs(torch_dev=True):
    """The uneven batch of 64 RS fields: 16 each of 256^2, 512^2, 1024^2 and 2048^2, random LG
    modes (p in [0, 3], l in [-4, 4], waist 0.1 * n * ds), ds 1, lambda 1.064, z ~ U[500, 5000]
    per field (seeded), one RSProp plan per size with per-field distances."""
    sizes, per, ds, lam = (256, 512, 1024, 2048), 16, 1.0, 1.064
    pairs, host = [], []
    for k, n in enumerate(sizes):
        ps = inputs.uniform(500 + 10 * k, (per,), 0, 4).astype(int)
        ls = inputs.uniform(501 + 10 * k, (per,), -4, 5).astype(int)
        zs = list(inputs.uniform(505 + 10 * k, (per,), 500.0, 5000.0))
        u0 = c64(np.stack([inputs.laguerre_gaussian(n, n, ds, ds, 0.1 * n * ds, int(p), int(l))
                           for p, l in zip(ps, ls)]))
        u = fo.ScalarField(to_dev(u0), (ds, ds), lam)
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            pairs.append((fo.RSProp(u, zs), u))
        host.append((n, u0, zs))
    return pairs, host


def test_uneven_batch_64_rs_fields_full_size():
    """The 64-field uneven batch at its full sizes, autotuned as the benchmark builds it: every
    field of both graph launches equals the eager launches of the same plans bit for bit, and
    two fields of each size (the first and the last, so two 2048^2 fields with 4096-point
    transforms) match the FP64 oracle with the reference's (2n - 1) padding in both directions
    (rayleigh_sommerfeld.jl:135-156)."""
    print("[group] 64 RS fields, 16 each of 256..2048, autotuned", flush=True)
    pairs, host = _config5_pairs()
    g = fo.PropagationGroup(pairs, autotune_reps=2)
    x0 = [u.electric.clone() for _, u in pairs]
    eager = {}
    for d in (fo.Forward, fo.Backward):
        eager[d] = [p.propagate(u, d).electric.clone() for p, u in pairs]
    for rep in range(2):
        for d in (fo.Forward, fo.Backward):
            for (_, u), x in zip(pairs, x0):
                u.electric.copy_(x)
            _ = torch.randn(1000, device="cuda").sum()
            g.propagate_(d)
            torch.cuda.synchronize()
            for (_, u), e, (n, _, _) in zip(pairs, eager[d], host):
                assert torch.equal(u.electric, e), f"{n}^2 group != eager, direction {d}, replay {rep}"
    assert g.graph_count() == 2
    for (n, u0, zs), ef, eb in zip(host, eager[fo.Forward], eager[fo.Backward]):
        for b in (0, len(zs) - 1):
            print(f"[group] oracle {n}^2 field {b} z={zs[b]:.1f}", flush=True)
            for d, e in ((1, ef), (-1, eb)):
                ref = O.rsprop(u0[b:b + 1], (1.0, 1.0), 1.064, [zs[b]], d, "2n-1")
                err = rel_l2(to_host(e[b:b + 1]), ref)
                assert err <= FIELD_TOL, (n, b, d, err)


def test_group_with_one_plan_for_several_fields():
    """One propagator listed for two same-shaped fields uses one plan (and its RS workspace)
    twice; the group runs those launches one after the other, so the results equal the eager
    launches and the oracle."""
    print("[group] one RSProp plan for two fields", flush=True)
    ds, lam, n = 1.0, 1.064, 128
    u1h = c64(lg_batch(n, 2, ds, 700))
    u2h = c64(lg_batch(n, 2, ds, 710))
    u3h = c64(lg_batch(64, 3, ds, 720))
    u1 = fo.ScalarField(to_dev(u1h), (ds, ds), lam)
    u2 = fo.ScalarField(to_dev(u2h), (ds, ds), lam)
    u3 = fo.ScalarField(to_dev(u3h), (ds, ds), lam)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        p = fo.RSProp(u1, 800.0)
        q = fo.ASProp(u3, 900.0)
    assert p.plan(u1).value == p.plan(u2).value
    e = [p.propagate(u1).electric.clone(), p.propagate(u2).electric.clone(), q.propagate(u3).electric.clone()]
    g = fo.PropagationGroup([(p, u1), (q, u3), (p, u2)])
    for rep in range(3):
        g.propagate_(fo.Forward)
        torch.cuda.synchronize()
        assert torch.equal(u1.electric, e[0]) and torch.equal(u2.electric, e[1]) and torch.equal(u3.electric, e[2]), rep
        u1.electric.copy_(to_dev(u1h)), u2.electric.copy_(to_dev(u2h)), u3.electric.copy_(to_dev(u3h))
    assert rel_l2(to_host(e[0]), O.rsprop(u1h, (ds, ds), lam, 800.0, 1, "2n-1")) <= FIELD_TOL
    assert rel_l2(to_host(e[1]), O.rsprop(u2h, (ds, ds), lam, 800.0, 1, "2n-1")) <= FIELD_TOL
    assert rel_l2(to_host(e[2]), O.asprop(u3h, (ds, ds), lam, 900.0)) <= FIELD_TOL
    # one buffer twice in one call is refused
    with pytest.raises(RuntimeError, match="same buffer"):
        g.propagate_(fo.Forward, [u1, u3, u1])
