"""Every launch variant of the three passes of a propagation (the candidates the autotuner may
keep, fh_plan_variants / fh_plan_set_variant) against the FP64 oracle, forward and backward:
ASProp at 2048 x 2048 (the y-pass length with the most variants, the wave-per-column pass among
them) with one wavelength, with a wavelength per field and with tilted fields (full
transfer-function tables), and RSProp on a 1024 grid (zero-padded 2048-point lines, of which the
y pass reads and writes half)."""

import ctypes
import warnings

import numpy as np
import pytest

import fluxoptics as fo
from fluxoptics import _lib
from oracle import fluxoptics_oracle as O

from _util import c64, rel_l2, speckle_batch, to_dev, to_host

pytestmark = pytest.mark.gpu

FIELD_TOL = 1e-5


def _variants(plan, pid):
    cnt, cur = ctypes.c_int(), ctypes.c_int()
    _lib.call("fh_plan_variants", plan, pid, ctypes.byref(cnt), ctypes.byref(cur))
    return cnt.value


def _sweep(prop, u, ref, tag):
    plan = prop.plan(u)
    for pid in (0, 1, 2):
        for k in range(_variants(plan, pid)):
            _lib.call("fh_plan_set_variant", plan, pid, k)
            print(f"[variants] {tag} pass {pid} variant {k}", flush=True)
            for d, s in ((fo.Forward, 1), (fo.Backward, -1)):
                v = to_host(prop.propagate(u, d).electric)
                err = rel_l2(v, ref[s])
                assert err <= FIELD_TOL, (tag, pid, k, s, err)
        _lib.call("fh_plan_set_variant", plan, pid, -1)


@pytest.mark.parametrize("case", ["one_lambda", "lambda_per_field", "tilted"])
def test_asprop_2048_every_variant(case):
    n, ds, z, n0 = 2048, (0.5, 0.5), 3000.0, 1.0
    lams = {"one_lambda": 0.8, "lambda_per_field": [0.8, 0.65], "tilted": [0.8, 0.8]}[case]
    tilts = ([0.01, -0.02], [0.015, 0.0]) if case == "tilted" else None
    u0 = c64(speckle_batch(2, n, n, ds[0], ds[1], 150.0, 41))
    u = fo.ScalarField(to_dev(u0), ds, lams, tilts)
    p = fo.ASProp(u, z, n0=n0)
    ref = {s: O.asprop(u0, ds, lams, z, n0, s, tilts=tilts) for s in (1, -1)}
    _sweep(p, u, ref, f"as {case}")


def test_rsprop_1024_every_variant():
    n, ds, lam, z = 1024, (0.5, 0.5), 0.8, 600.0
    u0 = c64(speckle_batch(2, n, n, ds[0], ds[1], 80.0, 43))
    u = fo.ScalarField(to_dev(u0), ds, lam)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")  # RSProp's sampling advice
        p = fo.RSProp(u, z)
    ref = {s: O.rsprop(u0, ds, lam, z, s, "2n-1") for s in (1, -1)}
    _sweep(p, u, ref, "rs")



def test_staged_x_passes_match_direct_bitwise():
    """The x passes whose blocked side moves through LDS (variant 2, the default at 2048) against
    the direct passes (variant 0) on the same inputs, bit for bit, repeated: the staging keeps each
    chunk's two 8-byte LDS accesses separate (merged into one instruction of adjacent slots, a few
    values per launch came back stale), so every repetition must agree exactly."""
    n, ds, lam, z, batch = 2048, (0.5, 0.5), 0.8, 3000.0, 8
    u0 = c64(speckle_batch(batch, n, n, ds[0], ds[1], 150.0, 47))
    u = fo.ScalarField(to_dev(u0), ds, lam)
    p = fo.ASProp(u, z)
    plan = p.plan(u)
    ptr = lambda f: ctypes.c_void_p(f.electric.data_ptr())
    for pid in (0, 2):
        outs = {}
        for k in (0, 2, 0, 2):
            _lib.call("fh_plan_set_variant", plan, pid, k)
            f = fo.ScalarField(to_dev(u0), ds, lam)
            _lib.call("fh_propagate_pass", plan, ptr(f), 1, pid)
            out = to_host(f.electric)
            if k in outs:
                assert np.array_equal(out, outs[k]), (pid, k)
            outs[k] = out
        _lib.call("fh_plan_set_variant", plan, pid, -1)
        print(f"[staged x pass] pass {pid}: variants 0 and 2 bitwise equal", flush=True)
        assert np.array_equal(outs[0], outs[2]), pid


def test_staged_x_pass_with_phase_prologue_bitwise():
    """A Phase -> ASProp -> SquaredIntensityDifference system at 2048^2 (the fused block: the first
    x pass carries the Phase as its prologue and stores the saved input) with the plan's x passes
    direct (variant 0) and staged (variant 2): loss and mask gradient bit for bit."""
    import math
    from fluxoptics import inputs
    n, ds, lam, z = 2048, (0.5, 0.5), 0.8, 3000.0
    u0 = c64(speckle_batch(2, n, n, ds[0], ds[1], 150.0, 49))
    tgt = (np.abs(c64(speckle_batch(2, n, n, ds[0], ds[1], 150.0, 50))) ** 2).astype(np.float32)
    u = fo.ScalarField(to_dev(u0), ds, lam)
    ph = fo.Phase(u, inputs.uniform(51, (n, n), 0, 2 * math.pi).astype(np.float32), trainable=True)
    p = fo.ASProp(u, z)
    system = fo.ScalarSource(u) | ph | p
    metric = fo.SquaredIntensityDifference((u, tgt))
    plan = p.plan(u)
    res = {}
    for k in (0, 2):
        for pid in (0, 2):
            _lib.call("fh_plan_set_variant", plan, pid, k)
        val, grads = system.withgradient(metric)
        res[k] = (float(val), grads[ph].cpu().numpy().copy())
    for pid in (0, 2):
        _lib.call("fh_plan_set_variant", plan, pid, -1)
    print(f"[staged x pass, Phase prologue] loss {res[2][0]:.9e}", flush=True)
    assert res[0][0] == res[2][0]
    assert np.array_equal(res[0][1], res[2][1])


@pytest.mark.parametrize("batch", [1, 3, 5])
def test_asprop_2048_default_y_pass_by_batch(batch):
    """The built-in y-pass default of a one-table 2048-point plan depends on the batch (one wave a
    column for one or two fields, the LDS-table pipelined pass up to four, the staged-tile pass
    from five on): each, as the default picks it, against the FP64 oracle in both directions."""
    n, ds, lam, z = 2048, (0.5, 0.5), 0.8, 3000.0
    u0 = c64(speckle_batch(batch, n, n, ds[0], ds[1], 150.0, 60 + batch))
    u = fo.ScalarField(to_dev(u0), ds, lam)
    p = fo.ASProp(u, z)
    print(f"[default y pass] batch {batch}", flush=True)
    for d, s in ((fo.Forward, 1), (fo.Backward, -1)):
        v = to_host(p.propagate(u, d).electric)
        err = rel_l2(v, O.asprop(u0, ds, lam, z, 1.0, s))
        assert err <= FIELD_TOL, (batch, s, err)
