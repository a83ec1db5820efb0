#!/usr/bin/env python3
"""Headline benchmark: 2048^2 ASProp forward + adjoint propagations per second on one MI355X.

Workload: 8 complex-FP32 fields on a 2048x2048 grid, ds = 0.5 um, lambda = 0.8 um,
z = 5000 um; each field is a Gaussian (waist 150 um) times exp(i*phi), phi uniform in
[0, 2*pi) per pixel (seeded).  One step = the forward propagation of the batch plus the
adjoint propagation of a cotangent batch, both in place and resident in HBM.  One unit
= one field's forward + adjoint (32*n^2 algorithmic bytes: read and write the field
twice).

Prints one JSON line.  Only one GPU is supported: --gpus above 1 exits with a message.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fluxoptics.jl_amd"))

HBM_PEAK = 8.0e12          # MI355X HBM3E, bytes/s
N, BATCH, DS, LAM, Z, W0 = 2048, 8, 0.5, 0.8, 5000.0, 150.0
METRIC = "2048² fwd+adjoint propagations/s on 1 MI355X; % HBM roofline; err vs FP64"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "asprop_2048x2048_b8_traffic.txt")


def parse():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=20,
                    help="untimed steps first (the clocks settle over the first ~20 steps)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain FP64 comparator")
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="bound on the CPU baseline sample's wall time")
    ap.add_argument("--no-workloads", action="store_true",
                    help="skip the secondary workloads (beam splitter, RGB cascade, AS_BPM volume, "
                         "uneven batch)")
    ap.add_argument("--workload", default=None,
                    help="only run the named secondary workload (e.g. train_step_2048x8) and print "
                         "its JSON object (for profiling it alone)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 512 MiB device copy (the reference kernel that calibrates "
                         "rocprofv3 FETCH_SIZE/WRITE_SIZE into bytes) and exit")
    return ap.parse_args()


def calibrate_copy():
    import torch
    a = torch.empty(64 * 1024 * 1024, dtype=torch.complex64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print("copied 512 MiB x5")


def read_traffic(pass_name):
    """HBM bytes per launch of a pass from the committed rocprofv3 summary (or None)."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    with open(TRAFFIC_FILE) as f:
        for line in f:
            parts = line.split()
            if len(parts) >= 3 and parts[0] == "traffic" and parts[1] == pass_name:
                return int(parts[2])
    return None


KERNELS_FILE = os.path.join(ROOT, "profiles", "asprop_2048x2048_b8_kernels.txt")
PASS_KERNELS = {"x_fft": "k_fast_fft_stg<Radices<16, 16, 8>, 4, false, NoOp, false>",
                "y_fft_mul_ifft": "k_fast_ystg<Radices<16, 16, 8>, true>", "x_ifft": "k_fast_fft_stg<Radices<16, 16, 8>, 4, true, NoOp, false>"}


def read_profiled_ms(pass_name):
    """Average ms per call of a pass's kernel in the committed rocprofv3 kernel-trace summary
    of this benchmark: the steady-state column (the calls of the window in which this script
    times its passes) where the summary has one, else the average over all calls (or None)."""
    if not os.path.exists(KERNELS_FILE):
        return None
    cols = None
    with open(KERNELS_FILE) as f:
        for line in f:
            if line.startswith("#") or not line.strip():
                continue
            if cols is None:  # the header: calls avg_ms [steady_ms] share% kernel
                cols = line.split()
                continue
            parts = line.split(None, len(cols) - 1)
            if len(parts) == len(cols) and PASS_KERNELS[pass_name] in parts[-1]:
                col = "steady_ms" if "steady_ms" in cols and parts[cols.index("steady_ms")] != "-" else "avg_ms"
                return float(parts[cols.index(col)])
    return None


def make_inputs():
    import numpy as np
    from fluxoptics import inputs
    g = inputs.gaussian(N, N, DS, DS, W0)
    u = np.empty((BATCH, N, N), np.complex64)
    for b in range(BATCH):
        u[b] = g * np.exp(1j * inputs.uniform(2000 + b, (N, N), 0.0, 2 * math.pi))
    t = (0.01 * (inputs.normal(77, (BATCH, N, N)) + 1j * inputs.normal(77 + 7919, (BATCH, N, N))))
    return u, t.astype(np.complex64)


def timed(fn, reps, torch):
    """Average ms of fn() over reps, HIP events on torch's current stream (the stream the
    library launches on)."""
    s = torch.cuda.Event(enable_timing=True)
    e = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    s.record()
    for _ in range(reps):
        fn()
    e.record()
    torch.cuda.synchronize()
    return s.elapsed_time(e) / reps


PASS_NAMES = ("x_fft", "y_fft_mul_ifft", "x_ifft")


def pass_times_in_step(torch, lib, ptr, plan, u, g, reps):
    """Average ms of each pass of the step (fh_propagate_pass 0, 1, 2 = fh_propagate), timed
    by HIP events recorded between the passes of reps steps, forward and adjoint pooled."""
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(7)] for _ in range(reps)]
    torch.cuda.synchronize()
    for r in range(reps):
        e = ev[r]
        e[0].record()
        for d, f, base in ((1, u, 0), (-1, g, 3)):
            for pid in range(3):
                lib.call("fh_propagate_pass", plan, ptr(f.electric), d, pid)
                e[base + pid + 1].record()
    torch.cuda.synchronize()
    out = {}
    for pid, name in enumerate(PASS_NAMES):
        t = [e[pid].elapsed_time(e[pid + 1]) + e[pid + 3].elapsed_time(e[pid + 4]) for e in ev]
        out[name] = sum(t) / (2 * reps)
    return out


def cpu_baseline(u0, dv0, v_dev0, du_dev0, budget_s):
    """The reference's CPU algorithm (FFT2 -> cached-kernel multiply -> IFFT2, complex
    FP32, pocketfft threads = all host cores) timed on a bounded sample, plus the error
    of the device outputs against the FP64 oracle."""
    import numpy as np
    from oracle import fluxoptics_oracle as O
    cores = len(os.sched_getaffinity(0))
    prop = O.CpuASPropC64(N, N, DS, DS, LAM, Z, 1.0, workers=cores)
    x = u0[0].copy()
    prop.forward(x)
    prop.adjoint(x)  # warm-up (plan creation), excluded like FFTW planning
    times = []
    t_end = time.perf_counter() + budget_s
    while time.perf_counter() < t_end and len(times) < 200:
        t0 = time.perf_counter()
        prop.forward(x)
        prop.adjoint(x)
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    v_ref = O.asprop(u0[0].astype(np.complex128), (DS, DS), LAM, Z, 1.0, 1)[0]
    du_ref = O.asprop(dv0[0].astype(np.complex128), (DS, DS), LAM, Z, 1.0, -1)[0]
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
    return ({"value": 1.0 / med, "unit": "propagations/s", "cores": cores, "kind": "port",
             "sample": f"1 field of {N}x{N} c64, forward+adjoint, {len(times)} runs (median), "
                       f"engine scipy.fft pocketfft workers={cores}"},
            {"field_rel_l2": rel(v_dev0, v_ref), "adjoint_rel_l2": rel(du_dev0, du_ref),
             "oracle": "FP64 numpy restatement, field 0"})


def cpu_baseline_beam_splitter(budget_s):
    """The beam-splitter configuration on the CPU in Float64, the reference's own precision there:
    512^2 Gaussian (w 30 um), RSProp(1500) -> Phase -> RSProp(2000) -> Phase -> RSProp(1500) with the
    reference's 2n-1 padding, intensity loss against two Gaussians at +-60 um, both mask gradients;
    pocketfft threads on all host cores; median of at least 5 iterations."""
    import numpy as np
    from fluxoptics import inputs
    from oracle import fluxoptics_oracle as O
    cores = len(os.sched_getaffinity(0))
    n, ds, lam = 512, 1.0, 1.064
    u0 = inputs.normalize_power(inputs.gaussian(n, n, ds, ds, 30.0), ds, ds)
    t = inputs.normalize_power(inputs.gaussian(n, n, ds, ds, 30.0, -60.0, 0.0)
                               + inputs.gaussian(n, n, ds, ds, 30.0, 60.0, 0.0), ds, ds)
    cpu = O.CpuRSPipelineF64(u0, np.abs(t) ** 2, (ds, ds), lam, [1500.0, 2000.0, 1500.0], workers=cores)
    phis = [np.zeros((n, n)), np.zeros((n, n))]
    cpu.iteration(phis)  # warm-up (FFT plans), excluded as FFTW planning is
    times = []
    t_end = time.perf_counter() + budget_s
    while len(times) < 5 or (time.perf_counter() < t_end and len(times) < 50):
        t0 = time.perf_counter()
        cpu.iteration(phis)
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    return {"value": 1.0 / med, "unit": "iterations/s", "cores": cores, "kind": "port",
            "sample": f"beam splitter 512^2 (3 RSProp with 2n-1 padding, 2 Phases, loss + 2 mask gradients), "
                      f"complex FP64, {len(times)} iterations (median), engine scipy.fft pocketfft workers={cores}"}


def _timed_host(fn, reps, torch):
    """Average ms of fn() over reps (host-driven sequences; synchronised on both sides)."""
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / reps


def workload_beam_splitter(torch, fo, inputs):
    """512x512 Gaussian (w 30 um), ds 1, lambda 1.064: RSProp(1500) -> Phase -> RSProp(2000)
    -> Phase -> RSProp(1500), intensity loss against two Gaussians at +-60 um, both mask
    gradients.  One iteration = value + gradients."""
    import numpy as np
    n, ds, lam = 512, 1.0, 1.064
    g0 = inputs.normalize_power(inputs.gaussian(n, n, ds, ds, 30.0), ds, ds)
    t = inputs.normalize_power(inputs.gaussian(n, n, ds, ds, 30.0, -60.0, 0.0)
                               + inputs.gaussian(n, n, ds, ds, 30.0, 60.0, 0.0), ds, ds)
    u = fo.ScalarField(torch.from_numpy(g0.astype(np.complex64)).cuda(), (ds, ds), lam)
    d1 = fo.Phase(u, np.zeros((n, n), np.float32), trainable=True, buffered=True)
    d2 = fo.Phase(u, np.zeros((n, n), np.float32), trainable=True, buffered=True)
    p1, p2 = fo.RSProp(u, 1500.0), fo.RSProp(u, 2000.0)
    system = fo.ScalarSource(u) | p1 | d1 | p2 | d2 | p1
    metric = fo.SquaredIntensityDifference((u, (np.abs(t) ** 2).astype(np.float32)))
    system.withgradient(metric)
    ms = _timed_host(lambda: system.withgradient(metric), 20, torch)
    step = system.capture_withgradient(metric)
    gms = _timed_host(step.replay, 50, torch)
    return {"ms_per_iter": gms, "iters_per_s": 1e3 / gms, "eager_ms_per_iter": ms,
            "what": "512^2 RS->Phase->RS->Phase->RS, loss + 2 mask gradients, RS padded to 1024^2; "
                    "one recorded graph per iteration (eager host-driven launches beside it)"}


def _published_splitter(torch, fo, inputs):
    """The reference's published benchmark system (docs/src/api/index.md:79-119): 250x250, ds
    10 um, lambda 0.7 um, Gaussian w0 150 um normalised to unit power; target the normalised sum
    of four Gaussians at (+-3.8 w0, +-3.8 w0); three zero trainable Phases with RSProp(2e5)
    after each; PowerCoupling; Adam(0.1) for every mask."""
    import numpy as np
    n, ds, lam, w0 = 250, 10.0, 0.7, 150.0
    u0 = inputs.normalize_power(inputs.gaussian(n, n, ds, ds, w0), ds, ds).astype(np.complex64)
    off = 3.8 * w0
    vf = sum(inputs.gaussian(n, n, ds, ds, w0, sx * off, sy * off) for sx in (-1, 1) for sy in (-1, 1))
    vf = inputs.normalize_power(vf, ds, ds).astype(np.complex64)
    u = fo.ScalarField(torch.from_numpy(u0).cuda(), (ds, ds), lam)
    v = fo.ScalarField(torch.from_numpy(vf).cuda(), (ds, ds), lam)
    does = [fo.Phase(u, np.zeros((n, n), np.float32), trainable=True, buffered=True) for _ in range(3)]
    prop = fo.RSProp(u, 2e5)
    system = fo.ScalarSource(u) | does[0] | prop | does[1] | prop | does[2] | prop
    metric = fo.PowerCoupling(v)
    return system, metric, does


def workload_published_splitter(torch, fo, inputs):
    """The reference's own headline benchmark (docs/src/api/index.md:79-190): 200 iterations of
    withgradient(sum(1 - PowerCoupling)) + update! with Adam(0.1) on the 250^2 three-mask RS beam
    splitter; published ~0.27 s on an RTX 4070 Super (GPU) and ~5 s on a CPU.  Timed from zeroed
    masks and fresh optimiser state, host-driven like the reference's loop."""
    import warnings
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        system, metric, does = _published_splitter(torch, fo, inputs)
    opt = fo.setup({}, system, fo.Adam(0.1))
    val, grads = system.withgradient(metric, weights=-1.0)   # warm-up (plans, tables)
    fo.update_(opt, system, grads)
    for d in does:
        d.phi.zero_()
    opt = fo.setup({}, system, fo.Adam(0.1))
    losses = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(200):
        val, grads = system.withgradient(metric, weights=-1.0)
        fo.update_(opt, system, grads)
        losses.append(val)
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    first, last = 1.0 + float(losses[0]), 1.0 + float(losses[-1])
    # the same 200 iterations, each one replay of a recorded iteration (gradient + Adam step)
    for d in does:
        d.phi.zero_()
    opt = fo.setup(fo.Adam(0.1), system)
    step = system.capture_train_step(metric, opt, weights=-1.0)
    glosses = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(200):
        val, _ = step.replay()
        glosses.append(val)
    torch.cuda.synchronize()
    gsec = time.perf_counter() - t0
    return {"seconds_200_iterations": sec, "ms_per_iter": sec * 5.0, "loss_first": first, "loss_last": last,
            "graph_seconds_200_iterations": gsec, "graph_loss_first": 1.0 + float(glosses[0]),
            "graph_loss_last": 1.0 + float(glosses[-1]),
            "published_gpu_seconds": 0.27, "published_hardware": "NVIDIA RTX 4070 Super",
            "speedup_vs_published": 0.27 / sec, "graph_speedup_vs_published": 0.27 / gsec,
            "what": "250^2 Phase->RSProp(2e5) x3, PowerCoupling loss sum(1-p), Adam(0.1), 200 iterations of "
                    "withgradient + update_ (RS padded to 512^2), host-driven as the reference's loop, and "
                    "as 200 replays of one recorded iteration"}


def workload_published_rgb(torch, fo, inputs):
    """The reference's RGB DOE notebook (notebooks/03_RGB_beam_shaping.ipynb, timed at its cell
    27: 200 iterations in 6.196 s on an unrecorded CUDA GPU): 800x800 (600x600 images padded by
    100), ds 0.5 um, lambda {0.640, 0.538, 0.455} as one 3-field batch with Gaussians (w0 40 um,
    red x2); RSProp(3000) -> TeaDOE -> ... (5 RSProps, 4 trainable TeaDOE with dn(lambda) =
    0.508/0.512/0.518); SquaredIntensityDifference against per-wavelength targets; ProxRule(Fista(200),
    ClampProx(-0.5, 0.5)).  The notebook's target images are replaced by synthetic intensity
    patterns of the same shape (normalised to unit sum, the head pattern x4)."""
    import warnings
    import numpy as np
    n, ds, lams = 800, 0.5, [0.640, 0.538, 0.455]
    dn_tab = {0.455: 1.518 - 1, 0.538: 1.512 - 1, 0.640: 1.508 - 1}
    dn = lambda lam: dn_tab[round(lam, 3)]
    g = inputs.gaussian(n, n, ds, ds, 40.0)
    u0 = np.stack([2 * g, g, g]).astype(np.complex64)
    x = inputs.grid(n, ds)
    r = np.hypot(x[None, :], x[:, None])
    head = np.exp(-((r - 120.0) / 12.0) ** 2)
    eyes = np.exp(-(np.hypot(x[None, :] - 45, x[:, None] - 40) / 15.0) ** 2) + \
        np.exp(-(np.hypot(x[None, :] + 45, x[:, None] - 40) / 15.0) ** 2)
    mouth = np.exp(-((r - 70.0) / 8.0) ** 2) * (x[:, None] < -20)
    tg = np.stack([4 * head / head.sum(), eyes / eyes.sum(), mouth / mouth.sum()]).astype(np.float32)
    u = fo.ScalarField(torch.from_numpy(u0).cuda(), (ds, ds), lams)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        p = fo.RSProp(u, 3000.0)
    does = [fo.TeaDOE(u, dn, trainable=True, buffered=True) for _ in range(4)]
    system = fo.ScalarSource(u) | p | does[0] | p | does[1] | p | does[2] | p | does[3] | p
    metric = fo.SquaredIntensityDifference((u, tg))
    for d in does:
        d.h.zero_()
    opt = fo.setup(fo.ProxRule(fo.Fista(200), fo.ClampProx(-0.5, 0.5)), system)
    _, grads = system.withgradient(metric)   # warm-up, as the notebook's cell 25
    fo.update_(opt, system, grads)
    losses = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(200):
        val, grads = system.withgradient(metric)
        fo.update_(opt, system, grads)
        losses.append(val)
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    # the same warm-up and 200 iterations, each one replay of a recorded iteration (gradient +
    # Fista step, its coefficients in device memory)
    for d in does:
        d.h.zero_()
    opt = fo.setup(fo.ProxRule(fo.Fista(200), fo.ClampProx(-0.5, 0.5)), system)
    _, grads = system.withgradient(metric)
    fo.update_(opt, system, grads)
    step = system.capture_train_step(metric, opt)
    glosses = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(200):
        val, _ = step.replay()
        glosses.append(val)
    torch.cuda.synchronize()
    gsec = time.perf_counter() - t0
    return {"seconds_200_iterations": sec, "ms_per_iter": sec * 5.0, "loss_first": float(losses[0]),
            "loss_last": float(losses[-1]), "graph_seconds_200_iterations": gsec,
            "graph_loss_first": float(glosses[0]), "graph_loss_last": float(glosses[-1]),
            "published_gpu_seconds": 6.196, "published_hardware": "CUDA GPU, model not recorded",
            "speedup_vs_published": 6.196 / sec, "graph_speedup_vs_published": 6.196 / gsec,
            "what": "800^2 x 3 lambda, 5 RSProp(3000) + 4 TeaDOE, intensity loss, Fista(200) + clamp, 200 "
                    "iterations of withgradient + update_, host-driven (RS padded to 2048^2), and as 200 replays "
                    "of one recorded iteration; synthetic targets"}


def tomography_setup(torch, fo, inputs):
    """The reference's waveguide-tomography notebook (notebooks/04_waveguide_tomography.ipynb):
    43 tilted 1-D fields of 1600 points (dx 0.0405 um), lambda 0.455 um, a 20 um sample of bulk
    index 1.52 in 300 slices (AS_BPM, trainable dn from zero), then ASProp(-6 um; n0 1.52) with the
    NA 1.4 diaphragm as its frequency filter; SquaredIntensityDifference against the measured
    intensity; ProxRule(Fista(0.01), IstaProx(3e-5)) on dn.  The measured intensity (a file this
    benchmark does not load) is replaced by the intensity the same system gives for a synthetic
    waveguide dn, normalised per angle to a mean of 1 as the notebook does (cell 7); the angles
    are synthetic, 43 of 49 spread over +-0.45 rad.  The notebook's trainable aberration
    (a Fourier-domain phase on a polynomial basis) is not on the path and is left out."""
    import numpy as np
    nx, dx, lam, nb, na, nz, L = 1600, 0.0405, 0.455, 1.52, 1.4, 300, 20.0
    th = list(np.linspace(-0.45, 0.45, 49)[5:48])
    b = len(th)
    dev = torch.device("cuda", 0)
    u0 = fo.ScalarField(torch.ones((b, nx), dtype=torch.complex64, device=dev), (dx,), lam, tilts=(th,))
    diaphragm = lambda fx: np.abs(fx) <= abs(na / lam)     # make_diaphragm(NA, lambda)
    x = (np.arange(nx) - (nx - 1) / 2) * dx
    zc = (np.arange(nz) + 0.5) * (L / nz)
    guide = 0.02 * np.exp(-((x[None, :] - 1.5 * np.sin(2 * np.pi * zc[:, None] / L)) / 1.2) ** 2)
    truth = fo.AS_BPM(u0, L, nb, guide.astype(np.float32))
    p = fo.ASProp(u0, -6.0, n0=nb, filter=diaphragm)  # a 1-D filter takes fx only
    meas = p.propagate(truth.propagate(u0))
    I = (meas.electric.abs() ** 2)
    I = (I * (nx / I.sum(dim=1, keepdim=True))).to(torch.float32).contiguous()
    sample = fo.AS_BPM(u0, L, nb, np.zeros((nz, nx), np.float32), trainable=True, buffered=True)
    system = fo.ScalarSource(u0) | sample | p
    metric = fo.SquaredIntensityDifference((u0, I))
    return system, sample, metric


def workload_published_tomography(torch, fo, inputs):
    """50 iterations of the tomography loop (withgradient + update_), host-driven as the notebook's
    cell 29 (published: 2.481 s on an unrecorded CUDA GPU, with the aberration basis), and as
    replays of one recorded iteration."""
    res = {}
    for mode in ("host", "graph"):
        system, sample, metric = tomography_setup(torch, fo, inputs)
        opt = fo.setup(fo.ProxRule(fo.Fista(0.01), fo.IstaProx(3e-5)), system)
        _, grads = system.withgradient(metric)   # warm-up, as the notebook's cell 27
        fo.update_(opt, system, grads)
        step = system.capture_train_step(metric, opt) if mode == "graph" else None
        losses = []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(50):
            if step is None:
                val, grads = system.withgradient(metric)
                fo.update_(opt, system, grads)
            else:
                val, _ = step.replay()
            losses.append(val)
        torch.cuda.synchronize()
        res[mode] = (time.perf_counter() - t0, float(losses[0]), float(losses[-1]))
    (sec, l0, l1), (gsec, g0, g1) = res["host"], res["graph"]
    return {"seconds_50_iterations": sec, "ms_per_iter": sec * 20.0, "loss_first": l0, "loss_last": l1,
            "graph_seconds_50_iterations": gsec, "graph_loss_first": g0, "graph_loss_last": g1,
            "published_gpu_seconds": 2.481, "published_hardware": "CUDA GPU, model not recorded",
            "speedup_vs_published": 2.481 / sec, "graph_speedup_vs_published": 2.481 / gsec,
            "what": "43 tilted 1-D fields x 1600, AS_BPM 300 slices (trainable dn, Fista + Ista prox) -> NA-filtered "
                    "ASProp(-6), intensity loss, 50 iterations; synthetic measured intensity and angles; the "
                    "notebook's aberration basis left out"}


def workload_train_step(torch, fo, inputs):
    """The headline grid as a training step: 8 fields of 2048x2048 (ds 1, lambda 0.8),
    Phase (trainable, shared) -> ASProp(5000), per-field intensity targets; value + mask
    gradient per iteration, the fused tail block (fh_sqint_block) against the same system run
    component by component, each as one recorded graph."""
    import numpy as np
    n, b, ds, lam, z = 2048, 8, 1.0, 0.8, 5000.0
    g = inputs.gaussian(n, n, ds, ds, 200.0)
    u0 = np.stack([g * np.exp(1j * inputs.uniform(171 + k, (n, n), 0, 0.5)) for k in range(b)]).astype(np.complex64)
    tgt = np.stack([np.abs(inputs.gaussian(n, n, ds, ds, 200.0, 40.0 * k, 0.0)) ** 2 for k in range(b)])
    u = fo.ScalarField(torch.from_numpy(u0).cuda(), (ds, ds), lam)
    ph = fo.Phase(u, inputs.uniform(170, (n, n), 0, 2 * math.pi).astype(np.float32), trainable=True,
                  buffered=True)
    system = fo.ScalarSource(u) | ph | fo.ASProp(u, z)
    metric = fo.SquaredIntensityDifference((u, tgt.astype(np.float32)))
    out = {}
    for fuse in (True, False):
        system.withgradient(metric, fuse=fuse)
        w = system._withgradient_launches  # record the chosen path as one graph
        system.withgradient(metric, fuse=fuse)
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            w(metric, None, None, fuse)
        out["fused" if fuse else "unfused"] = _timed_host(graph.replay, 30, torch)
    return {"ms_per_iter": out["fused"], "iters_per_s": 1e3 / out["fused"], "unfused_ms_per_iter": out["unfused"],
            "what": "8 x 2048^2 Phase->ASProp, loss + mask gradient; fused tail (5 passes + phase-gradient "
                    "epilogue) vs component by component (3 + metric + 3 passes), each one recorded graph"}


def workload_spectral(torch, fo, inputs):
    """The headline grid with a wavelength per field: 8 x 2048x2048 c64, lambda 0.45..0.80 um
    (one transfer function per field, read by the y pass for every field), ds 0.5 um, z 5000 um;
    one step = forward of the batch + adjoint of a cotangent batch, as the headline."""
    import numpy as np
    from fluxoptics import _lib
    from fluxoptics.components import _ptr
    lams = [float(x) for x in np.linspace(0.45, 0.80, BATCH)]
    u_np, t_np = make_inputs()
    u = fo.ScalarField(torch.from_numpy(u_np).cuda(), (DS, DS), lams)
    g = fo.ScalarField(torch.from_numpy(t_np).cuda(), (DS, DS), lams)
    prop = fo.ASProp(u, Z)
    plan = prop.plan(u)

    def step():
        prop.propagate_(u, fo.Forward)
        prop.propagate_(g, fo.Backward)

    for _ in range(3):
        step()
    ms = timed(step, 20, torch)
    passes = pass_times_in_step(torch, _lib, _ptr, plan, u, g, 10)
    return {"units_per_s": BATCH / (ms * 1e-3), "ms_per_step": ms,
            "pass_ms": {k: round(v, 4) for k, v in passes.items()},
            "y_pass_GBps": 16.0 * N * N * BATCH / (passes["y_fft_mul_ifft"] * 1e-3) / 1e9,
            "what": "8 x 2048^2 ASProp fwd+adj with 8 wavelengths (a transfer function per field)"}


def workload_rgb_cascade(torch, fo, inputs):
    """1024x1024 x 3 wavelengths, Phase -> ASProp(1e4) -> Phase -> ASProp(1e4) -> Phase,
    intensity loss, three mask gradients."""
    import numpy as np
    n, ds, lams, z = 1024, 1.0, [0.640, 0.538, 0.455], 1e4
    g = inputs.gaussian(n, n, ds, ds, 100.0)
    u0 = inputs.normalize_power(np.stack([g, g, g]), ds, ds).astype(np.complex64)
    tgt = np.stack([np.abs(inputs.normalize_power(inputs.gaussian(n, n, ds, ds, 100.0, c, 0.0), ds, ds)) ** 2
                    for c in (-200.0, 0.0, 200.0)]).astype(np.float32)
    u = fo.ScalarField(torch.from_numpy(u0).cuda(), (ds, ds), lams)
    does = [fo.Phase(u, inputs.uniform(31 + k, (n, n), 0, 2 * math.pi).astype(np.float32),
                     trainable=True, buffered=True) for k in range(3)]
    prop = fo.ASProp(u, z)
    system = fo.ScalarSource(u) | does[0] | prop | does[1] | prop | does[2]
    metric = fo.SquaredIntensityDifference((u, tgt))
    system.withgradient(metric)
    ms = _timed_host(lambda: system.withgradient(metric), 20, torch)
    step = system.capture_withgradient(metric)
    gms = _timed_host(step.replay, 50, torch)
    return {"ms_per_iter": gms, "iters_per_s": 1e3 / gms, "eager_ms_per_iter": ms,
            "what": "1024^2 x 3 lambda Phase/ASProp cascade, loss + 3 mask gradients; one recorded graph "
                    "per iteration (eager host-driven launches beside it)"}


def workload_bpm(torch, fo, inputs):
    """AS_BPM 512x512, ds 0.5, lambda 1.0, n 1.5, 400 slices of 1 um, checkpoint every 20:
    forward with checkpoints + adjoint with the dn gradient, and the peak device footprint."""
    import numpy as np
    n, ns, ds, lam = 512, 400, 0.5, 1.0
    dn = inputs.smooth_noise(ns, n, n, 3.0, 0.01, 4)  # Gaussian-filtered noise, sigma 3 px, peak 0.01
    x = inputs.grid(n, ds)[None, :]
    u0 = (inputs.gaussian(n, n, ds, ds, 80.0) * np.exp(2j * math.pi * math.sin(math.radians(2.0)) * x / lam))
    u = fo.ScalarField(torch.from_numpy(u0.astype(np.complex64)).cuda(), (ds, ds), lam)
    bpm = fo.AS_BPM(u, 400.0, 1.5, torch.from_numpy(dn).cuda(), trainable=True, buffered=True,
                    checkpoint_every=20)
    w = u.copy()

    def it():
        w.electric.copy_(u.electric)
        bpm.propagate_and_save_(w)
        bpm.backpropagate_with_gradient_(w)

    it()
    ms = _timed_host(it, 3, torch)
    # the same iteration recorded once as a HIP graph (about 2400 slice launches per replay)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        it()
    graph.replay()
    gms = _timed_host(graph.replay, 3, torch)
    ws = bpm.workspace_bytes(u)
    caller = 2 * dn.nbytes + 2 * u0.size * 8
    return {"ms_per_iter": ms, "slices_per_s": 2 * ns * 1e3 / ms, "graph_ms_per_iter": gms,
            "library_workspace_bytes": ws, "peak_device_bytes": ws + caller,
            "reference_saved_stack_bytes": ns * n * n * 8,
            "what": "512^2 x 400 slices, checkpoint every 20, forward + dn adjoint "
                    "(dn Gaussian-filtered white noise, sigma 3 px, peak 0.01); host-driven and as one recorded graph"}


def workload_uneven_batch(torch, fo, inputs):
    """64 fields, 16 each of 256^2, 512^2, 1024^2, 2048^2, random LG modes, ds 1, lambda
    1.064, RSProp (2n padding) with z ~ U[500, 5000]; forward + adjoint of all fields as one
    graph per direction, against 64 separate per-field launches."""
    import warnings
    import numpy as np
    sizes, per, ds, lam = (256, 512, 1024, 2048), 16, 1.0, 1.064
    pairs, singles = [], []
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        for k, n in enumerate(sizes):
            ps = inputs.uniform(500 + 10 * k, (per,), 0, 4).astype(int)
            ls = inputs.uniform(501 + 10 * k, (per,), -4, 5).astype(int)
            zs = list(inputs.uniform(505 + 10 * k, (per,), 500.0, 5000.0))
            u0 = np.stack([inputs.laguerre_gaussian(n, n, ds, ds, 0.1 * n * ds, int(p), int(l))
                           for p, l in zip(ps, ls)]).astype(np.complex64)
            u = fo.ScalarField(torch.from_numpy(u0).cuda(), (ds, ds), lam)
            pairs.append((fo.RSProp(u, zs), u))
            for b in range(per):
                ub = fo.ScalarField(u.electric[b].clone(), (ds, ds), lam)
                singles.append((fo.RSProp(ub, zs[b]), ub))
    g = fo.PropagationGroup(pairs, autotune_reps=3)

    def graph_step():
        g.propagate_(fo.Forward)
        g.propagate_(fo.Backward)

    def separate_step():
        for p, ub in singles:
            p.propagate_(ub, fo.Forward)
        for p, ub in singles:
            p.propagate_(ub, fo.Backward)

    graph_step()
    separate_step()
    gms = _timed_host(graph_step, 10, torch)
    sms = _timed_host(separate_step, 5, torch)
    alg = sum(32.0 * n * n * per for n in sizes)
    return {"graph_ms_per_step": gms, "separate_ms_per_step": sms, "speedup": sms / gms,
            "units_per_s": len(singles) * 1e3 / gms, "algorithmic_bytes_per_step": alg,
            "hbm_roofline_frac": alg / (gms * 1e-3) / HBM_PEAK,
            "what": "64 RS fields (16 each 256^2..2048^2, 2n padding, z~U[500,5000]), fwd+adj as 2 graph "
                    "launches vs 128 separate per-field propagations"}


def main():
    args = parse()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "the batch would shard across GPUs with no communication, which is not measured here.",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    import fluxoptics as fo
    from fluxoptics import _lib
    from fluxoptics.components import _ptr

    if args.workload:
        from fluxoptics import inputs as inp
        fns = {"published_splitter_250": workload_published_splitter, "published_rgb_800x3": workload_published_rgb, "beam_splitter_512": workload_beam_splitter, "train_step_2048x8": workload_train_step,
               "rgb_cascade_1024x3": workload_rgb_cascade, "as_bpm_512x400": workload_bpm,
               "uneven_batch_64": workload_uneven_batch, "published_tomography_1600x43": workload_published_tomography,
               "spectral_2048x8": workload_spectral}
        if args.workload not in fns:
            sys.exit(f"bench.py: unknown workload {args.workload!r} (one of {', '.join(fns)})")
        torch.cuda.set_device(0)
        print(json.dumps({args.workload: fns[args.workload](torch, fo, inp)}))
        return

    torch.cuda.set_device(0)
    dev = torch.device("cuda", 0)
    u_np, t_np = make_inputs()
    u = fo.ScalarField(torch.from_numpy(u_np).to(dev), (DS, DS), LAM)
    prop = fo.ASProp(u, Z)
    plan = prop.plan(u)

    # correctness sample (device outputs of field 0 before timing)
    v_dev = prop.propagate(u, fo.Forward)
    v0 = v_dev.electric[0].cpu().numpy().astype(np.complex128)
    g = fo.ScalarField(v_dev.electric - torch.from_numpy(t_np).to(dev), (DS, DS), LAM)
    dv0 = g.electric.cpu().numpy()
    gd = prop.propagate(g, fo.Backward)
    du0 = gd.electric[0].cpu().numpy().astype(np.complex128)
    del v_dev, gd

    def step():
        prop.propagate_(u, fo.Forward)
        prop.propagate_(g, fo.Backward)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ms = timed(step, args.steps, torch)
    wall = time.perf_counter() - t0
    units_per_s = BATCH / (ms * 1e-3)
    unit_bytes = 32.0 * N * N

    # per-kernel averages (each pass over the whole batch) inside the step's own sequence:
    # the forward's three passes on u, then the adjoint's on g, with a HIP event between
    # passes on the library's stream (torch's current stream), averaged over both directions
    passes = pass_times_in_step(torch, _lib, _ptr, plan, u, g, max(10, args.steps))
    dom = max(passes, key=passes.get)
    dom_bytes = 16.0 * N * N * BATCH     # read + write of the batch once
    achieved = dom_bytes / (passes[dom] * 1e-3)
    traffic = read_traffic(dom)

    out = {
        "metric": METRIC, "value": units_per_s, "unit": "propagations/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": f"ASProp forward+adjoint, batch {BATCH} x {N}x{N} complex-FP32, "
                               f"ds={DS}um, lambda={LAM}um, z={Z}um (one step = {BATCH} units)",
                   "global_batch": BATCH, "grid": N, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic, "kernel": dom,
                     "kernel_ms": passes[dom], "bytes_per_launch": dom_bytes,
                     "profiled_kernel_ms": read_profiled_ms(dom),
                     "profiled_frac": (None if read_profiled_ms(dom) is None
                                       else dom_bytes / (read_profiled_ms(dom) * 1e-3) / HBM_PEAK),
                     "profile": os.path.relpath(KERNELS_FILE, ROOT)},
        "path_roofline": {"achieved_GBps": units_per_s * unit_bytes / 1e9,
                          "frac": units_per_s * unit_bytes / HBM_PEAK,
                          "bytes_per_unit": unit_bytes},
        "pass_ms": passes, "wall_s_timed": wall,
    }

    if not args.no_plain:
        # plain GPU form: complex FP64, per field, unfused
        u128 = torch.from_numpy(u_np.astype(np.complex128)).to(dev)
        g128 = u128.clone()

        def plain_step():
            prop.propagate_plain_fp64_(u128, LAM, fo.Forward)
            prop.propagate_plain_fp64_(g128, LAM, fo.Backward)

        plain_step()
        pms = timed(plain_step, max(3, args.steps // 10), torch)
        out["plain_fp64"] = {"value": BATCH / (pms * 1e-3), "unit": "propagations/s",
                             "ms_per_step": pms, "speedup_of_fused": (BATCH / (ms * 1e-3)) / (BATCH / (pms * 1e-3))}
        del u128, g128

    if not args.no_workloads:
        from fluxoptics import inputs as inp
        out["workloads"] = {
            "published_splitter_250": workload_published_splitter(torch, fo, inp),
            "published_rgb_800x3": workload_published_rgb(torch, fo, inp),
            "beam_splitter_512": workload_beam_splitter(torch, fo, inp),
            "train_step_2048x8": workload_train_step(torch, fo, inp),
            "rgb_cascade_1024x3": workload_rgb_cascade(torch, fo, inp),
            "as_bpm_512x400": workload_bpm(torch, fo, inp),
            "uneven_batch_64": workload_uneven_batch(torch, fo, inp),
            "published_tomography_1600x43": workload_published_tomography(torch, fo, inp),
            "spectral_2048x8": workload_spectral(torch, fo, inp),
        }
        torch.cuda.empty_cache()

    if not args.no_cpu_baseline:
        cb, err = cpu_baseline(u_np, dv0, v0, du0, args.cpu_seconds)
        out["cpu_baseline"] = cb
        out["error_vs_fp64"] = err
        out["cpu_baseline_beam_splitter_fp64"] = cpu_baseline_beam_splitter(min(args.cpu_seconds, 8.0))

    print(json.dumps(out))


if __name__ == "__main__":
    main()
