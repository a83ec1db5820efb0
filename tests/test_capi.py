"""C-ABI checks that need no GPU: the library loads, exports every declared symbol,
and rejects bad arguments before any device work."""

import ctypes
import os
import re

import pytest

from conftest import ROOT
from fluxoptics import _lib

HEADER = os.path.join(ROOT, "include", "fluxhip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fh_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_functions()
    assert len(names) >= 15
    for name in names:
        assert hasattr(lib, name), f"{name} declared in fluxhip.h but not exported"
        assert name in _lib.SIGNATURES, f"{name} has no ctypes signature"
    assert set(_lib.SIGNATURES) == set(names)


def test_host_only_entry_points():
    lib = _lib.load()
    assert lib.fh_version() >= 100
    assert lib.fh_max_length() == 4096
    assert lib.fh_last_error() is not None


def test_argument_validation_before_device_work():
    lib = _lib.load()
    out = ctypes.c_void_p()
    lam = (ctypes.c_double * 1)(0.8)
    # null context
    rc = lib.fh_asprop_create(None, 64, 64, 1, 1.0, 1.0, lam, 1, 10.0, 1.0, ctypes.byref(out))
    assert rc == -1 and b"context" in lib.fh_last_error()
    rc = lib.fh_rsprop_create(None, 64, 64, 1, 1.0, 1.0, lam, 1, 10.0, ctypes.byref(out))
    assert rc == -1
    # null plan / field
    assert lib.fh_propagate(None, None, 1) == -1
    assert lib.fh_phase_apply(None, None, None, 4, 4, 1, 1, None) == -1
    assert lib.fh_propagate_plain_fp64(None, None, 1) == -1
    assert lib.fh_ctx_create(0, None, None) == -1
    with pytest.raises(_lib.FluxHipError):
        _lib.call("fh_propagate", None, None, 1)


def test_no_cpu_fallback_in_host_layer():
    """The host layer refuses CPU tensors instead of computing on the CPU."""
    import torch
    import fluxoptics as fo
    u = fo.ScalarField(torch.zeros((16, 16), dtype=torch.complex64), (1.0, 1.0), 0.8)
    p = fo.ASProp(u, 10.0)
    with pytest.raises(TypeError):
        p.propagate_(u)


def test_prop_desc_layout_matches_header(tmp_path):
    """The ctypes mirror of fh_prop_desc has the C compiler's size and field offsets."""
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("no host C compiler")
    names = [f[0] for f in _lib.PropDesc._fields_]
    src = tmp_path / "layout.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "fluxhip.h"\nint main(void) {\n'
                   '  printf("%zu", sizeof(fh_prop_desc));\n'
                   + "".join(f'  printf(" %zu", offsetof(fh_prop_desc, {n}));\n' for n in names)
                   + "  return 0;\n}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = [ctypes.sizeof(_lib.PropDesc)] + [getattr(_lib.PropDesc, n).offset for n in names]
    assert got == want


def _desc(nx=64, ny=64, batch=1, flags=0, z=None):
    d = _lib.PropDesc()
    d.nx, d.ny, d.batch, d.dx, d.dy = nx, ny, batch, 1.0, 1.0
    d._lam = (ctypes.c_double * 1)(0.8)
    d.lambdas, d.nlambdas = ctypes.cast(d._lam, _lib.c_double_p), 1
    if z is not None:
        d._z = (ctypes.c_double * 1)(z)
        d.z, d.nz = ctypes.cast(d._z, _lib.c_double_p), 1
    d.n0, d.flags = 1.0, flags
    return d


def test_new_entry_points_validate_arguments():
    """Collins, CZT, general BPM, fused metric block and offset_tilts reject bad arguments
    (with a message) before any device work."""
    lib = _lib.load()
    out = ctypes.c_void_p()
    ref = ctypes.byref
    # CollinsProp: B = 0, distances given, bad output sampling, unknown flags, no context
    assert lib.fh_collins_create(None, ref(_desc()), 1.0, 1.0, 1.0, 0.0, 1.0, ref(out)) == -1
    assert b"B != 0" in lib.fh_last_error()
    assert lib.fh_collins_create(None, ref(_desc(z=10.0)), 1.0, 1.0, 1.0, 100.0, 1.0, ref(out)) == -1
    assert lib.fh_collins_create(None, ref(_desc()), -1.0, 1.0, 1.0, 100.0, 1.0, ref(out)) == -1
    assert lib.fh_collins_create(None, ref(_desc(flags=1)), 1.0, 1.0, 1.0, 100.0, 1.0, ref(out)) == -1
    assert lib.fh_collins_create(None, ref(_desc()), 1.0, 1.0, 1.0, 100.0, 1.0, ref(out)) == -1
    assert b"context" in lib.fh_last_error()
    # CZT: dims outside {1, 2, 3}, non-positive magnitudes
    a = (ctypes.c_double * 4)(1.0, 0.0, 1.0, 0.0)
    w = (ctypes.c_double * 4)(1.0, -0.1, 1.0, -0.1)
    bad = (ctypes.c_double * 4)(0.0, 0.0, 1.0, 0.0)
    assert lib.fh_czt_create(None, 64, 64, 1, 0, a, w, 0, ref(out)) == -1
    assert b"dims" in lib.fh_last_error()
    assert lib.fh_czt_create(None, 64, 64, 1, 3, bad, w, 0, ref(out)) == -1
    assert b"magnitudes" in lib.fh_last_error()
    # general BPM: distances given, too few slices, shift with paraxial
    assert lib.fh_bpm_create_ex(None, ref(_desc(z=1.0)), 10.0, 4, 1, ref(out)) == -1
    assert lib.fh_bpm_create_ex(None, ref(_desc()), 10.0, 1, 1, ref(out)) == -1
    assert lib.fh_bpm_create_ex(None, ref(_desc(flags=4 | 1)), 10.0, 4, 1, ref(out)) == -1
    assert lib.fh_asprop_create_ex(None, ref(_desc(flags=4 | 1, z=1.0)), ref(out)) == -1
    # fused metric block and offset_tilts: null arguments
    assert lib.fh_sqint_block(None, None, None, 1, None, None, 1, None, None, None, 0) == -1
    assert lib.fh_sqint_block_src(None, None, None, None, 1, None, None, 1, None, None, 0, 1) == -1
    assert b"source" in lib.fh_last_error()
    assert lib.fh_window(None, None, 4, 4, None, 4, 4, 1, 0, 0, 0.0, 0.0) == -1
    assert b"destination" in lib.fh_last_error()
    lam1 = (ctypes.c_double * 1)(0.8)
    assert lib.fh_as_z_gradient(None, None, None, 4, 4, 1, 1.0, 1.0, lam1, 1, 1.0, 10.0, None) == -1
    assert b"spectra" in lib.fh_last_error()
    lam = (ctypes.c_double * 1)(0.8)
    t0 = (ctypes.c_double * 1)(0.0)
    assert lib.fh_offset_tilts(None, None, 4, 4, 1, 1.0, 1.0, lam, 1, t0, t0, t0, t0, 1) == -1
