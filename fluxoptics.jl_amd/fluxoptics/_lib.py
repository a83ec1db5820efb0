"""ctypes binding of libfluxhip.so (the C ABI declared in include/fluxhip.h).

The library is the only compute path: if it is missing or fails to load, every call
raises; there is no CPU fallback.
"""

from __future__ import annotations

import ctypes
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfluxhip.so")

FH_OK = 0
ERRORS = {-1: "invalid argument", -2: "HIP runtime error", -3: "out of device memory",
          -4: "unsupported"}

c_int, c_double, c_void_p, c_char_p = ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_char_p
c_float_p = ctypes.POINTER(ctypes.c_float)
c_double_p = ctypes.POINTER(ctypes.c_double)

FH_PROP_PARAXIAL = 1
FH_PROP_TRACK_TILTS = 2
FH_PROP_SHIFT = 4


class PropDesc(ctypes.Structure):
    """``fh_prop_desc`` of include/fluxhip.h."""
    _fields_ = [("nx", c_int), ("ny", c_int), ("batch", c_int), ("dx", c_double), ("dy", c_double),
                ("lambdas", c_double_p), ("nlambdas", c_int), ("z", c_double_p), ("nz", c_int),
                ("tilts_x", c_double_p), ("tilts_y", c_double_p), ("ntilts", c_int), ("n0", c_double),
                ("flags", c_int)]


# name -> (restype, argtypes); mirrors include/fluxhip.h one to one
SIGNATURES = {
    "fh_version": (c_int, []),
    "fh_last_error": (c_char_p, []),
    "fh_max_length": (c_int, []),
    "fh_ctx_create": (c_int, [c_int, c_void_p, ctypes.POINTER(c_void_p)]),
    "fh_ctx_set_stream": (c_int, [c_void_p, c_void_p]),
    "fh_ctx_synchronize": (c_int, [c_void_p]),
    "fh_ctx_destroy": (c_int, [c_void_p]),
    "fh_asprop_create": (c_int, [c_void_p, c_int, c_int, c_int, c_double, c_double, c_double_p,
                                 c_int, c_double, c_double, ctypes.POINTER(c_void_p)]),
    "fh_rsprop_create": (c_int, [c_void_p, c_int, c_int, c_int, c_double, c_double, c_double_p,
                                 c_int, c_double, ctypes.POINTER(c_void_p)]),
    "fh_asprop_create_ex": (c_int, [c_void_p, ctypes.POINTER(PropDesc), ctypes.POINTER(c_void_p)]),
    "fh_asprop_create_filtered": (c_int, [c_void_p, ctypes.POINTER(PropDesc), c_void_p, c_int,
                                          ctypes.POINTER(c_void_p)]),
    "fh_rsprop_create_ex": (c_int, [c_void_p, ctypes.POINTER(PropDesc), ctypes.POINTER(c_void_p)]),
    "fh_group_create": (c_int, [c_void_p, c_int, ctypes.POINTER(c_void_p), c_int, ctypes.POINTER(c_void_p)]),
    "fh_group_propagate": (c_int, [c_void_p, ctypes.POINTER(c_void_p), c_int]),
    "fh_group_graph_count": (c_int, [c_void_p, ctypes.POINTER(c_int)]),
    "fh_group_destroy": (c_int, [c_void_p]),
    "fh_plan_destroy": (c_int, [c_void_p]),
    "fh_propagate": (c_int, [c_void_p, c_void_p, c_int]),
    "fh_propagate_pass": (c_int, [c_void_p, c_void_p, c_int, c_int]),
    "fh_plan_variants": (c_int, [c_void_p, c_int, ctypes.POINTER(c_int), ctypes.POINTER(c_int)]),
    "fh_plan_set_variant": (c_int, [c_void_p, c_int, c_int]),
    "fh_autotune": (c_int, [c_void_p, c_void_p, c_int, c_char_p, c_int]),
    "fh_phase_propagate": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int]),
    "fh_phase_apply": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_void_p]),
    "fh_phase_backward": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int,
                                  c_int, c_int, c_int]),
    "fh_propagate_phase_backward": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                            c_int, c_int]),
    "fh_propagate_plain_fp64": (c_int, [c_void_p, c_void_p, c_int]),
    "fh_dot_product": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_double, c_void_p, c_void_p]),
    "fh_dot_product_pullback": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int,
                                        c_void_p]),
    "fh_dot_pullback_propagate": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int,
                                          c_void_p, c_void_p, c_void_p, c_int]),
    "fh_dot_matrix": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_double, c_void_p, c_void_p]),
    "fh_dot_matrix_pullback": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int,
                                       c_void_p]),
    "fh_offset_tilts": (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_double, c_double, c_double_p, c_int,
                                c_double_p, c_double_p, c_double_p, c_double_p, c_int]),
    "fh_mask_apply": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_void_p]),
    "fh_mask_backward": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_int,
                                 c_int]),
    "fh_tea_apply": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_int,
                             c_void_p]),
    "fh_tea_backward": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_void_p,
                                c_void_p, c_int, c_int]),
    "fh_optim_step": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, ctypes.c_long, c_double, c_double,
                              c_double, c_int, c_double, c_double]),
    "fh_fft2": (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_double]),
    "fh_sq_intensity_diff": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_void_p,
                                     c_void_p, c_void_p]),
    "fh_sq_intensity_diff_noalloc": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_void_p,
                                             c_void_p, c_void_p]),
    "fh_bpm_create": (c_int, [c_void_p, c_int, c_int, c_int, c_double, c_double, c_double_p, c_int,
                              c_double, c_double, c_int, c_int, ctypes.POINTER(c_void_p)]),
    "fh_collins_create": (c_int, [c_void_p, ctypes.POINTER(PropDesc), c_double, c_double, c_double, c_double,
                                  c_double, ctypes.POINTER(c_void_p)]),
    "fh_czt_create": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_double_p, c_double_p, c_int,
                              ctypes.POINTER(c_void_p)]),
    "fh_sqint_block": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_int, c_void_p,
                               c_void_p, c_void_p, c_int]),
    "fh_sqint_block_src": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_int,
                                   c_void_p, c_void_p, c_int, c_int]),
    "fh_as_z_gradient": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_double, c_double, c_void_p,
                                 c_int, c_double, c_double, c_void_p]),
    "fh_sq_field_diff": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p]),
    "fh_window": (c_int, [c_void_p, c_void_p, c_int, c_int, c_void_p, c_int, c_int, c_int, c_int, c_int, c_double,
                          c_double]),
    "fh_fista_step": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_long, c_double, c_void_p, c_int,
                              c_double, c_double]),
    "fh_adam_step": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_long, c_double, c_double, c_double,
                             c_double, c_void_p, c_int, c_double, c_double]),
    "fh_bpm_create_ex": (c_int, [c_void_p, ctypes.POINTER(PropDesc), c_double, c_int, c_int, ctypes.POINTER(c_void_p)]),
    "fh_bpm_destroy": (c_int, [c_void_p]),
    "fh_bpm_workspace_bytes": (c_int, [c_void_p, ctypes.POINTER(ctypes.c_size_t)]),
    "fh_bpm_forward": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int]),
    "fh_bpm_backward": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int]),
}


class FluxHipError(RuntimeError):
    """A libfluxhip call returned an error code."""

    def __init__(self, fn: str, code: int, msg: str):
        super().__init__(f"{fn} failed ({ERRORS.get(code, code)}): {msg}")
        self.code = code


_lock = threading.Lock()
_lib = None


def load(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load the library once and declare every exported signature."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(path):
            raise FileNotFoundError(
                f"{path} is missing: build it with `python fluxoptics.jl_amd/build.py` "
                "(there is no CPU fallback)")
        lib = ctypes.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


def call(name: str, *args) -> None:
    """Call ``name`` and raise FluxHipError on a nonzero status."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != FH_OK:
        raise FluxHipError(name, rc, lib.fh_last_error().decode(errors="replace"))
