"""ctypes binding of libpywave_amd.so, the HIP library (declared in include/pywave_amd.h).

The library is built in-tree by ``make -C pywave_amd/csrc`` (or ``__graft_entry__.build()``).
There is no CPU fallback: every entry point raises if the library or a GPU is missing.
"""
import ctypes
import os
import sys
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpywave_amd.so")

_lib = None
_ctx = {}
_lock = threading.Lock()

c_int, c_float, c_double, c_ll, c_vp = ctypes.c_int, ctypes.c_float, ctypes.c_double, ctypes.c_longlong, ctypes.c_void_p
_fp = ctypes.POINTER(ctypes.c_float)

# name -> (restype, argtypes); every exported symbol of include/pywave_amd.h
SIGNATURES = {
    "pw_create": (c_vp, [c_int]),
    "pw_destroy": (None, [c_vp]),
    "pw_last_error": (ctypes.c_char_p, [c_vp]),
    "pw_device_count": (c_int, []),
    "pw_set_stream": (c_int, [c_vp, c_vp]),
    "pw_use_own_stream": (c_int, [c_vp]),
    "pw_synchronize": (c_int, [c_vp]),
    "pw_stream_bandwidth": (c_int, [c_vp, c_ll, c_int, c_int, c_int, c_vp]),
    "pw_next_fast_size": (c_int, [c_int]),
    "pw_padded_sizes": (None, [c_int, c_int, c_int, ctypes.POINTER(c_int), ctypes.POINTER(c_int), ctypes.POINTER(c_int)]),
    "pw_abc_weights": (None, [c_int, c_float, c_vp]),
    "pw_ricker": (c_int, [c_int, c_float, c_float, c_float, c_float, c_vp]),
    "pw_afd2d": (c_int, [c_vp, c_vp, c_vp, c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_int,
                         c_int, c_float, c_float, c_float, c_float, c_float, c_float, c_vp, c_ll]),
    "pw_afd3d": (c_int, [c_vp, c_vp, c_vp, c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_int,
                         c_int, c_int, c_float, c_float, c_float, c_float, c_float, c_float, c_float,
                         c_float, c_vp, c_ll]),
    "pw_aps2d": (c_int, [c_vp, c_vp, c_vp, c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_int,
                         c_int, c_float, c_float, c_float, c_float, c_float, c_float, c_vp, c_ll]),
    "pw_aps3d": (c_int, [c_vp, c_vp, c_vp, c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_int,
                         c_int, c_int, c_float, c_float, c_float, c_float, c_float, c_float, c_float,
                         c_float, c_vp, c_ll]),
    "pw_afd2d_noalloc_capacity": (ctypes.c_size_t, []),
    "pw_afd2d_dev_noalloc": (c_int, [c_vp, c_vp, c_int, c_int, c_float, c_float, c_int, c_float, c_int, c_float,
                                     c_int, c_vp, c_vp, c_vp]),
    "pw_afd_output_len": (c_ll, [c_int, c_int, c_int, c_int, c_int]),
    "pw_afd_create": (c_vp, [c_vp, c_int, c_int, c_int, c_int, c_float, c_float, c_float, c_int, c_float,
                             c_int, c_float, c_int]),
    "pw_afd_set_velocity": (c_int, [c_vp, c_vp]),
    "pw_afd_set_sources": (c_int, [c_vp, c_int, c_vp, c_vp]),
    "pw_afd_set_tri_data": (c_int, [c_vp, c_vp]),
    "pw_afd_run": (c_int, [c_vp, c_int]),
    "pw_afd_run_timed": (c_int, [c_vp, c_int, c_int, ctypes.POINTER(c_double), ctypes.POINTER(c_double)]),
    "pw_afd_output": (c_int, [c_vp, c_vp, c_ll]),
    "pw_afd_final_field": (c_int, [c_vp, c_vp, c_ll]),
    "pw_afd_set_variant": (c_int, [c_vp, c_int, c_int]),
    "pw_afd_set_spectral": (c_int, [c_vp, c_int]),
    "pw_afd_variant_count": (c_int, []),
    "pw_afd_variant_info": (c_int, [c_int, c_vp]),
    "pw_afd_autotune": (c_int, [c_vp, c_int, ctypes.POINTER(c_int), ctypes.POINTER(c_int)]),
    "pw_afd_destroy": (None, [c_vp]),
    "pw_fwi_create": (c_vp, [c_vp, c_vp, c_int, c_int]),
    "pw_fwi_destroy": (None, [c_vp]),
    "pw_fwi_set_checkpointing": (c_int, [c_vp, c_int]),
    "pw_fwi_info": (c_int, [c_vp, c_vp]),
    "pw_fwi_set_wavelet": (c_int, [c_vp, c_vp]),
    "pw_fwi_set_observed": (c_int, [c_vp, c_vp]),
    "pw_fwi_set_source_field": (c_int, [c_vp, c_vp]),
    "pw_fwi_set_visco": (c_int, [c_vp, c_vp]),
    "pw_forward_modeling_c": (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_ll]),
    "pw_fwi_set_source_field_dev": (c_int, [c_vp, c_vp]),
    "pw_fwi_forward": (c_int, [c_vp, c_vp, c_vp]),
    "pw_fwi_gradient": (c_int, [c_vp, c_vp, c_int, c_vp, ctypes.POINTER(c_double), ctypes.POINTER(c_float)]),
    "pw_fwi_gradient_f64": (c_int, [c_vp, c_vp, c_int, c_vp, ctypes.POINTER(c_double)]),
    "pw_fwi_misfit_f64": (c_int, [c_vp, c_vp, ctypes.POINTER(c_double)]),
    "pw_fwi_forward_dev": (c_int, [c_vp, c_vp, c_vp]),
    "pw_fwi_adjoint_dev": (c_int, [c_vp, c_vp, c_vp, c_vp]),
    "pw_fwi_timing": (c_int, [c_vp, ctypes.POINTER(c_double), ctypes.POINTER(c_double)]),
    "pw_fwi_widths": (c_int, [c_vp, c_vp]),
    "pw_fwi_set_steps_per_launch": (c_int, [c_vp, c_int]),
    "pw_fwi_set_widths": (c_int, [c_vp, c_vp]),
    "pw_fwi_tuning_log": (c_int, [c_vp, c_vp, c_int]),
    "pw_fwi_lbfgs": (c_int, [c_vp, c_vp, c_int, c_vp, ctypes.POINTER(c_int)]),
    "pw_forward_modeling_ac": (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_ll]),
    "pw_fwic": (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_ll]),
    "pw_pfwic": (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_ll, c_ll]),
    "pw_lstri_create": (c_vp, [c_vp, c_vp]),
    "pw_lstri_destroy": (None, [c_vp]),
    "pw_lstri_set_model": (c_int, [c_vp, c_vp]),
    "pw_lstri_model": (c_int, [c_vp, c_vp, c_vp]),
    "pw_lstri_image": (c_int, [c_vp, c_vp, c_vp]),
    "pw_lstri_invert": (c_int, [c_vp, c_vp, c_int, c_vp, c_vp]),
    "pw_lstri_timing": (c_int, [c_vp, c_vp]),
    "pw_lstri_use_graph": (c_int, [c_vp, c_int]),
    "pw_lstri_steps_per_launch": (c_int, [c_vp, c_int]),
    "pw_lstri_launch_choice": (c_int, [c_vp, c_vp, c_vp, c_int]),
    "pw_lstri_set_tile_width": (c_int, [c_vp, c_int]),
    "pw_lstri_set_four_step_columns": (c_int, [c_vp, c_int]),
    "pw_lstric": (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_ll, c_ll]),
    "pw_tri_create": (c_vp, [c_vp, c_int, c_int, c_int, c_float, c_float, c_float, c_int, c_float, c_int, c_float,
                             c_int, c_int]),
    "pw_tri_set_velocity": (c_int, [c_vp, c_vp]),
    "pw_tri_set_data": (c_int, [c_vp, c_vp, c_vp]),
    "pw_tri_run": (c_int, [c_vp]),
    "pw_tri_run_timed": (c_int, [c_vp, ctypes.POINTER(c_double)]),
    "pw_tri_launch_choice": (c_int, [c_vp, c_vp, c_vp, c_int]),
    "pw_tri_set_launch": (c_int, [c_vp, c_int, c_int, c_int]),
    "pw_tri_image": (c_int, [c_vp, c_vp, c_ll]),
    "pw_tri_destroy": (None, [c_vp]),
    "pw_batch2d_create": (c_vp, [c_vp, c_int, c_vp, c_vp, c_int, c_float]),
    "pw_batch2d_set_shot": (c_int, [c_vp, c_int, c_vp, c_int, c_vp, c_vp]),
    "pw_batch2d_run": (c_int, [c_vp, c_int, ctypes.POINTER(c_double)]),
    "pw_batch2d_traces": (c_int, [c_vp, c_int, c_vp, c_ll]),
    "pw_batch2d_cell_updates": (c_ll, [c_vp]),
    "pw_batch2d_set_width": (c_int, [c_vp, c_int]),
    "pw_batch2d_width": (c_int, [c_vp, c_vp, c_int, ctypes.POINTER(c_int)]),
    "pw_batch2d_destroy": (None, [c_vp]),
}


class PywaveError(RuntimeError):
    """A HIP-side failure reported by libpywave_amd."""


class NonFiniteError(PywaveError):
    """A result (traces, wavefield, image, gradient or misfit) holds Inf or NaN: the run blew up,
    usually because the time step exceeds the scheme's CFL limit (PW_ERR_NONFINITE)."""


def _bind_torch_runtime():
    """One HIP runtime per process: torch ships its own libamdhip64 with the same soname as
    ROCm's, and whichever is loaded first serves both.  When torch is installed it is imported
    before this library loads, so the library runs on torch's runtime and torch keeps working
    (PYWAVE_AMD_NO_TORCH=1 skips this, for processes that never import torch)."""
    if "torch" in sys.modules or os.environ.get("PYWAVE_AMD_NO_TORCH") == "1":
        return
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def lib():
    """Load the in-tree library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                _bind_torch_runtime()
                if not os.path.exists(LIB_PATH):
                    raise ImportError("pywave_amd: %s is missing; build it with `make -C pywave_amd/csrc` "
                                      "(there is no CPU fallback)" % LIB_PATH)
                L = ctypes.CDLL(LIB_PATH)
                for name, (res, args) in SIGNATURES.items():
                    fn = getattr(L, name)
                    fn.restype = res
                    fn.argtypes = args
                _lib = L
    return _lib


def context(device=0):
    """Per-device library context (created on first use)."""
    c = _ctx.get(device)
    if c is None:
        with _lock:
            c = _ctx.get(device)
            if c is None:
                L = lib()
                if L.pw_device_count() <= device:
                    raise PywaveError("pywave_amd: no HIP device %d visible (the engine has no CPU path)" % device)
                c = L.pw_create(device)
                if not c:
                    raise PywaveError("pywave_amd: pw_create(%d) failed" % device)
                _ctx[device] = c
    return c


def check(rc, ctx, what):
    if rc != 0:
        msg = lib().pw_last_error(ctx).decode(errors="replace") if ctx else ""
        if rc == -1:
            raise ValueError("%s: %s" % (what, msg))
        if rc == -5:
            raise NonFiniteError("%s: %s" % (what, msg))
        raise PywaveError("%s failed (%d): %s" % (what, rc, msg))


def ptr(a):
    """Raw data pointer of a C-contiguous numpy array."""
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def f32c(a):
    return np.ascontiguousarray(a, dtype=np.float32)
