"""Loader of hpg_amd/libhpgx.so (the C ABI of include/hpgx.h).

There is no fallback: if the library is missing or fails to load, every
entry point raises, so a GPU run can never silently use another code path.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhpgx.so")
_lib = None

# (name, restype, argtypes) of every symbol declared in include/hpgx.h
vp, u, i, d, sz = (ctypes.c_void_p, ctypes.c_uint, ctypes.c_int, ctypes.c_double,
                   ctypes.c_size_t)
SIGNATURES = [
    ("hpgx_initialize", i, [i]),
    ("hpgx_finalize", None, []),
    ("hpgx_device_count", i, []),
    ("hpgx_version", ctypes.c_char_p, []),
    ("hpgx_create", i, [vp, u, sz, vp, vp, u, u, vp, vp, i, u, u, vp]),
    ("hpgx_clone", i, [vp, vp]),
    ("hpgx_destroy", None, [vp]),
    ("hpgx_error_message", sz, [vp, ctypes.c_char_p, sz]),
    ("hpgx_max_added_tasks", u, [vp]),
    ("hpgx_max_visibility_batch_size", sz, [vp]),
    ("hpgx_grid_size", None, [vp, vp]),
    ("hpgx_grid_scale", None, [vp, vp]),
    ("hpgx_num_polarizations", u, [vp]),
    ("hpgx_grid_mode", i, [vp]),
    ("hpgx_convolution_function_region_size", sz, [vp, u, u, vp]),
    ("hpgx_allocate_convolution_function_region", i, [vp, u, u, vp]),
    ("hpgx_set_convolution_function", i, [vp, u, u, vp, vp, i]),
    ("hpgx_set_model", i, [vp, vp, vp, i, i]),
    ("hpgx_reset_model", i, [vp]),
    ("hpgx_has_model", i, [vp]),
    ("hpgx_grid_visibilities", i, [vp, vp, sz, u, i, i, i, i, vp, i]),
    ("hpgx_grid_visibilities_noalloc", i, [vp, vp, sz, u, i]),
    ("hpgx_degrid_forward", i, [vp, vp, vp, sz, u, i]),
    ("hpgx_degrid_adjoint", i, [vp, vp, vp, sz, u, i]),
    ("hpgx_fence", i, [vp]),
    ("hpgx_normalize_by_weights", i, [vp, d]),
    ("hpgx_apply_grid_fft", i, [vp, d, i, i]),
    ("hpgx_apply_model_fft", i, [vp, d, i, i]),
    ("hpgx_shift_grid", i, [vp, i]),
    ("hpgx_shift_model", i, [vp, i]),
    ("hpgx_reset_grid", i, [vp]),
    ("hpgx_grid_values", i, [vp, vp, i]),
    ("hpgx_grid_weights", i, [vp, vp, i]),
    ("hpgx_model_values", i, [vp, vp, i]),
    ("hpgx_grid_window", i, [vp, u, u, u, u, u, u, vp]),
    ("hpgx_grid_device_ptr", i, [vp, vp, vp]),
    ("hpgx_event_record", i, [vp, i]),
    ("hpgx_event_elapsed", i, [vp, i, i, vp]),
]


def lib():
    global _lib
    if _lib is None:
        # torch bundles its own HIP runtime (same soname, libamdhip64.so.7).
        # Importing it first makes this library bind to that runtime instead
        # of loading a second HIP runtime into the process.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m hpg_amd.build` "
                "(there is no non-HIP fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def error_message(state):
    n = lib().hpgx_error_message(state, None, 0)
    buf = ctypes.create_string_buffer(n + 1)
    lib().hpgx_error_message(state, buf, n + 1)
    return buf.value.decode(errors="replace")


def last_create_error():
    return error_message(None)
