"""ctypes binding of libsles.so (the C ABI declared in include/sles.h).

The library is the only compute path: importing this module raises when the
shared object is missing, and every call raises ``RuntimeError`` with the
library's message when a call fails.  There is no CPU fallback.
"""

from __future__ import annotations

import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("SLES_LIB", os.path.join(_HERE, "libsles.so"))

SLES_FP64 = 0
SLES_FP32 = 1
SLES_LES_NONE = 0
SLES_LES_SMAGORINSKY = 1
SLES_LES_VREMAN = 2
TUNE_KEYS = {"strided_radix4": 0, "xfwd_radix4": 1, "z_radix4": 2, "xcd": 3, "fft_wide": 4, "fuse_x": 5, "pad": 6,
             "narrow": 7}

_vp = ctypes.c_void_p
_int = ctypes.c_int
_dbl = ctypes.c_double
_size = ctypes.c_size_t

# name -> (restype, argtypes)
_PROTOS = {
    "sles_last_error": (ctypes.c_char_p, []),
    "sles_version": (_int, []),
    "sles_plan_create": (_int, [ctypes.POINTER(_vp), _int, _int]),
    "sles_plan_destroy": (_int, [_vp]),
    "sles_plan_info": (_int, [_vp, ctypes.POINTER(_int), ctypes.POINTER(_int), ctypes.POINTER(_int)]),
    "sles_plan_set_tuning": (_int, [_vp, _int, _int]),
    "sles_plan_get_tuning": (_int, [_vp, _int, ctypes.POINTER(_int)]),
    "sles_fft_r2c": (_int, [_vp, _vp, _vp, _int, _dbl, _vp]),
    "sles_fft_c2r": (_int, [_vp, _vp, _vp, _vp, _int, _dbl, _vp]),
    "sles_cross_product": (_int, [_vp, _vp, _vp, _vp, _vp]),
    "sles_spectral_curl": (_int, [_vp, _vp, _vp, _vp]),
    "sles_spectral_gradient": (_int, [_vp, _vp, _vp, _vp]),
    "sles_smagorinsky_nu": (_int, [_vp, _vp, _vp, _dbl, _vp]),
    "sles_vreman_nu": (_int, [_vp, _vp, _vp, _vp, _dbl, _vp]),
    "sles_stress_products": (_int, [_vp, _vp, _vp, _vp, _vp]),
    "sles_stress_divergence": (_int, [_vp, _vp, _vp, _vp]),
    "sles_rhs_finish": (_int, [_vp, _vp, _vp, _vp, _vp, _vp, _dbl, _dbl, _vp]),
    "sles_mul_k2": (_int, [_vp, _vp, _vp, _int, _vp]),
    "sles_stage_update": (_int, [_vp, _int, _vp, _vp, _vp, _vp, _dbl, _dbl, _vp]),
    "sles_etd2_update": (_int, [_vp, _int, _vp, _vp, _vp, _dbl, _dbl, _vp]),
    "sles_reduce_work_bytes": (_size, []),
    "sles_sum_squares": (_int, [_vp, _vp, _int, _vp, _vp, _vp]),
    "sles_dot": (_int, [_vp, _vp, _vp, _int, _vp, _vp, _vp]),
    "sles_max": (_int, [_vp, _vp, _int, _vp, _vp, _vp]),
    "sles_spectrum_work_bytes": (_size, [_vp, _int]),
    "sles_spectrum": (_int, [_vp, _vp, _int, _vp, _vp, _vp, _vp, _vp]),
    "sles_shell_spectrum": (_int, [_vp, _vp, _int, _vp, _vp, _vp]),
    "sles_shell_spectrum_vjp": (_int, [_vp, _vp, _int, _vp, _vp, _vp]),
    "sles_scale": (_int, [_vp, _vp, _int, _dbl, _vp]),
    "sles_step_work_bytes": (_size, [_vp, _int]),
    "sles_adjoint_work_bytes": (_size, [_vp, _int]),
    "sles_adjoint_rhs": (_int, [_vp, _vp, _vp, _vp, _vp, _int, _int, _dbl, _vp, _vp, _vp]),
    "sles_axpby": (_int, [_vp, _int, _dbl, _vp, _dbl, _vp, _vp]),
    "sles_lincomb": (_int, [_vp, _int, _dbl, _vp, _dbl, _vp, _vp, _vp]),
    "sles_etd2_adjoint_combine": (_int, [_vp, _dbl, _dbl, _vp, _vp, _vp, _vp, _vp]),
    "sles_etd2_adjoint_rhs": (_int, [_vp, _vp, _vp, _vp, _vp, _vp, _dbl, _vp, _vp, _vp]),
    "sles_adjoint_rhs_combined": (_int, [_vp, _vp, _vp, _dbl, _vp, _dbl, _vp, _vp, _dbl, _vp, _vp, _vp]),
    "sles_sum_fields": (_int, [_vp, _int, _int, _vp, _vp, _vp]),
    "sles_profile_enable": (_int, [_int]),
    "sles_profile_read": (_int, [_vp, _vp, _int]),
    "sles_step": (_int, [_vp, _int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "sles_step_drive": (_int, [_vp, _int, _vp, _dbl, _int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "sles_step_batch_work_bytes": (_size, [_vp, _int, _int]),
    "sles_step_drive_batch": (_int, [_vp, _int, _vp, _int, _vp, _dbl, _int, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "sles_stream_probe": (_int, [_vp, _vp, _int, _vp, _int, _int, _vp]),
    "sles_step_noalloc": (_int, [_vp, _int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "sles_rk4_stage_states": (_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
}

OPTIONAL = set()  # symbols a given build may lack (none at present)


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"libsles.so not found at {LIB_PATH}; build it with "
            "`python spectral-les_amd/build_lib.py` (the HIP library is the only compute path)"
        )
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _PROTOS.items():
        try:
            fn = getattr(lib, name)
        except AttributeError:
            if name in OPTIONAL:
                continue
            raise
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def last_error() -> str:
    msg = lib.sles_last_error()
    return msg.decode() if msg else ""


def call(name: str, *args) -> None:
    rc = getattr(lib, name)(*args)
    if rc != 0:
        raise RuntimeError(f"{name} failed: {last_error()}")


def stream_handle(device=None) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def ptr(t) -> int:
    if t is None:
        return None
    return t.data_ptr()


def ptr_array(tensors):
    """(void* [k]) of device pointers, or None."""
    if tensors is None:
        return None
    arr = (ctypes.c_void_p * len(tensors))()
    for i, t in enumerate(tensors):
        arr[i] = t.data_ptr()
    return arr


class Plan:
    """Owns one sles_plan (twiddle table for an N^3 grid in one precision)."""

    def __init__(self, n: int, precision: int, device):
        self.n = int(n)
        self.nh = self.n // 2 + 1
        self.precision = precision
        self.device = torch.device(device)
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            call("sles_plan_create", ctypes.byref(h), self.n, precision)
        self.handle = h

    def set_tuning(self, **kw):
        for k, v in kw.items():
            call("sles_plan_set_tuning", self.handle, TUNE_KEYS[k], int(v))

    def get_tuning(self):
        out = {}
        for k, key in TUNE_KEYS.items():
            v = ctypes.c_int()
            call("sles_plan_get_tuning", self.handle, key, ctypes.byref(v))
            out[k] = v.value
        return out

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                lib.sles_plan_destroy(h)
            except Exception:  # pragma: no cover - interpreter shutdown
                pass
            self.handle = None
