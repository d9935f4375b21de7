"""The C-ABI library loads and exports every entry point include/*.h declares
(CPU only: no compute calls without a GPU)."""

import ctypes
import glob
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "spectral-les_amd", "sles", "libsles.so")


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(sles_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_header_declares_entry_points():
    names = declared_symbols()
    for n in ("sles_plan_create", "sles_fft_r2c", "sles_fft_c2r", "sles_rhs_finish", "sles_spectrum"):
        assert n in names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build libsles.so first (python spectral-les_amd/build_lib.py)"
    lib = ctypes.CDLL(LIB)
    missing = [n for n in sorted(declared_symbols()) if not hasattr(lib, n)]
    assert not missing, missing


def test_version_and_error_string_without_gpu():
    lib = ctypes.CDLL(LIB)
    lib.sles_version.restype = ctypes.c_int
    assert lib.sles_version() == 1
    # argument validation runs on the host and needs no device
    lib.sles_plan_create.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_int]
    lib.sles_last_error.restype = ctypes.c_char_p
    h = ctypes.c_void_p()
    assert lib.sles_plan_create(ctypes.byref(h), 100, 0) == -1
    assert b"unsupported grid size" in lib.sles_last_error()
    assert lib.sles_plan_create(ctypes.byref(h), 64, 7) == -1
    assert b"precision" in lib.sles_last_error()


def test_python_binding_matches_header():
    from sles import _lib

    assert set(_lib._PROTOS) <= declared_symbols()


def test_tuning_keys_match_header():
    """sles/_lib.py's tuning-knob names map to the SLES_TUNE_* values the
    header declares, one for one, and the autotuner only tunes known knobs."""
    from sles import _lib, autotune

    text = open(os.path.join(ROOT, "include", "sles.h")).read()
    enum = {m.group(1).lower(): int(m.group(2)) for m in re.finditer(r"\bSLES_TUNE_([A-Z0-9_]+)\s*=\s*(\d+)", text)}
    assert _lib.TUNE_KEYS == enum
    assert set(autotune.KNOBS) <= set(_lib.TUNE_KEYS)
