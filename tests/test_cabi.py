"""The C-ABI library loads and exports every entry point its headers declare;
host-side argument checks return status codes instead of exiting."""
import ctypes
import os
import re

import numpy as np
import pytest

import solargpu_amd as sg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            with open(os.path.join(ROOT, "include", fn)) as f:
                text = f.read()
            names.update(re.findall(r"^\s*(?:int|const char\s*\*)\s+(sgpu_\w+)\s*\(", text, re.M))
    return sorted(names)


def test_every_declared_symbol_is_exported():
    lib = ctypes.CDLL(sg.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_cpp_drop_in_symbols_exported():
    out = os.popen(f"nm -D --defined-only {sg.LIB_PATH}").read()
    assert "uncompressed_gpu_irradiance" in out and "compressed_gpu_irradiance" in out


def test_host_argument_errors_are_status_codes():
    with pytest.raises(sg.SolarGPUError) as e:
        sg.synth_dsm(1, 0, 10, 1.0, 0.3, 1)
    assert e.value.code == 1
    lib = ctypes.CDLL(sg.LIB_PATH)
    assert lib.sgpu_sun_positions(ctypes.c_float(0), ctypes.c_float(0), 0, None) == 1


def test_context_without_gpu_reports_error():
    lib = ctypes.CDLL(sg.LIB_PATH)
    h = ctypes.c_void_p()
    st = lib.sgpu_create(0, ctypes.byref(h))
    if st == 0:
        lib.sgpu_destroy(h)
        pytest.skip("a GPU is visible")
    assert st in (1, 2)
    lib.sgpu_last_error.restype = ctypes.c_char_p
    assert lib.sgpu_last_error()
