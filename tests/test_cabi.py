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


def test_headers_compile_link_and_run(tmp_path):
    """What a reference build does with the headers: include/sgpu.h compiled as C and
    include/solar_gpu_api.h + include/sun_position.h as C++, linked against libsolargpu.so, and
    the host-side entry points run (sun positions through both)."""
    import subprocess
    inc = os.path.join(ROOT, "include")
    libdir = os.path.join(ROOT, "solargpu_amd", "lib")
    link = ["-L" + libdir, "-Wl,-rpath," + libdir, "-lsolargpu", "-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib"]
    (tmp_path / "t.c").write_text(
        '#include "sgpu.h"\n'
        "static float az_alt[2 * SGPU_HOURS_PER_YEAR];\n"
        "int main(void) { return sgpu_sun_positions(15.65f, 46.56f, 1, az_alt) != 0 || az_alt[2 * 12 + 1] <= 0.0f; }\n")
    (tmp_path / "t.cpp").write_text(
        '#include <vector>\n#include "solar_gpu_api.h"\n#include "sun_position.h"\n'
        "int main() {\n"
        "    std::vector<solar_gpu::SunSample> s;\n"
        "    solar_gpu::SunPositionCalculator(15.65f, 46.56f, 1).calc(s);\n"
        "    auto f = &solar_gpu::compressed_gpu_irradiance;\n"
        "    auto g = &solar_gpu::uncompressed_gpu_irradiance;\n"
        "    return (s.size() != 8760 || !f || !g) ? 1 : 0;\n"
        "}\n")
    for cc, src, std in (("gcc", "t.c", "-std=c11"), ("g++", "t.cpp", "-std=c++17")):
        exe = str(tmp_path / (src + ".bin"))
        r = subprocess.run([cc, std, "-Wall", "-Werror", "-I" + inc, str(tmp_path / src), "-o", exe] + link,
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
        assert r.returncode == 0, (src, r.stdout, r.stderr)
