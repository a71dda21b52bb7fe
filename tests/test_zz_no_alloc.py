"""The allocation-free hour loop (sgpu_hour_loop_noalloc) against sgpu_hour_loop through the
C ABI: the same annual sums bit for bit, on two scenes in a row (the second smaller, so the
shared scratch array holds the first call's contents), in both list formats; a scene whose
scratch exceeds the array is refused, and hours_per_pass 0 halves the tuned value until the
scratch fits."""
import ctypes

import numpy as np
import pytest

import solargpu_amd as sg
from scenes import sky, synth_scene

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 128 << 20
HIP_D2H = 2


def _hip():
    # the HIP runtime libsolargpu.so is linked against (the one already loaded in this process)
    hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return hip


def _noalloc_annual(ctx, fmt, K, n):
    hip = _hip()
    ptr = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(ptr), n * 4) == 0
    try:
        ctx.hour_loop_noalloc(fmt, K, ptr.value)
        ctx.sync()
        out = np.empty(n, np.float32)
        assert hip.hipMemcpy(out.ctypes.data, ptr, n * 4, HIP_D2H) == 0
    finally:
        hip.hipFree(ptr)
    return out


def _ctx(sc):
    az_alt, dni, dhi, _ = sky()
    ctx = sg.Context(0)
    ctx.set_scene(sc)
    ctx.set_sky(az_alt, dni, dhi)
    return ctx


def test_noalloc_hour_loop_equals_hour_loop_on_two_scenes_in_a_row():
    for seed, w, h, K in ((41, 96, 72, 64), (42, 48, 40, 32)):   # the second scene is smaller
        sc = synth_scene(seed, w, h)
        print(f"noalloc: scene {seed} {w}x{h}, {sc.num_total} voxels, K={K}", flush=True)
        with _ctx(sc) as ctx:
            ctx.viewshed(sg.FMT_CSR | sg.FMT_BARLE)
            for fmt, queue in ((sg.FMT_BARLE, 0), (sg.FMT_CSR, 0), (sg.FMT_BARLE, 1)):
                ctx.set_tuning(16, K, 4, 16, 0, queue)
                ctx.hour_loop(fmt, K)
                want = ctx.irradiance()
                got = _noalloc_annual(ctx, fmt, K, sc.num_total)
                np.testing.assert_array_equal(got, want)
                assert np.isfinite(got).all() and got.min() > 0


def test_noalloc_scratch_bound_and_auto_hours_per_pass():
    side = 512                                   # 262,144 top cells (a flat plane: no walls)
    heights = np.full((side, side), 200.0, np.float32)
    sc = sg.assemble_scene(heights, 1.0)
    assert sc.num_total == sc.num_top == side * side
    need = lambda K: ((sc.num_top * K * 4 + 255) // 256 + (sc.num_total * 4 + 255) // 256) * 256 + 32768
    assert need(128) > SCRATCH_BYTES >= need(64)
    print(f"noalloc: flat {side}x{side}", flush=True)
    with _ctx(sc) as ctx:
        ctx.viewshed(sg.FMT_BARLE)
        with pytest.raises(sg.SolarGPUError) as e:
            _noalloc_annual(ctx, sg.FMT_BARLE, 128, sc.num_total)
        assert e.value.code == 1 and "scratch" in str(e.value)
        ctx.set_tuning(16, 256, 4, 0, 0)         # tuned K = 256: the call halves it to 64
        got = _noalloc_annual(ctx, sg.FMT_BARLE, 0, sc.num_total)
        ctx.hour_loop(sg.FMT_BARLE, 64)
        np.testing.assert_array_equal(got, ctx.irradiance())
