"""Autotuner: every launch shape it can pick gives identical results, and a
tuned key is served from the cache the second time."""
import numpy as np
import pytest

import solargpu_amd as sg
import oracle_lib as orc
from scenes import oracle_lists, sky, synth_scene

pytestmark = pytest.mark.gpu


def _ctx(sc):
    az_alt, dni, dhi, _ = sky()
    ctx = sg.Context(0)
    ctx.set_scene(sc)
    ctx.set_sky(az_alt, dni, dhi)
    return ctx


def test_every_shape_is_results_neutral():
    sc = synth_scene(12, 64, 64)
    ref = oracle_lists(12, 64, 64)
    az_alt, dni, dhi, bins = sky()
    exact = orc.annual_fp64(sc, bins, dni, dhi, ref["svf"], ref["off"], ref["ids"])
    with _ctx(sc) as ctx:
        for G, K, waves, band, group in ((16, 8, 2, 0, 0), (48, 32, 4, 16, 0), (32, 16, 2, 0, 4),
                                         (32, 64, 4, 0, 2)):
            ctx.set_tuning(G, K, waves, band, group)
            ctx.viewshed(sg.FMT_CSR | sg.FMT_BARLE)
            off, ids = ctx.csr()
            np.testing.assert_array_equal(ids, ref["ids"])
            np.testing.assert_array_equal(ctx.svf(), ref["svf"])
            ctx.hour_loop(sg.FMT_BARLE, 0)
            irr = ctx.irradiance()
            assert (np.abs(irr - exact) / np.maximum(exact, 1.0)).max() <= 1e-4


def test_autotune_measures_then_hits_the_cache(tmp_path, monkeypatch):
    monkeypatch.setenv("SGPU_AUTOTUNE_CACHE", str(tmp_path / "tune.txt"))
    sc = synth_scene(11, 48, 40)
    with _ctx(sc) as ctx:
        ctx.viewshed(sg.FMT_BARLE)
        t = ctx.autotune(sg.FMT_BARLE)
        assert t["from_cache"] == 0 and t["rays_per_s"] > 0 and t["voxel_hours_per_s"] > 0
        assert t["dirs_per_group"] in (16, 32, 48) and t["hours_per_pass"] in (64, 128, 256)
    with _ctx(sc) as ctx:
        ctx.viewshed(sg.FMT_BARLE)
        t2 = ctx.autotune(sg.FMT_BARLE)
        assert t2["from_cache"] == 1
        assert (t2["dirs_per_group"], t2["hours_per_pass"], t2["waves_per_group"]) == \
            (t["dirs_per_group"], t["hours_per_pass"], t["waves_per_group"])


def test_receiver_assignment_is_bit_identical():
    """Each voxel's reflection walk and Kahan sum are its own: consecutive receivers per wave,
    receiver groups over the tile-major order, XCD bands and the per-XCD receiver queue give
    bit-identical annual sums."""
    sc = synth_scene(13, 150, 120)
    with _ctx(sc) as ctx:
        ctx.viewshed(sg.FMT_CSR | sg.FMT_BARLE)
        out = []
        for fmt in (sg.FMT_CSR, sg.FMT_BARLE):
            for waves, band, group, queue in ((4, 0, 0, 0), (4, 16, 8, 0), (2, 0, 32, 0), (2, 64, 3, 0),
                                              (4, 0, 0, 1), (4, 16, 0, 1), (2, 32, 0, 1)):
                ctx.set_tuning(16, 32, waves, band, group, queue)
                ctx.hour_loop(fmt, 32)
                out.append(ctx.irradiance())
            for x in out[1:]:
                np.testing.assert_array_equal(x, out[0])
            out.clear()


def test_undersized_hourly_buffer_is_refused_before_any_launch(tmp_path, monkeypatch):
    """A pass of K hours writes and gathers K floats per hourly row.  The autotuner once
    ran K = 128 against a buffer sized for K = 64, and the reflection pass read past the buffer
    (a GPU memory fault); every pass now checks the capacity before it launches.  With the
    usable buffer capped at the K = 64 size, a K = 128 loop must be refused with SGPU_ERR_STATE
    and leave the context usable; the tuner's own sequence (a K = 64 loop, then trials of up to
    256 hours per pass) then runs, with results within 1e-4 of the FP64 oracle."""
    monkeypatch.setenv("SGPU_AUTOTUNE_CACHE", str(tmp_path / "tune.txt"))
    sc = synth_scene(11, 48, 40)
    ref = oracle_lists(11, 48, 40)
    az_alt, dni, dhi, bins = sky()
    exact = orc.annual_fp64(sc, bins, dni, dhi, ref["svf"], ref["off"], ref["ids"])
    rows = -(-sc.num_top // 1024) * 1024   # the swizzled hourly rows fill whole blocks of 1024
    with _ctx(sc) as ctx:
        ctx.viewshed(sg.FMT_BARLE)
        ctx.hour_loop(sg.FMT_BARLE, 64)
        ctx.set_hourly_limit(rows * 64)
        for K in (128, 256):
            with pytest.raises(sg.SolarGPUError) as e:
                ctx.hour_loop(sg.FMT_BARLE, K)
            assert e.value.code == 4 and "hourly irradiance buffer" in str(e.value)
        ctx.hour_loop(sg.FMT_BARLE, 64)          # fits the cap: runs
        assert (np.abs(ctx.irradiance() - exact) / np.maximum(exact, 1.0)).max() <= 1e-4
        ctx.set_hourly_limit(0)
        t = ctx.autotune(sg.FMT_BARLE)
        assert t["hours_per_pass"] in (64, 128, 256)
        ctx.hour_loop(sg.FMT_BARLE, 0)
        assert (np.abs(ctx.irradiance() - exact) / np.maximum(exact, 1.0)).max() <= 1e-4
