"""Persisted neighbour lists (sgpu_save_lists / sgpu_load_lists): a reloaded context reproduces
the viewshed's lists and SVF and the hour loop's annual sums bit for bit; files only load into
the scene and sky they were made for; damaged files are refused."""
import dataclasses

import numpy as np
import pytest

import solargpu_amd as sg
from scenes import sky, synth_scene

pytestmark = pytest.mark.gpu


def _ctx(sc, site_sky=None):
    az_alt, dni, dhi, _ = site_sky or sky()
    ctx = sg.Context(0)
    ctx.set_scene(sc)
    ctx.set_sky(az_alt, dni, dhi)
    return ctx


def test_round_trip_reproduces_lists_and_annual_sums(tmp_path):
    sc = synth_scene(31, 96, 72)
    path = str(tmp_path / "lists.sgl")
    with _ctx(sc) as a:
        a.viewshed(sg.FMT_CSR | sg.FMT_BARLE)
        a.save_lists(path)
        off, ids = a.csr()
        boff, data = a.barle()
        svf = a.svf()
        a.hour_loop(sg.FMT_BARLE, 64)
        want = a.irradiance()
    with _ctx(sc) as b:
        b.load_lists(path)
        o2, i2 = b.csr()
        np.testing.assert_array_equal(o2, off)
        np.testing.assert_array_equal(i2, ids)
        bo2, d2 = b.barle()
        np.testing.assert_array_equal(bo2, boff)
        np.testing.assert_array_equal(d2, data)
        np.testing.assert_array_equal(b.svf(), svf)
        b.hour_loop(sg.FMT_BARLE, 64)
        np.testing.assert_array_equal(b.irradiance(), want)


def test_barle_only_file_and_device_voxelized_scene(tmp_path):
    heights = sg.synth_dsm(32, 80, 64, 1.0, 0.35, 4)
    path = str(tmp_path / "barle.sgl")
    az_alt, dni, dhi, _ = sky()
    with sg.Context(0) as a:
        a.set_scene_heights(heights, 1.0)   # voxelized on the device
        a.set_sky(az_alt, dni, dhi)
        a.viewshed(sg.FMT_BARLE)
        a.save_lists(path, sg.FMT_BARLE)
        a.hour_loop(sg.FMT_BARLE, 64)
        want = a.irradiance()
    with _ctx(sg.assemble_scene(heights, 1.0)) as c:   # the same scene from the host path
        c.viewshed(sg.FMT_BARLE)
        boff, data = c.barle()
        c.hour_loop(sg.FMT_BARLE, 64)
        host_want = c.irradiance()
    with _ctx(sg.assemble_scene(heights, 1.0)) as b:
        b.load_lists(path)   # the scene check covers grid, parents and walls, not the angles
        with pytest.raises(sg.SolarGPUError):
            b.csr()
        bo2, d2 = b.barle()
        np.testing.assert_array_equal(bo2, boff)
        np.testing.assert_array_equal(d2, data)
        b.hour_loop(sg.FMT_BARLE, 64)
        np.testing.assert_array_equal(b.irradiance(), host_want)
    # The device scene's slopes and aspects come from the device atan2f (within 3 ulp of the
    # host's), so its annual sums agree with the host scene's to float rounding, not bit for bit.
    np.testing.assert_allclose(want, host_want, rtol=1e-5)


def test_files_load_only_into_their_scene_and_sky(tmp_path):
    sc = synth_scene(33, 64, 48)
    path = str(tmp_path / "lists.sgl")
    with _ctx(sc) as a:
        a.viewshed(sg.FMT_CSR)
        a.save_lists(path, sg.FMT_CSR)
    other = synth_scene(34, 64, 48)   # same dimensions, other heights
    with _ctx(other) as b, pytest.raises(sg.SolarGPUError, match="another scene"):
        b.load_lists(path)
    higher = dataclasses.replace(sc, max_height=sc.max_height + 5.0)   # rays leave at another height
    with _ctx(higher) as b, pytest.raises(sg.SolarGPUError, match="another scene"):
        b.load_lists(path)
    az_alt = sg.sun_positions(-70.0, -33.0, -4)   # another site: other sun bins
    dni, dhi = sg.synth_weather(5, az_alt)
    with _ctx(sc, (az_alt, dni, dhi, None)) as b, pytest.raises(sg.SolarGPUError, match="another sky"):
        b.load_lists(path)


def test_damaged_files_are_refused(tmp_path):
    sc = synth_scene(35, 48, 40)
    path = tmp_path / "lists.sgl"
    with _ctx(sc) as a:
        a.viewshed(sg.FMT_CSR | sg.FMT_BARLE)
        a.save_lists(str(path))
    raw = path.read_bytes()
    (tmp_path / "short.sgl").write_bytes(raw[: len(raw) - 7])
    (tmp_path / "magic.sgl").write_bytes(b"NOTLISTS" + raw[8:])
    with _ctx(sc) as b:
        for name in ("short.sgl", "magic.sgl", "missing.sgl"):
            with pytest.raises(sg.SolarGPUError):
                b.load_lists(str(tmp_path / name))
        b.load_lists(str(path))   # the intact file still loads
