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


def test_damaged_payload_sections_are_refused(tmp_path):
    """Damage inside the sections (not just the header or the length) is found on load, before
    the lists are used: an offset that breaks monotonicity, a CSR id outside the grid, a ba-RLE
    offset moved by one byte (two streams no longer end on a block boundary) and a ba-RLE width
    code of 5 are each refused with SGPU_ERR_FORMAT, and the context is left without lists."""
    sc = synth_scene(36, 48, 40)
    path = tmp_path / "lists.sgl"
    with _ctx(sc) as a:
        a.viewshed(sg.FMT_CSR | sg.FMT_BARLE)
        a.save_lists(str(path))
        off, ids = a.csr()
        boff, data = a.barle()
    raw = bytearray(path.read_bytes())
    n = sc.num_total
    svf_bytes = sky()[3]["svf_bytes"]
    csr_off_at = 80 + svf_bytes * n
    ids_at = csr_off_at + 8 * (n + 1)
    br_off_at = ids_at + 4 * ids.size
    br_at = br_off_at + 8 * (n + 1)
    assert br_at + data.size == len(raw)
    v = int(np.flatnonzero(np.diff(boff) > 8)[len(boff) // 3])   # a voxel in the middle with a stream

    def damaged(name, at, value, fmt):
        buf = bytearray(raw)
        buf[at:at + len(value)] = value
        p = tmp_path / name
        p.write_bytes(bytes(buf))
        return str(p)

    cases = [
        damaged("csr_off.sgl", csr_off_at + 8 * (v + 1), np.int64(off[v] - 1).tobytes(), "<q"),
        damaged("csr_id0.sgl", ids_at + 4 * 17, np.int32(0).tobytes(), "<i"),
        damaged("csr_idhi.sgl", ids_at + 4 * 23, np.int32(sc.num_top + 1).tobytes(), "<i"),
        damaged("br_off.sgl", br_off_at + 8 * v, np.int64(boff[v] + 1).tobytes(), "<q"),
        damaged("br_code.sgl", br_at + int(boff[v]), bytes([(data[boff[v]] & 0x1F) | (5 << 5)]), "B"),
    ]
    with _ctx(sc) as b:
        for p in cases:
            with pytest.raises(sg.SolarGPUError) as e:
                b.load_lists(p)
            assert e.value.code == 6 and "damaged" in str(e.value), (p, str(e.value))
            with pytest.raises(sg.SolarGPUError) as e:   # nothing half-loaded is left behind
                b.hour_loop(sg.FMT_BARLE, 64)
            assert e.value.code == 4
        b.load_lists(str(path))
        o2, i2 = b.csr()
        np.testing.assert_array_equal(i2, ids)
