"""Native GeoTIFF / TMY / voxel-dump I/O (the reference CLI's file contract)."""
import os

import numpy as np
import pytest

import solargpu_amd as sg

REF_DATA = "/root/reference/test_data"


@pytest.mark.skipif(not os.path.isdir(REF_DATA), reason="reference test data not mounted")
def test_reference_dsm_matches_pil_decode():
    """The reference's tiled, deflate + floating-point-predictor DSM, decoded natively,
    equals PIL's independent decode."""
    from PIL import Image
    path = os.path.join(REF_DATA, "dsm_heightmap.tiff")
    g = sg.GeoTiff(path)
    ours = g.read()
    pil = np.asarray(Image.open(path), dtype=np.float32)
    assert ours.shape == (320, 366)
    np.testing.assert_array_equal(ours, pil)
    assert g.scale == (1.0, 1.0) and g.compression in (8, 32946)


@pytest.mark.skipif(not os.path.isdir(REF_DATA), reason="reference test data not mounted")
def test_reference_tmy_series():
    d = sg.read_tmy(os.path.join(REF_DATA, "tmy_direct.txt"))
    f = sg.read_tmy(os.path.join(REF_DATA, "tmy_diffuse.txt"))
    assert d.size == f.size == 8760
    assert abs(d.sum() / 1000 - 719.9) < 0.1 and abs(f.sum() / 1000 - 558.8) < 0.1


def _tiny_tiff(tmp_path):
    """A float32 GeoTIFF written by PIL (strips, no compression) with a pixel-scale tag."""
    from PIL import Image
    from PIL.TiffImagePlugin import ImageFileDirectory_v2
    data = (200 + np.arange(12 * 9, dtype=np.float32).reshape(9, 12) * 0.25).astype(np.float32)
    ifd = ImageFileDirectory_v2()
    ifd[33550] = (0.5, 0.5, 0.0)
    ifd.tagtype[33550] = 12
    p = str(tmp_path / "tiny.tif")
    Image.fromarray(data, mode="F").save(p, tiffinfo=ifd)
    return p, data


def test_round_trip_through_writer(tmp_path):
    p, data = _tiny_tiff(tmp_path)
    g = sg.GeoTiff(p)
    np.testing.assert_array_equal(g.read(), data)
    assert g.scale == (0.5, 0.5)
    out = str(tmp_path / "out.tif")
    g.write(out, data * 2)
    h = sg.GeoTiff(out)
    np.testing.assert_array_equal(h.read(), data * 2)
    assert h.scale == (0.5, 0.5)
    from PIL import Image
    np.testing.assert_array_equal(np.asarray(Image.open(out), np.float32), data * 2)


def test_errors_are_status_codes(tmp_path):
    with pytest.raises(sg.SolarGPUError) as e:
        sg.GeoTiff(str(tmp_path / "missing.tif"))
    assert e.value.code == 5
    bad = tmp_path / "short.txt"
    bad.write_text("1 2 3\n")
    with pytest.raises(sg.SolarGPUError) as e:
        sg.read_tmy(str(bad))
    assert e.value.code == 6
    notiff = tmp_path / "x.tif"
    notiff.write_bytes(b"hello world")
    with pytest.raises(sg.SolarGPUError):
        sg.GeoTiff(str(notiff))


def test_voxel_dump_layout(tmp_path):
    h = np.full((4, 5), 200.0, np.float32)
    h[1, 1] = 205.0
    sc = sg.assemble_scene(h, 1.0)
    irr = np.arange(sc.num_total, dtype=np.float32)
    p = str(tmp_path / "v.bin")
    sg.write_voxel_dump(p, sc, irr)
    raw = open(p, "rb").read()
    n_total, n_top = np.frombuffer(raw[:8], np.int32)
    assert (n_total, n_top) == (sc.num_total, 20)
    rec = np.frombuffer(raw[8:], np.float32).reshape(-1, 4)
    top = sc.pixel_parent
    np.testing.assert_array_equal(rec[:, 0], top % 5)
    np.testing.assert_array_equal(rec[:, 1], top // 5)
    np.testing.assert_array_equal(rec[:20, 2], h.ravel())
    np.testing.assert_array_equal(rec[20:, 2], sc.wall_z)
    np.testing.assert_array_equal(rec[:, 3], irr)


@pytest.mark.gpu
def test_cli_matches_drop_in_entry_point(tmp_path):
    """The drop-in CLI with the reference's arguments (DSM, lon, lat, tz, direct and diffuse
    series) on a seeded DSM written as a GeoTIFF: its solar-potential raster and voxel dump hold
    the ba-RLE pipeline's annual sums, those of compressed_gpu_irradiance on the same scene (to
    FP32 regrouping of the hours, should the processes' cached launch shapes differ)."""
    import subprocess
    from PIL import Image
    from PIL.TiffImagePlugin import ImageFileDirectory_v2
    cli = os.path.join(os.path.dirname(sg.__file__), "bin", "solargpu")
    assert os.path.exists(cli), "the CLI is built with the library (make)"
    heights = sg.synth_dsm(41, 48, 40, 1.0, 0.35, 3)
    ifd = ImageFileDirectory_v2()
    ifd[33550] = (1.0, 1.0, 0.0)
    ifd.tagtype[33550] = 12
    dsm = str(tmp_path / "scene.tif")
    Image.fromarray(heights, mode="F").save(dsm, tiffinfo=ifd)
    lon, lat, tz = 15.65, 46.56, 1
    az_alt = sg.sun_positions(lon, lat, tz)
    dni, dhi = sg.synth_weather(41, az_alt)
    np.savetxt(tmp_path / "direct.txt", dni, fmt="%.9g")
    np.savetxt(tmp_path / "diffuse.txt", dhi, fmt="%.9g")
    r = subprocess.run([cli, dsm, str(lon), str(lat), str(tz), "direct.txt", "diffuse.txt"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    sc = sg.assemble_scene(heights, 1.0)
    want = sg.compressed_gpu_irradiance(sc, az_alt, dni, dhi)
    H, W = heights.shape
    raster = sg.GeoTiff(str(tmp_path / "scene.tif_solar_potential.tiff")).read()
    assert raster.shape == (H, W)
    np.testing.assert_allclose(raster.ravel(), want[: W * H], rtol=1e-5, atol=1e-3)
    raw = open(tmp_path / "scene.tif_voxels.bin", "rb").read()
    n_total, n_top = np.frombuffer(raw[:8], np.int32)
    assert (n_total, n_top) == (sc.num_total, W * H)
    rec = np.frombuffer(raw[8:], np.float32).reshape(-1, 4)
    np.testing.assert_allclose(rec[:, 3], want, rtol=1e-5, atol=1e-3)


def _tiff_bytes(w, h, strip_bytes, scale_type=12, extra=()):
    """A little-endian, uncompressed, single-strip float32 TIFF written field by field, with a
    ModelPixelScale tag of 1 m (stored as `scale_type`).  `strip_bytes` is what the
    StripByteCounts tag claims; `extra` adds (tag, type, count, value-or-offset) entries verbatim."""
    import struct
    pixels = np.arange(w * h, dtype="<f4").tobytes()
    entries = [(256, 4, 1, w), (257, 4, 1, h), (258, 3, 1, 32), (259, 3, 1, 1), (262, 3, 1, 1),
               (273, 4, 1, None), (277, 3, 1, 1), (278, 4, 1, h), (279, 4, 1, strip_bytes), (339, 3, 1, 3),
               (33550, scale_type, 3, None)]
    entries = sorted(entries + list(extra))
    ifd_at = 8
    scale_at = ifd_at + 2 + 12 * len(entries) + 4
    data_at = scale_at + 24
    out = bytearray(b"II" + struct.pack("<HI", 42, ifd_at))
    out += struct.pack("<H", len(entries))
    for tag, typ, count, val in entries:
        if tag == 273:
            val = data_at
        elif tag == 33550:
            val = scale_at
        out += struct.pack("<HHII", tag, typ, count, val)
    out += struct.pack("<I", 0)
    out += struct.pack("<3d", 1.0, 1.0, 0.0)
    out += pixels[:strip_bytes]
    return bytes(out)


def test_damaged_tiffs_are_refused_before_reading_past_the_file(tmp_path):
    good = tmp_path / "good.tif"
    good.write_bytes(_tiff_bytes(64, 64, 64 * 64 * 4))
    np.testing.assert_array_equal(sg.GeoTiff(str(good)).read().ravel(), np.arange(64 * 64, dtype=np.float32))
    short = tmp_path / "short.tif"          # StripByteCounts = 16 for a 16 KiB strip
    short.write_bytes(_tiff_bytes(64, 64, 16))
    with pytest.raises(sg.SolarGPUError) as e:
        sg.GeoTiff(str(short)).read()
    assert e.value.code == 6
    # GeoAsciiParams with an unknown type (13) and a count that reaches far past the file
    odd = tmp_path / "odd.tif"
    odd.write_bytes(_tiff_bytes(8, 8, 8 * 8 * 4, extra=[(34737, 13, 1 << 20, 0x10)]))
    with pytest.raises(sg.SolarGPUError) as e:
        sg.GeoTiff(str(odd)).read()
    assert e.value.code == 6
    # ModelPixelScale stored as LONG instead of DOUBLE
    bad_type = tmp_path / "bad_type.tif"
    bad_type.write_bytes(_tiff_bytes(8, 8, 8 * 8 * 4, scale_type=4))
    with pytest.raises(sg.SolarGPUError) as e:
        sg.GeoTiff(str(bad_type)).read()
    assert e.value.code == 6
