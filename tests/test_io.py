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
