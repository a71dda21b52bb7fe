"""Oracle and product host code against the golden vectors written from the
reference's own host sources (tests/golden/make_golden.py)."""
import hashlib
import os

import numpy as np
import pytest

import solargpu_amd as sg
import oracle_lib as orc

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SKY = np.load(os.path.join(GOLD, "sky.npz"))
VOX = np.load(os.path.join(GOLD, "voxelization.npz"))
BARLE = np.load(os.path.join(GOLD, "barle.npz"))
SITES = sorted(k[len("site_"):] for k in SKY.files if k.startswith("site_"))
DSMS = sorted({k.rsplit("_heights", 1)[0] for k in VOX.files if k.endswith("_heights")})


def test_fixture_hashes():
    with open(os.path.join(GOLD, "SHA256SUMS")) as f:
        for line in f:
            digest, name = line.split()
            with open(os.path.join(GOLD, name), "rb") as g:
                assert hashlib.sha256(g.read()).hexdigest() == digest, name


@pytest.mark.parametrize("site", SITES)
def test_sun_positions_bit_exact(site):
    lon, lat, tz = SKY[f"site_{site}"]
    gold = SKY[f"sun_{site}"]
    for got in (orc.sun_positions(lon, lat, int(tz)), sg.sun_positions(lon, lat, int(tz))):
        np.testing.assert_array_equal(got.view(np.uint32), gold.view(np.uint32))


@pytest.mark.parametrize("site", SITES)
def test_sky_bins_bit_exact(site):
    az = SKY[f"sun_{site}"]
    act, hgs = SKY[f"bins_{site}"]
    o = orc.sky_bins(az)
    p = sg.sky_bins(az)
    for f2d, sdi, a, b, nud in ((o["full_to_dense"], o["sun_dense_idx"], o["active"], o["svf_bytes"], o["nudged"]),
                               (p.full_to_dense, p.sun_dense_idx, p.active_bins, p.svf_bytes, p.nudged)):
        np.testing.assert_array_equal(f2d, SKY[f"f2d_{site}"])
        np.testing.assert_array_equal(sdi, SKY[f"sdi_{site}"])
        assert (a, b) == (act, hgs)
        assert hashlib.sha256(np.ascontiguousarray(nud, np.float32).tobytes()).digest() == \
            SKY[f"nudged_sha256_{site}"].tobytes()


def test_config_site_counts():
    """Sizing facts of the benchmark site: 265 sun bins, 34 SVF bytes, 4337 sunlit hours."""
    o = orc.sky_bins(SKY["sun_config"])
    assert (o["active"], o["svf_bytes"]) == (265, 34)
    assert orc.sunlit_hours(o["nudged"]).size == 4337


@pytest.mark.parametrize("name", DSMS)
def test_voxelization_bit_exact(name):
    h = VOX[f"{name}_heights"]
    sx, sy = VOX[f"{name}_scale"]
    gp, gz, ga = VOX[f"{name}_wall_parent"], VOX[f"{name}_wall_z"], VOX[f"{name}_wall_aspect"]
    gs, gasp = VOX[f"{name}_slope"], VOX[f"{name}_aspect"]
    p, z, a = orc.detect_walls(h, sx)
    np.testing.assert_array_equal(p, gp)
    np.testing.assert_array_equal(z.view(np.uint32), gz.view(np.uint32))
    np.testing.assert_array_equal(a.view(np.uint32), ga.view(np.uint32))
    s, asp = orc.top_slope_aspect(h, sx, sy)
    np.testing.assert_array_equal(s.view(np.uint32), gs.view(np.uint32))
    np.testing.assert_array_equal(asp.view(np.uint32), gasp.view(np.uint32))
    # product host path: the voxel arrays of main.cpp:112-127
    sc = sg.assemble_scene(h, sx, sy)
    nt = h.size
    np.testing.assert_array_equal(sc.pixel_parent[:nt], np.arange(nt))
    np.testing.assert_array_equal(sc.pixel_parent[nt:], gp)
    np.testing.assert_array_equal(sc.wall_z.view(np.uint32), gz.view(np.uint32))
    np.testing.assert_array_equal(sc.slopes[:nt].view(np.uint32), gs.view(np.uint32))
    np.testing.assert_array_equal(sc.aspects[:nt].view(np.uint32), gasp.view(np.uint32))
    np.testing.assert_array_equal(sc.aspects[nt:].view(np.uint32), ga.view(np.uint32))
    assert np.all(sc.slopes[nt:] == np.float32(np.float32(np.pi) * np.float32(0.5)))


def _barle_rows():
    raw, offs, data = BARLE["raw"], BARLE["offsets"], BARLE["bytes"]
    for i in range(raw.shape[0]):
        yield raw[i], data[offs[i]:offs[i + 1]]


def test_barle_encoder_bit_exact_on_golden_streams():
    n = 0
    for row, gold in _barle_rows():
        np.testing.assert_array_equal(orc.barle_encode_voxel(row), gold)
        n += 1
    assert n == BARLE["raw"].shape[0]


def test_barle_encode_all_matches_golden_layout():
    raw = BARLE["raw"]
    lists = []
    for row in raw:
        nz = np.flatnonzero(row == 0)
        lists.append(row[: nz[0]] if nz.size else row)
    off = np.zeros(len(lists) + 1, np.int64)
    off[1:] = np.cumsum([len(x) for x in lists])
    boff, data = orc.barle_encode_all(off, np.concatenate(lists))
    np.testing.assert_array_equal(boff, BARLE["offsets"].astype(np.int64))
    np.testing.assert_array_equal(data, BARLE["bytes"])


def test_barle_decoder_round_trip_on_golden_streams():
    for row, gold in _barle_rows():
        nz = np.flatnonzero(row == 0)
        ids = row[: nz[0]] if nz.size else row
        want = np.unique(ids.astype(np.uint32))[::-1]
        got = orc.barle_decode(gold)
        np.testing.assert_array_equal(got[got > 0], want)
