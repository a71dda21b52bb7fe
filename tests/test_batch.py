"""Batched tiles: several DSM tiles of different sizes, sites and skies through one
persistent viewshed launch per chunk and one hipGraph hour loop.  Each tile's lists,
SVF bits, ba-RLE streams and annual sums must equal those of a context of its own
(bit-exact), and the lists those of the oracle."""
import numpy as np
import pytest

import solargpu_amd as sg
import oracle_lib as orc
from scenes import synth_scene

pytestmark = pytest.mark.gpu

# (seed, width, height, lon, lat, tz): uneven sizes, one tile with no walls at all
TILES = [(31, 40, 37, 15.65, 46.56, 1), (32, 64, 64, -3.7, 40.4, 0), (33, 23, 71, 24.9, 60.2, 2),
         (34, 96, 50, 2.35, 48.86, 1), (35, 17, 19, 12.5, 41.9, 1)]


def _sky(seed, lon, lat, tz):
    az_alt = sg.sun_positions(lon, lat, tz)
    dni, dhi = sg.synth_weather(seed, az_alt)
    return az_alt, dni, dhi


def _tiles():
    out = []
    for seed, w, h, lon, lat, tz in TILES:
        sc = synth_scene(seed, w, h) if seed != 35 else sg.assemble_scene(np.full((h, w), 210.0, np.float32), 1.0)
        out.append((sc, _sky(seed, lon, lat, tz)))
    return out


@pytest.mark.parametrize("fmt,K,waves,queued", [(sg.FMT_BARLE, 64, 4, 1), (sg.FMT_CSR, 16, 2, 1),
                                                (sg.FMT_BARLE, 128, 2, 0), (sg.FMT_BARLE, 256, 4, 0),
                                                (sg.FMT_BARLE, 64, 4, 2), (sg.FMT_BARLE, 256, 2, 2)])
def test_batch_matches_single_tile_contexts(fmt, K, waves, queued):
    """Queued (receivers one by one from per-XCD queues over every tile) and static receiver
    assignments both give each tile's own annual sums bit for bit; so do queued receiver pairs
    against contexts that walk pairs too."""
    tiles = _tiles()
    with sg.Batch(0) as b:
        for sc, sky in tiles:
            b.add_tile(sc, *sky)
        b.viewshed()
        b.set_receiver_queue(queued)
        b.hour_loop(fmt, K, waves, graph=True)
        inf = b.info()
        assert inf["tiles"] == len(tiles) and inf["passes"] >= 1
        for t, (sc, (az, dni, dhi)) in enumerate(tiles):
            with sg.Context(0) as ctx:
                ctx.set_scene(sc)
                ctx.set_sky(az, dni, dhi)
                ctx.viewshed(sg.FMT_CSR | sg.FMT_BARLE)
                off, ids = b.csr(t)
                eoff, eids = ctx.csr()
                np.testing.assert_array_equal(off, eoff, err_msg=f"tile {t} offsets")
                np.testing.assert_array_equal(ids, eids, err_msg=f"tile {t} ids")
                np.testing.assert_array_equal(b.svf(t), ctx.svf(), err_msg=f"tile {t} svf")
                boff, data = b.barle(t)
                e_boff, e_data = ctx.barle()
                np.testing.assert_array_equal(boff, e_boff)
                np.testing.assert_array_equal(data, e_data)
                ctx.set_tuning(16, K, waves, 0, 0, 2 if queued == 2 else 0)
                ctx.hour_loop(fmt, K)
                np.testing.assert_array_equal(b.irradiance(t), ctx.irradiance(), err_msg=f"tile {t} annual")


def test_batch_lists_match_oracle_and_graph_matches_eager():
    tiles = _tiles()[:3]
    with sg.Batch(0) as b:
        for sc, sky in tiles:
            b.add_tile(sc, *sky)
        b.viewshed()
        for t, (sc, (az, dni, dhi)) in enumerate(tiles):
            ref = orc.viewshed(sc, orc.sky_bins(az))
            off, ids = b.csr(t)
            np.testing.assert_array_equal(off, ref["off"])
            np.testing.assert_array_equal(ids, ref["ids"])
            np.testing.assert_array_equal(b.svf(t), ref["svf"])
        b.hour_loop(sg.FMT_BARLE, 32, 4, graph=False)
        eager = [b.irradiance(t) for t in range(len(tiles))]
        for _ in range(2):
            b.hour_loop(sg.FMT_BARLE, 32, 4, graph=True)
            for t in range(len(tiles)):
                np.testing.assert_array_equal(b.irradiance(t), eager[t])
        sc, (az, dni, dhi) = tiles[0]
        bins = orc.sky_bins(az)
        ref = orc.viewshed(sc, bins)
        exact = orc.annual_fp64(sc, bins, dni, dhi, ref["svf"], ref["off"], ref["ids"])
        rel = np.abs(eager[0].astype(np.float64) - exact) / np.maximum(np.abs(exact), 1.0)
        assert rel.max() <= 1e-4
