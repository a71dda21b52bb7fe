"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Bars: neighbour lists, SVF bitmaps and ba-RLE streams bit-exact; per-hour
hourly/reflected irradiance within 1e-4 relative + 1e-3 W/m^2 (device fast
reciprocals and angle-addition cosine vs host libm);
annual irradiance within 1e-4 relative of the FP64 oracle (floor 1 Wh/m^2).
"""
import numpy as np
import pytest

import solargpu_amd as sg
import oracle_lib as orc
from scenes import oracle_lists, sky, synth_scene, scene_from_heights

pytestmark = pytest.mark.gpu

ANNUAL_RTOL = 1e-4
HOUR_RTOL = 1e-4


def _ctx(sc, formats=sg.FMT_CSR | sg.FMT_BARLE, chunk=0):
    az_alt, dni, dhi, _ = sky()
    ctx = sg.Context(0)
    ctx.set_scene(sc)
    ctx.set_sky(az_alt, dni, dhi)
    if chunk:
        ctx.set_chunk_voxels(chunk)
    if formats:
        ctx.viewshed(formats)
    return ctx


def _rel(a, b, floor=1.0):
    return np.abs(a.astype(np.float64) - b) / np.maximum(np.abs(b), floor)


@pytest.mark.parametrize("seed,w,h", [(11, 48, 40), (12, 64, 64), (13, 37, 53)])
def test_viewshed_lists_svf_bit_exact(seed, w, h):
    sc = synth_scene(seed, w, h)
    ref = oracle_lists(seed, w, h)
    with _ctx(sc) as ctx:
        off, ids = ctx.csr()
        svf = ctx.svf()
        np.testing.assert_array_equal(off, ref["off"])
        np.testing.assert_array_equal(ids, ref["ids"])
        np.testing.assert_array_equal(svf, ref["svf"])


@pytest.mark.parametrize("G,chunk", [(32, 1024), (48, 0), (16, 1536), (32, 777)])
def test_trace_tile_shapes_and_lane_order_are_results_neutral(G, chunk):
    """Every directions-per-workgroup shape of the pooled trace, with chunks that are whole
    lane-order blocks (height-sorted lanes) and one that is not (lanes in id order)."""
    sc = synth_scene(12, 64, 64)
    ref = oracle_lists(12, 64, 64)
    az_alt, dni, dhi, _ = sky()
    with sg.Context(0) as ctx:
        ctx.set_scene(sc)
        ctx.set_sky(az_alt, dni, dhi)
        ctx.set_tuning(G, 64, 4, 0, 0)
        if chunk:
            ctx.set_chunk_voxels(chunk)
        ctx.viewshed(sg.FMT_CSR)
        off, ids = ctx.csr()
        np.testing.assert_array_equal(off, ref["off"])
        np.testing.assert_array_equal(ids, ref["ids"])
        np.testing.assert_array_equal(ctx.svf(), ref["svf"])


def test_c1_configuration_end_to_end_vs_oracle():
    """The 256^2 / 1 m configuration (fBm terrain, 35% buildings, 40 trees; ~165k voxels) end to
    end: every CSR list, SVF byte and ba-RLE stream bit-exact against the oracle, and every
    voxel's annual irradiance from the ba-RLE hour loop within 1e-4 of the FP64 oracle."""
    sc = synth_scene(20260201, 256, 256, 1.0, 0.35, 40)
    az_alt, dni, dhi, bins = sky()
    ref = orc.viewshed(sc, bins, nthreads=16)
    eoff, edata = orc.barle_encode_all(ref["off"], ref["ids"], nthreads=16)
    exact = orc.annual_fp64(sc, bins, dni, dhi, ref["svf"], ref["off"], ref["ids"], nthreads=16)
    with _ctx(sc) as ctx:
        off, ids = ctx.csr()
        np.testing.assert_array_equal(off, ref["off"])
        np.testing.assert_array_equal(ids, ref["ids"])
        np.testing.assert_array_equal(ctx.svf(), ref["svf"])
        boff, data = ctx.barle()
        np.testing.assert_array_equal(boff, eoff)
        np.testing.assert_array_equal(data, edata)
        ctx.set_tuning(32, 128, 4, 256, 32)
        ctx.hour_loop(sg.FMT_BARLE, 0)
        irr = ctx.irradiance()
    assert sc.num_total > 150_000 and ids.size > 10 * sc.num_total
    assert _rel(irr, exact).max() <= ANNUAL_RTOL


def test_grid_beyond_32bit_byte_offsets_is_rejected():
    """Grids with W (H + 2) > 2^30 are refused with a status code before any array is read."""
    import ctypes
    g = np.zeros(4, np.float32)
    pp = np.zeros(4, np.int32)
    with sg.Context(0) as ctx:
        for w, h in ((32768, 32767), (1 << 20, 1023)):
            rc = sg._lib.sgpu_set_scene(ctx._h, sg._ptr(g), w, h, ctypes.c_int64(w * h), ctypes.c_int64(w * h),
                                        sg._ptr(pp), None, ctypes.c_float(1.0), sg._ptr(g), sg._ptr(g))
            assert rc != 0 and b"2^30" in sg._lib.sgpu_last_error()


def test_hourly_buffer_beyond_4gib_matches_32bit_rows():
    """2048 x 2100 cells at K = 256 hours per pass need 4.4 GB of hourly rows, addressed with
    64-bit offsets; K = 128 (2.2 GB) uses 32-bit row offsets.  The annual sums agree to FP32
    regrouping of the hours (well inside the 1e-4 bar)."""
    sc = synth_scene(20260203, 2048, 2100, 0.5, 0.45, 40)
    assert sc.num_top * 256 * 4 > (1 << 32) > sc.num_top * 128 * 4
    az_alt, dni, dhi, _ = sky()
    with sg.Context(0) as ctx:
        ctx.set_scene(sc)
        ctx.set_sky(az_alt, dni, dhi)
        ctx.viewshed(sg.FMT_BARLE)
        ctx.set_tuning(32, 128, 4, 256, 32)
        ctx.hour_loop(sg.FMT_BARLE, 128, 8)
        narrow = ctx.irradiance()
        ctx.hour_loop(sg.FMT_BARLE, 256, 8)
        wide = ctx.irradiance()
    assert _rel(wide, narrow).max() <= 1e-5


def test_trace_raw_matches_oracle():
    sc = synth_scene(12, 64, 64)
    bins = sky()[3]
    picks = [0, 1, 63, 64 * 32 + 17, sc.num_top - 1, sc.num_top, sc.num_total - 1]
    with _ctx(sc, formats=0) as ctx:
        for v in picks:
            np.testing.assert_array_equal(ctx.trace_raw(v, 1), orc.trace_raw(sc, bins, v, 1), err_msg=f"voxel {v}")


def test_barle_gpu_encoder_bit_exact():
    sc = synth_scene(12, 64, 64)
    ref = oracle_lists(12, 64, 64)
    eoff, edata = orc.barle_encode_all(ref["off"], ref["ids"])
    with _ctx(sc) as ctx:
        boff, data = ctx.barle()
        np.testing.assert_array_equal(boff, eoff)
        np.testing.assert_array_equal(data, edata)


@pytest.mark.parametrize("passes", ["1", "2"])
def test_barle_encoder_edge_lists(passes, monkeypatch):
    """Crafted lists through the GPU encoder (set_csr -> barle_encode) vs the oracle encoder, in
    one pass into per-voxel slots (the default) and in a size pass + write pass (the path taken
    when the slots do not fit)."""
    monkeypatch.setenv("SGPU_BARLE_PASSES", passes)
    sc = synth_scene(11, 48, 40)
    rng = np.random.default_rng(5)
    lists = [[], [1], [15], [16], [255, 256], [70000, 3], list(range(600, 0, -1)),
             [1 << 25, (1 << 25) - 1, 5, 4], list(range(1, 2592 * 3, 3))[:2592]]
    # the encoder's three list classes: <= 512 ids, 513..1024, longer (and their boundaries)
    for n in (512, 513, 1000, 1024, 1025, 2047):
        lists.append(list(rng.choice(1 << 24, size=n, replace=False) + 1))
    for _ in range(sc.num_total - len(lists)):
        n = int(rng.integers(0, 60))
        lists.append(list(rng.choice(sc.num_top, size=n, replace=False) + 1))
    lists = lists[: sc.num_total]
    off = np.zeros(sc.num_total + 1, np.int64)
    off[1:] = np.cumsum([len(x) for x in lists])
    ids = np.concatenate([np.asarray(x, np.int32) for x in lists]) if off[-1] else np.zeros(0, np.int32)
    eoff, edata = orc.barle_encode_all(off, ids)
    with _ctx(sc, formats=0) as ctx:
        ctx.set_csr(off, ids)
        ctx.barle_encode()
        boff, data = ctx.barle()
    np.testing.assert_array_equal(boff, eoff)
    np.testing.assert_array_equal(data, edata)


def test_repeated_viewsheds_reuse_the_workspace():
    """Viewsheds on one context keep their workspace (raw table, counts, offsets) between calls
    and the encoder borrows its raw table for its slots; a smaller chunk reuses it, a larger one
    replaces it: the lists and streams stay identical."""
    sc = synth_scene(13, 37, 53)
    with _ctx(sc, chunk=512) as ctx:
        off0, ids0 = ctx.csr()
        boff0, data0 = ctx.barle()
        for chunk in (256, 2048, 512):
            ctx.set_chunk_voxels(chunk)
            ctx.viewshed(sg.FMT_CSR | sg.FMT_BARLE)
            off, ids = ctx.csr()
            boff, data = ctx.barle()
            np.testing.assert_array_equal(off, off0)
            np.testing.assert_array_equal(ids, ids0)
            np.testing.assert_array_equal(boff, boff0)
            np.testing.assert_array_equal(data, data0)


def test_chunking_is_results_neutral():
    sc = synth_scene(13, 37, 53)
    ref = oracle_lists(13, 37, 53)
    with _ctx(sc, chunk=777) as ctx:
        assert ctx.info()["viewshed_chunks"] > 1
        off, ids = ctx.csr()
        np.testing.assert_array_equal(off, ref["off"])
        np.testing.assert_array_equal(ids, ref["ids"])
        np.testing.assert_array_equal(ctx.svf(), ref["svf"])
        boff, data = ctx.barle()
        eoff, edata = orc.barle_encode_all(ref["off"], ref["ids"])
        np.testing.assert_array_equal(data, edata)


@pytest.mark.parametrize("fmt", [sg.FMT_CSR, sg.FMT_BARLE])
def test_single_hour_matches_oracle(fmt):
    sc = synth_scene(11, 48, 40)
    ref = oracle_lists(11, 48, 40)
    az_alt, dni, dhi, bins = sky()
    hours = orc.sunlit_hours(bins["nudged"])
    boff, data = orc.barle_encode_all(ref["off"], ref["ids"])
    with _ctx(sc) as ctx:
        for hi in [hours[0], hours[len(hours) // 3], hours[len(hours) // 2], hours[-1]]:
            hv, rv = ctx.hour_single(fmt, int(hi))
            eh = orc.solar_hour(sc, bins, dni, dhi, ref["svf"], int(hi))
            er = (orc.reflect_hour_csr(sc, ref["off"], ref["ids"], eh) if fmt == sg.FMT_CSR
                  else orc.reflect_hour_barle(sc, boff, data, eh))
            assert (np.abs(hv - eh) <= HOUR_RTOL * np.abs(eh) + 1e-3).all(), f"hourly, hour {hi}"
            assert (np.abs(rv - er) <= HOUR_RTOL * np.abs(er) + 1e-4).all(), f"reflected, hour {hi}"


@pytest.mark.parametrize("fmt,K", [(sg.FMT_CSR, 1), (sg.FMT_CSR, 16), (sg.FMT_CSR, 64), (sg.FMT_CSR, 128),
                                   (sg.FMT_BARLE, 1), (sg.FMT_BARLE, 8), (sg.FMT_BARLE, 32), (sg.FMT_BARLE, 64),
                                   (sg.FMT_BARLE, 128), (sg.FMT_BARLE, 256)])
def test_annual_within_1e4_of_fp64(fmt, K):
    sc = synth_scene(12, 64, 64)
    ref = oracle_lists(12, 64, 64)
    az_alt, dni, dhi, bins = sky()
    exact = orc.annual_fp64(sc, bins, dni, dhi, ref["svf"], ref["off"], ref["ids"])
    with _ctx(sc) as ctx:
        ctx.hour_loop(fmt, K)
        inf = ctx.info()
        assert inf["hours_computed"] == len(orc.sunlit_hours(bins["nudged"]))
        irr = ctx.irradiance()
    assert _rel(irr, exact).max() <= ANNUAL_RTOL


@pytest.mark.parametrize("queue", [1, 2])
def test_graph_replays_match_eager_loop(queue):
    """Captured hour loops (receivers one by one, or in pairs with the own-irradiance pass) replay
    to the eager loop's sums bit for bit."""
    sc = synth_scene(11, 48, 40)
    with _ctx(sc) as ctx:
        ctx.set_tuning(32, 16, 4, 32, 0, queue)
        ctx.hour_loop(sg.FMT_BARLE, 16)
        eager = ctx.irradiance()
        for _ in range(3):
            ctx.hour_loop_graph(sg.FMT_BARLE, 16, replays=2)
            np.testing.assert_array_equal(ctx.irradiance(), eager)
            ctx.sync()


def test_drop_in_entry_points():
    sc = synth_scene(11, 48, 40)
    ref = oracle_lists(11, 48, 40)
    az_alt, dni, dhi, bins = sky()
    exact = orc.annual_fp64(sc, bins, dni, dhi, ref["svf"], ref["off"], ref["ids"])
    u = sg.uncompressed_gpu_irradiance(sc, az_alt, dni, dhi)
    c = sg.compressed_gpu_irradiance(sc, az_alt, dni, dhi)
    assert _rel(u, exact).max() <= ANNUAL_RTOL
    assert _rel(c, exact).max() <= ANNUAL_RTOL


def test_flat_plane_has_no_upper_hits():
    """Flat ground: no upward ray can hit; the SVF is all clear; lists hold only ground cells."""
    sc = scene_from_heights(np.full((32, 32), 205.0, np.float32))
    with _ctx(sc) as ctx:
        assert not ctx.svf().any()
        off, ids = ctx.csr()
        ref = orc.viewshed(sc, sky()[3])
        np.testing.assert_array_equal(ids, ref["ids"])


@pytest.mark.parametrize("fmt", [sg.FMT_CSR, sg.FMT_BARLE])
def test_host_streamed_lists_match_device_resident(fmt):
    """Out-of-core mode: lists built chunk by chunk into pinned host memory and streamed
    back per pass (double-buffered copies) give the same lists and bit-identical sums."""
    sc = synth_scene(13, 37, 53)
    ref = oracle_lists(13, 37, 53)
    with _ctx(sc) as ctx:
        ctx.hour_loop(fmt, 8)
        resident = ctx.irradiance()
    with _ctx(sc, formats=0, chunk=500) as ctx:
        ctx.set_list_residency(True)
        ctx.viewshed(sg.FMT_CSR | sg.FMT_BARLE)
        assert ctx.info()["viewshed_chunks"] > 1
        off, ids = ctx.csr()
        np.testing.assert_array_equal(off, ref["off"])
        np.testing.assert_array_equal(ids, ref["ids"])
        boff, data = ctx.barle()
        eoff, edata = orc.barle_encode_all(ref["off"], ref["ids"])
        np.testing.assert_array_equal(boff, eoff)
        np.testing.assert_array_equal(data, edata)
        ctx.hour_loop(fmt, 8)
        np.testing.assert_array_equal(ctx.irradiance(), resident)


@pytest.mark.parametrize("waves", [2, 4])
def test_long_barle_streams_match_cpu_decoder(waves):
    """Injected lists of distinct random cells: 3000 voxels with 0..2592 cells each, 10% of the
    rest with 0..60, the others empty; and 200 voxels with 2000 cells spaced 300 apart (runs of
    2-byte deltas: ~4 KB ba-RLE blocks).  The streams run from nothing to several KB, so the
    reflection kernel slides its LDS stream window many times per wave and decodes the blocks
    larger than a window straight from global memory.  The single-hour reflected irradiance
    matches the CPU ba-RLE decoder + reflection, and the annual sums of the ba-RLE and CSR walks
    agree."""
    sc = synth_scene(14, 1024, 1024)
    rng = np.random.default_rng(7)
    cells = sc.width * sc.height
    n = sc.num_total
    counts = np.where(rng.random(n) < 0.1, rng.integers(0, 60, n), 0)
    longv = rng.choice(n, 3000, replace=False)
    counts[longv] = rng.integers(0, 2593, 3000)   # at most one hit per ray direction
    lists = {int(v): rng.choice(cells, int(counts[v]), replace=False) + 1 for v in np.flatnonzero(counts)}
    for v in longv[:200]:
        lists[int(v)] = 1 + int(rng.integers(0, 1000)) + 300 * np.arange(2000)
    counts = np.zeros(n, np.int64)
    for v, x in lists.items():
        counts[v] = x.size
    off = np.zeros(n + 1, np.int64)
    off[1:] = np.cumsum(counts)
    ids = np.concatenate([lists[v] for v in sorted(lists)]).astype(np.int32)
    boff, data = orc.barle_encode_all(off, ids)
    sizes = np.diff(boff)
    window = 8192 // waves
    assert (sizes > window).sum() > 100 and ((sizes > 0) & (sizes <= window - 15)).sum() > 100_000
    assert (sizes[longv[:200]] >= 2000 * 2).all()   # blocks of 8 x 255 two-byte deltas
    az_alt, dni, dhi, bins = sky()
    hours = orc.sunlit_hours(bins["nudged"])
    with _ctx(sc, formats=sg.FMT_CSR) as ctx:
        svf = ctx.svf()
        ctx.set_csr(off, ids)
        ctx.set_barle(boff, data)
        ctx.set_tuning(32, 16, waves, 64 if waves == 4 else 0)   # XCDs side by side / contiguous
        hi = int(hours[len(hours) // 2])
        hv, rv = ctx.hour_single(sg.FMT_BARLE, hi)
        eh = orc.solar_hour(sc, bins, dni, dhi, svf, hi)
        er = orc.reflect_hour_barle(sc, boff, data, eh)
        assert (np.abs(rv - er) <= HOUR_RTOL * np.abs(er) + 1e-4).all()
        ctx.hour_loop(sg.FMT_BARLE, 0)
        a_barle = ctx.irradiance()
        ctx.hour_loop(sg.FMT_CSR, 0)
        a_csr = ctx.irradiance()
    assert _rel(a_barle, a_csr).max() <= 1e-5


def test_barle_only_build_encodes_chunk_by_chunk():
    """A ba-RLE-only viewshed encodes every chunk's lists at once and keeps no CSR: the
    streams still equal the oracle encoder's byte for byte, the annual sums equal those of the
    combined CSR + ba-RLE build, and asking for the CSR lists is an error."""
    sc = synth_scene(13, 37, 53)
    ref = oracle_lists(13, 37, 53)
    eoff, edata = orc.barle_encode_all(ref["off"], ref["ids"])
    with _ctx(sc) as ctx:
        ctx.hour_loop(sg.FMT_BARLE, 16)
        combined = ctx.irradiance()
    with _ctx(sc, formats=0, chunk=300) as ctx:
        ctx.viewshed(sg.FMT_BARLE)
        assert ctx.info()["viewshed_chunks"] > 1 and ctx.info()["nnz"] == ref["ids"].size
        boff, data = ctx.barle()
        np.testing.assert_array_equal(boff, eoff)
        np.testing.assert_array_equal(data, edata)
        np.testing.assert_array_equal(ctx.svf(), ref["svf"])
        ctx.hour_loop(sg.FMT_BARLE, 16)
        np.testing.assert_array_equal(ctx.irradiance(), combined)
        with pytest.raises(sg.SolarGPUError):
            ctx.csr()
