"""The step-table trace (k_trace_table + k_trace_pending) against the pooled walk.

The tables rest on one fact of float arithmetic: for a cell coordinate g two or more from both
ends of its binade [2^e, 2^(e+1)), e >= 3, the ray origin g + 2c rounds to g + (2c rounded to the
binade's spacing), so every such g has the same set-up (start cell - g, first-edge distances) along
every direction.  The CPU test checks that over all 2592 directions and every such g up to 2^12;
the GPU tests check that the viewshed by tables equals the pooled walk's bit for bit (lists, SVF,
ba-RLE streams with the skips of the max maps; step counts, which walk every step), over one and
several chunks, and through the overflow path.
"""
import numpy as np
import pytest

import solargpu_amd as sg
from scenes import sky, synth_scene


def _classes(e):
    lo, hi = 1 << e, (2 << e) - 1
    return np.arange(lo + 2, hi - 1, dtype=np.int64)   # g - lo >= 2 and hi - g >= 2


def test_ray_setup_is_shared_within_a_binade_class():
    cdir, _ = sg.directions()
    c2 = (cdir[:, :2].astype(np.float32) * np.float32(2.0))          # exact doubling
    for e in range(3, 12):
        g = _classes(e).astype(np.float32)
        for axis in range(2):
            ox = g[None, :] + c2[:, axis:axis + 1]                   # float32 sums, rounded once
            assert ox.dtype == np.float32
            ix = np.floor(ox)
            frac_hi = (ix + np.float32(1.0)) - ox                   # the set-up numerators
            frac_lo = ox - ix
            for v in (ix - g[None, :], frac_hi, frac_lo):
                assert np.all(v == v[:, :1]), f"binade {e}: set-up differs across the class"
        # a cell next to the binade end is outside the class: some direction sets it up differently
        lo = np.float32((1 << e) + 1)
        dx = (lo + c2[:, 0]) - lo
        dref = (g[0] + c2[:, 0]) - g[0]
        assert np.any(dx != dref)


def _run(sc, tables, chunk=0, cap=0, steps=False, shape=(4, 255, True)):
    az_alt, dni, dhi, _ = sky()
    ctx = sg.Context(0)
    ctx.set_scene(sc)
    ctx.set_sky(az_alt, dni, dhi)
    ctx.set_trace_tables(tables, cap)
    ctx.set_trace_table_shape(*shape)
    ctx.set_step_counting(steps)
    if chunk:
        ctx.set_chunk_voxels(chunk)
    ctx.viewshed(sg.FMT_CSR | sg.FMT_BARLE)
    out = dict(csr=ctx.csr(), barle=ctx.barle(), svf=ctx.svf(), info=ctx.info())
    ctx.close()
    return out


def _same(a, b):
    for x, y in zip(a["csr"], b["csr"]):
        np.testing.assert_array_equal(x, y)
    for x, y in zip(a["barle"], b["barle"]):
        np.testing.assert_array_equal(x, y)
    np.testing.assert_array_equal(a["svf"], b["svf"])


@pytest.mark.gpu
@pytest.mark.parametrize("seed,w,h,cell,density,trees", [
    (31, 300, 200, 1.0, 0.35, 8),      # widths across the 256 binade end, rows across 128
    (32, 97, 45, 1.0, 0.45, 2),        # small: most cells irregular
    (33, 1100, 70, 0.5, 0.45, 20),     # columns across 1024, long rays at 0.5 m
])
def test_tables_equal_pooled_walk(seed, w, h, cell, density, trees):
    """With skipping (the lists) and without (the step counts: counting walks every step)."""
    sc = synth_scene(seed, w, h, cell, density, trees)
    print(f"scene seed {seed} {w}x{h}: {sc.num_total} voxels", flush=True)
    p = _run(sc, False, steps=True)
    for shape in [(4, 255, True), (4, 510, True), (8, 126, False)]:
        t = _run(sc, True, shape=shape)
        _same(t, p)
        ti = t["info"]
        assert ti["table_chunks"] == ti["viewshed_chunks"] >= 1 and p["info"]["table_chunks"] == 0
        assert ti["table_fallbacks"] == 0
        assert ti["table_irregular_voxels"] > 0
    t = _run(sc, True, steps=True)
    _same(t, p)
    assert t["info"]["dda_steps"] == p["info"]["dda_steps"] > 0


@pytest.mark.gpu
def test_tables_over_chunks_with_pending_rays():
    """Several chunks of one class block each, and rays longer than a table (pending)."""
    sc = synth_scene(34, 400, 400, 0.5, 0.45, 30)
    print(f"scene 400x400 @ 0.5 m: {sc.num_total} voxels", flush=True)
    assert sc.num_total > 3 * 65536
    p = _run(sc, False, chunk=65536, steps=True)
    t = _run(sc, True, chunk=65536)
    _same(t, p)
    ti = t["info"]
    assert ti["viewshed_chunks"] == ti["table_chunks"] == -(-sc.num_total // 65536)
    assert ti["table_pending_rays"] > 0 and ti["table_fallbacks"] == 0
    t = _run(sc, True, chunk=65536, steps=True)
    _same(t, p)
    assert t["info"]["dda_steps"] == p["info"]["dda_steps"]


@pytest.mark.gpu
def test_pending_overflow_traces_the_chunk_again():
    """A pending capacity of one ray overflows on every chunk: each is traced again by the pooled
    walk (with the step count restored), and the lists are unchanged."""
    sc = synth_scene(34, 400, 400, 0.5, 0.45, 30)
    p = _run(sc, False, chunk=65536, steps=True)
    for steps in (False, True):
        t = _run(sc, True, chunk=65536, cap=1, steps=steps)
        _same(t, p)
        ti = t["info"]
        assert ti["table_fallbacks"] == ti["table_chunks"] == ti["viewshed_chunks"]
        if steps:
            assert ti["dda_steps"] == p["info"]["dda_steps"]


@pytest.mark.gpu
def test_unaligned_chunks_take_the_pooled_walk():
    sc = synth_scene(35, 300, 300, 1.0, 0.35, 8)
    t = _run(sc, True, chunk=100000)
    p = _run(sc, False)
    _same(t, p)
    assert t["info"]["table_chunks"] == 0
