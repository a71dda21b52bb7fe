"""The 2048 x 2048 / 0.5 m workload the benchmark measures (seeded DSM, 45% buildings,
11.5M voxels, 2.6e9 neighbour ids), checked against the oracle on a random sample of voxels:
their neighbour lists and SVF bits equal the oracle's first-hit traces, their ba-RLE streams
equal the oracle encoder's bytes for the same lists (and decode back to them), and their
annual sums lie within 1e-4 of the FP64 oracle."""
import numpy as np
import pytest

import solargpu_amd as sg
import oracle_lib as orc

pytestmark = pytest.mark.gpu

SEED, SITE = 20260102, (15.65, 46.56, 1)
W, CELL, DENSITY = 2048, 0.5, 0.45
ANNUAL_RTOL = 1e-4
SAMPLE = 2048


def _first_occurrences(row):
    hits = row[row != 0]
    _, first = np.unique(hits, return_index=True)
    return hits[np.sort(first)]


def test_bench_workload_sampled_voxels_match_oracle():
    trees = int(round(40 * (W * CELL / 256.0) ** 2))
    heights = sg.synth_dsm(SEED, W, W, CELL, DENSITY, trees)
    sc = sg.assemble_scene(heights, CELL)
    az_alt = sg.sun_positions(*SITE)
    dni, dhi = sg.synth_weather(SEED, az_alt)
    bins = orc.sky_bins(az_alt)
    _, _, dense = orc.directions(bins["full_to_dense"])
    rng = np.random.default_rng(SEED)
    picks = np.sort(np.concatenate([rng.choice(sc.num_top, SAMPLE // 2, replace=False),
                                    rng.choice(np.arange(sc.num_top, sc.num_total), SAMPLE // 2, replace=False)]))
    with sg.Context(0) as ctx:
        ctx.set_scene(sc)
        ctx.set_sky(az_alt, dni, dhi)
        ctx.viewshed(sg.FMT_CSR | sg.FMT_BARLE)
        off, ids = ctx.csr()
        boff, data = ctx.barle()
        svf = ctx.svf()
        ctx.set_tuning(32, 128, 4, 256, 32)
        ctx.hour_loop(sg.FMT_BARLE, 128)
        irr = ctx.irradiance()
    N, nb = sc.num_total, bins["svf_bytes"]
    assert off.size == N + 1 and boff.size == N + 1 and svf.size == nb * N
    for v in picks:
        raw = orc.trace_raw(sc, bins, int(v), 1)[0]
        row = ids[off[v]:off[v + 1]]
        np.testing.assert_array_equal(row, _first_occurrences(raw), err_msg=f"voxel {v}")
        want = np.zeros(nb, np.uint8)
        for d in np.nonzero((raw != 0) & (dense >= 0))[0]:
            want[dense[d] >> 3] |= np.uint8(1 << (dense[d] & 7))
        np.testing.assert_array_equal(svf[np.arange(nb) * N + v], want, err_msg=f"SVF of voxel {v}")
        stream = data[boff[v]:boff[v + 1]]
        np.testing.assert_array_equal(stream, orc.barle_encode_voxel(row), err_msg=f"ba-RLE of voxel {v}")
        np.testing.assert_array_equal(np.sort(orc.barle_decode(stream).astype(np.int64)),
                                      np.sort(row.astype(np.int64)), err_msg=f"decode of voxel {v}")
    exact = orc.annual_fp64(sc, bins, dni, dhi, svf, off, ids, voxels=picks.astype(np.int32))
    rel = np.abs(irr[picks].astype(np.float64) - exact) / np.maximum(np.abs(exact), 1.0)
    assert rel.max() <= ANNUAL_RTOL, (rel.max(), int(picks[rel.argmax()]))
