"""Pins of the oracle where the reference offers no golden vectors (its CUDA
kernels cannot be built here): an independent vectorised float32 restatement
of the DDA, closed forms (flat plane, a single box, two-cell reflection, a
flat unshaded voxel), the FP64 linearity identity against a per-hour FP64 sum,
and an independent solar-position formula."""
import numpy as np
import pytest
from hypothesis import given, settings, strategies as st

import solargpu_amd as sg
import oracle_lib as orc
from scenes import SITE, scene_from_heights, sky, synth_scene

F32 = np.float32


def numpy_dda(grid, maxz, gx, gy, bz, cdir, upper):
    """Vectorised float32 DDA over many rays (numpy float32 arithmetic is IEEE
    single without contraction).  Returns hit id + 1 (0 = none)."""
    H, W = grid.shape
    flat = grid.ravel()
    cx, cy, cz = (cdir[:, 0].astype(F32), cdir[:, 1].astype(F32), cdir[:, 2].astype(F32))
    ox = gx.astype(F32) + cx * F32(2.0)
    oy = gy.astype(F32) + cy * F32(2.0)
    oz = bz.astype(F32) + cz * F32(2.0)
    ix = np.floor(ox).astype(np.int64)
    iy = np.floor(oy).astype(np.int64)
    ax = np.maximum(np.abs(cx), F32(1e-8))
    ay = np.maximum(np.abs(cy), F32(1e-8))
    tmx = np.where(cx >= 0, (ix.astype(F32) + F32(1.0)) - ox, ox - ix.astype(F32)) / ax
    tmy = np.where(cy >= 0, (iy.astype(F32) + F32(1.0)) - oy, oy - iy.astype(F32)) / ay
    rx, ry = F32(1.0) / ax, F32(1.0) / ay
    sx = np.where(cx >= 0, 1, -1)
    sy = np.where(cy >= 0, 1, -1)
    hit = np.zeros(len(cx), np.int64)
    live = np.ones(len(cx), bool)
    while live.any():
        i = np.flatnonzero(live)
        stepx = tmx[i] < tmy[i]
        t = np.where(stepx, tmx[i], tmy[i])
        ix[i] += np.where(stepx, sx[i], 0)
        iy[i] += np.where(stepx, 0, sy[i])
        tmx[i] = np.where(stepx, tmx[i] + rx[i], tmx[i])
        tmy[i] = np.where(stepx, tmy[i], tmy[i] + ry[i])
        out = (ix[i] < 0) | (ix[i] >= W) | (iy[i] < 0) | (iy[i] >= H)
        zr = oz[i] + cz[i] * t
        stop = out | np.where(upper[i] == 1, zr > F32(maxz), zr < F32(0.0))
        cell = np.where(out, 0, ix[i] + iy[i] * W)
        h = ~stop & ((flat[cell] - zr) > F32(0.01))
        hit[i[h]] = cell[h] + 1
        live[i[stop | h]] = False
    return hit


def test_oracle_dda_matches_independent_float32_restatement():
    sc = synth_scene(21, 20, 17)
    bins = sky()[3]
    cdir, up, _ = orc.directions(bins["full_to_dense"])
    rng = np.random.default_rng(3)
    voxels = np.unique(np.concatenate([rng.choice(sc.num_total, 40, replace=False), [0, sc.num_top - 1,
                                                                                       sc.num_total - 1]]))
    raw = np.concatenate([orc.trace_raw(sc, bins, int(v), 1) for v in voxels])
    top = sc.pixel_parent[voxels]
    gx, gy = top % sc.width, top // sc.width
    bz = np.where(voxels < sc.num_top, sc.grid.ravel()[top], sc.wall_z[np.maximum(voxels - sc.num_top, 0)])
    nv = len(voxels)
    got = numpy_dda(sc.grid, sc.max_height, np.repeat(gx, 2592), np.repeat(gy, 2592), np.repeat(bz, 2592),
                    np.tile(cdir, (nv, 1)), np.tile(up, nv))
    np.testing.assert_array_equal(raw.ravel(), got)


def test_ray_setup_division_is_exact_for_every_direction():
    """The GPU sets up tMax by RN(1/|c|) and two FMAs instead of an IEEE division; for all 2592
    directions, both axes and every origin column/row below 32768 the quotients are the IEEE
    ones bit for bit (an exhaustive check over the domain the kernels serve)."""
    cdir, _, _ = orc.directions(np.full(72 * 18, -1, np.int32))
    assert orc.check_setup_division(np.asarray(cdir).reshape(-1), 32768) == 0


def test_oracle_lists_are_first_occurrence_of_raw_hits():
    sc = synth_scene(22, 16, 14)
    bins = sky()[3]
    ref = orc.viewshed(sc, bins)
    raw = orc.trace_raw(sc, bins, 0, sc.num_total)
    _, _, dense = orc.directions(bins["full_to_dense"])
    for v in range(sc.num_total):
        row = raw[v]
        seen, lst = set(), []
        for h in row:
            if h and h not in seen:
                seen.add(h)
                lst.append(h)
        np.testing.assert_array_equal(ref["ids"][ref["off"][v]:ref["off"][v + 1]], lst)
        # SVF bit set iff the ray of that dense sun bin hit (every hit, duplicates included)
        for d in np.flatnonzero(dense >= 0):
            b = dense[d]
            bit = (ref["svf"][(b // 8) * sc.num_total + v] >> (b % 8)) & 1
            assert bit == (row[d] != 0)


def test_flat_plane_closed_form():
    z0 = 205.0
    sc = scene_from_heights(np.full((20, 24), z0, np.float32))
    bins = sky()[3]
    cdir, up, _ = orc.directions(bins["full_to_dense"])
    ref = orc.viewshed(sc, bins)
    assert not ref["svf"].any()
    raw = orc.trace_raw(sc, bins, 0, sc.num_total)
    assert not raw[:, up == 1].any(), "an upward ray hit a flat plane"
    # every downward ray hits the first cell the DDA enters from the origin cell
    for v in range(sc.num_total):
        gx, gy = v % sc.width, v // sc.width
        ox = F32(gx) + cdir[:, 0] * F32(2)
        oy = F32(gy) + cdir[:, 1] * F32(2)
        cx, cy = np.floor(ox).astype(int), np.floor(oy).astype(int)
        low = np.flatnonzero(up == 0)
        for d in low:
            h = raw[v, d]
            if h == 0:
                continue
            hx, hy = (h - 1) % sc.width, (h - 1) // sc.width
            assert abs(hx - cx[d]) + abs(hy - cy[d]) == 1
        inside = (cx[low] >= 1) & (cx[low] < sc.width - 1) & (cy[low] >= 1) & (cy[low] < sc.height - 1)
        assert (raw[v, low][inside] != 0).all()


def test_single_box_shadow_closed_form():
    """A 50 m box north of a ground voxel: every low-elevation ray towards azimuth
    bins 0 and 71 (due north) hits the box; the matching sun bins are obstructed."""
    h = np.full((40, 40), 200.0, np.float32)
    h[5:18, 12:28] = 250.0
    sc = scene_from_heights(h)
    bins = sky()[3]
    cdir, up, dense = orc.directions(bins["full_to_dense"])
    v = 20 * 40 + 20     # ground cell (20, 20), two cells south of the box's south face (y = 17)
    raw = orc.trace_raw(sc, bins, v, 1)[0]
    ref = orc.viewshed(sc, bins)
    for i in (0, 71):
        for j in range(0, 12):      # elevations 2.5 .. 57.5 deg: rise over <= 15 cells < 50 m
            d = i * 36 + j
            hcell = raw[d] - 1
            assert raw[d] != 0
            hx, hy = hcell % 40, hcell // 40
            assert 12 <= hx < 28 and 5 <= hy < 18
            if dense[d] >= 0:
                b = dense[d]
                assert (ref["svf"][(b // 8) * sc.num_total + v] >> (b % 8)) & 1


def test_reflection_two_cell_closed_form():
    h = np.full((8, 8), 200.0, np.float32)
    h[0, 3] = 210.0
    sc = scene_from_heights(h)
    n = sc.num_total
    off = np.zeros(n + 1, np.int64)
    off[1:] = 1                       # voxel 0 sees cell (3, 0)
    ids = np.array([3 + 1], np.int32)
    hourly = np.zeros(sc.num_top, np.float32)
    hourly[3] = 500.0
    slopes = sc.slopes.copy()
    slopes[0] = 0.0
    sc.slopes[:] = slopes
    r = orc.reflect_hour_csr(sc, off, ids, hourly)
    d2 = 3.0 ** 2 + 10.0 ** 2
    d = np.sqrt(d2)
    want = 0.2 * 500.0 * (10.0 / d) * max(10.0 / d, 0.1) / (np.pi * d2)
    assert abs(r[0] - want) <= 1e-5 * want
    assert np.all(r[1:] == 0)


def _perez_f64(cos_z, cos_aoi, dhi, dni, slope, day):
    """Perez 1990 anisotropic diffuse as the reference applies it, in float64."""
    F = np.array([[-0.0083117, 0.1299457, 0.3296958, 0.5682053, 0.8730280, 1.1326077, 1.0601591, 0.6777470],
                  [0.5877285, 0.6825954, 0.4868735, 0.1874525, -0.3920403, -1.2367284, -1.5999137, -0.3272588],
                  [-0.0620636, -0.1513752, -0.2210958, -0.2951290, -0.3616149, -0.4118494, -0.3589221, -0.2504286],
                  [-0.0596012, -0.0189325, 0.0554140, 0.1088631, 0.2255647, 0.2877813, 0.2642124, 0.1561313],
                  [0.0721249, 0.0659650, -0.0639588, -0.1519229, -0.4620442, -0.8230357, -1.1272340, -1.3765031],
                  [-0.0220216, -0.0288748, -0.0260542, -0.0139754, 0.0012448, 0.0558651, 0.1310694, 0.2506212]])
    z = np.arccos(cos_z)
    am = 1.0 / (cos_z + 0.50572 * (96.07995 - np.degrees(z)) ** -1.6364)
    eps = ((dhi + dni / np.cos(z)) / dhi + 1.041 * z ** 3) / (1 + 1.041 * z ** 3)
    B = 2 * np.pi * day / 365.0
    e0 = 1367 * (1.00011 + 0.034221 * np.cos(B) + 0.00128 * np.sin(B) + 0.000719 * np.cos(2 * B)
                 + 0.000077 * np.sin(2 * B))
    delta = am * dhi / e0
    k = int(np.searchsorted([1.065, 1.230, 1.500, 1.950, 2.800, 4.500, 6.200], max(eps, 1.0), side="right"))
    f1 = max(0.0, F[0, k] + F[1, k] * delta + F[2, k] * z)
    f2 = F[3, k] + F[4, k] * delta + F[5, k] * z
    return dhi * ((1 - f1) * (1 + np.cos(slope)) / 2 + f1 * max(cos_aoi, 0) / max(np.cos(np.radians(85)), cos_z)
                  + f2 * np.sin(slope))


def test_solar_closed_form_flat_and_shaded():
    h = np.full((6, 6), 200.0, np.float32)
    sc = scene_from_heights(h)
    az_alt, dni, dhi, bins = sky()
    svf_clear = np.zeros(bins["svf_bytes"] * sc.num_total, np.uint8)
    svf_dark = np.full_like(svf_clear, 255)
    hours = orc.sunlit_hours(bins["nudged"])
    for hi in hours[:: max(len(hours) // 25, 1)]:
        alt = np.float64(bins["nudged"][hi, 1])
        cos_z = np.cos(np.radians(90.0 - alt))
        diff = _perez_f64(cos_z, cos_z, float(dhi[hi]), float(dni[hi]), 0.0, hi // 24)
        direct = max(float(dni[hi]), 1e-3) * cos_z if bins["sun_dense_idx"][hi] >= 0 else 0.0
        lit = orc.solar_hour(sc, bins, dni, dhi, svf_clear, int(hi))[0]
        dark = orc.solar_hour(sc, bins, dni, dhi, svf_dark, int(hi))[0]
        assert abs(lit - (direct + diff)) <= 2e-5 * (direct + diff) + 1e-3, hi
        assert abs(dark - diff) <= 2e-5 * diff + 1e-3, hi


def test_fp64_linearity_matches_per_hour_fp64_sum():
    sc = synth_scene(23, 10, 9)
    az_alt, dni, dhi, bins = sky()
    ref = orc.viewshed(sc, bins)
    a = orc.annual_fp64(sc, bins, dni, dhi, ref["svf"], ref["off"], ref["ids"])
    b = orc.annual_fp64_direct(sc, bins, dni, dhi, ref["svf"], ref["off"], ref["ids"])
    np.testing.assert_allclose(a, b, rtol=1e-12)
    sel = np.array([0, 5, sc.num_total - 1], np.int32)
    np.testing.assert_allclose(orc.annual_fp64(sc, bins, dni, dhi, ref["svf"], ref["off"], ref["ids"], sel), a[sel],
                               rtol=1e-13)


@pytest.mark.parametrize("mode", [0, 1])
def test_fp32_oracle_within_1e4_of_fp64(mode):
    sc = synth_scene(24, 14, 12)
    # walls (slope pi/2) facing every quadrant carry the FP64 tilt terms into the comparison
    walls = sc.aspects[sc.num_top:]
    assert sc.num_total > 2 * sc.num_top and np.all(sc.slopes[sc.num_top:] == np.float32(np.pi / 2))
    assert np.unique(np.floor(np.mod(walls, 2 * np.pi) / (np.pi / 2))).size == 4
    az_alt, dni, dhi, bins = sky()
    ref = orc.viewshed(sc, bins)
    boff, data = orc.barle_encode_all(ref["off"], ref["ids"])
    exact = orc.annual_fp64(sc, bins, dni, dhi, ref["svf"], ref["off"], ref["ids"])
    got, n = orc.annual_fp32(sc, bins, dni, dhi, ref["svf"], ref["off"], ref["ids"], boff, data, mode=mode)
    assert n == orc.sunlit_hours(bins["nudged"]).size
    assert (np.abs(got - exact) / np.maximum(exact, 1.0)).max() <= 1e-4


def _noaa_alt(lon, lat, tz):
    """NOAA general solar position (Spencer Fourier series), independent of PSA."""
    out = np.zeros(8760)
    for k in range(8760):
        day, hour = k // 24, k % 24
        g = 2 * np.pi / 365 * (day + (hour - 12) / 24)
        eot = 229.18 * (0.000075 + 0.001868 * np.cos(g) - 0.032077 * np.sin(g) - 0.014615 * np.cos(2 * g)
                        - 0.040849 * np.sin(2 * g))
        decl = (0.006918 - 0.399912 * np.cos(g) + 0.070257 * np.sin(g) - 0.006758 * np.cos(2 * g)
                + 0.000907 * np.sin(2 * g) - 0.002697 * np.cos(3 * g) + 0.00148 * np.sin(3 * g))
        tst = hour * 60 + eot + 4 * lon - 60 * tz
        ha = np.radians(tst / 4 - 180)
        la = np.radians(lat)
        cz = np.sin(la) * np.sin(decl) + np.cos(la) * np.cos(decl) * np.cos(ha)
        out[k] = 90 - np.degrees(np.arccos(np.clip(cz, -1, 1)))
    return out


def test_sun_position_against_noaa():
    lon, lat, tz = SITE
    psa = orc.sun_positions(lon, lat, tz)
    noaa = _noaa_alt(lon, lat, tz)
    up = noaa > 5
    # Spencer's series is good to ~0.5 deg of altitude; a wrong zone sign, hour or
    # latitude would be off by degrees.
    assert np.abs(psa[up, 1] - noaa[up]).max() < 0.6
    assert np.abs(psa[up, 1] - noaa[up]).mean() < 0.3


@settings(max_examples=150, deadline=None)
@given(st.lists(st.integers(min_value=1, max_value=(1 << 26)), min_size=0, max_size=400))
def test_barle_round_trip(ids):
    enc = orc.barle_encode_voxel(np.array(ids, np.int32))
    dec = orc.barle_decode(enc)
    np.testing.assert_array_equal(dec[dec > 0], np.unique(np.array(ids, np.uint32))[::-1])


def test_generators_are_deterministic_and_in_spec():
    a = sg.synth_dsm(7, 128, 96, 1.0, 0.35, 10)
    b = sg.synth_dsm(7, 128, 96, 1.0, 0.35, 10)
    assert np.array_equal(a, b)
    assert 200.0 <= a.min() and a.max() <= 220.0 + 40.0 + 12.0 + 1e-3
    az_alt = sg.sun_positions(*SITE)
    dni, dhi = sg.synth_weather(7, az_alt)
    lit = az_alt[:, 1] - np.float32(1e-8) >= 1.0
    assert (dhi[lit] > 0).all() and (dni[~lit] == 0).all() and (dhi[~lit] == 0).all()
    d2, h2 = sg.synth_weather(7, az_alt)
    assert np.array_equal(dni, d2) and np.array_equal(dhi, h2)


PEREZ_EDGES = np.array([1.065, 1.230, 1.500, 1.950, 2.800, 4.500, 6.200], np.float32).astype(np.float64)


@pytest.mark.parametrize("site,seed", [(SITE, 20260101), ((15.65, 46.56, 1), 20260102), ((-3.7, 40.4, 0), 7)])
def test_perez_epsilon_bins_agree_with_fp64_away_from_edges(site, seed):
    """Every sunlit hour's Perez clearness epsilon in FP64 (zenith in radians from the nudged
    altitude, no float rounding): hours within 1e-5 relative of a bin edge are reported, and at
    every other hour the FP32 bin (the reference's arithmetic, which the library's hour constants
    and the FP64 oracle share) is the bin of the FP64 epsilon."""
    az_alt, dni, dhi, bins = sky(site, seed)
    hours = orc.sunlit_hours(bins["nudged"])
    assert hours.size > 4000
    alt = bins["nudged"][hours, 1].astype(np.float64)
    dif, dr = dhi[hours].astype(np.float64), dni[hours].astype(np.float64)
    assert (dif > 0).all()   # the generator keeps DHI > 0 in sunlit hours (no Perez 0/0)
    za = np.radians(90.0 - alt)
    eps64 = np.fmax(1.0, ((dif + dr / np.cos(za)) / dif + 1.041 * za ** 3) / (1.0 + 1.041 * za ** 3))
    bin64 = np.searchsorted(PEREZ_EDGES, eps64, side="right")
    bin32 = np.array([orc.hour_eps(a, f, d)[1] for a, f, d in zip(alt, dif, dr)])
    near = (np.abs(eps64[:, None] / PEREZ_EDGES[None, :] - 1.0) <= 1e-5).any(axis=1)
    print(f"site {site}: {hours.size} sunlit hours, {int(near.sum())} within 1e-5 of a Perez bin edge "
          f"(hours {hours[near].tolist()})")
    np.testing.assert_array_equal(bin32[~near], bin64[~near])


def test_subset_forms_equal_the_whole_scene_forms():
    """The CPU baseline's random-sample forms (viewshed and FP32 hour loop over a voxel list)
    give the whole-scene results at those voxels."""
    sc = synth_scene(25, 20, 16)
    az_alt, dni, dhi, bins = sky()
    ref = orc.viewshed(sc, bins)
    vox = np.random.default_rng(5).choice(sc.num_total, 60, replace=False).astype(np.int32)
    sub = orc.viewshed_subset(sc, bins, vox, with_svf=True)
    lens = np.diff(ref["off"])[vox]
    np.testing.assert_array_equal(np.diff(sub["off"]), lens)
    want_ids = np.concatenate([ref["ids"][ref["off"][v]:ref["off"][v + 1]] for v in vox])
    np.testing.assert_array_equal(sub["ids"], want_ids)
    N, n, hgs = sc.num_total, vox.size, bins["svf_bytes"]
    np.testing.assert_array_equal(sub["svf"].reshape(hgs, n), ref["svf"].reshape(hgs, N)[:, vox])
    np.testing.assert_array_equal(
        orc.annual_fp64_subset(sc, bins, dni, dhi, ref["svf"], vox, sub["off"], sub["ids"]),
        orc.annual_fp64(sc, bins, dni, dhi, ref["svf"], ref["off"], ref["ids"], voxels=vox))
    full, _ = orc.annual_fp32(sc, bins, dni, dhi, ref["svf"], ref["off"], ref["ids"], stride=7)
    part, nh, _, _ = orc.annual_fp32_subset(sc, bins, dni, dhi, ref["svf"], vox, sub["off"], sub["ids"], stride=7)
    assert nh == len(range(0, orc.sunlit_hours(bins["nudged"]).size, 7))
    np.testing.assert_array_equal(part, full[vox])
