"""Voxelization on the GPU (sgpu_set_scene_heights) against the reference's own voxelization
(golden vectors of voxelization.cpp) and the host restatement: parents and wall heights bit for
bit; slopes and aspects within 3 ulp (the device atan2f's accuracy against glibc's; measured at most
3 ulp, 4.8e-7 rad, on a million cells); lists from either scene identical."""
import os

import numpy as np
import pytest

import solargpu_amd as sg
import oracle_lib as orc
from scenes import sky

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
VOX = np.load(os.path.join(GOLDEN, "voxelization.npz"))
ATAN2_ULPS = 3
DSMS = sorted({k.rsplit("_heights", 1)[0] for k in VOX.files if k.endswith("_heights")})


def _ulps(a, b):
    """Distance in units of the last place of b (float32)."""
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    return np.abs(a.astype(np.float64) - b.astype(np.float64)) / np.spacing(np.maximum(np.abs(b), 1e-30))


def _device_scene(heights, sx, sy):
    with sg.Context(0) as ctx:
        ctx.set_scene_heights(heights, sx, sy)
        return ctx.voxels(heights, sx)


@pytest.mark.gpu
@pytest.mark.parametrize("name", DSMS)
def test_device_voxelization_matches_reference_golden(name):
    h = VOX[f"{name}_heights"]
    sx, sy = VOX[f"{name}_scale"]
    gp, gz, ga = VOX[f"{name}_wall_parent"], VOX[f"{name}_wall_z"], VOX[f"{name}_wall_aspect"]
    gs, gasp = VOX[f"{name}_slope"], VOX[f"{name}_aspect"]
    sc = _device_scene(h, sx, sy)
    nt = h.size
    np.testing.assert_array_equal(sc.pixel_parent[:nt], np.arange(nt))
    np.testing.assert_array_equal(sc.pixel_parent[nt:], gp)
    np.testing.assert_array_equal(sc.wall_z.view(np.uint32), gz.view(np.uint32))
    assert np.all(sc.slopes[nt:] == np.float32(np.float32(np.pi) * np.float32(0.5)))
    assert _ulps(sc.slopes[:nt], gs).max() <= ATAN2_ULPS
    assert _ulps(sc.aspects[:nt], gasp).max() <= ATAN2_ULPS
    assert _ulps(sc.aspects[nt:], ga).max() <= ATAN2_ULPS


@pytest.mark.gpu
@pytest.mark.parametrize("seed,w,h,cell", [(21, 64, 48, 1.0), (22, 300, 257, 0.5), (23, 1024, 1024, 0.5)])
def test_device_voxelization_matches_host(seed, w, h, cell):
    heights = sg.synth_dsm(seed, w, h, cell, 0.45, 8)
    host = sg.assemble_scene(heights, cell)
    with sg.Context(0) as ctx:
        ctx.set_scene_heights(heights, cell, cell)
        dev = ctx.voxels(heights, cell)
        order_d, lanes_d = ctx.orders()
        ctx.set_scene(host)   # the same scene through the host path
        order_h, lanes_h = ctx.orders()
    np.testing.assert_array_equal(order_d, order_h)
    np.testing.assert_array_equal(lanes_d, lanes_h)
    assert dev.num_total == host.num_total and dev.max_height == host.max_height
    np.testing.assert_array_equal(dev.pixel_parent, host.pixel_parent)
    np.testing.assert_array_equal(dev.wall_z.view(np.uint32), host.wall_z.view(np.uint32))
    assert _ulps(dev.slopes, host.slopes).max() <= ATAN2_ULPS
    assert _ulps(dev.aspects, host.aspects).max() <= ATAN2_ULPS


@pytest.mark.gpu
def test_lists_from_device_scene_equal_oracle():
    heights = sg.synth_dsm(24, 96, 80, 1.0, 0.35, 4)
    host = sg.assemble_scene(heights, 1.0)
    az_alt, dni, dhi, bins = sky()
    ref = orc.viewshed(host, bins)
    exact = orc.annual_fp64(host, bins, dni, dhi, ref["svf"], ref["off"], ref["ids"])
    with sg.Context(0) as ctx:
        ctx.set_scene_heights(heights, 1.0)
        ctx.set_sky(az_alt, dni, dhi)
        ctx.viewshed(sg.FMT_CSR | sg.FMT_BARLE)
        off, ids = ctx.csr()
        np.testing.assert_array_equal(off, ref["off"])
        np.testing.assert_array_equal(ids, ref["ids"])
        np.testing.assert_array_equal(ctx.svf(), ref["svf"])
        ctx.hour_loop(sg.FMT_BARLE, 64)
        irr = ctx.irradiance()
    rel = np.abs(irr.astype(np.float64) - exact) / np.maximum(np.abs(exact), 1.0)
    assert rel.max() <= 1e-4


@pytest.mark.gpu
def test_non_finite_heights_are_refused():
    heights = sg.synth_dsm(5, 32, 24, 1.0, 0.3, 2)
    with sg.Context(0) as ctx:
        for bad in (np.nan, np.inf, -np.inf):
            h = heights.copy()
            h[7, 11] = bad
            with pytest.raises(sg.SolarGPUError, match="finite"):
                ctx.set_scene_heights(h, 1.0)
        ctx.set_scene_heights(heights, 1.0)   # the context stays usable
