"""Receiver pairs in the reflection pass (receiver_queue 2): one wave walks two consecutive
receivers of the processing order, merges their kept neighbour lists and gathers each row once.
Only the order of the additions differs from the receiver-by-receiver walk, so the annual sums stay
within 1e-4 of the FP64 oracle; a pair's sums do not depend on which wave takes it (bit-identical
across launch shapes); pairs over the pair limit are walked receiver by receiver, bit for bit as
receiver_queue 1 does; CSR lists and passes of fewer than 16 hours take the receiver-by-receiver
kernel (bit-identical to receiver_queue 1)."""
import numpy as np
import pytest

import solargpu_amd as sg
import oracle_lib as orc
from scenes import sky, synth_scene

pytestmark = pytest.mark.gpu

# 0.5 m cells: buildings give stacks of wall voxels, the receivers that share most neighbours
SCENE = (41, 120, 100, 0.5, 0.45, 6)


def _ctx(sc):
    az_alt, dni, dhi, _ = sky()
    ctx = sg.Context(0)
    ctx.set_scene(sc)
    ctx.set_sky(az_alt, dni, dhi)
    return ctx


def _annual(ctx, fmt, K, queue, waves=4, band=16, per_cu=0):
    ctx.set_tuning(16, K, waves, band, per_cu, queue)
    ctx.hour_loop(fmt, K)
    return ctx.irradiance()


def _exact(sc, ctx):
    _, dni, dhi, bins = sky()
    off, ids = ctx.csr()
    return orc.annual_fp64(sc, bins, dni, dhi, ctx.svf(), off, ids)


def _rel(x, exact):
    return float((np.abs(x - exact) / np.maximum(exact, 1.0)).max())


def test_groups_match_the_fp64_oracle_and_the_single_walk():
    sc = synth_scene(*SCENE)
    print(f"scene {SCENE}: {sc.num_total} voxels", flush=True)
    with _ctx(sc) as ctx:
        ctx.viewshed(sg.FMT_CSR | sg.FMT_BARLE)
        exact = _exact(sc, ctx)
        for K in (32, 256):
            single = _annual(ctx, sg.FMT_BARLE, K, 1)
            assert _rel(single, exact) <= 1e-4
            print(f"K {K} pairs", flush=True)
            paired = _annual(ctx, sg.FMT_BARLE, K, 2)
            assert _rel(paired, exact) <= 1e-4
            # the same terms added in another order: FP32 rounding only
            assert (np.abs(paired - single) / np.maximum(np.abs(single), 1.0)).max() <= 1e-5
        with pytest.raises(sg.SolarGPUError):
            ctx.set_tuning(16, 32, 4, 16, 0, 4)


def test_pair_sums_do_not_depend_on_the_launch_shape():
    sc = synth_scene(*SCENE)
    with _ctx(sc) as ctx:
        ctx.viewshed(sg.FMT_BARLE)
        for K in (64, 256):
            out = [_annual(ctx, sg.FMT_BARLE, K, 2, waves, band, per_cu)
                   for waves, band, per_cu in ((4, 16, 0), (2, 32, 0), (4, 0, 0), (2, 0, 3), (4, 64, 2))]
            for x in out[1:]:
                np.testing.assert_array_equal(x, out[0])


def test_pairs_over_the_limit_are_walked_receiver_by_receiver():
    """With a limit of 64 kept entries per pair most pairs take the receiver-by-receiver routine,
    whose sums are those of receiver_queue 1 bit for bit; the others are merged as before; every
    sum stays within the oracle's tolerance and independent of the launch shape."""
    sc = synth_scene(*SCENE)
    with _ctx(sc) as ctx:
        ctx.viewshed(sg.FMT_CSR | sg.FMT_BARLE)
        exact = _exact(sc, ctx)
        single = _annual(ctx, sg.FMT_BARLE, 32, 1)
        full = _annual(ctx, sg.FMT_BARLE, 32, 2)
        ctx.set_pair_limit(64)
        small = [_annual(ctx, sg.FMT_BARLE, 32, 2, waves, band) for waves, band in ((4, 16), (2, 0))]
        np.testing.assert_array_equal(small[0], small[1])
        assert _rel(small[0], exact) <= 1e-4
        assert (np.abs(small[0] - full) / np.maximum(np.abs(full), 1.0)).max() <= 1e-5
        walked = small[0] == single
        assert walked.mean() > 0.5   # the pairs over the limit
        assert not np.array_equal(small[0], single)   # and some merged ones
        with pytest.raises(sg.SolarGPUError):
            ctx.set_pair_limit(-1)
        ctx.set_pair_limit(0)
        np.testing.assert_array_equal(_annual(ctx, sg.FMT_BARLE, 32, 2), full)


def test_csr_and_short_passes_take_the_single_walk():
    sc = synth_scene(*SCENE)
    with _ctx(sc) as ctx:
        ctx.viewshed(sg.FMT_CSR | sg.FMT_BARLE)
        np.testing.assert_array_equal(_annual(ctx, sg.FMT_CSR, 32, 2), _annual(ctx, sg.FMT_CSR, 32, 1))
        np.testing.assert_array_equal(_annual(ctx, sg.FMT_BARLE, 8, 2), _annual(ctx, sg.FMT_BARLE, 8, 1))
