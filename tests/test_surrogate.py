"""Surrogate path on the GPU: its sun-bin SVF equals the plain tracer's bits,
its exact direct + diffuse and its labels agree with the FP64 oracle and the
plain hour loop, the fallback reproduces the plain method, and a trained
network beats the method without it."""
import numpy as np
import pytest

import solargpu_amd as sg
import solargpu_amd.surrogate as sm
import oracle_lib as orc
from scenes import SITE, oracle_lists, scene_from_heights, sky, synth_scene

pytestmark = pytest.mark.gpu


def _ctx(sc, lists=True):
    az_alt, dni, dhi, _ = sky()
    ctx = sg.Context(0)
    ctx.set_scene(sc)
    ctx.set_sky(az_alt, dni, dhi)
    if lists:
        ctx.viewshed(sg.FMT_CSR | sg.FMT_BARLE)
    return ctx


def _rel(a, b, floor=1.0):
    return np.abs(np.asarray(a, np.float64) - b) / np.maximum(np.abs(b), floor)


def test_sun_bin_svf_equals_plain_svf():
    sc = synth_scene(12, 64, 64)
    with _ctx(sc) as ctx:
        plain = ctx.svf()
        sm.prepare(ctx, SITE[1])
        np.testing.assert_array_equal(ctx.svf(), plain)


def test_annual_direct_diffuse_vs_fp64():
    sc = synth_scene(11, 48, 40)
    ref = oracle_lists(11, 48, 40)
    az_alt, dni, dhi, bins = sky()
    no_lists = np.zeros(sc.num_total + 1, np.int64)
    exact = orc.annual_fp64(sc, bins, dni, dhi, ref["svf"], no_lists, np.zeros(1, np.int32))
    with _ctx(sc, lists=False) as ctx:
        sm.prepare(ctx, SITE[1])
        X, dd, diff = sm.features(ctx)
    assert _rel(dd, exact).max() <= 1e-4
    assert (diff <= dd * (1 + 1e-6)).all() and (diff > 0).all()


def test_labels_equal_plain_reflected():
    sc = synth_scene(12, 64, 64)
    with _ctx(sc) as ctx:
        ctx.hour_loop(sg.FMT_BARLE, 16)
        total = ctx.irradiance()
        sm.prepare(ctx, SITE[1])
        _, dd, _ = sm.features(ctx)
        refl = sm.labels(ctx, sg.FMT_CSR)
    assert (np.abs(dd + refl - total) <= 1e-4 * total).all()


def test_flat_plane_horizon_is_level():
    sc = scene_from_heights(np.full((24, 24), 210.0, np.float32))
    with _ctx(sc, lists=False) as ctx:
        sm.prepare(ctx, SITE[1])
        X, _, _ = sm.features(ctx)
    assert (sm.from_bf16(X[:, :32]) == 0).all()
    assert (sm.from_bf16(X[:, 37]) == 1.0).all()        # horizon sky-view factor of open ground


def test_box_raises_the_horizon_towards_it():
    h = np.full((40, 40), 200.0, np.float32)
    h[5:15, 10:30] = 240.0                                # 40 m box north of the voxel
    sc = scene_from_heights(h)
    v = 20 * 40 + 20                                      # ground cell 5 cells south of the box face
    with _ctx(sc, lists=False) as ctx:
        sm.prepare(ctx, SITE[1])
        X, _, _ = sm.features(ctx)
    ang = sm.from_bf16(X[v, :32]) * 90.0                  # degrees
    north = np.r_[ang[:2], ang[-2:]]                      # sectors within +-22.5 deg of north
    south = ang[14:18]
    assert north.min() > 60.0 and np.abs(south).max() < 1.0


@pytest.mark.parametrize("fallback,queue", [(16, 1), (16, 2), (sm.FALLBACK_ANNUAL, 1)])
def test_all_fallback_reproduces_plain(fallback, queue):
    """Every voxel gated: the per-hour fallback loop (its receivers walked one by one, or in pairs
    over the gated subset) and the one-walk annual fallback all give the plain method's annual
    totals (FP64 oracle, 1e-4)."""
    sc = synth_scene(11, 48, 40)
    ref = oracle_lists(11, 48, 40)
    az_alt, dni, dhi, bins = sky()
    exact = orc.annual_fp64(sc, bins, dni, dhi, ref["svf"], ref["off"], ref["ids"])
    mlp = sm.MLP(seed=2)
    with _ctx(sc, lists=False) as ctx:
        ctx.set_tuning(32, 128, 4, 32, 0, queue)
        ctx_lat = SITE[1]
        sm.prepare(ctx, ctx_lat)
        out, fb, st = sm.run(ctx, mlp, sm.Gate(-1.0, 1.0, 1.0, 0.0), fallback)
    assert fb.all() and st["fallback"] == sc.num_total
    assert _rel(out, exact).max() <= 1e-4


def test_no_fallback_is_direct_diffuse_plus_prediction():
    sc = synth_scene(12, 64, 64)
    mlp = sm.MLP(seed=3)
    with _ctx(sc, lists=False) as ctx:
        sm.prepare(ctx, SITE[1])
        X, dd, _ = sm.features(ctx)
        out, fb, st = sm.run(ctx, mlp, sm.Gate(1e30, 1.0, 1.0, 0.0), 16)
    assert not fb.any() and st["fallback"] == 0
    Y = mlp.forward(X)
    want = dd * (1.0 + np.maximum(Y[:, 0], 0.0))
    np.testing.assert_allclose(out, want, rtol=1e-5)


def test_error_network_gate_sends_high_predicted_error_to_the_exact_path():
    """With an error network the gate falls back exactly where 10^(its first output) exceeds tau,
    and those voxels get the plain method's annual totals."""
    sc = synth_scene(11, 48, 40)
    ref = oracle_lists(11, 48, 40)
    az_alt, dni, dhi, bins = sky()
    exact = orc.annual_fp64(sc, bins, dni, dhi, ref["svf"], ref["off"], ref["ids"])
    mlp = sm.MLP(seed=3)
    err = sm.MLP(seed=4)
    with _ctx(sc, lists=False) as ctx:
        sm.prepare(ctx, SITE[1])
        X, dd, _ = sm.features(ctx)
        p = err.forward(X)[:, 0]
        tau = float(10.0 ** np.percentile(p, 70))        # about 30% of the voxels above it
        out, fb, st = sm.run(ctx, mlp, sm.Gate(tau, 0.0, 0.0, 0.0), sm.FALLBACK_ANNUAL, err)
    want_fb = ~(10.0 ** p.astype(np.float64) <= tau * (1 + 1e-6))
    near = np.abs(p - np.log10(tau)) < 1e-5                 # exp10f vs numpy at the threshold
    np.testing.assert_array_equal(fb.astype(bool)[~near], want_fb[~near])
    assert 0.2 < fb.mean() < 0.4 and st["fallback"] == fb.sum()
    assert _rel(out[fb == 1], exact[fb == 1]).max() <= 1e-4
    Y = mlp.forward(X)
    np.testing.assert_allclose(out[fb == 0], (dd * (1.0 + np.maximum(Y[:, 0], 0.0)))[fb == 0], rtol=1e-5)


def test_trained_network_beats_the_method_without_it():
    """Train on small seeded tiles labelled by the plain method; on held-out tiles the
    trained surrogate's p95 error of the annual total must beat (a) dropping the
    reflection term and (b) the untrained network."""
    from solargpu_amd import train_surrogate as ts
    rng = np.random.default_rng(0)
    Xs, Ts = [], []
    for k in range(16):
        X, T, _, _ = ts.labelled(ts.TRAIN_SEED + 1000 * k, 96, None, rng)
        Xs.append(X)
        Ts.append(T)
    X = np.concatenate(Xs)
    T = np.concatenate(Ts)
    untrained = sm.MLP(seed=9)
    mlp = sm.MLP(seed=9)
    mlp.set_loss_weights(100.0, 1.0)
    losses = mlp.train(X, T, epochs=10, batch=1024, lr=1e-3)
    assert losses[-1] < 0.5 * losses[0]
    errs = {"trained": [], "untrained": [], "no_reflection": []}
    for k in range(2):
        Xh, Th, dd, refl = ts.labelled(ts.CALIB_SEED + 1000 * k, 128, None, rng)
        exact = dd.astype(np.float64) + refl
        for name, net in (("trained", mlp), ("untrained", untrained)):
            Y = net.forward(Xh)
            errs[name].append(_rel(dd * (1 + np.maximum(Y[:, 0], 0)), exact))
        errs["no_reflection"].append(_rel(dd, exact))
    p95 = {k: float(np.percentile(np.concatenate(v), 95)) for k, v in errs.items()}
    assert p95["trained"] < 0.7 * p95["no_reflection"], p95
    assert p95["trained"] < 0.7 * p95["untrained"], p95
