"""Train and calibrate the per-voxel reflection surrogate.

Training data are seeded synthetic tiles labelled by the plain method (the
GPU pipeline's annual reflected irradiance); features, training (forward,
backward, Adam) and evaluation all run on the GPU through libsolargpu.

    python -m solargpu_amd.train_surrogate --tiles 200 --size 512 \
        --out solargpu_amd/weights/surrogate.npz

Tile k: seed = base + 1000 k, building density U(10 %, 60 %), latitude
U(30, 60), longitude U(-10, 30), cells of 0.5 m or 1 m (even odds; the cell size
is a feature), synthetic weather of the same seed.
The error network (same shape, trained by the same kernels) learns log10 of the
per-voxel relative error the surrogate makes on a third set of tiles, never seen
by the surrogate's own training; its threshold tau is calibrated on separate
1024^2 tiles so that the gate catches a set share (--recall) of the voxels whose
true error exceeds 2%, with tau at most 2%.  --calibrate-only keeps the
surrogate network of a weights file and trains and calibrates the error network.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

import solargpu_amd as sg
from solargpu_amd import surrogate as sm

TRAIN_SEED = 20260400
CALIB_SEED = 20260700
ERROR_SEED = 20260550      # the error network's tiles: seed + 1000 k (disjoint from the others)
ERR_LIMIT = 0.02           # the error the gate must catch: 2% of the annual total


def tile_site(seed: int):
    rng = np.random.default_rng(seed)
    density = float(rng.uniform(0.10, 0.60))
    lat = float(rng.uniform(30.0, 60.0))
    lon = float(rng.uniform(-10.0, 30.0))
    return density, lat, lon, int(round(lon / 15.0))


def train_cell(seed: int) -> float:
    return 0.5 if np.random.default_rng(seed + 17).random() < 0.5 else 1.0


def make_tile(seed: int, size: int, cell: float = 1.0, lists: bool = True, site=None):
    """Scene + context of one seeded tile; lists=True also runs the plain viewshed.
    site = (density, lat, lon, tz) overrides the seeded draw."""
    density, lat, lon, tz = site if site is not None else tile_site(seed)
    trees = int(round(40 * (size * cell / 256.0) ** 2))
    heights = sg.synth_dsm(seed, size, size, cell, density, trees)
    sc = sg.assemble_scene(heights, cell)
    az = sg.sun_positions(lon, lat, tz)
    dni, dhi = sg.synth_weather(seed, az)
    ctx = sg.Context(0)
    ctx.set_scene(sc)
    ctx.set_sky(az, dni, dhi)
    if lists:
        ctx.viewshed(sg.FMT_CSR | sg.FMT_BARLE)
    return ctx, sc, lat


def labelled(seed: int, size: int, sample: int | None, rng):
    """Features and targets of one tile (optionally a random voxel sample)."""
    cell = train_cell(seed)
    ctx, sc, lat = make_tile(seed, size, cell)
    sm.prepare(ctx, lat, cell)
    X, dd, diff = sm.features(ctx)
    refl = sm.labels(ctx, sg.FMT_CSR)
    ctx.close()
    T = sm.targets(refl, dd, diff)
    if sample is not None and sample < X.shape[0]:
        idx = rng.choice(X.shape[0], sample, replace=False)
        return X[idx], T[idx], dd[idx], refl[idx]
    return X, T, dd, refl


def gate_scores(Y, T_diff, dd):
    share = np.maximum(Y[:, 0], 0.0)
    frac = share / (1.0 + share)
    return frac, np.abs(Y[:, 1] - T_diff)


def calibrate(Y, T, target_p95: float, diff_miss_weight: float = 1.0):
    """Smallest fallback share whose p95 per-voxel error of the annual total meets the target."""
    share_pred = np.maximum(Y[:, 0], 0.0)
    err = np.abs(share_pred - T[:, 0]) / (1.0 + T[:, 0])        # relative error of dd * (1 + share)
    frac, dmiss = gate_scores(Y, T[:, 1], None)
    score = frac + diff_miss_weight * dmiss
    order = np.argsort(-score)
    n = err.size
    e_sorted = err[order]
    best = None
    for f in np.linspace(0.0, 0.5, 201):
        k = int(round(f * n))
        rest = e_sorted[k:]
        p95 = np.percentile(np.concatenate([np.zeros(k), rest]), 95) if n else 0.0
        if p95 <= target_p95:
            best = (f, k)
            break
    if best is None:
        best = (1.0, n)
    f, k = best
    tau = float(score[order[k - 1]]) if k > 0 else float(score.max() + 1.0)
    # gate: err_est = refl_rel_err * frac + diff_miss_weight * dmiss + share_abs_err * dd/total > tau
    return sm.Gate(tau, 1.0, diff_miss_weight, 0.0), f


def train_error_net(net, tiles: int, size: int, sample: int, epochs: int, batch: int, lr: float, rng, log=print):
    """The error network: on `tiles` seeded tiles the surrogate's own per-voxel relative error of
    the annual total is the target (log10, floored at 1e-4); same shape and kernels as the
    surrogate network."""
    Xs, Es = [], []
    for k in range(tiles):
        X, T, dd, refl = labelled(ERROR_SEED + 1000 * k, size, sample, rng)
        Xs.append(X)
        Es.append(sm.error_target(net.forward(X), dd, refl))
        if k % 20 == 0:
            e = 10.0 ** Es[-1][:, 0]
            log(f"error tile {k}: {X.shape[0]} samples, share above {ERR_LIMIT:.0%}: {(e > ERR_LIMIT).mean():.4f}")
    X = np.concatenate(Xs)
    E = np.concatenate(Es)
    del Xs, Es
    err = sm.MLP(seed=23)
    err.set_loss_weights(1.0, 0.0)
    phases = [epochs - 2 * (epochs // 3), epochs // 3, epochs // 3]
    losses = np.concatenate([err.train(X, E, epochs=e, batch=batch, lr=lr * f, seed=31 + i)
                             for i, (e, f) in enumerate(zip(phases, (1.0, 1.0 / 3.0, 0.1))) if e > 0])
    return err, losses, int(X.shape[0])


def gate_report(pred_log_err, true_err, tau):
    """What the error-network gate does with threshold tau on voxels of known true error: the
    fallback share, the share of voxels above ERR_LIMIT it catches, and the errors left
    (fallback voxels are exact)."""
    fb = pred_log_err > np.log10(tau)
    pos = true_err > ERR_LIMIT
    left = np.where(fb, 0.0, true_err)
    return {"tau": float(tau), "fallback": float(fb.mean()),
            "recall_above_2pct": float((fb & pos).sum() / max(pos.sum(), 1)),
            "precision_above_2pct": float((fb & pos).sum() / max(fb.sum(), 1)),
            "share_above_2pct": float(pos.mean()),
            "p95": float(np.percentile(left, 95)), "max_non_fallback": float(left.max()),
            "p95_no_gate": float(np.percentile(true_err, 95)), "max_no_gate": float(true_err.max())}


def calibrate_error_gate(pred_log_err, true_err, recall: float):
    """tau such that the gate catches `recall` of the voxels whose true error exceeds 2%,
    capped at 2% (a voxel predicted above 2% always falls back)."""
    pos = true_err > ERR_LIMIT
    if pos.any():
        t = float(np.quantile(pred_log_err[pos], 1.0 - recall))
        tau = min(10.0 ** t, ERR_LIMIT)
    else:
        tau = ERR_LIMIT
    return sm.Gate(tau, 0.0, 0.0, 0.0), gate_report(pred_log_err, true_err, tau)


def errors_of(total, exact):
    rel = np.abs(total.astype(np.float64) - exact) / np.maximum(np.abs(exact), 1.0)
    return {"p50": float(np.percentile(rel, 50)), "p95": float(np.percentile(rel, 95)), "max": float(rel.max())}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--tiles", type=int, default=200)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--sample", type=int, default=98304, help="voxels per training tile")
    ap.add_argument("--calib-tiles", type=int, default=10)
    ap.add_argument("--calib-size", type=int, default=1024, help="calibration tile side (0 = --size)")
    ap.add_argument("--calibrate-only", metavar="WEIGHTS", default=None,
                    help="keep the network of WEIGHTS, fit only the gate, write --out")
    ap.add_argument("--epochs", type=int, default=32, help="split in three phases at lr, lr/3, lr/10")
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--target-p95", type=float, default=0.016, help="score gate (no error network)")
    ap.add_argument("--error-tiles", type=int, default=60, help="tiles of the error network (0 = score gate)")
    ap.add_argument("--error-size", type=int, default=512)
    ap.add_argument("--recall", type=float, default=0.8,
                    help="share of the calibration voxels above 2%% error the gate must catch")
    ap.add_argument("--out", default=os.path.join(os.path.dirname(__file__), "weights", "surrogate.npz"))
    args = ap.parse_args(argv)
    rng = np.random.default_rng(1)
    t0 = time.perf_counter()
    mlp = sm.MLP(seed=11)
    if args.calibrate_only:
        params, _, old_meta = sm.load(args.calibrate_only)
        mlp.set_params(params)
        meta = {"network_from": old_meta}
        t1 = t2 = time.perf_counter()
    else:
        Xs, Ts = [], []
        for k in range(args.tiles):
            X, T, _, _ = labelled(TRAIN_SEED + 1000 * k, args.size, args.sample, rng)
            Xs.append(X)
            Ts.append(T)
            if k % 20 == 0:
                print(f"tile {k}: {X.shape[0]} samples, mean reflected share {T[:, 0].mean():.4f} "
                      f"({time.perf_counter() - t0:.0f} s)", flush=True)
        X = np.concatenate(Xs)
        T = np.concatenate(Ts)
        del Xs, Ts
        t1 = time.perf_counter()
        mlp.set_loss_weights(100.0, 1.0)
        phases = [args.epochs - 2 * (args.epochs // 3), args.epochs // 3, args.epochs // 3]
        losses = np.concatenate([mlp.train(X, T, epochs=e, batch=args.batch, lr=args.lr * f, seed=5 + i)
                                 for i, (e, f) in enumerate(zip(phases, (1.0, 1.0 / 3.0, 0.1))) if e > 0])
        t2 = time.perf_counter()
        print(f"trained on {X.shape[0]} samples: epoch losses {np.round(losses, 6).tolist()} ({t2 - t1:.1f} s)",
              flush=True)
    if not args.calibrate_only:
        meta = {"tiles": args.tiles, "size": args.size, "samples": int(X.shape[0]), "epochs": args.epochs,
                "losses": [float(x) for x in losses], "train_s": t2 - t1, "label_s": t1 - t0}
    err_net = None
    if args.error_tiles > 0:
        t3 = time.perf_counter()
        err_net, err_losses, n_err = train_error_net(mlp, args.error_tiles, args.error_size, args.sample,
                                                     args.epochs, args.batch, args.lr, rng,
                                                     log=lambda m: print(m, flush=True))
        meta.update({"error_tiles": args.error_tiles, "error_size": args.error_size, "error_samples": n_err,
                     "error_losses": [float(x) for x in err_losses], "error_s": time.perf_counter() - t3})
        print(f"error network: {n_err} samples, epoch losses {np.round(err_losses, 4).tolist()}", flush=True)
    Yc, Tc, Pc, Ec = [], [], [], []
    for k in range(args.calib_tiles):
        Xk, Tk, dd, refl = labelled(CALIB_SEED + 1000 * k, args.calib_size or args.size, None, rng)
        Yk = mlp.forward(Xk)
        Yc.append(Yk)
        Tc.append(Tk)
        if err_net is not None:
            Pc.append(err_net.forward(Xk)[:, 0])
            Ec.append(10.0 ** sm.error_target(Yk, dd, refl)[:, 0])
    Yc = np.concatenate(Yc)
    Tc = np.concatenate(Tc)
    share_err = np.abs(np.maximum(Yc[:, 0], 0) - Tc[:, 0])
    meta.update({"calib_tiles": args.calib_tiles, "calib_size": args.calib_size or args.size,
                 "calib_share_abs_err_p95": float(np.percentile(share_err, 95))})
    if err_net is not None:
        gate, rep = calibrate_error_gate(np.concatenate(Pc), np.concatenate(Ec), args.recall)
        meta.update({"gate": "error network", "recall_target": args.recall, "calib": rep})
    else:
        gate, f = calibrate(Yc, Tc, args.target_p95)
        meta.update({"gate": "score", "target_p95": args.target_p95, "calib_fallback": f})
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    sm.save(args.out, mlp.params(), gate, meta, err_net.params() if err_net is not None else None)
    print(json.dumps({"out": args.out, "gate": [gate.tau, gate.refl_rel_err, gate.diff_miss_weight], **meta}),
          flush=True)


if __name__ == "__main__":
    sys.exit(main())
