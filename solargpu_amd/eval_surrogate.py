"""Evaluate the reflection surrogate against the plain method on held-out tiles.

For every tile the plain pipeline (2592-direction viewshed, ba-RLE lists, the
full sunlit-hour loop) gives the exact annual irradiance and its GPU time; the
surrogate path (sun-bin SVF, exact direct + diffuse, MLP, gate, exact fallback
for the gated voxels) runs in a fresh context on the same inputs.  Reported:
per-voxel relative error of the annual total (p50 / p95 / max, floor 1 Wh/m^2),
the fallback share and the GPU-seconds of both forms (HIP events), for both
exact fallbacks (the per-hour loop and the one-walk annual form), plus the
one-walk exact path applied to every voxel.

    python -m solargpu_amd.eval_surrogate --tiles 20 --size 1024 \
        --weights solargpu_amd/weights/surrogate.npz
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
from solargpu_amd.train_surrogate import make_tile

EVAL_SEED = 20261000
DEFAULT_WEIGHTS = os.path.join(os.path.dirname(__file__), "weights", "surrogate.npz")


def rel_errors(total: np.ndarray, exact: np.ndarray) -> np.ndarray:
    return np.abs(total.astype(np.float64) - exact) / np.maximum(np.abs(exact.astype(np.float64)), 1.0)


def plain_annual(ctx, tuning=None) -> tuple[np.ndarray, float]:
    """Exact annual irradiance (ba-RLE hour loop on the context's lists) and the plain GPU ms
    (viewshed + hour loop).  tuning: fixed launch shapes instead of the autotuner."""
    if tuning:
        ctx.set_tuning(*tuning)
    else:
        ctx.viewshed(sg.FMT_BARLE)
        ctx.autotune(sg.FMT_BARLE)    # launch shapes for this tile size (cached after the first)
    ctx.viewshed(sg.FMT_BARLE)
    t_vs = ctx.info()["t_viewshed_ms"]
    ctx.hour_loop(sg.FMT_BARLE, 0)
    t_hr = ctx.info()["t_hours_ms"]
    return ctx.irradiance(), t_vs + t_hr


def surrogate_annual(seed: int, size: int, cell: float, mlp, gate, K: int, site=None, tuning=None, err_mlp=None):
    ctx, sc, lat = make_tile(seed, size, cell, lists=False, site=site)
    with ctx:
        if tuning:
            ctx.set_tuning(*tuning)
        sm.prepare(ctx, lat, cell)     # sets the site features; run() repeats it inside its timing
        out, fb, st = sm.run(ctx, mlp, gate, K, err_mlp)
    return out, fb, st


EVERYWHERE = sm.Gate(-1.0, 1.0, 1.0, 0.0)   # every voxel to the exact path
NOWHERE = sm.Gate(3.0e38, 0.0, 0.0, 0.0)    # no voxel to the exact path (the network alone)
ERR_LIMIT = 0.02


def gate_quality(fb, raw_err, err):
    """The gate against the network's own error: the share of voxels above 2% it sends to the
    exact path, the share of its fallback that was above 2%, and the worst error it lets pass."""
    fb = fb.astype(bool)
    pos = raw_err > ERR_LIMIT
    return {"above_2pct": int(pos.sum()), "caught": int((fb & pos).sum()),
            "recall_above_2pct": float((fb & pos).sum() / max(pos.sum(), 1)),
            "precision_above_2pct": float((fb & pos).sum() / max(fb.sum(), 1)),
            "max_non_fallback": float(err[~fb].max()) if (~fb).any() else 0.0}


def evaluate(seeds, size: int, cell: float, mlp, gate, K: int, log=print, site=None, tuning=None,
             exact_all: bool = True, err_mlp=None):
    """Per tile: the plain pipeline (viewshed + sunlit-hour loop), the network alone (untimed: the
    per-voxel error the gate must catch), the surrogate with the per-hour fallback (K hours per
    pass), the surrogate with the one-walk annual fallback, and the exact one-walk path for every
    voxel (the same arithmetic the annual fallback uses)."""
    per_tile = []
    errs = {"hourly": [], "annual": []}
    raws, fbs = [], []
    t = {"plain": 0.0, "hourly": 0.0, "annual": 0.0, "exact_annual": 0.0}
    n_fb = n_all = 0
    worst_exact = 0.0
    for k, seed in enumerate(seeds):
        ctx, sc, lat = make_tile(seed, size, cell, lists=False, site=site)
        with ctx:
            exact, ms_plain = plain_annual(ctx, tuning)
        log(f"tile {k}: plain {ms_plain:.0f} ms ({sc.num_total} voxels)")
        row = {"seed": seed, "voxels": int(sc.num_total), "plain_ms": ms_plain}
        t["plain"] += ms_plain
        raw, _, _ = surrogate_annual(seed, size, cell, mlp, NOWHERE, sm.FALLBACK_ANNUAL, site, tuning)
        raw_err = rel_errors(raw, exact)
        raws.append(raw_err)
        row["network_alone"] = {"p95": float(np.percentile(raw_err, 95)), "max": float(raw_err.max()),
                                "share_above_2pct": float((raw_err > ERR_LIMIT).mean())}
        for mode, kk in (("hourly", K), ("annual", sm.FALLBACK_ANNUAL)):
            out, fb, st = surrogate_annual(seed, size, cell, mlp, gate, kk, site, tuning, err_mlp)
            e = rel_errors(out, exact)
            errs[mode].append(e)
            t[mode] += st["t_total_ms"]
            row[mode] = {"p50": float(np.percentile(e, 50)), "p95": float(np.percentile(e, 95)),
                         "max": float(e.max()), "fallback": float(fb.mean()), "ms": st["t_total_ms"],
                         **gate_quality(fb, raw_err, e),
                         "parts_ms": {k2: st[k2] for k2 in ("t_svf_ms", "t_dd_ms", "t_features_ms", "t_mlp_ms",
                                                            "t_fallback_ms")}}
            if mode == "hourly":
                n_fb += int(fb.sum())
                n_all += fb.size
                fbs.append(fb)
        if exact_all:
            out, fb, st = surrogate_annual(seed, size, cell, mlp, EVERYWHERE, sm.FALLBACK_ANNUAL, site, tuning)
            ex_err = float(rel_errors(out, exact).max())
            worst_exact = max(worst_exact, ex_err)
            row["exact_annual"] = {"ms": st["t_total_ms"], "max_rel_vs_plain": ex_err}
            t["exact_annual"] += st["t_total_ms"]
        per_tile.append(row)
        log(f"tile {k} (seed {seed}, {sc.num_total} voxels): p95 {row['hourly']['p95']:.4f} "
            f"fallback {row['hourly']['fallback']:.3f}; plain {ms_plain:.0f} ms, surrogate "
            f"{row['hourly']['ms']:.0f} ms (hourly fallback) / {row['annual']['ms']:.0f} ms (annual fallback)")
    raw_all = np.concatenate(raws)
    summary = {"tiles": len(per_tile), "size": size, "cell_m": cell, "fallback": n_fb / max(n_all, 1),
               "plain_gpu_s": t["plain"] / 1e3,
               "network_alone": {"p95": float(np.percentile(raw_all, 95)), "max": float(raw_all.max()),
                                 "share_above_2pct": float((raw_all > ERR_LIMIT).mean())},
               "gate": "error network" if err_mlp is not None else "score"}
    for mode in ("hourly", "annual"):
        e = np.concatenate(errs[mode])
        summary[mode + "_fallback"] = {
            "voxels": int(e.size), "p50": float(np.percentile(e, 50)), "p95": float(np.percentile(e, 95)),
            "max": float(e.max()), "surrogate_gpu_s": t[mode] / 1e3,
            "gpu_seconds_ratio": t["plain"] / max(t[mode], 1e-9),
            **(gate_quality(np.concatenate(fbs), raw_all, e) if mode == "hourly" else {})}
    if exact_all:
        summary["exact_annual_gpu_s"] = t["exact_annual"] / 1e3
        summary["exact_annual_max_rel_vs_plain"] = worst_exact
    return summary, per_tile


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--tiles", type=int, default=20)
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--cell", type=float, default=1.0)
    ap.add_argument("--hours-per-pass", type=int, default=0, help="per-hour fallback K (0 = tuned)")
    ap.add_argument("--weights", default=DEFAULT_WEIGHTS)
    ap.add_argument("--out", default=None, help="write the JSON summary here")
    ap.add_argument("--seed", type=int, default=EVAL_SEED, help="tile k uses seed + 1000 k")
    ap.add_argument("--site", default=None, help="density,lat,lon,tz for every tile (default: seeded draw)")
    ap.add_argument("--tuning", default=None, help="G,K,waves,band,group launch shapes (default: autotune)")
    ap.add_argument("--no-exact-all", action="store_true", help="skip the one-walk exact path on every voxel")
    args = ap.parse_args(argv)
    site = None
    if args.site:
        d, la, lo, tz = args.site.split(",")
        site = (float(d), float(la), float(lo), int(tz))
    tuning = tuple(int(x) for x in args.tuning.split(",")) if args.tuning else None
    mlp, err_mlp, gate, meta = sm.load_models(args.weights)
    t0 = time.perf_counter()
    seeds = [args.seed + 1000 * k for k in range(args.tiles)]
    summary, per_tile = evaluate(seeds, args.size, args.cell, mlp, gate, args.hours_per_pass,
                                 log=lambda s: print(s, flush=True), site=site, tuning=tuning,
                                 exact_all=not args.no_exact_all, err_mlp=err_mlp)
    summary["wall_s"] = time.perf_counter() - t0
    summary["gate"] = [gate.tau, gate.refl_rel_err, gate.diff_miss_weight, gate.share_abs_err]
    print(json.dumps(summary), flush=True)
    if args.out:
        with open(args.out, "w") as f:
            json.dump({"summary": summary, "tiles": per_tile}, f, indent=1)
    mlp.close()
    if err_mlp is not None:
        err_mlp.close()


if __name__ == "__main__":
    sys.exit(main())
