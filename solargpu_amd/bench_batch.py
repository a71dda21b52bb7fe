"""Many tiles at once: the batched pipeline against one context per tile.

Workload: 64 seeded tiles, side log-uniform in [128, 1024] cells of 1 m,
building density U(10 %, 60 %), latitude U(30, 60), longitude U(-10, 30),
synthetic weather per tile.  Batched: one persistent viewshed launch per voxel
chunk over all tiles and a hipGraph hour loop with one solar and one reflection
launch per pass (queued receiver pairs on ba-RLE lists; static blocks of the
concatenated tiles and receivers one by one from per-XCD queues are timed too).
Sequential: a context per tile (viewshed + hour loop), the same launch shapes.
Both on the GPU (HIP events); prints one JSON line.

    python -m solargpu_amd.bench_batch [--tiles 64] [--fmt barle]
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

import solargpu_amd as sg

SEED = 20260103


def c3_tiles(n: int, seed: int = SEED):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        side = int(round(np.exp(rng.uniform(np.log(128), np.log(1024)))))
        density = float(rng.uniform(0.10, 0.60))
        lat = float(rng.uniform(30.0, 60.0))
        lon = float(rng.uniform(-10.0, 30.0))
        tz = int(round(lon / 15.0))
        s = seed + 1000 * k
        trees = int(round(40 * (side / 256.0) ** 2))
        sc = sg.assemble_scene(sg.synth_dsm(s, side, side, 1.0, density, trees), 1.0)
        az = sg.sun_positions(lon, lat, tz)
        dni, dhi = sg.synth_weather(s, az)
        out.append((sc, az, dni, dhi))
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--tiles", type=int, default=64)
    ap.add_argument("--fmt", choices=("barle", "csr"), default="barle")
    ap.add_argument("--hours-per-pass", type=int, default=256)
    ap.add_argument("--waves", type=int, default=4)
    ap.add_argument("--no-sequential", action="store_true")
    args = ap.parse_args(argv)
    fmt = sg.FMT_BARLE if args.fmt == "barle" else sg.FMT_CSR
    K, waves = args.hours_per_pass, args.waves
    t0 = time.perf_counter()
    tiles = c3_tiles(args.tiles)
    n_vox = sum(sc.num_total for sc, *_ in tiles)
    print(f"{len(tiles)} tiles, {n_vox} voxels, built in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)

    with sg.Batch(0) as b:
        for sc, az, dni, dhi in tiles:
            b.add_tile(sc, az, dni, dhi)
        b.viewshed()
        b.hour_loop(fmt, K, waves, graph=True)   # warm-up (graph build, first touch)
        by_mode = {}
        for mode in (0, 1, 2):   # static blocks, receivers from per-XCD queues, receiver pairs
            b.set_receiver_queue(mode)
            b.hour_loop(fmt, K, waves, graph=True)
            by_mode[mode] = b.info()["t_hours_ms"]
        mode = 2 if fmt == sg.FMT_BARLE and K >= 16 else 0   # pairs walk ba-RLE lists at K >= 16
        b.set_receiver_queue(mode)
        b.viewshed()
        b.hour_loop(fmt, K, waves, graph=True)
        bi = b.info()
    print(f"batched: viewshed {bi['t_viewshed_ms']:.0f} ms, hours {bi['t_hours_ms']:.0f} ms", file=sys.stderr,
          flush=True)
    res = {"workload": f"{len(tiles)} tiles, side log-U(128, 1024) at 1 m, density U(10%, 60%), "
                       f"lat U(30, 60), lon U(-10, 30)",
           "voxels": n_vox, "nnz": bi["nnz"], "barle_bytes": bi["barle_bytes"], "format": args.fmt,
           "hours_per_pass": K, "waves": waves,
           "batched": {"viewshed_ms": bi["t_viewshed_ms"], "trace_ms": bi["t_trace_ms"],
                       "lists_ms": bi["t_lists_ms"], "encode_ms": bi["t_encode_ms"], "hours_ms": bi["t_hours_ms"],
                       "receiver_assignment": {0: "static blocks", 1: "queued receivers", 2: "queued pairs"}[mode],
                       "dda_rays_per_s": n_vox * 2592 / (bi["t_viewshed_ms"] / 1e3),
                       "voxel_hours_per_s": bi["voxel_hours"] / (bi["t_hours_ms"] / 1e3),
                       "hours_ms_static_blocks": by_mode[0], "hours_ms_queued_receivers": by_mode[1],
                       "hours_ms_queued_pairs": by_mode[2]}}
    if not args.no_sequential:
        tv = th = 0.0
        for k, (sc, az, dni, dhi) in enumerate(tiles):
            with sg.Context(0) as ctx:
                ctx.set_scene(sc)
                ctx.set_sky(az, dni, dhi)
                ctx.set_tuning(32, K, waves, 256, 0, max(mode, 1))   # the batched launches' shapes (queued)
                ctx.viewshed(sg.FMT_CSR | sg.FMT_BARLE)
                tv += ctx.info()["t_viewshed_ms"]
                ctx.hour_loop(fmt, K)
                th += ctx.info()["t_hours_ms"]
        res["sequential"] = {"viewshed_ms": tv, "hours_ms": th,
                             "dda_rays_per_s": n_vox * 2592 / (tv / 1e3),
                             "voxel_hours_per_s": bi["voxel_hours"] / (th / 1e3)}
        res["speedup_viewshed"] = tv / bi["t_viewshed_ms"]
        res["speedup_hours"] = th / bi["t_hours_ms"]
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    sys.exit(main())
