#!/usr/bin/env python3
"""SolarGPU on MI355X: voxel-hours/s and DDA rays/s of the plain pipeline.

Workload (one GPU): seeded 2048 x 2048 float32 urban DSM at 0.5 m cells,
45% building cover, site lat 46.56 / lon 15.65 / tz +1, synthetic hourly
DNI/DHI.  One step = the whole hot path over that input, resident in HBM:
stage 1 (2592-direction DDA viewshed, SVF bitmap, CSR lists, GPU ba-RLE
encode) and stage 2 (every sunlit hour: direct + Perez diffuse, ba-RLE
inter-reflection decoded in-kernel, FP32 Kahan annual sums).

value = voxel-hours/s = N_voxels x sunlit hours / stage-2 seconds (HIP
events).  Extra keys: DDA rays/s (N x 2592 / stage-1 seconds), the CSR
variant, the reflection kernel's HBM roofline, the CPU baseline and the
surrogate form on the same scene (trained weights under
solargpu_amd/weights/): its GPU-seconds against the plain form's and its
per-voxel annual error against the plain result.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
PROFILES = os.path.join(ROOT, "profiles")

HBM_PEAK_GBPS = 8000.0
WEIGHTS = os.path.join(ROOT, "solargpu_amd", "weights", "surrogate.npz")
METRIC = "voxel-hours/s + DDA rays/s, plain vs surrogate at <=2% p95 annual error"
SITE = (15.65, 46.56, 1)
SEED = 20260102


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def reflect_pass_bytes(n_total, n_top, stream_bytes, nh):
    """Compulsory HBM bytes of one reflection/accumulate pass over nh hours (ba-RLE or CSR):
    neighbour stream + int64 offsets + receiver state (pix_parent, slope, aspect, z: 16 B)
    + one SVF byte per voxel-hour + Kahan sum/compensation read and write (16 B)
    + one read of the gathered top-cell hourly values (4 nh B) and heights (4 B)."""
    return (stream_bytes + 8 * (n_total + 1) + 16 * n_total + nh * n_total + 16 * n_total
            + (4 * nh + 4) * n_top)


def read_traffic():
    """HBM bytes/launch of the reflection kernel from the committed rocprofv3 PMC summary."""
    path = os.path.join(PROFILES, "reflect_barle_c2_traffic.txt")
    if not os.path.exists(path):
        return None
    vals = {}
    with open(path) as f:
        for line in f:
            if "=" in line and not line.lstrip().startswith("#"):
                k, v = line.split("=", 1)
                vals[k.strip()] = v.strip()
    try:
        return float(vals["traffic_bytes_per_launch"])
    except (KeyError, ValueError):
        return None


def cpu_baseline(sg, orc, ctx, sc, az_alt, dni, dhi, threads, budget_s):
    """Oracle (scalar C++ port, `threads` threads) on a bounded sample of the same scene."""
    bins = orc.sky_bins(az_alt)
    svf = ctx.svf()
    N, ntop = sc.num_total, sc.num_top
    hours_total = orc.sunlit_hours(bins["nudged"]).size
    # stage 2 sample: 50000 top + 50000 wall voxels, every 40th sunlit hour
    stride = 40
    rates = []
    t_top = t_rng = 0.0
    nvox = 0
    for v0 in (ntop // 2, ntop + (N - ntop) // 2):
        cnt = min(50000, N - v0)
        off, ids = ctx.csr_range(v0, cnt)
        _, nh, tt, tr = orc.annual_fp32_range(sc, bins, dni, dhi, svf, v0, off, ids, stride=stride,
                                              nthreads=threads)
        t_top += tt
        t_rng += tr
        nvox += cnt
    # extrapolate: per sampled hour, all-top solar once + per-voxel cost for all N voxels
    per_hour = t_top / (2 * nh) + (t_rng / nh) * (N / nvox)
    vh_rate = N / per_hour
    # stage 1 sample: DDA rays of 4 x 40000 voxels (all 2592 directions)
    t0 = time.perf_counter()
    for v0 in (ntop // 3, 2 * ntop // 3, ntop + (N - ntop) // 3, ntop + 2 * (N - ntop) // 3):
        orc.trace_raw(sc, bins, v0, 40000, nthreads=threads)
    t1 = time.perf_counter()
    rays_rate = 160000 * 2592 / (t1 - t0)
    sample = (f"oracle FP32 hour loop on 2 x 50000 voxels (top, wall) x {nh} sunlit hours (every {stride}th of "
              f"{hours_total}), all-top solar per hour, extrapolated linearly to {N} voxels; "
              f"DDA on 160000 voxels x 2592 rays; {t_top + t_rng + (t1 - t0):.1f} s wall on {threads} threads")
    return {"value": vh_rate, "unit": "voxel-hours/s", "cores": threads, "kind": "port", "sample": sample,
            "rays_per_s": rays_rate}


def surrogate_leg(sg, sc, az_alt, dni, dhi, cell, tuning, exact, plain_s, hours):
    """The surrogate form on the bench scene in a fresh context, with both exact fallbacks."""
    from solargpu_amd import surrogate as sm
    params, gate, _ = sm.load(WEIGHTS)
    mlp = sm.MLP(seed=1)
    mlp.set_params(params)
    out = {"weights": os.path.relpath(WEIGHTS, ROOT), "plain_gpu_s": plain_s}
    with sg.Context(0) as c2:
        c2.set_scene(sc)
        c2.set_sky(az_alt, dni, dhi)
        c2.set_tuning(tuning["dirs_per_group"], tuning["hours_per_pass"], tuning["waves_per_group"],
                      tuning["xcd_band"], tuning["group_blocks"])
        sm.prepare(c2, SITE[1], cell)
        for name, k in (("annual_fallback", sm.FALLBACK_ANNUAL), ("hourly_fallback", 0)):
            tot, fb, st = sm.run(c2, mlp, gate, k)
            rel = np.abs(tot.astype(np.float64) - exact) / np.maximum(np.abs(exact.astype(np.float64)), 1.0)
            gs = st["t_total_ms"] / 1e3
            out[name] = {"gpu_s": gs, "gpu_seconds_ratio": plain_s / gs,
                         "voxel_hours_per_s": sc.num_total * hours / gs, "fallback_share": float(fb.mean()),
                         "p50": float(np.percentile(rel, 50)), "p95": float(np.percentile(rel, 95)),
                         "max": float(rel.max())}
            log(f"surrogate ({name}): {gs * 1e3:.0f} ms, p95 {out[name]['p95']:.4f}, "
                f"fallback {out[name]['fallback_share']:.3f}")
    mlp.close()
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=2048)
    ap.add_argument("--cell", type=float, default=0.5)
    ap.add_argument("--density", type=float, default=0.45)
    ap.add_argument("--hours-per-pass", type=int, default=0, help="0 = autotuned")
    ap.add_argument("--hour-stride", type=int, default=1,
                    help="every n-th sunlit hour only (profiling runs; the reported rate stays per computed hour)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-csr", action="store_true", help="skip the CSR hour-loop report")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--tuning", default="",
                    help="G,K,waves,band,group: use this launch shape instead of autotuning (profiling runs)")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--no-dropin", action="store_true", help="skip the host-buffer drop-in timing")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (see DESIGN.md, "
              "'How the path would shard'); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import solargpu_amd as sg

    W = args.size
    cell = args.cell
    trees = int(round(40 * (W * cell / 256.0) ** 2))
    t0 = time.perf_counter()
    heights = sg.synth_dsm(SEED, W, W, cell, args.density, trees)
    sc = sg.assemble_scene(heights, cell)
    az_alt = sg.sun_positions(*SITE)
    dni, dhi = sg.synth_weather(SEED, az_alt)
    log(f"scene {W}x{W} @ {cell} m: {sc.num_total} voxels ({sc.num_top} top) in {time.perf_counter() - t0:.1f} s")

    ctx = sg.Context(0)
    ctx.set_scene(sc)
    ctx.set_sky(az_alt, dni, dhi)
    K = args.hours_per_pass
    fmt = sg.FMT_BARLE

    def step():
        ctx.viewshed(sg.FMT_CSR | sg.FMT_BARLE)
        a = ctx.info()
        ctx.hour_loop(fmt, K, args.hour_stride)
        b = ctx.info()
        return a, b

    tuning = None
    if args.tuning:
        G, Kt, wv, band, group = (int(x) for x in args.tuning.split(","))
        ctx.set_tuning(G, Kt, wv, band, group)
        tuning = ctx.tuning()
        log(f"tuning: {tuning}")
    elif not args.no_autotune:
        ctx.viewshed(sg.FMT_CSR | sg.FMT_BARLE)
        tuning = ctx.autotune(fmt)
        log(f"autotune: {tuning}")
    for i in range(args.warmup):
        a, b = step()
        log(f"warmup {i}: viewshed {a['t_viewshed_ms']:.0f} ms, hours {b['t_hours_ms']:.0f} ms")

    ctx.set_kernel_timing(True)
    ctx.sync()
    t_start = time.perf_counter()
    recs = []
    for i in range(args.steps):
        a, b = step()
        recs.append((a, b))
        log(f"step {i}: viewshed {a['t_viewshed_ms']:.0f} ms (trace {a['t_trace_ms']:.0f}, lists {a['t_lists_ms']:.0f}), "
            f"hours {b['t_hours_ms']:.0f} ms, K={b['hours_per_pass']}")
    ctx.sync()
    wall = time.perf_counter() - t_start
    ctx.set_kernel_timing(False)

    inf = recs[-1][1]
    N, ntop, H = sc.num_total, sc.num_top, inf["hours_computed"]
    t_vs = sum(a["t_viewshed_ms"] for a, _ in recs) / 1e3
    t_tr = sum(a["t_trace_ms"] for a, _ in recs) / 1e3
    t_hr = sum(b["t_hours_ms"] for _, b in recs) / 1e3
    vh = N * H * args.steps / t_hr
    rays = N * 2592 * args.steps / t_vs

    # roofline of the dominant kernel: reflection + Kahan accumulate (ba-RLE decode)
    Kp = inf["hours_per_pass"]
    passes = inf["hour_passes"]
    nh_list = [min(Kp, H - p * Kp) for p in range(passes)]
    bytes_all = sum(reflect_pass_bytes(N, ntop, inf["barle_bytes"], nh) for nh in nh_list)
    t_refl = sum(b["t_refl_kernel_ms"] for _, b in recs) / 1e3
    launches = sum(b["timed_launches"] for _, b in recs)
    avg_launch = t_refl / launches
    achieved = bytes_all * args.steps / t_refl / 1e9
    traffic = read_traffic()

    extra = {}
    exact = ctx.irradiance()
    # DDA cell steps of the same viewshed (a counting run after the timed steps; counters cost time)
    ctx.set_step_counting(True)
    ctx.viewshed(sg.FMT_CSR | sg.FMT_BARLE)
    dda_steps = ctx.info()["dda_steps"]
    ctx.set_step_counting(False)
    if not args.no_surrogate and os.path.exists(WEIGHTS) and args.hour_stride == 1:
        extra["surrogate"] = surrogate_leg(sg, sc, az_alt, dni, dhi, cell, ctx.tuning(), exact,
                                           (t_vs + t_hr) / args.steps, H)
    copy_gbps = ctx.copy_bandwidth(1 << 30, 20)
    if not args.no_dropin:
        # the drop-in entry point end to end: host arrays in (PCIe), viewshed, hour loop, host out
        t0 = time.perf_counter()
        out = sg.compressed_gpu_irradiance(sc, az_alt, dni, dhi)
        dw = time.perf_counter() - t0
        extra["dropin"] = {"entry": "compressed_gpu_irradiance", "wall_s": dw, "voxel_hours_per_s": N * H / dw,
                           "max_rel_vs_bench": float((np.abs(out.astype(np.float64) - exact)
                                                      / np.maximum(np.abs(exact.astype(np.float64)), 1.0)).max())}
    if not args.no_csr:
        tc = ctx.autotune(sg.FMT_CSR) if not args.no_autotune else ctx.tuning()
        ctx.hour_loop(sg.FMT_CSR, 0, args.hour_stride)
        c = ctx.info()
        extra["csr_voxel_hours_per_s"] = N * c["hours_computed"] / (c["t_hours_ms"] / 1e3)
        extra["csr_tuning"] = {k: tc[k] for k in ("hours_per_pass", "waves_per_group", "xcd_band", "group_blocks")}
    if tuning is not None:
        extra["autotune"] = tuning
    res = {
        "metric": METRIC,
        "value": vh,
        "unit": "voxel-hours/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded DSM + weather generators)",
        "config": {"workload": f"plain method at scale: {W}x{W} DSM at {cell} m, {args.density:.0%} buildings, "
                               f"{N} voxels ({ntop} top), {H} sunlit hours, ba-RLE neighbour streams",
                   "grid": W, "cell_m": cell, "voxels": N, "top_voxels": ntop, "sunlit_hours": H,
                   "hours_per_pass": Kp, "nnz": inf["nnz"], "barle_bytes": inf["barle_bytes"]},
        "dda_rays_per_s": rays,
        "dda_rays_per_s_trace_kernel": N * 2592 * args.steps / t_tr,
        "dda_cell_steps": dda_steps,
        "dda_cell_steps_per_s_trace_kernel": dda_steps * args.steps / t_tr,
        "stage1_ms": t_vs * 1e3 / args.steps,
        "stage2_ms": t_hr * 1e3 / args.steps,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBPS, "traffic": traffic,
                     "kernel": f"k_reflect<K={Kp}, ba-RLE>", "avg_launch_ms": avg_launch * 1e3,
                     "algorithmic_bytes_per_launch": bytes_all / passes,
                     # what binds the batched gather: its L2-miss traffic (the PMC figure above) per
                     # launch time, against the copy bandwidth measured in this run
                     "traffic_gbps": (traffic / avg_launch / 1e9) if traffic else None,
                     "traffic_frac_of_copy": (traffic / avg_launch / 1e9 / copy_gbps) if traffic else None},
        "measured_copy_gbps": copy_gbps,
    }
    res.update(extra)
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as orc
        threads = min(16, os.cpu_count() or 1)
        res["cpu_baseline"] = cpu_baseline(sg, orc, ctx, sc, az_alt, dni, dhi, threads, args.cpu_budget)
    ctx.close()
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
