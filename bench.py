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
import subprocess
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


def read_profile(name, key):
    """A figure (key = value line) from a committed rocprofv3 summary under profiles/."""
    path = os.path.join(PROFILES, name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if "=" in line and not line.lstrip().startswith("#"):
                k, v = line.split("=", 1)
                if k.strip() == key:
                    try:
                        return float(v.strip())
                    except ValueError:
                        return None
    return None


def read_traffic():
    """HBM bytes/launch of the reflection kernel from the committed rocprofv3 PMC summary."""
    return read_profile("reflect_barle_c2_traffic.txt", "traffic_bytes_per_launch")


# VALU issue of one wave64 instruction per CU per cycle: 256 CUs x 2.4 GHz (4 SIMD16 x 4 cycles)
VALU_NOMINAL_PEAK = 256 * 2.4e9


def host_cores():
    """Host cores this run may use: what `nproc` reports (the CPU affinity, bounded by
    OMP_NUM_THREADS where the host sets it -- a shared machine's CPU share), and the machine's
    logical CPU count.  Read before the GPU is touched."""
    try:
        n = int(subprocess.run(["nproc"], capture_output=True, text=True, check=True).stdout.strip())
    except (OSError, ValueError, subprocess.CalledProcessError):
        n = len(os.sched_getaffinity(0))
    return max(n, 1), os.cpu_count() or n


CPU_SAMPLE_SEED = 20260199
CPU_SAMPLE_FRACTION = 0.01
CPU_SAMPLE_BLOCK = 512
CPU_HOUR_STRIDE = 24


def cpu_baseline(sg, orc, ctx, sc, az_alt, dni, dhi, threads, host_cpus):
    """The oracle (a multithreaded C++ port of the reference method) on a fixed random 1% of the
    scene's voxels (random 512-voxel blocks): stage 1 (the 2592-direction DDA + first-hit lists) of the sample, then the FP32
    hour loop of the sample (its own lists) on every 24th sunlit hour, with the solar term of every
    top cell per sampled hour (the neighbours' values).  Rates are extrapolated linearly to all
    voxels and hours."""
    bins = orc.sky_bins(az_alt)
    N = sc.num_total
    hours_total = orc.sunlit_hours(bins["nudged"]).size
    # 1% of the voxels in randomly chosen blocks of 512 consecutive voxels: unbiased over the scene,
    # with the memory locality of a full in-order run (single random voxels would each miss the
    # CPU caches and understate the port)
    rng = np.random.default_rng(CPU_SAMPLE_SEED)
    nblk = (N + CPU_SAMPLE_BLOCK - 1) // CPU_SAMPLE_BLOCK
    pick = np.sort(rng.choice(nblk, max(int(nblk * CPU_SAMPLE_FRACTION), 1), replace=False))
    vox = (pick[:, None] * CPU_SAMPLE_BLOCK + np.arange(CPU_SAMPLE_BLOCK)[None, :]).ravel()
    vox = vox[vox < N].astype(np.int32)
    n = vox.size
    t0 = time.perf_counter()
    sub = orc.viewshed_subset(sc, bins, vox, nthreads=threads)
    t1 = time.perf_counter()
    rays_rate = n * 2592 / (t1 - t0)
    svf = ctx.svf()   # the SVF bits of every voxel (the sample's own equal the oracle's; tested)
    _, nh, t_top, t_vox = orc.annual_fp32_subset(sc, bins, dni, dhi, svf, vox, sub["off"], sub["ids"],
                                                 stride=CPU_HOUR_STRIDE, nthreads=threads)
    # per sampled hour: the all-top solar term once + the per-voxel cost scaled to all N voxels
    per_hour = t_top / nh + (t_vox / nh) * (N / n)
    vh_rate = N / per_hour
    sample = (f"oracle on a fixed random {CPU_SAMPLE_FRACTION:.0%} of the voxels ({n} of {N} in {pick.size} "
              f"random blocks of {CPU_SAMPLE_BLOCK}, seed {CPU_SAMPLE_SEED}): DDA + lists on all 2592 directions ({n * 2592:.3g} rays, {sub['ids'].size} ids), "
              f"then the FP32 hour loop of the sample on every {CPU_HOUR_STRIDE}th sunlit hour ({nh} of "
              f"{hours_total}) with the all-top solar term per hour; extrapolated linearly to {N} voxels x "
              f"{hours_total} hours; {t1 - t0 + t_top + t_vox:.1f} s wall on {threads} threads "
              f"(nproc; the host has {host_cpus} logical CPUs)")
    return {"value": vh_rate, "unit": "voxel-hours/s", "cores": threads, "kind": "port", "sample": sample,
            "rays_per_s": rays_rate, "host_logical_cpus": host_cpus,
            "extrapolated_year_s": per_hour * hours_total + N * 2592 / rays_rate}


def surrogate_leg(sg, sc, az_alt, dni, dhi, cell, tuning, exact, plain_s, hours):
    """The surrogate form on the bench scene in a fresh context, with both exact fallbacks; the
    network alone (untimed) gives the per-voxel error the gate has to catch."""
    from solargpu_amd import surrogate as sm
    from solargpu_amd.eval_surrogate import NOWHERE, gate_quality, rel_errors
    mlp, err_mlp, gate, _ = sm.load_models(WEIGHTS)
    out = {"weights": os.path.relpath(WEIGHTS, ROOT), "plain_gpu_s": plain_s,
           "gate": "error network" if err_mlp is not None else "score", "tau": gate.tau,
           "network": f"{sm.FEATURES} features (padded to {sm.IN_PAD}) -> 4 x {sm.HIDDEN} GELU -> {sm.OUTPUTS}, "
                      "bf16 weights/activations, fp32 accumulation",
           "inputs_vs_config": f"{sm.FEATURES} inputs where the surrogate configuration names 48: its listed "
                               "inputs (32 horizon-sector angles, normal xyz, height, a neighbour count -- here the "
                               "log count of sun-bin directions that hit --, sky-view factor, latitude, albedo: 40 "
                               "values) plus 13 more: wall flag, slope, the exact annual direct + diffuse and its "
                               "diffuse share, the direct + diffuse of the horizon cells summed over four quadrants, "
                               "a Lambertian estimate over the horizon cells, the cell size, for walls the direct "
                               "+ diffuse of the ground cell in front and the height above it, and the mean direct "
                               "+ diffuse of the 8 surrounding cells"}
    with sg.Context(0) as c2:
        c2.set_scene(sc)
        c2.set_sky(az_alt, dni, dhi)
        c2.set_tuning(tuning["dirs_per_group"], tuning["hours_per_pass"], tuning["waves_per_group"],
                      tuning["xcd_band"], tuning["group_blocks"], tuning["receiver_queue"])
        sm.prepare(c2, SITE[1], cell)
        raw, _, _ = sm.run(c2, mlp, NOWHERE, sm.FALLBACK_ANNUAL)
        raw_err = rel_errors(raw, exact)
        out["network_alone"] = {"p95": float(np.percentile(raw_err, 95)), "max": float(raw_err.max()),
                                "share_above_2pct": float((raw_err > 0.02).mean())}
        for name, k in (("annual_fallback", sm.FALLBACK_ANNUAL), ("hourly_fallback", 0)):
            tot, fb, st = sm.run(c2, mlp, gate, k, err_mlp)
            rel = rel_errors(tot, exact)
            gs = st["t_total_ms"] / 1e3
            out[name] = {"gpu_s": gs, "gpu_seconds_ratio": plain_s / gs,
                         "voxel_hours_per_s": sc.num_total * hours / gs, "fallback_share": float(fb.mean()),
                         "p50": float(np.percentile(rel, 50)), "p95": float(np.percentile(rel, 95)),
                         "max": float(rel.max()), **gate_quality(fb, raw_err, rel)}
            log(f"surrogate ({name}): {gs * 1e3:.0f} ms, p95 {out[name]['p95']:.4f}, "
                f"fallback {out[name]['fallback_share']:.3f}, recall {out[name]['recall_above_2pct']:.3f}")
    mlp.close()
    if err_mlp is not None:
        err_mlp.close()
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
                    help="G,K,waves,band,group[,queue]: use this launch shape instead of autotuning (profiling runs)")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--no-dropin", action="store_true", help="skip the host-buffer drop-in timing")
    ap.add_argument("--trace-tables", action="store_true",
                    help="trace the viewsheds by step tables (sgpu_set_trace_tables) instead of the pooled walk")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    cores, host_cpus = host_cores()   # before anything touches the GPU
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
    if args.trace_tables:
        ctx.set_trace_tables(True)
    if args.tuning:
        vals = [int(x) for x in args.tuning.split(",")]
        ctx.set_tuning(*vals)
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
    rq = ctx.tuning()["receiver_queue"]
    refl_kernel = f"k_reflect_pair<K={Kp}, ba-RLE>" if rq == 2 and Kp >= 16 else f"k_reflect<K={Kp}, ba-RLE>"
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
    l2_gbps = ctx.l2_bandwidth(2 << 20, 3)
    # the reflection pass's row gathers: L1 -> L2 line requests per launch (committed PMC summary of the
    # same workload and shape) over the live launch time, against the measured L2 read bandwidth
    l2_req = read_profile("reflect_barle_c2_traffic.txt", "l1_to_l2_read_requests_per_launch")
    gather = {"bound": "l2", "unit": "GB/s", "kernel": refl_kernel, "peak": l2_gbps,
              "peak_source": "measured: every wave streams 2 MiB (L2-resident) as 1 KB rows of 16-byte lane loads",
              "l1_to_l2_read_bytes_per_launch": (l2_req * 128) if l2_req else None,
              "l2_hit_rate": read_profile("reflect_barle_c2_traffic.txt", "l2_hit_rate"),
              "achieved": (l2_req * 128 / avg_launch / 1e9) if l2_req else None}
    gather["frac"] = gather["achieved"] / l2_gbps if gather["achieved"] else None
    # the rate its hit rate allows: hits at the L2-resident rate, misses at the rate of the same 1 KB
    # row stream over 32 MiB (beyond the L2, in the MALL / HBM), blended by the profiled hit rate
    beyond_gbps = ctx.l2_bandwidth(32 << 20, 3)
    hit = gather["l2_hit_rate"]
    if hit is not None:
        gather["peak_at_hit_rate"] = 1.0 / (hit / l2_gbps + (1.0 - hit) / beyond_gbps)
        gather["peak_beyond_l2"] = beyond_gbps
        gather["frac_at_hit_rate"] = (gather["achieved"] / gather["peak_at_hit_rate"]) if gather["achieved"] else None
    valu_fma, valu_mix, valu_cmp = ctx.valu_rate(0), ctx.valu_rate(1), ctx.valu_rate(2)
    valu_peak = max(valu_fma, valu_mix, valu_cmp)
    # DDA roofline: VALU issue.  Instructions per trace launch from the committed PMC summary of the
    # same workload; the launch time live (HIP events around each chunk's trace launch).
    chunks = recs[-1][0]["viewshed_chunks"]
    avg_trace = t_tr / (chunks * args.steps)
    # (the committed summary is the pooled walk's; with --trace-tables the profile figures do not apply)
    valu_per_launch = None if args.trace_tables else read_profile("trace_pool_c2_valu.txt", "valu_insts_per_launch")
    dda_roof = {"bound": "valu", "unit": "wave64 VALU instructions/s", "peak": valu_peak,
                "peak_probes": {"fp32_fma": valu_fma, "int_add_compare_select": valu_mix,
                                "int_add_compare_to_sgpr": valu_cmp},
                "peak_nominal": VALU_NOMINAL_PEAK, "kernel": f"k_trace_pool<G={ctx.tuning()['dirs_per_group']}>",
                "avg_launch_ms": avg_trace * 1e3, "launches_per_viewshed": chunks,
                "valu_insts_per_launch": valu_per_launch,
                "achieved": (valu_per_launch / avg_trace) if valu_per_launch else None,
                "rays_per_s": N * 2592 * args.steps / t_tr,
                "cell_steps_per_s": dda_steps * args.steps / t_tr,
                "l2_hit_rate": read_profile("trace_pool_c2_valu.txt", "l2_hit_rate"),
                "hbm_traffic_gbps_profiled": read_profile("trace_pool_c2_valu.txt", "hbm_traffic_gbps")}
    dda_roof["frac"] = dda_roof["achieved"] / valu_peak if dda_roof["achieved"] else None
    dda_roof["frac_nominal"] = dda_roof["achieved"] / VALU_NOMINAL_PEAK if dda_roof["achieved"] else None
    # the DDA's second bound: L1 tag lookups of its scattered height gathers (one per lane whose cell
    # shares no line with its quad's), per launch from the same PMC summary, against the probe's rate
    # for lanes on distinct lines of an L1-resident buffer
    l1_rate = ctx.l1_gather_rate(16 << 10, 5)
    l1_per_launch = None if args.trace_tables else read_profile("trace_pool_c2_valu.txt", "l1_accesses_per_launch")
    dda_roof["l1"] = {"bound": "l1 tag lookups", "unit": "lookups/s", "peak": l1_rate,
                      "peak_source": "measured: every lane gathers 4-byte values from its own 128-byte line of a "
                                     "16 KiB (L1-resident) buffer",
                      "lookups_per_launch": l1_per_launch,
                      "achieved": (l1_per_launch / avg_trace) if l1_per_launch else None}
    dda_roof["l1"]["frac"] = dda_roof["l1"]["achieved"] / l1_rate if dda_roof["l1"]["achieved"] else None
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
                   "hours_per_pass": Kp, "nnz": inf["nnz"], "barle_bytes": inf["barle_bytes"],
                   "trace": "step tables" if args.trace_tables else "pooled walk"},
        "dda_rays_per_s": rays,
        "dda_rays_per_s_trace_kernel": N * 2592 * args.steps / t_tr,
        "dda_cell_steps": dda_steps,
        "dda_cell_steps_per_s_trace_kernel": dda_steps * args.steps / t_tr,
        "stage1_ms": t_vs * 1e3 / args.steps,
        "stage2_ms": t_hr * 1e3 / args.steps,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBPS, "traffic": traffic,
                     "kernel": refl_kernel, "avg_launch_ms": avg_launch * 1e3,
                     "algorithmic_bytes_per_launch": bytes_all / passes,
                     # what binds the batched gather: its L2-miss traffic (the PMC figure above) per
                     # launch time, against the copy bandwidth measured in this run
                     "traffic_gbps": (traffic / avg_launch / 1e9) if traffic else None,
                     "traffic_frac_of_copy": (traffic / avg_launch / 1e9 / copy_gbps) if traffic else None},
        "roofline_gather": gather,
        "roofline_dda": dda_roof,
        "measured_copy_gbps": copy_gbps,
        "measured_l2_read_gbps": l2_gbps,
        "measured_l1_gather_lookups_per_s": l1_rate,
        "measured_valu_wave_insts_per_s": {"fp32_fma": valu_fma, "int_add_compare_select": valu_mix,
                                           "int_add_compare_to_sgpr": valu_cmp},
    }
    res.update(extra)
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as orc
        res["cpu_baseline"] = cpu_baseline(sg, orc, ctx, sc, az_alt, dni, dhi, cores, host_cpus)
    ctx.close()
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
