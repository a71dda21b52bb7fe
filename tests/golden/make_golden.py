"""Writes the golden vectors under tests/golden/ from the reference's own
host-only sources, compiled into oracle/_ref/libsolarref.so by
`make -C oracle ref` (sun_position.h, sky_bins.cpp, voxelization.cpp,
barle_compression.cpp of SAMPA-Project/SolarGPU).

Run in the build container (where the reference sources exist):
    make -C oracle ref && python tests/golden/make_golden.py
The fixtures are data only (inputs and expected outputs).
"""
from __future__ import annotations

import ctypes
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))

from oracle_lib import HOURS, load_ref  # noqa: E402

P = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731

SITES = {
    "config": (15.65, 46.56, 1),
    "cli_default": (15.592, 46.543, 1),
    "south_west": (-9.5, 30.5, -1),
    "north_east": (29.0, 59.5, 2),
}


def golden_dsms():
    """Small hand-shaped and seeded DSMs that exercise the wall/slope rules."""
    rng = np.random.default_rng(20260199)
    d = {}
    a = np.full((24, 32), 210.0, np.float32)
    a[5:12, 6:15] = 231.5          # box with a drop of 21.5 m
    a[0:4, 28:32] = 225.25         # box touching the north/east border
    a[15:24, 0:3] = 212.75         # low box touching west/south borders
    a[18, 20] = 260.0              # single tall pillar
    d["boxes_1m"] = (a, 1.0, 1.0)
    b = (205.0 + 3.0 * rng.random((40, 36))).astype(np.float32)
    for _ in range(12):
        x, y = rng.integers(0, 30), rng.integers(0, 34)
        b[y:y + rng.integers(2, 7), x:x + rng.integers(2, 7)] += np.float32(rng.uniform(4, 30))
    d["random_boxes_05m"] = (b, 0.5, 0.5)
    c = np.linspace(200, 230, 30 * 20, dtype=np.float32).reshape(20, 30)
    c[8:12, 10:20] += np.float32(12.0)
    d["slope_aniso"] = (c, 0.75, 1.25)
    return d


def barle_cases():
    rng = np.random.default_rng(20260200)
    cases = [
        [], [1], [2], [15], [16], [17], [255], [256], [257], [65535], [65536], [16777215], [16777216],
        [1 << 26], [3, 2, 1], [1, 2, 3], [5, 5, 5], [100, 90, 80, 70, 60, 50, 40, 30, 20, 10, 1],
        list(range(1, 300)), list(range(1, 600)), list(range(1000, 0, -2)),
        [1000, 999, 700, 699, 698, 500, 10, 9, 8, 7],
        [70000, 69999, 69990, 69700, 50000, 49999, 3, 2, 1],
        [20000000, 19999999, 300000, 299999, 1000, 999, 40, 1],
        [1 << 25, (1 << 25) - 1, (1 << 25) - 17, (1 << 25) - 300, 5, 4, 3],
        list(range(16, 16 * 40, 16)), list(range(17, 17 * 40, 17)),
        [260, 244, 243, 242, 10, 9, 8],
    ]
    # exactly 8 and 9 runs, runs longer than 255 with a width change, alternating widths
    alt = [20000]
    for k in range(30):
        alt.append(alt[-1] - (3 if k % 2 else 40))
    cases.append(alt)
    cases.append(list(np.cumsum(np.tile([1, 300, 1, 20, 1, 70000], 5))[::-1] + 1))
    for n in (1, 2, 7, 8, 9, 31, 64, 200, 900, 2592):
        cases.append(list(rng.choice(4_000_000, size=n, replace=False) + 1))
        cases.append(list(np.unique(rng.integers(1, 3000, size=n))))
    # a raw row in reference form: trace order with duplicates, zero-terminated
    cases.append([17, 16, 18, 17, 400, 16, 1, 0, 99, 98])
    return cases


def main():
    ref = load_ref()
    if ref is None:
        sys.exit("oracle/_ref/libsolarref.so not built: make -C oracle ref")
    out = {}
    for name, (lon, lat, tz) in SITES.items():
        az = np.zeros((HOURS, 2), np.float32)
        ref.ref_sun_positions(ctypes.c_float(lon), ctypes.c_float(lat), tz, P(az))
        nud = np.zeros((HOURS, 2), np.float32)
        f2d = np.zeros(72 * 18, np.int32)
        sdi = np.zeros(HOURS, np.int32)
        hgs = ctypes.c_int()
        act = ref.ref_sky_bins(P(az), P(nud), P(f2d), P(sdi), ctypes.byref(hgs))
        out[f"sun_{name}"] = az
        out[f"nudged_sha256_{name}"] = np.frombuffer(hashlib.sha256(nud.tobytes()).digest(), np.uint8)
        out[f"f2d_{name}"] = f2d
        out[f"sdi_{name}"] = sdi
        out[f"bins_{name}"] = np.array([act, hgs.value], np.int32)
        out[f"site_{name}"] = np.array([lon, lat, tz], np.float64)
    np.savez_compressed(os.path.join(HERE, "sky.npz"), **out)

    vox = {}
    for name, (h, sx, sy) in golden_dsms().items():
        H, W = h.shape
        n = ref.ref_detect_walls(P(h), W, H, ctypes.c_double(sx), ctypes.c_double(sy), ctypes.c_float(sx),
                                 None, None, None, ctypes.c_int64(0))
        parent = np.zeros(n, np.int32)
        z = np.zeros(n, np.float32)
        asp = np.zeros(n, np.float32)
        ref.ref_detect_walls(P(h), W, H, ctypes.c_double(sx), ctypes.c_double(sy), ctypes.c_float(sx), P(parent),
                             P(z), P(asp), ctypes.c_int64(n))
        s = np.zeros(H * W, np.float32)
        a = np.zeros(H * W, np.float32)
        ref.ref_top_slope_aspect(P(h), W, H, ctypes.c_double(sx), ctypes.c_double(sy), P(s), P(a))
        vox[f"{name}_heights"] = h
        vox[f"{name}_scale"] = np.array([sx, sy], np.float64)
        vox[f"{name}_wall_parent"] = parent
        vox[f"{name}_wall_z"] = z
        vox[f"{name}_wall_aspect"] = asp
        vox[f"{name}_slope"] = s
        vox[f"{name}_aspect"] = a
    np.savez_compressed(os.path.join(HERE, "voxelization.npz"), **vox)

    cases = barle_cases()
    width = max(len(c) for c in cases) + 1
    raw = np.zeros((len(cases), width), np.int32)        # reference raw layout: zero-padded rows
    for i, c in enumerate(cases):
        raw[i, : len(c)] = c
    offs = np.zeros(len(cases) + 1, np.int32)
    cap = 64 * raw.size
    flat = np.zeros(cap, np.uint8)
    total = ref.ref_barle_compress_all(P(raw), len(cases), width, P(offs), P(flat), ctypes.c_int64(cap))
    assert total >= 0
    np.savez_compressed(os.path.join(HERE, "barle.npz"), raw=raw, offsets=offs, bytes=flat[:total])

    with open(os.path.join(HERE, "SHA256SUMS"), "w") as f:
        for fn in ("sky.npz", "voxelization.npz", "barle.npz"):
            with open(os.path.join(HERE, fn), "rb") as g:
                f.write(f"{hashlib.sha256(g.read()).hexdigest()}  {fn}\n")
    print("golden vectors written:", {k: v.shape for k, v in out.items() if k.startswith("bins")}, total, "ba-RLE bytes")


if __name__ == "__main__":
    main()
