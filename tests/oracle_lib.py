"""ctypes binding of the CPU oracle (oracle/liboracle.so) -- the checker.

Test infrastructure: only tests, the smoke check and the benchmark's CPU
baseline use it.  Also loads oracle/_ref/libsolarref.so (the reference's
own host-only sources, compiled from /root/reference) when it was built.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libsolarref.so")

HOURS = 8760
NDIRS = 2592

_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_F = ctypes.c_float


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_P)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def load_oracle():
    if not os.path.exists(ORACLE_SO):
        raise ImportError(f"{ORACLE_SO} missing: run make -C oracle")
    lib = ctypes.CDLL(ORACLE_SO)
    for name in ("orc_viewshed", "orc_viewshed_subset", "orc_detect_walls", "orc_barle_encode_voxel",
                 "orc_barle_encode_all", "orc_check_setup_division"):
        getattr(lib, name).restype = ctypes.c_int64
    return lib


def load_ref():
    if not os.path.exists(REF_SO):
        return None
    lib = ctypes.CDLL(REF_SO)
    for name in ("ref_detect_walls", "ref_barle_compress_voxel", "ref_barle_compress_all"):
        getattr(lib, name).restype = ctypes.c_int64
    return lib


O = load_oracle()


def sun_positions(lon, lat, tz):
    out = np.zeros((HOURS, 2), np.float32)
    O.orc_sun_positions(_F(lon), _F(lat), tz, _ptr(out))
    return out


def sky_bins(az_alt):
    az_alt = _c(az_alt, np.float32)
    nud = np.zeros((HOURS, 2), np.float32)
    f2d = np.zeros(72 * 18, np.int32)
    sdi = np.zeros(HOURS, np.int32)
    hs, vs, hgs = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    act = O.orc_sky_bins(_ptr(az_alt), _F(5.0), _F(5.0), _ptr(nud), _ptr(f2d), _ptr(sdi), ctypes.byref(hs),
                         ctypes.byref(vs), ctypes.byref(hgs))
    return dict(nudged=nud, full_to_dense=f2d, sun_dense_idx=sdi, active=act, svf_bytes=hgs.value)


def directions(full_to_dense):
    cdir = np.zeros((NDIRS, 3), np.float32)
    up = np.zeros(NDIRS, np.int32)
    dense = np.zeros(NDIRS, np.int32)
    O.orc_directions(72, 18, _F(5.0), _F(5.0), _ptr(_c(full_to_dense, np.int32)), _ptr(cdir), _ptr(up), _ptr(dense))
    return cdir, up, dense


def detect_walls(h, voxel):
    h = _c(h, np.float32)
    H, W = h.shape
    n = O.orc_detect_walls(_ptr(h), W, H, _F(voxel), None, None, None, _I64(0))
    p = np.zeros(n, np.int32)
    z = np.zeros(n, np.float32)
    a = np.zeros(n, np.float32)
    O.orc_detect_walls(_ptr(h), W, H, _F(voxel), _ptr(p), _ptr(z), _ptr(a), _I64(n))
    return p, z, a


def top_slope_aspect(h, sx, sy):
    h = _c(h, np.float32)
    H, W = h.shape
    s = np.zeros(H * W, np.float32)
    a = np.zeros(H * W, np.float32)
    O.orc_top_slope_aspect(_ptr(h), W, H, ctypes.c_double(sx), ctypes.c_double(sy), _ptr(s), _ptr(a))
    return s, a


def check_setup_division(cdir, gmax):
    """Set-up quotients (all directions, both axes, origins < gmax) where the GPU's reciprocal +
    two-FMA form differs from IEEE division; 0 pins the product's set-up."""
    c = _c(cdir, np.float32)
    return O.orc_check_setup_division(_ptr(c), len(c) // 3, int(gmax))


def trace_ray(grid, maxz, gx, gy, base_z, cdir3, upper):
    grid = _c(grid, np.float32)
    H, W = grid.shape
    steps = ctypes.c_int()
    c = _c(cdir3, np.float32)
    hit = O.orc_trace_ray(_ptr(grid), W, H, _F(maxz), _F(gx), _F(gy), _F(base_z), _ptr(c), int(upper),
                          ctypes.byref(steps))
    return hit, steps.value


def viewshed(sc, bins, nthreads=0):
    """CSR lists (first-occurrence direction order) + word-major SVF of scene `sc`."""
    cdir, up, dense = directions(bins["full_to_dense"])
    N = sc.num_total
    hgs = bins["svf_bytes"]
    svf = np.zeros(max(hgs * N, 1), np.uint8)
    off = np.zeros(N + 1, np.int64)
    steps = ctypes.c_int64()
    cap = 64 * N + 1024
    while True:
        ids = np.zeros(cap, np.int32)
        nnz = O.orc_viewshed(_ptr(sc.grid), sc.width, sc.height, N, sc.num_top, _ptr(sc.pixel_parent),
                             _ptr(sc.wall_z) if sc.wall_z.size else None, _F(sc.max_height), _ptr(cdir), _ptr(up),
                             _ptr(dense), NDIRS, hgs, nthreads, _ptr(svf), _ptr(off), _ptr(ids), _I64(cap),
                             ctypes.byref(steps))
        if nnz >= 0:
            return dict(svf=svf[: hgs * N], off=off, ids=ids[:nnz].copy(), steps=steps.value)
        cap = -nnz


def viewshed_subset(sc, bins, voxels, nthreads=0, with_svf=False):
    """CSR lists (by position) of the voxels `voxels`, optionally with their SVF bits
    (word-major over the subset); also the DDA step count."""
    cdir, up, dense = directions(bins["full_to_dense"])
    vox = _c(voxels, np.int32)
    n = vox.size
    hgs = bins["svf_bytes"]
    svf = np.zeros(max(hgs * n, 1), np.uint8) if with_svf else None
    off = np.zeros(n + 1, np.int64)
    steps = ctypes.c_int64()
    cap = 600 * n + 1024
    while True:
        ids = np.zeros(cap, np.int32)
        nnz = O.orc_viewshed_subset(_ptr(sc.grid), sc.width, sc.height, sc.num_top, _ptr(sc.pixel_parent),
                                    _ptr(sc.wall_z) if sc.wall_z.size else None, _F(sc.max_height), _ptr(cdir),
                                    _ptr(up), _ptr(dense), NDIRS, hgs, _ptr(vox), _I64(n), nthreads, _ptr(svf),
                                    _ptr(off), _ptr(ids), _I64(cap), ctypes.byref(steps))
        if nnz >= 0:
            return dict(svf=None if svf is None else svf[: hgs * n], off=off, ids=ids[:nnz].copy(),
                        steps=steps.value)
        cap = -nnz


def trace_raw(sc, bins, v0, count, nthreads=0):
    cdir, up, _ = directions(bins["full_to_dense"])
    raw = np.zeros((count, NDIRS), np.int32)
    O.orc_trace_raw(_ptr(sc.grid), sc.width, sc.height, sc.num_top, _ptr(sc.pixel_parent),
                    _ptr(sc.wall_z) if sc.wall_z.size else None, _F(sc.max_height), _ptr(cdir), _ptr(up), NDIRS,
                    v0, count, nthreads, _ptr(raw))
    return raw


def barle_encode_voxel(ids):
    ids = _c(ids, np.int32)
    cap = 16 * max(ids.size, 1) + 64
    out = np.zeros(cap, np.uint8)
    n = O.orc_barle_encode_voxel(_ptr(ids), ids.size, _ptr(out), _I64(cap))
    assert n >= 0
    return out[:n].copy()


def barle_encode_all(off, ids, nthreads=0):
    off = _c(off, np.int64)
    ids = _c(ids, np.int32)
    N = off.size - 1
    cap = 8 * max(ids.size, 1) + 64
    while True:
        out = np.zeros(cap, np.uint8)
        boff = np.zeros(N + 1, np.int64)
        n = O.orc_barle_encode_all(_ptr(off), _ptr(ids) if ids.size else None, N, nthreads, _ptr(boff), _ptr(out),
                                   _I64(cap))
        if n >= 0:
            return boff, out[:n].copy()
        cap = -n


def barle_decode(data):
    data = _c(data, np.uint8)
    cap = 8 * max(data.size, 1) + 64
    out = np.zeros(cap, np.uint32)
    n = O.orc_barle_decode(_ptr(data), _I64(data.size), _ptr(out), cap)
    if n < 0:
        raise ValueError("malformed ba-RLE stream")
    return out[:n].copy()


def sunlit_hours(nudged):
    out = np.zeros(HOURS, np.int32)
    n = O.orc_sunlit_hours(_ptr(_c(nudged, np.float32)), _ptr(out))
    return out[:n].copy()


def hour_eps(alt, dif, dir_):
    """FP32 Perez clearness epsilon of one hour and its bin."""
    e = ctypes.c_float()
    b = O.orc_hour_eps(_F(alt), _F(dif), _F(dir_), ctypes.byref(e))
    return e.value, b


def solar_hour(sc, bins, dni, dhi, svf, hour_index, nthreads=0):
    out = np.zeros(sc.num_total, np.float32)
    O.orc_solar_hour(_ptr(bins["nudged"]), _ptr(bins["sun_dense_idx"]), _ptr(_c(dni, np.float32)),
                     _ptr(_c(dhi, np.float32)), hour_index, sc.num_total, _ptr(sc.slopes), _ptr(sc.aspects),
                     _ptr(_c(svf, np.uint8)), sc.num_total, nthreads, _ptr(out))
    return out


def _scene_args(sc):
    return (_ptr(sc.grid), sc.width, sc.height, sc.num_total, sc.num_top, _ptr(sc.pixel_parent),
            _ptr(sc.wall_z) if sc.wall_z.size else None, _ptr(sc.slopes), _ptr(sc.aspects))


def reflect_hour_csr(sc, off, ids, hourly, nthreads=0):
    out = np.zeros(sc.num_total, np.float32)
    ids = _c(ids, np.int32)
    O.orc_reflect_hour_csr(*_scene_args(sc), _ptr(_c(off, np.int64)), _ptr(ids) if ids.size else None,
                           _ptr(_c(hourly, np.float32)), nthreads, _ptr(out))
    return out


def reflect_hour_barle(sc, boff, data, hourly, nthreads=0):
    out = np.zeros(sc.num_total, np.float32)
    data = _c(data, np.uint8)
    O.orc_reflect_hour_barle(*_scene_args(sc), _ptr(_c(boff, np.int64)), _ptr(data) if data.size else None,
                             _ptr(_c(hourly, np.float32)), nthreads, _ptr(out))
    return out


def annual_fp32(sc, bins, dni, dhi, svf, off, ids, boff=None, data=None, mode=0, stride=1, kahan=1, nthreads=0):
    out = np.zeros(sc.num_total, np.float32)
    ids = _c(ids, np.int32)
    n = O.orc_annual_fp32(*_scene_args(sc), _ptr(bins["nudged"]), _ptr(bins["sun_dense_idx"]),
                          _ptr(_c(dni, np.float32)), _ptr(_c(dhi, np.float32)), _ptr(_c(svf, np.uint8)),
                          _ptr(_c(off if mode == 0 else boff, np.int64)), _ptr(ids) if ids.size else None,
                          _ptr(_c(data, np.uint8)) if data is not None else None, mode, stride, kahan, nthreads,
                          _ptr(out))
    return out, n


def annual_fp32_range(sc, bins, dni, dhi, svf, v0, off_slice, ids, stride=1, nthreads=0):
    """Reference-structured FP32 hour loop for voxels [v0, v0 + len(off_slice) - 1)."""
    count = off_slice.size - 1
    out = np.zeros(count, np.float32)
    ids = _c(ids, np.int32)
    tt, tr = ctypes.c_double(), ctypes.c_double()
    n = O.orc_annual_fp32_range(*_scene_args(sc), _ptr(bins["nudged"]), _ptr(bins["sun_dense_idx"]),
                                _ptr(_c(dni, np.float32)), _ptr(_c(dhi, np.float32)), _ptr(_c(svf, np.uint8)),
                                _I64(v0), _I64(count), _ptr(_c(off_slice, np.int64)), _ptr(ids), stride, nthreads,
                                _ptr(out), ctypes.byref(tt), ctypes.byref(tr))
    return out, n, tt.value, tr.value


def annual_fp32_subset(sc, bins, dni, dhi, svf, voxels, off_sub, ids, stride=1, nthreads=0):
    """Reference-structured FP32 hour loop for the voxels `voxels` with their lists by position;
    returns (annual, hours done, all-top solar seconds, per-voxel seconds)."""
    vox = _c(voxels, np.int32)
    out = np.zeros(vox.size, np.float32)
    ids = _c(ids, np.int32)
    tt, tr = ctypes.c_double(), ctypes.c_double()
    n = O.orc_annual_fp32_subset(*_scene_args(sc), _ptr(bins["nudged"]), _ptr(bins["sun_dense_idx"]),
                                 _ptr(_c(dni, np.float32)), _ptr(_c(dhi, np.float32)), _ptr(_c(svf, np.uint8)),
                                 _ptr(vox), _I64(vox.size), _ptr(_c(off_sub, np.int64)),
                                 _ptr(ids) if ids.size else None, stride, nthreads, _ptr(out), ctypes.byref(tt),
                                 ctypes.byref(tr))
    return out, n, tt.value, tr.value


def annual_fp64(sc, bins, dni, dhi, svf, off, ids, voxels=None, nthreads=0):
    ids = _c(ids, np.int32)
    sel = None if voxels is None else _c(voxels, np.int32)
    n_out = sc.num_total if sel is None else sel.size
    out = np.zeros(n_out, np.float64)
    O.orc_annual_fp64(*_scene_args(sc), _ptr(bins["nudged"]), _ptr(bins["sun_dense_idx"]),
                      _ptr(_c(dni, np.float32)), _ptr(_c(dhi, np.float32)), _ptr(_c(svf, np.uint8)),
                      _ptr(_c(off, np.int64)), _ptr(ids) if ids.size else None, _ptr(sel),
                      0 if sel is None else sel.size, nthreads, _ptr(out))
    return out


def annual_fp64_subset(sc, bins, dni, dhi, svf, voxels, off_sub, ids_sub, nthreads=0):
    """FP64 annual totals of the voxels `voxels` from their lists by position."""
    vox = _c(voxels, np.int32)
    ids = _c(ids_sub, np.int32)
    out = np.zeros(vox.size, np.float64)
    O.orc_annual_fp64_subset(*_scene_args(sc), _ptr(bins["nudged"]), _ptr(bins["sun_dense_idx"]),
                             _ptr(_c(dni, np.float32)), _ptr(_c(dhi, np.float32)), _ptr(_c(svf, np.uint8)),
                             _ptr(vox), _I64(vox.size), _ptr(_c(off_sub, np.int64)), _ptr(ids) if ids.size else None,
                             nthreads, _ptr(out))
    return out


def annual_fp64_direct(sc, bins, dni, dhi, svf, off, ids):
    ids = _c(ids, np.int32)
    out = np.zeros(sc.num_total, np.float64)
    O.orc_annual_fp64_direct(*_scene_args(sc), _ptr(bins["nudged"]), _ptr(bins["sun_dense_idx"]),
                             _ptr(_c(dni, np.float32)), _ptr(_c(dhi, np.float32)), _ptr(_c(svf, np.uint8)),
                             _ptr(_c(off, np.int64)), _ptr(ids) if ids.size else None, _ptr(out))
    return out
