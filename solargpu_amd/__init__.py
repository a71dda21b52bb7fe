"""Python binding of libsolargpu (the MI355X SolarGPU pipeline) over its C ABI.

The compute path is the HIP library ``solargpu_amd/lib/libsolargpu.so``; this
module only moves numpy arrays across ``include/sgpu.h``.  Importing fails
loudly when the library has not been built -- there is no CPU fallback.

Reference entry points mirrored here (SAMPA-Project/SolarGPU):
``uncompressed_gpu_irradiance`` / ``compressed_gpu_irradiance``
(include/solar_gpu_api.h:37-54), ``SunPositionCalculator.calc``
(include/sun_position.h:24-40), ``compute_sky_bins`` (src/sky_bins.cpp:9-52),
``detect_wall_voxels`` / ``compute_top_slope_aspect`` (src/voxelization.cpp:10-97)
and the voxel assembly of src/main.cpp:112-127.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

HOURS = 8760
NDIRS = 2592
FMT_CSR = 1
FMT_BARLE = 2

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libsolargpu.so")


class SolarGPUError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"sgpu status {code}: {msg}")
        self.code = code


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    return ctypes.CDLL(LIB_PATH)


_lib = _load()
_lib.sgpu_last_error.restype = ctypes.c_char_p
_lib.sgpu_version.restype = ctypes.c_char_p

_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_F = ctypes.c_float


class Info(ctypes.Structure):
    _fields_ = [
        ("grid_width", ctypes.c_int32), ("grid_height", ctypes.c_int32),
        ("num_total", ctypes.c_int64), ("num_top", ctypes.c_int64),
        ("active_bins", ctypes.c_int32), ("svf_bytes_per_voxel", ctypes.c_int32),
        ("sunlit_hours", ctypes.c_int32), ("viewshed_chunks", ctypes.c_int32),
        ("chunk_voxels", ctypes.c_int64), ("nnz", ctypes.c_int64), ("barle_bytes", ctypes.c_int64),
        ("t_trace_ms", ctypes.c_double), ("t_lists_ms", ctypes.c_double),
        ("t_viewshed_ms", ctypes.c_double), ("t_hours_ms", ctypes.c_double),
        ("hours_computed", ctypes.c_int32), ("hour_passes", ctypes.c_int32),
        ("hours_per_pass", ctypes.c_int32), ("timed_launches", ctypes.c_int32),
        ("t_refl_kernel_ms", ctypes.c_double), ("t_solar_kernel_ms", ctypes.c_double),
        ("dda_steps", ctypes.c_int64),
        ("table_pending_rays", ctypes.c_int64), ("table_irregular_voxels", ctypes.c_int64),
        ("table_chunks", ctypes.c_int32), ("table_fallbacks", ctypes.c_int32),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class Tuning(ctypes.Structure):
    _fields_ = [("dirs_per_group", ctypes.c_int32), ("hours_per_pass", ctypes.c_int32),
                ("waves_per_group", ctypes.c_int32), ("xcd_band", ctypes.c_int32), ("group_blocks", ctypes.c_int32),
                ("from_cache", ctypes.c_int32),
                ("rays_per_s", ctypes.c_double), ("voxel_hours_per_s", ctypes.c_double),
                ("receiver_queue", ctypes.c_int32)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


def _check(code: int) -> None:
    if code != 0:
        raise SolarGPUError(code, _lib.sgpu_last_error().decode(errors="replace"))


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(_P)


def _c(a, dtype) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=dtype)


def version() -> str:
    return _lib.sgpu_version().decode()


# ---------------- host-side contract ----------------

def sun_positions(lon: float, lat: float, tz: int) -> np.ndarray:
    """(8760, 2) float32 (azimuth, altitude) degrees, PSA, year 2026."""
    out = np.zeros((HOURS, 2), np.float32)
    _check(_lib.sgpu_sun_positions(_F(lon), _F(lat), ctypes.c_int(tz), _ptr(out)))
    return out


@dataclass
class SkyBins:
    nudged: np.ndarray          # (8760, 2)
    full_to_dense: np.ndarray   # (72*18,)
    sun_dense_idx: np.ndarray   # (8760,)
    active_bins: int
    svf_bytes: int


def sky_bins(az_alt: np.ndarray) -> SkyBins:
    az_alt = _c(az_alt, np.float32)
    nud = np.zeros((HOURS, 2), np.float32)
    f2d = np.zeros(72 * 18, np.int32)
    sdi = np.zeros(HOURS, np.int32)
    act = ctypes.c_int32()
    hgs = ctypes.c_int32()
    _check(_lib.sgpu_sky_bins(_ptr(az_alt), _ptr(nud), _ptr(f2d), _ptr(sdi), ctypes.byref(act), ctypes.byref(hgs)))
    return SkyBins(nud, f2d, sdi, act.value, hgs.value)


def directions() -> tuple[np.ndarray, np.ndarray]:
    cdir = np.zeros((NDIRS, 3), np.float32)
    up = np.zeros(NDIRS, np.int32)
    _check(_lib.sgpu_directions(_ptr(cdir), _ptr(up)))
    return cdir, up


def detect_walls(heights: np.ndarray, voxel: float):
    h = _c(heights, np.float32)
    H, W = h.shape
    cnt = _I64()
    _check(_lib.sgpu_detect_walls(_ptr(h), W, H, _F(voxel), None, None, None, _I64(0), ctypes.byref(cnt)))
    n = cnt.value
    parent = np.zeros(n, np.int32)
    z = np.zeros(n, np.float32)
    asp = np.zeros(n, np.float32)
    _check(_lib.sgpu_detect_walls(_ptr(h), W, H, _F(voxel), _ptr(parent), _ptr(z), _ptr(asp), _I64(n),
                                  ctypes.byref(cnt)))
    return parent, z, asp


def top_slope_aspect(heights: np.ndarray, sx: float, sy: float):
    h = _c(heights, np.float32)
    H, W = h.shape
    s = np.zeros(H * W, np.float32)
    a = np.zeros(H * W, np.float32)
    _check(_lib.sgpu_top_slope_aspect(_ptr(h), W, H, ctypes.c_double(sx), ctypes.c_double(sy), _ptr(s), _ptr(a)))
    return s, a


@dataclass
class Scene:
    """Voxelized DSM in the layout of the reference's main.cpp:112-127."""
    grid: np.ndarray          # (H, W) float32
    pixel_parent: np.ndarray  # (N,) int32
    slopes: np.ndarray        # (N,) float32
    aspects: np.ndarray       # (N,) float32
    wall_z: np.ndarray        # (N - W*H,) float32
    max_height: float
    cell: float

    @property
    def width(self) -> int:
        return int(self.grid.shape[1])

    @property
    def height(self) -> int:
        return int(self.grid.shape[0])

    @property
    def num_top(self) -> int:
        return int(self.grid.size)

    @property
    def num_total(self) -> int:
        return int(self.pixel_parent.size)


def assemble_scene(heights: np.ndarray, cell: float, cell_y: float | None = None) -> Scene:
    h = _c(heights, np.float32)
    H, W = h.shape
    sy = cell if cell_y is None else cell_y
    n = _I64()
    _check(_lib.sgpu_assemble_voxels(_ptr(h), W, H, ctypes.c_double(cell), ctypes.c_double(sy), _I64(0),
                                     ctypes.byref(n), None, None, None, None))
    N = n.value
    pp = np.zeros(N, np.int32)
    sl = np.zeros(N, np.float32)
    asp = np.zeros(N, np.float32)
    wz = np.zeros(max(N - W * H, 0), np.float32)
    _check(_lib.sgpu_assemble_voxels(_ptr(h), W, H, ctypes.c_double(cell), ctypes.c_double(sy), _I64(N),
                                     ctypes.byref(n), _ptr(pp), _ptr(sl), _ptr(asp), _ptr(wz)))
    return Scene(h, pp, sl, asp, wz, float(h.max()), float(cell))


class GeoTiff:
    """A float32 GeoTIFF read by the library's native reader (no libtiff)."""

    def __init__(self, path: str):
        self._h = ctypes.c_void_p()
        _check(_lib.sgpu_geotiff_open(path.encode(), ctypes.byref(self._h)))
        w, h, comp = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        sx, sy = ctypes.c_double(), ctypes.c_double()
        _check(_lib.sgpu_geotiff_info(self._h, ctypes.byref(w), ctypes.byref(h), ctypes.byref(sx),
                                      ctypes.byref(sy), ctypes.byref(comp)))
        self.width, self.height = w.value, h.value
        self.scale = (sx.value, sy.value)
        self.compression = comp.value

    def read(self) -> np.ndarray:
        out = np.zeros((self.height, self.width), np.float32)
        _check(_lib.sgpu_geotiff_read(self._h, _ptr(out)))
        return out

    def write(self, path: str, data: np.ndarray) -> None:
        """Write `data` (H x W float32) with this file's geo tags and compression."""
        d = _c(data, np.float32)
        _check(_lib.sgpu_geotiff_write(self._h, path.encode(), _ptr(d), d.shape[1], d.shape[0]))

    def close(self) -> None:
        if self._h:
            _lib.sgpu_geotiff_close(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        self.close()


def read_tmy(path: str, n: int = HOURS) -> np.ndarray:
    out = np.zeros(n, np.float32)
    _check(_lib.sgpu_read_tmy(path.encode(), _ptr(out), n))
    return out


def write_voxel_dump(path: str, sc: "Scene", irradiance: np.ndarray) -> None:
    irr = _c(irradiance, np.float32)
    _check(_lib.sgpu_write_voxel_dump(path.encode(), sc.num_total, sc.num_top, sc.width, _ptr(sc.pixel_parent),
                                      _ptr(sc.grid), _ptr(sc.wall_z) if sc.wall_z.size else None, _ptr(irr)))


def synth_dsm(seed: int, width: int, height: int, cell: float, density: float, trees: int) -> np.ndarray:
    out = np.zeros((height, width), np.float32)
    _check(_lib.sgpu_synth_dsm(ctypes.c_uint64(seed), width, height, _F(cell), _F(density), trees, _ptr(out)))
    return out


def synth_weather(seed: int, az_alt: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    az_alt = _c(az_alt, np.float32)
    dni = np.zeros(HOURS, np.float32)
    dhi = np.zeros(HOURS, np.float32)
    _check(_lib.sgpu_synth_weather(ctypes.c_uint64(seed), _ptr(az_alt), _ptr(dni), _ptr(dhi)))
    return dni, dhi


# ---------------- device pipeline ----------------

class Context:
    """One GPU pipeline context (sgpu_ctx): device buffers live until close()."""

    def __init__(self, device: int = 0):
        self._h = ctypes.c_void_p()
        _check(_lib.sgpu_create(device, ctypes.byref(self._h)))

    def close(self) -> None:
        if self._h:
            _check(_lib.sgpu_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> dict:
        i = Info()
        _check(_lib.sgpu_get_info(self._h, ctypes.byref(i)))
        return i.as_dict()

    def set_scene(self, sc: Scene) -> None:
        self._scene = sc
        _check(_lib.sgpu_set_scene(self._h, _ptr(sc.grid), sc.width, sc.height, _I64(sc.num_total),
                                   _I64(sc.num_top), _ptr(sc.pixel_parent), _ptr(sc.wall_z) if sc.wall_z.size else None,
                                   _F(sc.max_height), _ptr(sc.slopes), _ptr(sc.aspects)))

    def set_scene_heights(self, heights: np.ndarray, cell: float, cell_y: float | None = None) -> None:
        """Set the scene from a heightmap voxelized on the GPU (the arrays stay on the device)."""
        h = _c(heights, np.float32)
        self._scene = None
        _check(_lib.sgpu_set_scene_heights(self._h, _ptr(h), h.shape[1], h.shape[0], ctypes.c_double(cell),
                                           ctypes.c_double(cell if cell_y is None else cell_y)))

    def save_lists(self, path: str, fmt: int = FMT_CSR | FMT_BARLE) -> None:
        """Write the SVF bitmap and the lists of `fmt` to `path` (reruns skip the viewshed)."""
        _check(_lib.sgpu_save_lists(self._h, path.encode(), ctypes.c_uint(fmt)))

    def load_lists(self, path: str) -> None:
        """Read lists saved for this scene and sky, in place of viewshed()."""
        _check(_lib.sgpu_load_lists(self._h, path.encode()))

    def orders(self) -> tuple[np.ndarray, np.ndarray]:
        """(reflection-pass processing order, DDA lane order) of the context's scene."""
        n = self.info()["num_total"]
        order = np.zeros(n, np.int32)
        lanes = np.zeros(n, np.int32)
        _check(_lib.sgpu_get_orders(self._h, _ptr(order), _ptr(lanes)))
        return order, lanes

    def voxels(self, heights: np.ndarray, cell: float) -> Scene:
        """The context's voxel arrays as a Scene (heights and cell as given to set_scene_heights)."""
        h = _c(heights, np.float32)
        n = self.info()["num_total"]
        pp = np.zeros(n, np.int32)
        sl = np.zeros(n, np.float32)
        asp = np.zeros(n, np.float32)
        wz = np.zeros(max(n - h.size, 0), np.float32)
        _check(_lib.sgpu_get_voxels(self._h, _ptr(pp), _ptr(sl), _ptr(asp), _ptr(wz) if wz.size else None))
        return Scene(h, pp, sl, asp, wz, float(h.max()), float(cell))

    def set_sky(self, az_alt: np.ndarray, dni: np.ndarray, dhi: np.ndarray) -> None:
        self._az_alt = _c(az_alt, np.float32)
        _check(_lib.sgpu_set_sky(self._h, _ptr(self._az_alt), _ptr(_c(dni, np.float32)), _ptr(_c(dhi, np.float32))))

    def set_chunk_voxels(self, n: int) -> None:
        _check(_lib.sgpu_set_chunk_voxels(self._h, _I64(n)))

    def set_list_residency(self, host: bool) -> None:
        """host=True: lists in pinned host memory, streamed to the GPU per pass."""
        _check(_lib.sgpu_set_list_residency(self._h, 1 if host else 0))

    def set_kernel_timing(self, on: bool) -> None:
        _check(_lib.sgpu_set_kernel_timing(self._h, int(on)))

    def set_hourly_limit(self, floats: int) -> None:
        """Diagnostic: cap the hourly buffer the hour passes may use (0 = all of it)."""
        _check(_lib.sgpu_set_hourly_limit(self._h, ctypes.c_int64(floats)))

    def set_step_counting(self, on: bool) -> None:
        _check(_lib.sgpu_set_step_counting(self._h, int(on)))

    def set_trace_tables(self, on: bool, pending_cap: int = 0) -> None:
        """Whole-scene viewsheds by step tables or by the pooled walk alone (default); pending_cap
        > 0 caps the pending rays per chunk (forces the overflow path in tests)."""
        _check(_lib.sgpu_set_trace_tables(self._h, int(on), _I64(pending_cap)))

    def set_trace_table_shape(self, dirs: int = 4, steps: int = 255, skip: bool = True) -> None:
        _check(_lib.sgpu_set_trace_table_shape(self._h, int(dirs), int(steps), int(skip)))

    def set_pair_limit(self, entries: int = 0) -> None:
        """Diagnostic: receiver pairs whose kept lists hold more than `entries` entries together are
        walked receiver by receiver (0 = only pairs that overflow the pair scratch)."""
        _check(_lib.sgpu_set_pair_limit(self._h, int(entries)))

    def viewshed(self, formats: int = FMT_CSR | FMT_BARLE) -> None:
        _check(_lib.sgpu_viewshed(self._h, ctypes.c_uint(formats)))

    def svf(self) -> np.ndarray:
        inf = self.info()
        out = np.zeros(inf["svf_bytes_per_voxel"] * inf["num_total"], np.uint8)
        _check(_lib.sgpu_get_svf(self._h, _ptr(out)))
        return out

    def csr(self) -> tuple[np.ndarray, np.ndarray]:
        inf = self.info()
        off = np.zeros(inf["num_total"] + 1, np.int64)
        ids = np.zeros(max(inf["nnz"], 1), np.int32)
        _check(_lib.sgpu_get_csr(self._h, _ptr(off), _ptr(ids)))
        return off, ids[: inf["nnz"]]

    def csr_range(self, v0: int, count: int) -> tuple[np.ndarray, np.ndarray]:
        off = np.zeros(count + 1, np.int64)
        cap = count * NDIRS
        ids = np.zeros(max(cap, 1), np.int32)
        _check(_lib.sgpu_get_csr_range(self._h, _I64(v0), _I64(count), _ptr(off), _ptr(ids), _I64(cap)))
        return off, ids[: off[-1]].copy()

    def barle_range(self, v0: int, count: int) -> tuple[np.ndarray, np.ndarray]:
        """ba-RLE streams of voxels [v0, v0 + count), offsets rebased to 0."""
        off = np.zeros(count + 1, np.int64)
        cap = count * (6 * NDIRS + 8)          # a stream of n ids takes at most 5.375 n + 3 bytes
        b = np.zeros(max(cap, 1), np.uint8)
        _check(_lib.sgpu_get_barle_range(self._h, _I64(v0), _I64(count), _ptr(off), _ptr(b), _I64(cap)))
        return off, b[: off[-1]].copy()

    def barle(self) -> tuple[np.ndarray, np.ndarray]:
        inf = self.info()
        off = np.zeros(inf["num_total"] + 1, np.int64)
        b = np.zeros(max(inf["barle_bytes"], 1), np.uint8)
        _check(_lib.sgpu_get_barle(self._h, _ptr(off), _ptr(b)))
        return off, b[: inf["barle_bytes"]]

    def trace_raw(self, v0: int, count: int) -> np.ndarray:
        out = np.zeros((count, NDIRS), np.int32)
        _check(_lib.sgpu_trace_raw(self._h, _I64(v0), _I64(count), _ptr(out)))
        return out

    def set_svf(self, svf: np.ndarray) -> None:
        _check(_lib.sgpu_set_svf(self._h, _ptr(_c(svf, np.uint8))))

    def set_csr(self, off: np.ndarray, ids: np.ndarray) -> None:
        ids = _c(ids, np.int32)
        if ids.size == 0:
            ids = np.zeros(1, np.int32)
        _check(_lib.sgpu_set_csr(self._h, _ptr(_c(off, np.int64)), _ptr(ids)))

    def set_barle(self, off: np.ndarray, data: np.ndarray) -> None:
        data = _c(data, np.uint8)
        if data.size == 0:
            data = np.zeros(1, np.uint8)
        _check(_lib.sgpu_set_barle(self._h, _ptr(_c(off, np.int64)), _ptr(data)))

    def barle_encode(self) -> None:
        _check(_lib.sgpu_barle_encode(self._h))

    def hour_loop(self, fmt: int, hours_per_pass: int = 0, hour_stride: int = 1) -> None:
        _check(_lib.sgpu_hour_loop(self._h, ctypes.c_uint(fmt), hours_per_pass, hour_stride))

    def hour_loop_noalloc(self, fmt: int, hours_per_pass: int, out_device_ptr: int) -> None:
        """Allocation-free hour loop into a caller-owned device buffer of num_total floats
        (enqueued on the context stream; call sync() before reading the buffer)."""
        _check(_lib.sgpu_hour_loop_noalloc(self._h, ctypes.c_uint(fmt), hours_per_pass,
                                           ctypes.c_void_p(out_device_ptr)))

    def hour_loop_graph(self, fmt: int, hours_per_pass: int = 0, replays: int = 1) -> None:
        _check(_lib.sgpu_hour_loop_graph(self._h, ctypes.c_uint(fmt), hours_per_pass, replays))

    def irradiance(self) -> np.ndarray:
        out = np.zeros(self.info()["num_total"], np.float32)
        _check(_lib.sgpu_get_irradiance(self._h, _ptr(out)))
        return out

    def hour_single(self, fmt: int, hour_index: int) -> tuple[np.ndarray, np.ndarray]:
        n = self.info()["num_total"]
        hv = np.zeros(n, np.float32)
        rv = np.zeros(n, np.float32)
        _check(_lib.sgpu_hour_single(self._h, ctypes.c_uint(fmt), hour_index, _ptr(hv), _ptr(rv)))
        return hv, rv

    def sync(self) -> None:
        _check(_lib.sgpu_sync(self._h))

    def autotune(self, fmt: int = FMT_BARLE) -> dict:
        t = Tuning()
        _check(_lib.sgpu_autotune(self._h, ctypes.c_uint(fmt), ctypes.byref(t)))
        return t.as_dict()

    def tuning(self) -> dict:
        t = Tuning()
        _check(_lib.sgpu_get_tuning(self._h, ctypes.byref(t)))
        return t.as_dict()

    def set_tuning(self, dirs_per_group: int = 16, hours_per_pass: int = 64, waves_per_group: int = 4,
                   xcd_band: int = 0, group_blocks: int = 0, receiver_queue: int = 0) -> None:
        t = Tuning(dirs_per_group, hours_per_pass, waves_per_group, xcd_band, group_blocks, 0, 0.0, 0.0,
                   receiver_queue)
        _check(_lib.sgpu_set_tuning(self._h, ctypes.byref(t)))

    def copy_bandwidth(self, nbytes: int, reps: int = 10) -> float:
        """GB/s (read + write) of a plain float4 device copy of nbytes."""
        g = ctypes.c_double()
        _check(_lib.sgpu_copy_bandwidth(self._h, _I64(nbytes), reps, ctypes.byref(g)))
        return g.value

    def l2_bandwidth(self, footprint: int = 2 << 20, reps: int = 3) -> float:
        """GB/s of L2-resident reads: every wave streams a footprint-byte buffer as 1 KB rows."""
        g = ctypes.c_double()
        _check(_lib.sgpu_l2_bandwidth(self._h, _I64(footprint), reps, ctypes.byref(g)))
        return g.value

    def l1_gather_rate(self, footprint: int = 16 << 10, reps: int = 5) -> float:
        """L1 tag lookups/s of scattered 4-byte gathers, every lane on its own 128-byte line."""
        g = ctypes.c_double()
        _check(_lib.sgpu_l1_gather_rate(self._h, _I64(footprint), reps, ctypes.byref(g)))
        return g.value

    def valu_rate(self, mix: int = 0) -> float:
        """Measured VALU issue rate, wave64 instructions/s, of a probe filling every SIMD:
        mix 0 = FP32 FMAs, mix 1 = integer add / compare / select (the DDA's classes)."""
        r = ctypes.c_double()
        _check(_lib.sgpu_valu_rate(self._h, mix, ctypes.byref(r)))
        return r.value


class BatchInfo(ctypes.Structure):
    _fields_ = [("tiles", ctypes.c_int32), ("chunks", ctypes.c_int32), ("voxels", ctypes.c_int64),
                ("nnz", ctypes.c_int64), ("barle_bytes", ctypes.c_int64), ("hours_per_pass", ctypes.c_int32),
                ("passes", ctypes.c_int32), ("voxel_hours", ctypes.c_int64), ("t_viewshed_ms", ctypes.c_double),
                ("t_trace_ms", ctypes.c_double), ("t_lists_ms", ctypes.c_double), ("t_encode_ms", ctypes.c_double),
                ("t_hours_ms", ctypes.c_double)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class Batch:
    """Many DSM tiles (own sizes and sites) through one batched pipeline (sgpu_batch_*)."""

    def __init__(self, device: int = 0):
        self._h = ctypes.c_void_p()
        self._keep = []
        _check(_lib.sgpu_batch_create(device, ctypes.byref(self._h)))

    def close(self) -> None:
        if self._h:
            _check(_lib.sgpu_batch_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_tile(self, sc: Scene, az_alt, dni, dhi) -> int:
        idx = ctypes.c_int32()
        _check(_lib.sgpu_batch_add_tile(self._h, _ptr(sc.grid), sc.width, sc.height, _I64(sc.num_total),
                                        _I64(sc.num_top), _ptr(sc.pixel_parent),
                                        _ptr(sc.wall_z) if sc.wall_z.size else None, _F(sc.max_height),
                                        _ptr(sc.slopes), _ptr(sc.aspects), _ptr(_c(az_alt, np.float32)),
                                        _ptr(_c(dni, np.float32)), _ptr(_c(dhi, np.float32)), ctypes.byref(idx)))
        return idx.value

    def viewshed(self) -> None:
        _check(_lib.sgpu_batch_viewshed(self._h))

    def set_receiver_queue(self, on) -> None:
        """Receiver assignment of the batched reflection pass: False / 0 static blocks (default),
        True / 1 queued receivers, 2 queued receiver pairs."""
        _check(_lib.sgpu_batch_set_receiver_queue(self._h, int(on)))

    def hour_loop(self, fmt: int = FMT_BARLE, hours_per_pass: int = 64, waves: int = 4, graph: bool = True) -> None:
        _check(_lib.sgpu_batch_hour_loop(self._h, ctypes.c_uint(fmt), hours_per_pass, waves, int(graph)))

    def info(self) -> dict:
        i = BatchInfo()
        _check(_lib.sgpu_batch_get_info(self._h, ctypes.byref(i)))
        return i.as_dict()

    def _sizes(self, t: int):
        n, nnz, nb = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        sb = ctypes.c_int32()
        _check(_lib.sgpu_batch_tile_sizes(self._h, t, ctypes.byref(n), ctypes.byref(nnz), ctypes.byref(nb),
                                          ctypes.byref(sb)))
        return n.value, nnz.value, nb.value, sb.value

    def svf(self, t: int) -> np.ndarray:
        n, _, _, sb = self._sizes(t)
        out = np.zeros(max(sb * n, 1), np.uint8)
        _check(_lib.sgpu_batch_get_svf(self._h, t, _ptr(out)))
        return out[: sb * n]

    def csr(self, t: int):
        n, nnz, _, _ = self._sizes(t)
        off = np.zeros(n + 1, np.int64)
        ids = np.zeros(max(nnz, 1), np.int32)
        _check(_lib.sgpu_batch_get_csr(self._h, t, _ptr(off), _ptr(ids)))
        return off, ids[:nnz]

    def barle(self, t: int):
        n, _, nb, _ = self._sizes(t)
        off = np.zeros(n + 1, np.int64)
        data = np.zeros(max(nb, 1), np.uint8)
        _check(_lib.sgpu_batch_get_barle(self._h, t, _ptr(off), _ptr(data)))
        return off, data[:nb]

    def irradiance(self, t: int) -> np.ndarray:
        n = self._sizes(t)[0]
        out = np.zeros(n, np.float32)
        _check(_lib.sgpu_batch_get_irradiance(self._h, t, _ptr(out)))
        return out


def _whole(fn, sc: Scene, az_alt, dni, dhi) -> np.ndarray:
    out = np.zeros(sc.num_total, np.float32)
    _check(fn(_ptr(sc.grid), sc.width, sc.height, sc.num_total, sc.num_top, _ptr(sc.pixel_parent),
              _ptr(sc.wall_z) if sc.wall_z.size else None, _F(sc.max_height), _ptr(_c(az_alt, np.float32)),
              _ptr(_c(dni, np.float32)), _ptr(_c(dhi, np.float32)), _ptr(sc.slopes), _ptr(sc.aspects), _ptr(out)))
    return out


def uncompressed_gpu_irradiance(sc: Scene, az_alt, dni, dhi) -> np.ndarray:
    """solar_gpu_api.h:37-42 -- CSR neighbour lists."""
    return _whole(_lib.sgpu_uncompressed_irradiance, sc, az_alt, dni, dhi)


def compressed_gpu_irradiance(sc: Scene, az_alt, dni, dhi) -> np.ndarray:
    """solar_gpu_api.h:49-54 -- ba-RLE neighbour streams."""
    return _whole(_lib.sgpu_compressed_irradiance, sc, az_alt, dni, dhi)
