"""Per-voxel surrogate for the reflection pass: the MLP binding (sgpu_mlp_* of
include/sgpu.h) and the surrogate pipeline helpers.

The network (53 features padded to 64 -> 4 x 256 GELU -> 2, BF16 weights/activations, FP32
accumulation on MFMA) is trained by the library's own forward, backward and
Adam kernels on labels produced by the plain pipeline.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _check, _lib, _ptr, _c, _I64

FALLBACK_ANNUAL = -1   # run(): exact fallback by one walk over the neighbours' annual totals
IN_PAD = 64
FEATURES = 53
HIDDEN = 256
OUTPUTS = 2

_lib.sgpu_mlp_last_error.restype = ctypes.c_char_p


def _mcheck(code: int) -> None:
    if code != 0:
        from . import SolarGPUError
        raise SolarGPUError(code, _lib.sgpu_mlp_last_error().decode(errors="replace"))


def to_bf16(x: np.ndarray) -> np.ndarray:
    """float32 -> bf16 bit patterns (uint16), round to nearest even."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def from_bf16(h: np.ndarray) -> np.ndarray:
    return (np.asarray(h, np.uint16).astype(np.uint32) << 16).view(np.float32)


def layer_shapes():
    ins = [IN_PAD, HIDDEN, HIDDEN, HIDDEN, HIDDEN]
    outs = [HIDDEN, HIDDEN, HIDDEN, HIDDEN, OUTPUTS]
    return list(zip(outs, ins))


def unpack_params(p: np.ndarray):
    """Flat parameter vector -> [(W[out][in], b[out]), ...]."""
    out, pos = [], 0
    for o, i in layer_shapes():
        W = p[pos:pos + o * i].reshape(o, i)
        pos += o * i
        b = p[pos:pos + o]
        pos += o
        out.append((W, b))
    return out


def pack_params(layers) -> np.ndarray:
    return np.concatenate([np.concatenate([W.ravel(), b.ravel()]) for W, b in layers]).astype(np.float32)


class MLP:
    def __init__(self, seed: int = 1, device: int = 0):
        self._h = ctypes.c_void_p()
        _mcheck(_lib.sgpu_mlp_create(device, ctypes.c_uint64(seed), ctypes.byref(self._h)))
        self.nparams = _lib.sgpu_mlp_param_count()

    def close(self):
        if self._h:
            _mcheck(_lib.sgpu_mlp_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self) -> np.ndarray:
        p = np.zeros(self.nparams, np.float32)
        _mcheck(_lib.sgpu_mlp_get_params(self._h, _ptr(p)))
        return p

    def set_params(self, p: np.ndarray) -> None:
        p = _c(p, np.float32)
        assert p.size == self.nparams
        _mcheck(_lib.sgpu_mlp_set_params(self._h, _ptr(p)))

    def set_loss_weights(self, w0: float, w1: float) -> None:
        _mcheck(_lib.sgpu_mlp_set_loss_weights(self._h, ctypes.c_float(w0), ctypes.c_float(w1)))

    def set_inference(self, rows: int) -> None:
        """Inference kernel: 32 or 64 voxels per workgroup of the fused kernel, 0 = layer by layer."""
        _mcheck(_lib.sgpu_mlp_set_inference(self._h, rows))

    def forward(self, X_bf16: np.ndarray) -> np.ndarray:
        X = _c(X_bf16, np.uint16)
        n = X.shape[0]
        Y = np.zeros((n, OUTPUTS), np.float32)
        _mcheck(_lib.sgpu_mlp_forward(self._h, _ptr(X), _I64(n), _ptr(Y)))
        return Y

    def train(self, X_bf16, T, epochs: int, batch: int, lr: float, seed: int = 1) -> np.ndarray:
        X = _c(X_bf16, np.uint16)
        T = _c(T, np.float32)
        losses = np.zeros(epochs, np.float64)
        _mcheck(_lib.sgpu_mlp_train(self._h, _ptr(X), _ptr(T), _I64(X.shape[0]), epochs, batch,
                                    ctypes.c_float(lr), ctypes.c_uint64(seed), _ptr(losses)))
        return losses

    def gradient(self, X_bf16, T) -> tuple[np.ndarray, float]:
        X = _c(X_bf16, np.uint16)
        T = _c(T, np.float32)
        g = np.zeros(self.nparams, np.float32)
        loss = ctypes.c_double()
        _mcheck(_lib.sgpu_mlp_gradient(self._h, _ptr(X), _ptr(T), X.shape[0], _ptr(g), ctypes.byref(loss)))
        return g, loss.value


# ---------------- surrogate pipeline ----------------

class Gate(ctypes.Structure):
    """Fallback rule: exact path where the predicted relative error of the annual total exceeds
    tau.  With an error network the prediction is 10^(its first output); without one it is
    refl_rel_err * predicted reflected fraction + diff_miss_weight * |diffuse-share miss|
    + share_abs_err * direct_diffuse / total."""
    _fields_ = [("tau", ctypes.c_float), ("refl_rel_err", ctypes.c_float),
                ("diff_miss_weight", ctypes.c_float), ("share_abs_err", ctypes.c_float)]


class Stats(ctypes.Structure):
    _fields_ = [("fallback", ctypes.c_int64), ("t_svf_ms", ctypes.c_double), ("t_dd_ms", ctypes.c_double),
                ("t_features_ms", ctypes.c_double), ("t_mlp_ms", ctypes.c_double),
                ("t_fallback_ms", ctypes.c_double), ("t_total_ms", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def prepare(ctx, lat: float, cell_m: float = 1.0) -> dict:
    st = Stats()
    _check(_lib.sgpu_surrogate_prepare(ctx._h, ctypes.c_float(lat), ctypes.c_float(cell_m), ctypes.byref(st)))
    return st.as_dict()


def features(ctx) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    n = ctx.info()["num_total"]
    X = np.zeros((n, IN_PAD), np.uint16)
    dd = np.zeros(n, np.float32)
    diff = np.zeros(n, np.float32)
    _check(_lib.sgpu_surrogate_get(ctx._h, _ptr(X), _ptr(dd), _ptr(diff)))
    return X, dd, diff


def labels(ctx, fmt: int = 1) -> np.ndarray:
    """Annual reflected irradiance of the plain method (needs ctx.viewshed())."""
    out = np.zeros(ctx.info()["num_total"], np.float32)
    _check(_lib.sgpu_surrogate_labels(ctx._h, ctypes.c_uint(fmt), _ptr(out)))
    return out


def targets(refl, dd, diff) -> np.ndarray:
    """Network targets: reflected share and diffuse share of the exact direct + diffuse."""
    scale = np.maximum(dd, 1e-3)
    return np.stack([refl / scale, diff / scale], 1).astype(np.float32)


def run(ctx, mlp: MLP, gate: Gate, hours_per_pass: int = 0, error_mlp: MLP | None = None):
    """Surrogate annual totals.  hours_per_pass > 0: per-hour exact fallback with K hours per
    pass (0 = the context's tuned K); FALLBACK_ANNUAL: one-walk annual fallback.  error_mlp:
    the error network (first output = log10 of the predicted relative error of the surrogate
    total); voxels whose predicted error exceeds gate.tau go to the exact path."""
    n = ctx.info()["num_total"]
    out = np.zeros(n, np.float32)
    fb = np.zeros(n, np.int32)
    st = Stats()
    _check(_lib.sgpu_surrogate_run(ctx._h, mlp._h, error_mlp._h if error_mlp is not None else None,
                                   ctypes.byref(gate), hours_per_pass, _ptr(out), _ptr(fb), ctypes.byref(st)))
    return out, fb, st.as_dict()


# error network target: log10 of the relative error of the annual total, floored here
ERR_FLOOR = 1e-4


def error_target(Y: np.ndarray, dd: np.ndarray, refl: np.ndarray) -> np.ndarray:
    """Per-voxel relative error of the surrogate total dd (1 + max(share, 0)) against the plain
    total dd + refl (floor 1 Wh/m^2), as the error network's targets [log10(err), 0]."""
    total = dd.astype(np.float64) * (1.0 + np.maximum(Y[:, 0], 0.0))
    exact = dd.astype(np.float64) + refl
    err = np.abs(total - exact) / np.maximum(np.abs(exact), 1.0)
    return np.stack([np.log10(np.maximum(err, ERR_FLOOR)), np.zeros_like(err)], 1).astype(np.float32)


def save(path: str, params: np.ndarray, gate: Gate, meta: dict, error_params: np.ndarray | None = None) -> None:
    extra = {} if error_params is None else {"error_params": error_params.astype(np.float32)}
    np.savez(path, params=params.astype(np.float32),
             gate=np.array([gate.tau, gate.refl_rel_err, gate.diff_miss_weight, gate.share_abs_err], np.float32),
             meta=np.frombuffer(repr(meta).encode(), np.uint8), **extra)


def load(path: str) -> tuple[np.ndarray, Gate, str]:
    z = np.load(path, allow_pickle=False)
    g = z["gate"]
    return z["params"], Gate(*[float(x) for x in g]), z["meta"].tobytes().decode()


def load_error_params(path: str) -> np.ndarray | None:
    """The error network's parameters of a weights file (None when it has none)."""
    z = np.load(path, allow_pickle=False)
    return z["error_params"] if "error_params" in z.files else None


def load_models(path: str, device: int = 0):
    """(network, error network or None, gate, meta) of a weights file, on the device."""
    params, gate, meta = load(path)
    net = MLP(seed=1, device=device)
    net.set_params(params)
    ep = load_error_params(path)
    err = None
    if ep is not None:
        err = MLP(seed=2, device=device)
        err.set_params(ep)
    return net, err, gate, meta
