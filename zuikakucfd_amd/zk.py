"""ctypes binding of libzk.so (the C ABI declared in include/zk.h).

This is the Python side of the drop-in boundary: `Solver` owns one zk handle
(one whole-domain state on one GPU).  Arrays cross the boundary as host numpy
buffers that the library copies; nothing here computes the update itself.
If libzk.so is missing or no GPU is visible the constructor raises -- there is
no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libzk.so"

METHOD_WENO, METHOD_DG = 0, 1
PREC_FP64, PREC_FP32_BULK = 0, 1
FLUX_HLLC, FLUX_HLL = 0, 1
LIMITER_NONE, LIMITER_PERSSON_ZS, LIMITER_TVB, LIMITER_LEARNED, LIMITER_TVB_LINEAR = 0, 1, 2, 3, 4
RECON_WENO5, RECON_MUSCL_MINMOD, RECON_MUSCL_VANLEER, RECON_MUSCL_SUPERBEE = 0, 1, 2, 3

ZK_OK = 0
_ERR_NAMES = {1: "ZK_ERR_ARG", 2: "ZK_ERR_HIP", 3: "ZK_ERR_NUMERIC", 4: "ZK_ERR_CFL"}

EXPORTED = [
    "zk_create", "zk_destroy", "zk_last_error", "zk_load", "zk_set_state", "zk_get_state",
    "zk_get_conserved", "zk_max_speed", "zk_step", "zk_step_cfl", "zk_sync", "zk_step_timed",
    "zk_get_flags", "zk_updated_cells", "zk_profile", "zk_profile_report", "zk_set_tuning",
    "zk_set_reconstruction",
    "zk_set_slip_walls",
    "zk_ensemble_create", "zk_ensemble_add", "zk_ensemble_finalize", "zk_ensemble_set_tuning",
    "zk_ensemble_step", "zk_ensemble_sync", "zk_ensemble_size", "zk_ensemble_get",
    "zk_ensemble_last_error", "zk_ensemble_destroy",
    "zk_mlp_create", "zk_mlp_destroy", "zk_mlp_last_error", "zk_mlp_param_count", "zk_mlp_get_params",
    "zk_mlp_set_params", "zk_mlp_grad", "zk_mlp_infer", "zk_mlp_dataset", "zk_mlp_clear_dataset", "zk_mlp_train",
    "zk_harvest", "zk_attach_mlp", "zk_learned_stats", "zk_mlp_features",
    "zk_weno_step_noalloc", "zk_stream_bandwidth",
]
_NO_STATUS = ("zk_last_error", "zk_ensemble_last_error", "zk_mlp_last_error", "zk_mlp_param_count")


class ZkConfig(C.Structure):
    _fields_ = [
        ("method", C.c_int),
        ("P", C.c_int),
        ("precision", C.c_int),
        ("flux", C.c_int),
        ("limiter", C.c_int),
        ("tvb_M", C.c_double),
        ("device", C.c_int),
        ("positivity", C.c_int),
    ]


class ZkError(RuntimeError):
    """A failed zk_* call; `status` holds the ZK_ERR_* code (0 when none)."""

    def __init__(self, msg: str, status: int = 0):
        super().__init__(msg)
        self.status = status


class ZkNumericError(ZkError):
    """ZK_ERR_NUMERIC: the state is no longer finite (the message names the
    variable, the cell and the step count of the check that found it)."""


class ZkCflError(ZkError):
    """ZK_ERR_CFL: a DG CFL number above the stable limit 0.9 / (2 (2P + 1))."""


def _raise(st: int, msg: str):
    cls = ZkNumericError if st == 3 else ZkCflError if st == 4 else ZkError
    raise cls(msg, st)


_lib = None


def load_library(path: str | os.PathLike | None = None) -> C.CDLL:
    """Load libzk.so (raises if it has not been built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise ZkError(f"{p} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = C.CDLL(str(p))
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int32)
    vp = C.c_void_p
    lib.zk_create.argtypes = [C.POINTER(ZkConfig), C.POINTER(vp)]
    lib.zk_destroy.argtypes = [vp]
    lib.zk_last_error.argtypes = [vp]
    lib.zk_last_error.restype = C.c_char_p
    lib.zk_load.argtypes = [vp, C.c_int, C.c_int, C.c_double, C.c_double, dp, dp, dp, dp, ip]
    lib.zk_set_state.argtypes = [vp, dp]
    lib.zk_get_state.argtypes = [vp, dp]
    lib.zk_get_conserved.argtypes = [vp, dp, dp, dp, dp]
    lib.zk_max_speed.argtypes = [vp, dp]
    lib.zk_step.argtypes = [vp, C.c_double, C.c_int]
    lib.zk_step_cfl.argtypes = [vp, C.c_double, C.c_double, C.c_int, dp, C.POINTER(C.c_int)]
    lib.zk_sync.argtypes = [vp]
    lib.zk_step_timed.argtypes = [vp, C.c_double, C.c_int, C.POINTER(C.c_float)]
    lib.zk_get_flags.argtypes = [vp, C.POINTER(C.c_uint8)]
    lib.zk_updated_cells.argtypes = [vp, C.POINTER(C.c_int64)]
    lib.zk_stream_bandwidth.argtypes = [C.c_int, C.c_size_t, dp, dp]
    lib.zk_profile.argtypes = [vp, C.c_int]
    lib.zk_profile_report.argtypes = [vp, C.c_char_p, C.c_size_t]
    lib.zk_set_tuning.argtypes = [vp, C.c_int, C.c_int, C.c_int]
    lib.zk_set_reconstruction.argtypes = [vp, C.c_int]
    lib.zk_set_slip_walls.argtypes = [vp, C.c_int]
    lib.zk_ensemble_create.argtypes = [C.POINTER(ZkConfig), C.POINTER(vp)]
    lib.zk_ensemble_add.argtypes = [vp, C.c_int, C.c_int, C.c_double, C.c_double, dp, dp, dp, dp, ip, C.c_double]
    lib.zk_ensemble_finalize.argtypes = [vp]
    lib.zk_ensemble_set_tuning.argtypes = [vp, C.c_int, C.c_int]
    lib.zk_ensemble_step.argtypes = [vp, C.c_int]
    lib.zk_ensemble_sync.argtypes = [vp]
    lib.zk_ensemble_size.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int64)]
    lib.zk_ensemble_get.argtypes = [vp, C.c_int, dp, dp]
    lib.zk_ensemble_last_error.argtypes = [vp]
    lib.zk_ensemble_last_error.restype = C.c_char_p
    lib.zk_ensemble_destroy.argtypes = [vp]
    u16p = C.POINTER(C.c_uint16)
    fp = C.POINTER(C.c_float)
    lib.zk_mlp_create.argtypes = [C.c_int, C.c_uint64, C.c_int, C.POINTER(vp)]
    lib.zk_mlp_destroy.argtypes = [vp]
    lib.zk_mlp_last_error.argtypes = [vp]
    lib.zk_mlp_last_error.restype = C.c_char_p
    lib.zk_mlp_param_count.argtypes = []
    lib.zk_mlp_param_count.restype = C.c_int
    lib.zk_mlp_get_params.argtypes = [vp, fp]
    lib.zk_mlp_set_params.argtypes = [vp, fp]
    lib.zk_mlp_grad.argtypes = [vp, u16p, fp, C.c_int, fp, fp]
    lib.zk_mlp_infer.argtypes = [vp, u16p, C.c_int, fp]
    lib.zk_mlp_dataset.argtypes = [vp, C.POINTER(C.c_int), u16p, fp, C.c_int]
    lib.zk_mlp_clear_dataset.argtypes = [vp]
    lib.zk_mlp_train.argtypes = [vp, C.c_int, C.c_int, C.c_float, C.c_uint64, fp]
    lib.zk_harvest.argtypes = [vp, vp, C.c_float, C.c_float]
    lib.zk_attach_mlp.argtypes = [vp, vp, C.c_float, C.c_float, C.c_float]
    lib.zk_learned_stats.argtypes = [vp, C.POINTER(C.c_int64), C.c_int]
    lib.zk_mlp_features.argtypes = [vp, vp, C.POINTER(C.c_int32), C.c_int, u16p, fp]
    lib.zk_weno_step_noalloc.argtypes = [C.POINTER(ZkConfig), C.c_int, C.c_int, C.c_double, C.c_double, vp, vp,
                                         C.c_double, C.c_int, vp]
    for name in EXPORTED:
        if name not in _NO_STATUS:
            getattr(lib, name).restype = C.c_int
    if path is None:
        _lib = lib
    return lib


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Solver:
    """One whole-domain explicit Euler state on one GPU.

    Mirrors the reference's per-rank `Mesh` + `updateMesh` pair
    (src_WENO/fluid.h:26-80,218; src_DG/fluid.h:44-110,160) behind the C ABI.
    """

    def __init__(self, method: int = METHOD_WENO, P: int = 1, precision: int = PREC_FP64,
                 flux: int = FLUX_HLLC, limiter: int = LIMITER_PERSSON_ZS, tvb_M: float = 50.0,
                 device: int = 0, positivity: bool = False):
        self.lib = load_library()
        self.cfg = ZkConfig(method, P, precision, flux, limiter, float(tvb_M), device, 1 if positivity else 0)
        h = C.c_void_p()
        st = self.lib.zk_create(C.byref(self.cfg), C.byref(h))
        if st != ZK_OK:
            raise ZkError(f"zk_create: {_ERR_NAMES.get(st, st)}: {self.lib.zk_last_error(None).decode()}")
        self.h = h
        self.method = method
        self.P = P
        self.nm = (P + 1) ** 2 if method == METHOD_DG else 1
        self.ny = self.nx = 0

    # -- lifetime -------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.zk_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, st: int, what: str) -> None:
        if st != ZK_OK:
            msg = self.lib.zk_last_error(self.h).decode()
            _raise(st, f"{what}: {_ERR_NAMES.get(st, st)}: {msg}")

    # -- data -----------------------------------------------------------
    def load(self, da: float, gamma: float, rho, u, v, p, bctype) -> None:
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        ny, nx = rho.shape
        arrs = [np.ascontiguousarray(a, dtype=np.float64).reshape(ny, nx) for a in (u, v, p)]
        bc = np.ascontiguousarray(bctype, dtype=np.int32).reshape(ny, nx)
        self._check(self.lib.zk_load(self.h, ny, nx, float(da), float(gamma), _dptr(rho), *map(_dptr, arrs),
                                     bc.ctypes.data_as(C.POINTER(C.c_int32))), "zk_load")
        self.ny, self.nx = ny, nx
        self.da, self.gamma = float(da), float(gamma)

    def set_state(self, Q: np.ndarray) -> None:
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        if Q.size != 4 * self.nm * self.ny * self.nx:
            raise ValueError("state has the wrong size")
        self._check(self.lib.zk_set_state(self.h, _dptr(Q)), "zk_set_state")

    def get_state(self) -> np.ndarray:
        Q = np.empty((4, self.nm, self.ny, self.nx), dtype=np.float64)
        self._check(self.lib.zk_get_state(self.h, _dptr(Q)), "zk_get_state")
        return Q

    def get_conserved(self) -> np.ndarray:
        U = np.empty((4, self.ny, self.nx), dtype=np.float64)
        self._check(self.lib.zk_get_conserved(self.h, *(_dptr(U[c]) for c in range(4))), "zk_get_conserved")
        return U

    def get_flags(self) -> np.ndarray:
        f = np.zeros((self.ny, self.nx), dtype=np.uint8)
        self._check(self.lib.zk_get_flags(self.h, f.ctypes.data_as(C.POINTER(C.c_uint8))), "zk_get_flags")
        return f

    def updated_cells(self) -> int:
        n = C.c_int64()
        self._check(self.lib.zk_updated_cells(self.h, C.byref(n)), "zk_updated_cells")
        return n.value

    # -- compute --------------------------------------------------------
    def max_speed(self) -> float:
        out = C.c_double()
        self._check(self.lib.zk_max_speed(self.h, C.byref(out)), "zk_max_speed")
        return out.value

    def step(self, dt: float, nsteps: int = 1) -> None:
        self._check(self.lib.zk_step(self.h, float(dt), int(nsteps)), "zk_step")

    def step_cfl(self, cfl: float, t_end: float, nsteps: int) -> tuple[float, int]:
        t = C.c_double()
        n = C.c_int()
        self._check(self.lib.zk_step_cfl(self.h, float(cfl), float(t_end), int(nsteps), C.byref(t), C.byref(n)),
                    "zk_step_cfl")
        return t.value, n.value

    def step_timed(self, dt: float, nsteps: int) -> float:
        ms = C.c_float()
        self._check(self.lib.zk_step_timed(self.h, float(dt), int(nsteps), C.byref(ms)), "zk_step_timed")
        return ms.value

    def set_tuning(self, fv_tile: int = 0, dg_variant: int = 0, use_graph: bool = True) -> None:
        self._check(self.lib.zk_set_tuning(self.h, int(fv_tile), int(dg_variant), 1 if use_graph else 0),
                    "zk_set_tuning")

    def set_reconstruction(self, kind: int) -> None:
        """WENO face reconstruction: RECON_WENO5 (default) or RECON_MUSCL_*."""
        self._check(self.lib.zk_set_reconstruction(self.h, int(kind)), "zk_set_reconstruction")

    def set_slip_walls(self, enable: bool) -> None:
        """WENO: bctype-1 cells become mirror slip walls (off by default, as in
        the reference)."""
        self._check(self.lib.zk_set_slip_walls(self.h, 1 if enable else 0), "zk_set_slip_walls")

    def attach_mlp(self, mlp: "Mlp", lo: float = 0.2, hi: float = 0.8, audit: float = 0.02) -> None:
        self._mlp = mlp  # keep the network alive while the handle uses it
        self._check(self.lib.zk_attach_mlp(self.h, mlp.h, float(lo), float(hi), float(audit)), "zk_attach_mlp")

    def harvest(self, mlp: "Mlp | None", keep_pos: float = 1.0, keep_neg: float = 0.05) -> None:
        self._harvest = mlp
        self._check(self.lib.zk_harvest(self.h, mlp.h if mlp is not None else None, float(keep_pos), float(keep_neg)),
                    "zk_harvest")

    def learned_stats(self, reset: bool = False) -> dict:
        v = (C.c_int64 * 8)()
        self._check(self.lib.zk_learned_stats(self.h, v, 1 if reset else 0), "zk_learned_stats")
        keys = ("audited", "agree", "miss", "false_flag", "net_flagged", "in_band", "candidates", "audited_positive")
        return dict(zip(keys, list(v)))

    def mlp_features(self, cells, mlp: "Mlp | None" = None) -> tuple[np.ndarray, np.ndarray | None]:
        cells = np.ascontiguousarray(cells, dtype=np.int32)
        n = cells.size
        X = np.zeros((n, 192), dtype=np.uint16)
        sc = np.zeros(n, dtype=np.float32)
        self._check(self.lib.zk_mlp_features(self.h, mlp.h if mlp is not None else None,
                                             cells.ctypes.data_as(C.POINTER(C.c_int32)), n,
                                             X.ctypes.data_as(C.POINTER(C.c_uint16)),
                                             sc.ctypes.data_as(C.POINTER(C.c_float)) if mlp is not None else None),
                    "zk_mlp_features")
        return X, (sc if mlp is not None else None)

    def profile(self, enable: bool = True) -> None:
        self._check(self.lib.zk_profile(self.h, 1 if enable else 0), "zk_profile")

    def profile_report(self) -> dict:
        import json

        buf = C.create_string_buffer(1 << 16)
        self._check(self.lib.zk_profile_report(self.h, buf, len(buf)), "zk_profile_report")
        return {k: {"launches": v[0], "total_ms": v[1]} for k, v in json.loads(buf.value.decode()).items()}

    def sync(self) -> None:
        self._check(self.lib.zk_sync(self.h), "zk_sync")


class Ensemble:
    """Independent WENO5 problems of uneven sizes advanced together on one GPU,
    each with its own on-device CFL time step (zk_ensemble_* in include/zk.h)."""

    def __init__(self, precision: int = PREC_FP64, flux: int = FLUX_HLLC, positivity: bool = False,
                 device: int = 0):
        self.lib = load_library()
        self.cfg = ZkConfig(METHOD_WENO, 0, precision, flux, 0, 0.0, device, 1 if positivity else 0)
        h = C.c_void_p()
        st = self.lib.zk_ensemble_create(C.byref(self.cfg), C.byref(h))
        if st != ZK_OK:
            raise ZkError(f"zk_ensemble_create failed ({_ERR_NAMES.get(st, st)})")
        self.h = h
        self.shapes = []

    def _check(self, st: int, what: str) -> None:
        if st != ZK_OK:
            _raise(st, f"{what}: {_ERR_NAMES.get(st, st)}: {self.lib.zk_ensemble_last_error(self.h).decode()}")

    def add(self, da, gamma, rho, u, v, p, bctype, cfl: float) -> int:
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        ny, nx = rho.shape
        arrs = [np.ascontiguousarray(a, dtype=np.float64).reshape(ny, nx) for a in (u, v, p)]
        bc = np.ascontiguousarray(bctype, dtype=np.int32).reshape(ny, nx)
        self._check(self.lib.zk_ensemble_add(self.h, ny, nx, float(da), float(gamma), _dptr(rho), *map(_dptr, arrs),
                                             bc.ctypes.data_as(C.POINTER(C.c_int32)), float(cfl)), "zk_ensemble_add")
        self.shapes.append((ny, nx))
        return len(self.shapes) - 1

    def finalize(self) -> None:
        self._check(self.lib.zk_ensemble_finalize(self.h), "zk_ensemble_finalize")

    def set_tuning(self, tile: int = 1, use_graph: bool = True) -> None:
        self._check(self.lib.zk_ensemble_set_tuning(self.h, int(tile), 1 if use_graph else 0), "zk_ensemble_set_tuning")

    def step(self, nsteps: int) -> None:
        self._check(self.lib.zk_ensemble_step(self.h, int(nsteps)), "zk_ensemble_step")

    def sync(self) -> None:
        self._check(self.lib.zk_ensemble_sync(self.h), "zk_ensemble_sync")

    def size(self) -> tuple[int, int]:
        n, c = C.c_int(), C.c_int64()
        self._check(self.lib.zk_ensemble_size(self.h, C.byref(n), C.byref(c)), "zk_ensemble_size")
        return n.value, c.value

    def get(self, idx: int) -> tuple[np.ndarray, float]:
        ny, nx = self.shapes[idx]
        U = np.empty((4, ny, nx), dtype=np.float64)
        t = C.c_double()
        self._check(self.lib.zk_ensemble_get(self.h, int(idx), _dptr(U), C.byref(t)), "zk_ensemble_get")
        return U, t.value

    def close(self) -> None:
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.zk_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def weno_step_noalloc(Q_ptr: int, bctype_ptr: int, ny: int, nx: int, da: float, gamma: float, dt: float,
                      nsteps: int, stream: int = 0, precision: int = PREC_FP64, flux: int = FLUX_HLLC,
                      device: int = 0) -> None:
    """zk_weno_step_noalloc on device buffers given by address: Q_ptr (4
    row-major ny*nx float64 planes, advanced in place) and bctype_ptr (ny*nx
    int32), enqueued on `stream` (a hipStream_t address, 0 for the default
    stream).  The library allocates nothing."""
    lib = load_library()
    cfg = ZkConfig(METHOD_WENO, 0, precision, flux, 0, 0.0, device, 0)
    st = lib.zk_weno_step_noalloc(C.byref(cfg), int(ny), int(nx), float(da), float(gamma), C.c_void_p(Q_ptr),
                                  C.c_void_p(bctype_ptr), float(dt), int(nsteps), C.c_void_p(stream or None))
    if st != ZK_OK:
        _raise(st, f"zk_weno_step_noalloc: {_ERR_NAMES.get(st, st)}: {lib.zk_last_error(None).decode()}")


def stream_bandwidth(array_bytes: int, device: int = 0) -> tuple[float, float]:
    """zk_stream_bandwidth: GB/s of a copy (1 read + 1 write) and of an add
    (2 reads + 1 write) over arrays of array_bytes on `device`."""
    lib = load_library()
    cp, ad = C.c_double(), C.c_double()
    st = lib.zk_stream_bandwidth(int(device), int(array_bytes), C.byref(cp), C.byref(ad))
    if st != ZK_OK:
        _raise(st, f"zk_stream_bandwidth: {_ERR_NAMES.get(st, st)}: {lib.zk_last_error(None).decode()}")
    return cp.value, ad.value


def _f32p(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _u16p(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint16))


class Mlp:
    """The troubled-cell network (zk_mlp_* in include/zk.h): FP32 master
    parameters, BF16 MFMA forward/backward, Adam, and a device-side dataset
    filled by Solver.harvest."""

    def __init__(self, seed: int = 0, capacity: int = 0, device: int = 0):
        self.lib = load_library()
        h = C.c_void_p()
        st = self.lib.zk_mlp_create(int(device), int(seed), int(capacity), C.byref(h))
        if st != ZK_OK:
            raise ZkError(f"zk_mlp_create failed ({_ERR_NAMES.get(st, st)})")
        self.h = h
        self.nparams = self.lib.zk_mlp_param_count()

    def _check(self, st: int, what: str) -> None:
        if st != ZK_OK:
            raise ZkError(f"{what}: {_ERR_NAMES.get(st, st)}: {self.lib.zk_mlp_last_error(self.h).decode()}")

    def get_params(self) -> np.ndarray:
        w = np.zeros(self.nparams, dtype=np.float32)
        self._check(self.lib.zk_mlp_get_params(self.h, _f32p(w)), "zk_mlp_get_params")
        return w

    def set_params(self, w) -> None:
        w = np.ascontiguousarray(w, dtype=np.float32)
        if w.size != self.nparams:
            raise ValueError("wrong parameter count")
        self._check(self.lib.zk_mlp_set_params(self.h, _f32p(w)), "zk_mlp_set_params")

    def grad(self, Xbits: np.ndarray, y: np.ndarray) -> tuple[np.ndarray, float]:
        Xbits = np.ascontiguousarray(Xbits, dtype=np.uint16)
        y = np.ascontiguousarray(y, dtype=np.float32)
        g = np.zeros(self.nparams, dtype=np.float32)
        loss = C.c_float()
        self._check(self.lib.zk_mlp_grad(self.h, _u16p(Xbits), _f32p(y), y.size, _f32p(g), C.byref(loss)),
                    "zk_mlp_grad")
        return g, loss.value

    def infer(self, Xbits: np.ndarray) -> np.ndarray:
        Xbits = np.ascontiguousarray(Xbits, dtype=np.uint16)
        out = np.zeros(Xbits.shape[0], dtype=np.float32)
        self._check(self.lib.zk_mlp_infer(self.h, _u16p(Xbits), Xbits.shape[0], _f32p(out)), "zk_mlp_infer")
        return out

    def dataset(self, max_samples: int = 0) -> tuple[int, np.ndarray, np.ndarray]:
        n = C.c_int()
        X = np.zeros((max(max_samples, 0), 192), dtype=np.uint16)
        y = np.zeros(max(max_samples, 0), dtype=np.float32)
        self._check(self.lib.zk_mlp_dataset(self.h, C.byref(n), _u16p(X) if max_samples else None,
                                            _f32p(y) if max_samples else None, int(max_samples)), "zk_mlp_dataset")
        k = min(n.value, max_samples)
        return n.value, X[:k], y[:k]

    def clear_dataset(self) -> None:
        self._check(self.lib.zk_mlp_clear_dataset(self.h), "zk_mlp_clear_dataset")

    def train(self, epochs: int, batch: int = 256, lr: float = 1e-3, seed: int = 0) -> np.ndarray:
        losses = np.zeros(max(epochs, 1), dtype=np.float32)
        self._check(self.lib.zk_mlp_train(self.h, int(epochs), int(batch), float(lr), int(seed), _f32p(losses)),
                    "zk_mlp_train")
        return losses[:epochs]

    def close(self) -> None:
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.zk_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
