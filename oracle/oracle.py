"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the zuikakucfd_amd package.
Arrays are numpy float64 with the oracle's layout (4*nm row-major planes).
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
_lib = None


def use_library(path) -> None:
    """Load the oracle from another build of zk_oracle.c (for instance one
    compiled with -march=native for CPU timing); before the first call only."""
    global LIB
    if _lib is not None:
        raise RuntimeError("the oracle library is already loaded")
    LIB = Path(path)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            subprocess.run(["make", "-s", "-C", str(HERE), "liboracle.so"], check=True)
        L = C.CDLL(str(LIB))
        dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        up = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
        i, d = C.c_int, C.c_double
        L.zko_weno5.argtypes = [dp, C.POINTER(d), C.POINTER(d)]
        L.zko_muscl.argtypes = [dp, i, C.POINTER(d), C.POINTER(d)]
        for n in ("zko_hllc_fv", "zko_hllc_dg", "zko_hll"):
            getattr(L, n).argtypes = [dp, dp, d, dp]
        L.zko_hllc_speeds.argtypes = [i, dp, dp, d, dp]
        L.zko_fv_fill_ghosts.argtypes = [i, i, ip, dp]
        L.zko_fv_fill_ghosts_slip.argtypes = [i, i, ip, dp]
        L.zko_fv_rhs.argtypes = [i, i, d, d, ip, dp, d, i, dp]
        L.zko_fv_step.argtypes = [i, i, d, d, ip, dp, d, i]
        L.zko_fv_step.restype = i
        L.zko_fv_max_speed.argtypes = [i, i, d, dp]
        L.zko_fv_max_speed.restype = d
        L.zko_prim_to_cons.argtypes = [i, d, dp, dp, dp, dp, dp]
        L.zko_legP.argtypes = [i, d]; L.zko_legP.restype = d
        L.zko_legD.argtypes = [i, d]; L.zko_legD.restype = d
        L.zko_gl_points.argtypes = [i]; L.zko_gl_points.restype = C.POINTER(d)
        L.zko_gl_weights.argtypes = [i]; L.zko_gl_weights.restype = C.POINTER(d)
        L.zko_dg_fill_ghosts.argtypes = [i, i, i, ip, dp]
        L.zko_dg_rhs.argtypes = [i, i, i, d, d, ip, dp, i, dp]
        L.zko_dg_limit_persson_zs.argtypes = [i, i, i, d, ip, dp]
        L.zko_dg_limit_tvb.argtypes = [i, i, i, d, d, ip, dp, up]
        L.zko_dg_limit_tvb_linear.argtypes = [i, i, i, d, d, ip, dp, up]
        L.zko_dg_limit_zs.argtypes = [i, i, i, d, dp]
        L.zko_dg_limit_zs_ext.argtypes = [i, i, i, d, dp]
        L.zko_dg_step.argtypes = [i, i, i, d, d, ip, dp, d, i, i, d, up]
        L.zko_dg_max_speed.argtypes = [i, i, i, d, ip, dp]
        L.zko_dg_max_speed.restype = d
        L.zko_dg_cfl_limit.argtypes = [i]
        L.zko_dg_cfl_limit.restype = d
        _lib = L
    return _lib


def weno5(s) -> tuple[float, float]:
    s = np.ascontiguousarray(s, dtype=np.float64)
    a, b = C.c_double(), C.c_double()
    lib().zko_weno5(s, C.byref(a), C.byref(b))
    return a.value, b.value


def flux(kind: str, L, R, gamma: float) -> np.ndarray:
    F = np.zeros(4)
    fn = {"hllc_fv": "zko_hllc_fv", "hllc_dg": "zko_hllc_dg", "hll": "zko_hll"}[kind]
    getattr(lib(), fn)(np.ascontiguousarray(L, dtype=np.float64), np.ascontiguousarray(R, dtype=np.float64),
                       float(gamma), F)
    return F


def hllc_speeds(kind: int, L, R, gamma: float) -> np.ndarray:
    out = np.zeros(3)
    lib().zko_hllc_speeds(int(kind), np.ascontiguousarray(L, dtype=np.float64),
                          np.ascontiguousarray(R, dtype=np.float64), float(gamma), out)
    return out


def prim_to_cons(gamma, rho, u, v, p) -> np.ndarray:
    ny, nx = rho.shape
    U = np.zeros((4, ny, nx))
    c = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    lib().zko_prim_to_cons(ny * nx, float(gamma), c(rho), c(u), c(v), c(p), U)
    return U


POSITIVITY = 16
# DG limiter bit: the positivity scaling also at the face points and corners (zk_oracle.h)
DG_POSITIVITY = 8
# flux bits selecting MUSCL instead of WENO5 (zk_oracle.h ZKO_RECON_SHIFT)
MUSCL_MINMOD, MUSCL_VANLEER, MUSCL_SUPERBEE = 1 << 5, 2 << 5, 3 << 5
# flux bit making bctype-1 cells mirror slip walls (zk_oracle.h ZKO_FV_SLIP)
FV_SLIP = 1 << 7


def muscl(s, limiter: int):
    """(UL, UR) of zko_muscl on the four cells s = (UL2, UL1, UR1, UR2)."""
    a = np.ascontiguousarray(s, dtype=np.float64)
    ul, ur = C.c_double(), C.c_double()
    lib().zko_muscl(a, int(limiter), C.byref(ul), C.byref(ur))
    return ul.value, ur.value


class FV:
    """Finite-volume oracle state: U has shape (4, ny, nx).  flux may be or-ed
    with POSITIVITY to apply the face positivity check and the cell
    positivity fallback, with one of the MUSCL_* bits to reconstruct with
    MUSCL instead of WENO5, and with FV_SLIP to make bctype-1 cells mirror
    slip walls; `fallbacks` counts the bad cells the fallback met."""

    def __init__(self, da, gamma, U, bctype, flux: int = 0):
        self.U = np.ascontiguousarray(U, dtype=np.float64).copy()
        self.bct = np.ascontiguousarray(bctype, dtype=np.int32)
        self.ny, self.nx = self.bct.shape
        self.da, self.gamma, self.flux = float(da), float(gamma), flux
        self.fallbacks = 0

    def step(self, dt: float, n: int = 1) -> None:
        for _ in range(n):
            self.fallbacks += lib().zko_fv_step(self.ny, self.nx, self.da, self.gamma, self.bct, self.U, float(dt),
                                                self.flux)

    def rhs(self, dt: float) -> np.ndarray:
        dU = np.zeros_like(self.U)
        lib().zko_fv_rhs(self.ny, self.nx, self.da, self.gamma, self.bct, self.U, float(dt), self.flux, dU)
        return dU

    def fill(self) -> None:
        f = lib().zko_fv_fill_ghosts_slip if self.flux & FV_SLIP else lib().zko_fv_fill_ghosts
        f(self.ny, self.nx, self.bct, self.U)

    def max_speed(self) -> float:
        return lib().zko_fv_max_speed(self.ny, self.nx, self.gamma, self.U)


class DG:
    """Modal DG oracle state: Q has shape (4, nm, ny, nx)."""

    def __init__(self, P, da, gamma, Q, bctype, flux: int = 0, limiter: int = 1, tvb_M: float = 50.0):
        self.P = P
        self.nm = (P + 1) ** 2
        self.Q = np.ascontiguousarray(Q, dtype=np.float64).copy()
        self.bct = np.ascontiguousarray(bctype, dtype=np.int32)
        self.ny, self.nx = self.bct.shape
        self.da, self.gamma, self.flux, self.limiter, self.M = float(da), float(gamma), flux, limiter, float(tvb_M)
        self.flags = np.zeros((self.ny, self.nx), dtype=np.uint8)

    @classmethod
    def from_prims(cls, P, da, gamma, rho, u, v, p, bctype, **kw):
        U = prim_to_cons(gamma, rho, u, v, p)
        ny, nx = rho.shape
        Q = np.zeros((4, (P + 1) ** 2, ny, nx))
        Q[:, 0] = U
        return cls(P, da, gamma, Q, bctype, **kw)

    def step(self, dt: float, n: int = 1) -> None:
        for _ in range(n):
            lib().zko_dg_step(self.ny, self.nx, self.P, self.da, self.gamma, self.bct, self.Q, float(dt),
                              self.flux, self.limiter, self.M, self.flags)

    def rhs(self) -> np.ndarray:
        dQ = np.zeros_like(self.Q)
        lib().zko_dg_rhs(self.ny, self.nx, self.P, self.da, self.gamma, self.bct, self.Q, self.flux, dQ)
        return dQ

    def fill(self) -> None:
        lib().zko_dg_fill_ghosts(self.ny, self.nx, self.P, self.bct, self.Q)

    def limit_tvb(self, linear_only: bool = False) -> np.ndarray:
        f = np.zeros((self.ny, self.nx), dtype=np.uint8)
        fn = lib().zko_dg_limit_tvb_linear if linear_only else lib().zko_dg_limit_tvb
        fn(self.ny, self.nx, self.P, self.da, self.M, self.bct, self.Q, f)
        return f

    def limit_persson_zs(self) -> None:
        lib().zko_dg_limit_persson_zs(self.ny, self.nx, self.P, self.gamma, self.bct, self.Q)

    def limit_zs(self, ext: bool = False) -> None:
        (lib().zko_dg_limit_zs_ext if ext else lib().zko_dg_limit_zs)(self.ny, self.nx, self.P, self.gamma, self.Q)

    def max_speed(self) -> float:
        return lib().zko_dg_max_speed(self.ny, self.nx, self.P, self.gamma, self.bct, self.Q)


def dg_cfl_limit(P: int) -> float:
    return lib().zko_dg_cfl_limit(P)
