"""zk_weno_step_noalloc (the entry point that allocates nothing) against the
handle path's zk_step, bit for bit, through the C ABI, on two calls in a row,
the second with a smaller grid.  The test's own device buffers and stream come
from the HIP runtime libzk.so itself uses (through ctypes); the library
allocates nothing for the call."""
import ctypes as C

import numpy as np
import pytest

from zuikakucfd_amd import meshio
from zuikakucfd_amd import zk

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[case] {msg}", flush=True)


def obstacle_mesh(ny, nx, seed):
    """Four-quadrant data on a ny x nx grid with interior -1 obstacles, a
    bctype-1 block (never filled on the finite-volume path) and fixed cells."""
    m = meshio.riemann_2d(max(ny, nx))
    m.rho, m.u, m.v, m.p = (a[:ny, :nx].copy() for a in (m.rho, m.u, m.v, m.p))
    b = meshio.ring_bctype(ny, nx, -1)
    rng = np.random.default_rng(seed)
    for _ in range(6):
        i, j = rng.integers(8, ny - 12), rng.integers(8, nx - 12)
        b[i:i + 3, j:j + 4] = -1
    b[ny // 2:ny // 2 + 2, 10:12] = 1
    b[20:24, :3] = 0  # fixed inflow-like cells in the ring
    m.bctype, m.ny, m.nx = b, ny, nx
    return m


def periodic_mesh(n, seed):
    m = meshio.random_riemann(seed, n)
    m.bctype = meshio.ring_bctype(n, n, -2)
    return m


class Hip:
    """The few HIP runtime calls the test needs, from the runtime libzk.so is
    linked against (already loaded by zk.load_library, found by its soname)."""

    def __init__(self):
        zk.load_library()
        self.h = C.CDLL("libamdhip64.so.7")
        self.bufs = []

    def check(self, e, what):
        assert e == 0, f"{what}: hip error {e}"

    def upload(self, a: np.ndarray) -> int:
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        self.check(self.h.hipMalloc(C.byref(p), C.c_size_t(a.nbytes)), "hipMalloc")
        self.bufs.append(p)
        self.check(self.h.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1), "hipMemcpy H2D")
        return p.value

    def download(self, ptr: int, like: np.ndarray) -> np.ndarray:
        out = np.empty_like(like)
        self.check(self.h.hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(ptr), C.c_size_t(out.nbytes), 2),
                   "hipMemcpy D2H")
        return out

    def stream(self) -> int:
        s = C.c_void_p()
        self.check(self.h.hipStreamCreate(C.byref(s)), "hipStreamCreate")
        self.streams = getattr(self, "streams", []) + [s]
        return s.value

    def sync(self, s: int):
        self.check(self.h.hipStreamSynchronize(C.c_void_p(s)), "hipStreamSynchronize")

    def close(self):
        for p in self.bufs:
            self.h.hipFree(p)
        for s in getattr(self, "streams", []):
            self.h.hipStreamDestroy(s)


def handle_result(m, dt, steps, precision, flux):
    s = zk.Solver(method=zk.METHOD_WENO, precision=precision, flux=flux)
    s.load(m.da, m.gamma, m.rho, m.u, m.v, m.p, m.bctype)
    Q0 = s.get_state().reshape(4, m.ny, m.nx).copy()
    s.step(dt, steps)
    Q = s.get_state().reshape(4, m.ny, m.nx).copy()
    s.close()
    return Q0, Q


def test_noalloc_matches_zk_step():
    cases = [
        ("obstacles 120x136 fp64 HLLC", obstacle_mesh(120, 136, 1), 5, zk.PREC_FP64, zk.FLUX_HLLC),
        ("periodic 40x40 fp32-bulk HLL", periodic_mesh(40, 7), 3, zk.PREC_FP32_BULK, zk.FLUX_HLL),
    ]
    hip = Hip()
    stream = hip.stream()
    try:
        for name, m, steps, prec, flux in cases:
            say(f"noalloc {name}")
            dt = 2e-4
            Q0, ref = handle_result(m, dt, steps, prec, flux)
            q = hip.upload(Q0)
            b = hip.upload(np.ascontiguousarray(m.bctype, dtype=np.int32))
            zk.weno_step_noalloc(q, b, m.ny, m.nx, m.da, m.gamma, dt, steps, stream=stream, precision=prec,
                                 flux=flux)
            hip.sync(stream)
            got = hip.download(q, Q0)
            assert np.isfinite(got).all()
            np.testing.assert_array_equal(got, ref)
    finally:
        hip.close()


def test_noalloc_rejects_too_large_grid():
    """More cells than ZK_NOALLOC_MAX_CELLS is an argument error, decided on
    the host before anything is enqueued."""
    say("noalloc size limit")
    with pytest.raises(zk.ZkError, match="exceed the maximum"):
        zk.weno_step_noalloc(1, 1, 1025, 1024, 1e-3, 1.4, 1e-4, 1)
