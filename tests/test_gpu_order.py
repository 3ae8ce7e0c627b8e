"""Order of accuracy of the GPU DG path on the smooth density wave of the C2
configuration (rho = 1 + 0.2 sin(2 pi (x + y)), u = v = 1, p = 1, periodic),
at grids of 128^2, 256^2 and 512^2 updated cells, with the TVB limiter (M = 50)
on: the observed order of the L2 error of the density polynomial (every mode,
by Legendre orthogonality) is at least P + 0.5 between successive grids, and
the limiter flags no cell at any checkpoint.  (The cell means alone
superconverge and reach round-off at 512^2 for P >= 2.)

The exact solution is the wave translated by (t, t); both the initial modes
and the reference are its L2 projection (meshio.smooth_wave_modes).  At P = 3
the time step shrinks as h^(4/3), so that SSP-RK3's third-order time error
does not mask the fourth-order spatial error.
"""
import numpy as np
import pytest

from zuikakucfd_amd import meshio
from zuikakucfd_amd import zk

pytestmark = pytest.mark.gpu

T = 0.05


def say(msg):
    print(f"[case] {msg}", flush=True)


def run(P, n_upd):
    n = n_upd + 6
    m = meshio.smooth_wave(n)
    s = zk.Solver(method=zk.METHOD_DG, P=P, limiter=zk.LIMITER_TVB, tvb_M=50.0)
    s.load(m.da, m.gamma, m.rho, m.u, m.v, m.p, m.bctype)
    s.set_state(meshio.smooth_wave_modes(n, P))
    lam = s.max_speed()
    cfl = 0.9 / (2.0 * (2 * P + 1))
    dt0 = 0.5 * cfl * m.da / lam
    if P == 3:
        dt0 *= (m.da * 128.0) ** (1.0 / 3.0)
    k = int(np.ceil(T / dt0))
    dt = T / k
    flagged = 0
    chunks = 10
    done = 0
    for c in range(chunks):
        todo = k * (c + 1) // chunks - done
        s.step(dt, todo)
        done += todo
        flagged += int(s.get_flags().sum())
    Q = s.get_state()
    s.close()
    Qe = meshio.smooth_wave_modes(n, P, t=T)
    sl = slice(3, n - 3)
    NQ = P + 1
    # |P_px P_py|^2 averages to 1/((2px+1)(2py+1)) over a cell
    w = np.array([1.0 / ((2 * (mm // NQ) + 1) * (2 * (mm % NQ) + 1)) for mm in range(NQ * NQ)])
    d = Q[0][:, sl, sl] - Qe[0][:, sl, sl]
    err = float(np.sqrt(np.einsum("m,mij->", w, d * d) / d[0].size))
    return err, m.da, flagged, k


@pytest.mark.parametrize("P", [1, 2, 3])
def test_dg_tvb_order_smooth_wave(P):
    say(f"dg P={P} smooth-wave order, 128^2 -> 512^2, TVB on")
    res = [run(P, n) for n in (128, 256, 512)]
    errs = [r[0] for r in res]
    rates = [np.log(errs[i] / errs[i + 1]) / np.log(res[i][1] / res[i + 1][1]) for i in range(2)]
    print(f"[order P={P}] errors {errs}, rates {rates}, steps {[r[3] for r in res]}, "
          f"flagged {[r[2] for r in res]}", flush=True)
    assert all(r[2] == 0 for r in res), [r[2] for r in res]
    assert min(rates) >= P + 0.5, rates
