"""Mirror slip walls of the WENO path (an option of this project: the
reference's finite-volume fill leaves bctype-1 cells untouched).  The oracle's
fill (oracle/zk_oracle.c, fill_slip_mirror) is pinned by hand-placed marker
values, by a closed box that must conserve mass and energy, by a uniform
flow along the walls that must stay uniform, and by the wall reproducing the
upper half of a doubled domain whose lower half is the mirror image."""
import numpy as np
import pytest

import oracle
from zuikakucfd_amd import meshio

G = 1.4


def fill(U, bct):
    o = oracle.FV(1.0, G, U, bct, flux=oracle.FV_SLIP)
    o.fill()
    return o.U


def test_mirror_fill_marker_values():
    """Bottom wall under interior rows 3..8 (column 5): ghost rows 9, 10, 11
    take rows 8, 7, 6 (1, 3, 5 cells away); a wall two cells thick sitting on
    a one-row interior takes that row for both layers; rho*v flips sign,
    rho*u does not."""
    ny, nx = 12, 11
    bct = meshio.ring_bctype(ny, nx)
    bct[9:, :] = 1
    U = np.zeros((4, ny, nx))
    for c in range(4):
        U[c] = np.arange(ny * nx).reshape(ny, nx) + 1000 * c
    out = fill(U, bct)
    j = 5
    for gi, si in ((9, 8), (10, 7), (11, 6)):
        assert out[0, gi, j] == U[0, si, j] and out[3, gi, j] == U[3, si, j]
        assert out[1, gi, j] == U[1, si, j] and out[2, gi, j] == -U[2, si, j]
    # row pass: a left wall (columns 0..2 slip) mirrors columns 5, 4, 3 into 0, 1, 2 and flips rho*u
    b2 = meshio.ring_bctype(ny, nx)
    b2[:, :3] = 1
    out = fill(U, b2)
    for gj, sj in ((2, 3), (1, 4), (0, 5)):
        assert out[0, 6, gj] == U[0, 6, sj] and out[1, 6, gj] == -U[1, 6, sj] and out[2, 6, gj] == U[2, 6, sj]
    # a thin interior: rows 3 only (rows 4.. slip); the mirrored rows 2 (k=2) and 1 (k=3)
    # are not updated, so those layers take row 3 itself
    b3 = meshio.ring_bctype(ny, nx)
    b3[4:, :] = 1
    out = fill(U, b3)
    for gi in (4, 5, 6):
        assert out[0, gi, j] == U[0, 3, j] and out[2, gi, j] == -U[2, 3, j]
    assert out[0, 7, j] == U[0, 7, j]  # 4 cells from the interior: left alone
    # the reference's fill (option off) leaves slip cells alone
    o = oracle.FV(1.0, G, U, bct)
    o.fill()
    assert np.array_equal(o.U[:, 9:, :], U[:, 9:, :])


def random_state(rng, ny, nx):
    rho = rng.uniform(0.5, 1.5, (ny, nx)); p = rng.uniform(0.5, 1.5, (ny, nx))
    u = rng.uniform(-0.5, 0.5, (ny, nx)); v = rng.uniform(-0.5, 0.5, (ny, nx))
    return rho, u, v, p


def test_closed_box_conserves_mass_and_energy():
    """Slip walls all round (a 3-cell slip ring) with an obstacle inside, 7
    cells thick so that no wall cell serves two opposite faces: the wall
    faces carry mirror states, the x faces from the row mirror and the y
    faces from the column mirror (also at the obstacle's corners, where a
    wall cell is near both), so no mass or energy crosses them."""
    n = 40
    m = meshio.riemann_2d(n, split=(0.6, 0.6))
    bct = meshio.ring_bctype(n, n, 1)
    bct[15:22, 12:25] = 1
    U = oracle.prim_to_cons(G, m.rho, m.u, m.v, m.p)
    o = oracle.FV(1.0 / 34, G, U, bct, flux=oracle.FV_SLIP)
    upd = bct == 0
    upd[:3] = upd[-3:] = False
    upd[:, :3] = upd[:, -3:] = False
    m0, e0 = o.U[0][upd].sum(), o.U[3][upd].sum()
    dt = 0.3 / 34 / o.max_speed()
    o.step(dt, 30)
    assert abs(o.U[0][upd].sum() - m0) < 1e-13 * m0
    assert abs(o.U[3][upd].sum() - e0) < 1e-13 * e0
    # with the option off the slip cells are never filled: mass leaks
    o2 = oracle.FV(1.0 / 34, G, U, bct)
    o2.step(dt, 30)
    assert abs(o2.U[0][upd].sum() - m0) > 1e-6 * m0


def test_uniform_flow_along_walls_stays_uniform():
    n = 30
    bct = meshio.ring_bctype(n, n)
    bct[:3, :] = 1
    bct[-3:, :] = 1  # channel: slip walls top and bottom, -1 inflow/outflow ends
    one = np.ones((n, n))
    U = oracle.prim_to_cons(G, one, 0.7 * one, 0.0 * one, one)
    o = oracle.FV(1.0 / 24, G, U, bct, flux=oracle.FV_SLIP)
    o.step(1e-3, 20)
    for c in range(4):
        np.testing.assert_allclose(o.U[c, 3:-3, 3:-3], U[c, 3:-3, 3:-3], rtol=0, atol=1e-14)


@pytest.mark.parametrize("axis", [0, 1])
def test_wall_reproduces_the_mirrored_doubled_domain(axis):
    """Half domain with a slip wall against the upper half of a domain twice
    as tall whose lower half is the mirror image (normal momentum negated).
    Away from the doubled domain's far boundary (reached after H/9 steps) the
    two agree to rounding: a wrong layer index or sign at the wall does not.
    axis 1 transposes the set-up (wall on the right, the row pass)."""
    rng = np.random.default_rng(11 + axis)
    H, W, steps = 48, 20, 4
    rho, u, v, p = random_state(rng, H, W + 6)
    # doubled domain F: interior rows 3 .. 3+2H-1, mirror line between 3+H-1 and 3+H
    ny_f = 2 * H + 6
    F = [np.ones((ny_f, W + 6)) for _ in range(4)]
    for f, a, sgn in zip(F, (rho, u, v, p), (1, 1, -1, 1)):
        f[3:3 + H] = a
        f[3 + H:3 + 2 * H] = sgn * a[::-1]
    bf = meshio.ring_bctype(ny_f, W + 6)
    # half domain A: the upper half, then 3 slip rows
    ny_a = H + 6
    A = [f[:ny_a].copy() for f in F]
    ba = meshio.ring_bctype(ny_a, W + 6)
    ba[3 + H:, :] = 1
    if axis == 1:
        F = [F[0].T, F[2].T, F[1].T, F[3].T]
        A = [A[0].T, A[2].T, A[1].T, A[3].T]
        bf, ba = bf.T.copy(), ba.T.copy()
    h = 1.0 / W
    of = oracle.FV(h, G, oracle.prim_to_cons(G, *F), bf)
    oa = oracle.FV(h, G, oracle.prim_to_cons(G, *A), ba, flux=oracle.FV_SLIP)
    dt = 0.3 * h / of.max_speed()
    of.step(dt, steps)
    oa.step(dt, steps)
    if axis == 1:
        ua, uf = oa.U[:, 3:-3, 3:3 + H], of.U[:, 3:-3, 3:3 + H]
    else:
        ua, uf = oa.U[:, 3:3 + H, 3:-3], of.U[:, 3:3 + H, 3:-3]
    np.testing.assert_allclose(ua, uf, rtol=1e-12, atol=1e-13)
