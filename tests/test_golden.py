"""Golden checks against the reference's own data files.

The reference ships two input fixtures (2D-Riemann/, HighMachJet/) and no
expected outputs.  Their sha256 hashes are committed in
tests/golden/fixture_hashes.json; the generators must reproduce them byte for
byte, which pins the input conventions (y-down rows, v_file = -v_phys, the
3-cell ghost ring, the fixed-inflow bctype-0 strip, the text layouts).
"""
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

from zuikakucfd_amd import meshio

GOLD = json.loads((Path(__file__).parent / "golden" / "fixture_hashes.json").read_text())
REF = Path("/root/reference")
FILES = ["params.txt", "rho.dat", "u.dat", "v.dat", "p.dat", "bctype.dat"]


def _sha(p: Path) -> str:
    return hashlib.sha256(p.read_bytes()).hexdigest()


def test_riemann_generator_reproduces_shipped_fixture(tmp_path):
    meshio.write_mesh_folder(meshio.riemann_2d(256, split=(0.8, 0.8)), tmp_path)
    for f in FILES:
        assert _sha(tmp_path / f) == GOLD["2D-Riemann"][f], f


def test_jet_generator_reproduces_shipped_fixture(tmp_path):
    m = meshio.jet(400, 200, da=0.0025, inflow=(5.0, 800.0, 0.0, 0.4127), fixed_left=True)
    meshio.write_mesh_folder(m, tmp_path, joined=True)
    for f in FILES:
        assert _sha(tmp_path / f) == GOLD["HighMachJet"][f], f


def test_reader_on_generated_fixture(tmp_path):
    m = meshio.riemann_2d(256)
    meshio.write_mesh_folder(m, tmp_path)
    r = meshio.read_mesh_folder(tmp_path)
    assert (r.nx, r.ny, r.gamma) == (256, 256, 1.4)
    assert r.da == pytest.approx(1 / 256, rel=1e-10)
    np.testing.assert_array_equal(r.bctype, m.bctype)
    # quadrant split at column 205 / row 51, v_file = -v_phys below the split
    assert r.rho[50, 205] == 1.5 and r.rho[51, 205] == 0.5323 and r.rho[50, 204] == 0.5323
    assert r.v[100, 100] == -1.206 and r.v[10, 10] == 0.0


@pytest.mark.skipif(not REF.exists(), reason="reference checkout not present")
def test_hashes_match_reference_checkout():
    for d in ("2D-Riemann", "HighMachJet"):
        for f in FILES:
            assert _sha(REF / d / f) == GOLD[d][f]


def test_config3_jet_layout():
    """2048x1024 jet: inflow rows |y_c| < 0.05 are rows 461..562 (102 rows)."""
    m = meshio.jet()
    rows = np.nonzero(m.bctype[:, 0] == 0)[0]
    assert rows.min() == 461 and rows.max() == 562 and rows.size == 102
    assert m.u[461, 0] == 30.0 and m.u[461, 3] == 0.0
