"""The closed-form finite-volume ghost map (zuikakucfd_amd/csrc/fv_map.h),
which builds the map on the device one cell at a time, against the oracle's
fill (the reference's loops, src_WENO/fluid.cpp:338-370) with marker values
on random masks; and the ghost-ring order as a bijection.  CPU only: runs the
host check program tests/_bin/fvmap_check."""
import subprocess
from pathlib import Path

import pytest

EXE = Path(__file__).resolve().parent / "_bin" / "fvmap_check"


def test_closed_form_map_matches_oracle_fill():
    if not EXE.exists():
        pytest.fail(f"{EXE} is missing: run the build (zuikakucfd_amd/build.py)")
    r = subprocess.run([str(EXE)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert ", 0 mismatches" in r.stdout
