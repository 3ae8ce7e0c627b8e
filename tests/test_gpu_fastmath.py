"""The flux routines divide and take square roots with refined hardware
estimates (rcp_nz, div_nz, sqrt_pos in zuikakucfd_amd/csrc/zk_device.h)
instead of the IEEE sequences.  On normal operands they must give the
correctly rounded results: tests/fastmath_check.hip compares them with the
IEEE operations on 2.7e8 pseudo-random operands over 16 binades either side
of 1, both signs for divisors."""
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu
EXE = Path(__file__).resolve().parent / "_bin" / "fastmath_check"


def test_flux_arithmetic_is_correctly_rounded():
    assert EXE.exists(), "build() compiles tests/fastmath_check.hip into tests/_bin/"
    print("[case] fastmath_check, 2^28 operands", flush=True)
    out = subprocess.run([str(EXE)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    words = out.stdout.split()
    res = {words[i]: int(words[i + 1]) for i in range(0, len(words) - 1, 2)}
    assert res["samples"] == 4096 * 256 * 256, res
    assert res["rcp_ulp"] == 0 and res["div_ulp"] == 0 and res["sqrt_ulp"] == 0, res
