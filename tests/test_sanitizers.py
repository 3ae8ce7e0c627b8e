"""AddressSanitizer + UndefinedBehaviorSanitizer builds of the host code (CPU
only): the oracle (oracle/zk_oracle.c) driven through every entry point by
oracle/sanitize_check.c, and the drop-in executable's input parsing
(zuikakucfd_amd/csrc/host/zk_cli.cpp) on a missing folder, a short file, bad
numbers and a bad option.  Both are plain executables built with the
sanitizers linked in; any report fails the test."""
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from zuikakucfd_amd import meshio

ROOT = Path(__file__).resolve().parents[1]
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]
ENV = {"ASAN_OPTIONS": "detect_leaks=0:abort_on_error=0:halt_on_error=1",
       "UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1", "OMP_NUM_THREADS": "4"}


def build(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr


def run(cmd, cwd=None):
    import os

    env = dict(os.environ)
    env.update(ENV)
    return subprocess.run(cmd, capture_output=True, text=True, cwd=cwd, env=env, timeout=300)


def clean(r):
    out = r.stdout + r.stderr
    return "AddressSanitizer" not in out and "runtime error" not in out and "LeakSanitizer" not in out


@pytest.fixture(scope="module")
def outdir(tmp_path_factory):
    if not shutil.which("gcc") or not shutil.which("g++"):
        pytest.skip("no gcc/g++")
    return tmp_path_factory.mktemp("san")


def test_oracle_under_asan_ubsan(outdir):
    exe = outdir / "sanitize_check"
    build(["gcc", *SAN, "-fopenmp", "-std=gnu11", f"-I{ROOT / 'oracle'}", str(ROOT / "oracle" / "sanitize_check.c"),
           str(ROOT / "oracle" / "zk_oracle.c"), "-lm", "-o", str(exe)])
    r = run([str(exe)])
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert r.returncode == 0
    assert clean(r)
    assert "sanitize_check:" in r.stdout


def test_cli_input_parsing_under_asan_ubsan(outdir, tmp_path):
    exe = outdir / "zk_weno_san"
    build(["g++", *SAN, "-std=c++17", "-DZK_CLI_WENO", f"-I{ROOT / 'include'}",
           str(ROOT / "zuikakucfd_amd" / "csrc" / "host" / "zk_cli.cpp"), f"-L{ROOT / 'zuikakucfd_amd'}", "-lzk",
           f"-Wl,-rpath,{ROOT / 'zuikakucfd_amd'}", "-Wl,-rpath,/opt/rocm/lib", "-o", str(exe)])
    m = meshio.riemann_2d(16)
    good = tmp_path / "good"
    meshio.write_mesh_folder(m, good)
    short = tmp_path / "short"
    meshio.write_mesh_folder(m, short)
    txt = (short / "rho.dat").read_text().split()
    (short / "rho.dat").write_text(" ".join(txt[: len(txt) // 2]))
    cases = [
        ([], "usage"),
        ([str(tmp_path / "missing"), "1e-4", "10"], None),
        ([str(short), "1e-4", "10"], None),
        ([str(good), "abc", "10"], "bad numeric"),
        ([str(good), "1e-4", "10", "--flux", "nosuch"], "unknown flux"),
        ([str(good), "1e-4", "10", "--M", "abc"], "bad numeric option"),
        ([str(good), "1e-4", "10", "--limiter"], "unknown option"),
    ]
    for args, needle in cases:
        r = run([str(exe), *args], cwd=tmp_path)
        print(args, r.returncode, r.stderr[-500:])
        assert r.returncode == 1, (args, r.returncode, r.stderr[-2000:])
        assert clean(r), r.stderr[-4000:]
        if needle:
            assert needle in (r.stdout + r.stderr)
