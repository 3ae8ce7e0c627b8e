"""Build recipe for the native parts of zuikakucfd_amd.

* ``libzk.so``   -- the HIP library behind include/zk.h, compiled for gfx950
                    with hipcc (one object per .hip file, linked together).
* ``bin/zk_weno``, ``bin/zk_dg{2,3,4}`` -- the drop-in host executables
                    (C++17, g++) that read params.txt/rho,u,v,p,bctype.dat and
                    drive libzk.so with the reference's command line.
* ``tests/_bin/*_check`` -- check programs the tests run: device arithmetic
                    (``tests/*_check.hip``) and host checks of the library's
                    headers against the oracle (``tests/*_check.cpp``).
* the oracle     -- ``oracle/liboracle.so`` via ``make -C oracle`` (test
                    infrastructure; built here so a checkout has one entry).

Everything is built in-tree so the shared objects travel with the sources.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "libzk.so"
BIN = PKG / "bin"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-command-line-argument",
    f"-I{ROOT / 'include'}",
]
HIP_SOURCES = ["zk_api.hip", "fv_kernels.hip", "dg_kernels.hip", "ens_kernels.hip", "mlp_kernels.hip",
               "noalloc.hip", "stream.hip"]


def _run(cmd: list[str], cwd: Path | None = None) -> None:
    proc = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{proc.stdout}")
    # hipcc's host pass warns that it does not know the packed-fp32 feature; keep other output
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if out:
        print(out, file=sys.stderr)


def build_lib(verbose: bool = True) -> Path:
    OBJ.mkdir(exist_ok=True)
    srcs = [s for s in HIP_SOURCES if (CSRC / s).exists()]

    def one(src: str) -> Path:
        obj = OBJ / (Path(src).stem + ".o")
        _run([HIPCC, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
        return obj

    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(one, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}", file=sys.stderr)
    return LIB


def build_bins() -> list[Path]:
    BIN.mkdir(exist_ok=True)
    src = CSRC / "host" / "zk_cli.cpp"
    if not src.exists():
        return []
    outs = []
    common = ["g++", "-O2", "-std=c++17", "-Wall", "-Wextra", f"-I{ROOT / 'include'}", str(src),
              f"-L{PKG}", "-lzk", "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath,/opt/rocm/lib"]
    targets = [("zk_weno", "-DZK_CLI_WENO"), ("zk_dg2", "-DZK_CLI_DG_P=1"), ("zk_dg3", "-DZK_CLI_DG_P=2"),
               ("zk_dg4", "-DZK_CLI_DG_P=3")]
    for name, define in targets:
        out = BIN / name
        _run([*common[:1], define, *common[1:], "-o", str(out)])
        outs.append(out)
    return outs


def build_checks() -> list[Path]:
    """Check programs the tests run (tests/_bin/): device-arithmetic checks
    (tests/*_check.hip, hipcc) and host checks of the library's headers against
    the oracle (tests/*_check.cpp, g++, linked with oracle/liboracle.so)."""
    out_dir = ROOT / "tests" / "_bin"
    out_dir.mkdir(exist_ok=True)
    outs = []
    for src in sorted((ROOT / "tests").glob("*_check.hip")):
        out = out_dir / src.stem
        _run([HIPCC, *HIP_FLAGS, f"-I{CSRC}", str(src), "-o", str(out)])
        outs.append(out)
    for src in sorted((ROOT / "tests").glob("*_check.cpp")):
        out = out_dir / src.stem
        _run(["g++", "-O2", "-std=c++17", "-Wall", "-Wextra", f"-I{CSRC}", f"-I{ROOT / 'oracle'}", str(src),
              f"-L{ROOT / 'oracle'}", "-loracle", "-Wl,-rpath,$ORIGIN/../../oracle", "-o", str(out)])
        outs.append(out)
    return outs


def build_oracle() -> Path:
    make = shutil.which("make") or "make"
    _run([make, "-s", "-C", str(ROOT / "oracle"), "-B", "liboracle.so"])
    return ROOT / "oracle" / "liboracle.so"


def build_all() -> None:
    build_lib()
    build_bins()
    build_oracle()
    build_checks()


if __name__ == "__main__":
    build_all()
