"""The XCD-aware workgroup-to-tile map is a bijection for every grid size
(compiled from the kernel's own header with the host compiler)."""
import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def test_xcd_tile_is_a_bijection(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text(
        '#include "tile_map.h"\n#include <cstdio>\n#include <vector>\n'
        "int main(){for(int nb=1;nb<5000;++nb){std::vector<int> seen(nb,0);"
        "for(int b=0;b<nb;++b){int t=zk::xcd_tile(b,nb);if(t<0||t>=nb||seen[t]++)return 1;}}"
        "std::puts(\"ok\");return 0;}\n")
    exe = tmp_path / "t"
    subprocess.run(["g++", "-O1", "-std=c++17", f"-I{ROOT / 'zuikakucfd_amd' / 'csrc'}", str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0 and out.stdout.strip() == "ok"
