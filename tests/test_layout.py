"""Host-side check of the stage's launch layouts (no GPU needed).

layout_check.cpp includes the library's own layout code (marching regions,
chunk rows, fast-block test, edge-tile rectangles from
navierstokes_amd/csrc/ns_kernels.h) and checks over ~30k grid / buffer /
rows-per-step / chunk-height combinations that the marching blocks and the
edge tiles cover every point exactly once, that every block's U window and
level-1 rows stay inside the grid with interior stencils only, and that
blocks taking the fast per-point code keep their level-1 halo in the fast
zone.  Dropping the row alignment of the general region with fast blocks,
or widening the fast-block test by 3 columns, makes it fail.
"""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def _hipcc():
    for cand in ("/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    return None


@pytest.mark.skipif(_hipcc() is None, reason="hipcc (host compiler for the HIP headers) not available")
def test_stage_layouts_cover_the_grid_once(tmp_path):
    exe = tmp_path / "layout_check"
    subprocess.run([_hipcc(), "-std=c++17", "-O1", "-I", str(ROOT / "navierstokes_amd" / "csrc"), "-I",
                    str(ROOT / "include"), str(ROOT / "tests" / "layout_check.cpp"), "-o", str(exe)],
                   check=True, capture_output=True)
    res = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    print(res.stdout)
    assert res.returncode == 0, res.stdout
    assert res.stdout.startswith("ok ")
    assert int(res.stdout.split()[1]) > 10000
