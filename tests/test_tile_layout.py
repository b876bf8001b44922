"""Host check of the tile kernels' LDS layouts (tests/cxx/tile_layout_check.cpp
compiled with g++ against salvus_amd/csrc/tile_layout.h, the header the
kernels use): regions inside the layout and apart, the 160 KiB fit, the
padded layout only where it keeps the workgroups per CU, the padded slot
mapping one to one, and the padded layout's bank spread."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_tile_layouts(tmp_path):
    exe = tmp_path / "tile_layout_check"
    subprocess.run(["g++", "-O1", "-std=c++17", "-Wall", "-I", str(ROOT / "salvus_amd" / "csrc"),
                    "-o", str(exe), str(ROOT / "tests" / "cxx" / "tile_layout_check.cpp")],
                   check=True, timeout=300)
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    print(run.stdout)
    assert run.returncode == 0, run.stdout[-3000:]
    assert "all layouts ok" in run.stdout
    # the order-4 3-D elastic 4x2x2 tile takes the padded layout
    assert "dim 3 order 4 elastic 4x2x2 normal: 20442 LDS words (2 per CU), padded 1" in run.stdout
