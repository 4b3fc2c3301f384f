"""The host's choice of the giant-island grid (csrc/giant_grid.hpp), checked on the CPU.

k_t_giant and k_t_color_giant meet at grid-wide barriers, so the engine launches them only with a
grid whose workgroups can all be resident at once: the preferred 64 workgroups, lowered to the
occupancy of the two kernels times the compute units, and 0 (the engine refuses, before any kernel
runs) when not one workgroup fits. The header is plain C++; this test compiles it with g++."""
import ctypes as C
import subprocess
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent
HDR = HERE.parent / "3fold-motion-engine_amd" / "csrc" / "giant_grid.hpp"


@pytest.fixture(scope="module")
def grid(tmp_path_factory):
    d = tmp_path_factory.mktemp("giant_grid")
    src = d / "g.cpp"
    src.write_text(f'#include "{HDR}"\n'
                   'extern "C" int gg(int p, int s, int c, int n) { return me::giant_grid_size(p, s, c, n); }\n')
    lib = d / "libg.so"
    r = subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-o", str(lib), str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    L = C.CDLL(str(lib))
    L.gg.argtypes = [C.c_int] * 4
    return L.gg


def test_preferred_grid_when_it_fits(grid):
    # MI355X: 256 CUs, one 1024-thread workgroup per CU for either kernel
    assert grid(64, 1, 1, 256) == 64
    assert grid(64, 2, 1, 256) == 64


def test_lowered_to_what_is_resident(grid):
    assert grid(64, 1, 1, 40) == 40          # fewer CUs than the preferred grid
    assert grid(64, 1, 2, 20) == 20          # the sweep kernel is the tighter one
    assert grid(64, 3, 1, 30) == 30          # the colouring kernel is the tighter one
    assert grid(64, 2, 2, 20) == 40


def test_refused_when_nothing_fits(grid):
    assert grid(64, 0, 1, 256) == 0
    assert grid(64, 1, 0, 256) == 0
    assert grid(64, 1, 1, 0) == 0
    assert grid(0, 1, 1, 256) == 0
