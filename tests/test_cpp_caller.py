"""The C++ API as a reference caller uses it: a program written like the
reference benchmark (benchmark/pbf_benchmark.cu:121-211), compiled with the
system C++ compiler against include/ and linked to libpbf_amd.so."""
from __future__ import annotations

import os
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

REPO = Path(__file__).resolve().parents[1]
SRC = REPO / "tests" / "cpp" / "reference_caller.cpp"
LIBDIR = REPO / "fluid-simulator_amd" / "pbf_amd"


def _build(tmp_path: Path) -> Path:
    cxx = shutil.which("g++") or shutil.which("c++")
    if cxx is None:
        pytest.skip("no host C++ compiler")
    if not (LIBDIR / "libpbf_amd.so").exists():
        pytest.fail("libpbf_amd.so is not built")
    exe = tmp_path / "reference_caller"
    subprocess.run([cxx, "-std=c++17", "-O2", "-Wall", "-Werror", f"-I{REPO / 'include'}",
                    str(SRC), f"-L{LIBDIR}", "-lpbf_amd", "-o", str(exe)], check=True)
    return exe


def test_reference_shaped_caller_compiles_and_links(tmp_path):
    exe = _build(tmp_path)
    assert exe.exists()


@pytest.mark.gpu
def test_reference_shaped_caller_matches_the_binding(tmp_path):
    import pbf_amd
    from pbf_amd import scenes

    exe = _build(tmp_path)
    out = tmp_path / "state.bin"
    env = dict(os.environ, LD_LIBRARY_PATH=f"{LIBDIR}:{os.environ.get('LD_LIBRARY_PATH', '')}")
    r = subprocess.run([str(exe), "12", "10", "8", "3", str(out)], env=env, capture_output=True,
                       text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    state = np.fromfile(out, dtype=np.float32).reshape(2, -1, 4)

    sc = scenes.small_block(12, 10, 8)
    s = pbf_amd.PBFSolver()
    s.initialize(sc.n, sc.minBounds, sc.maxBounds, sc.params)
    s.setParticles(sc.positions, sc.velocities)
    s.run(3)
    assert np.array_equal(state[0], s.copyPositionsToHost())
    assert np.array_equal(state[1], s.copyVelocitiesToHost())
