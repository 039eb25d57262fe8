"""The C ABI library loads and exports every entry point include/pbf_amd.h
declares; the Python binding declares a signature for each.  No GPU needed:
loading the library does not touch a device."""
from __future__ import annotations

import ctypes
import re
import subprocess
from pathlib import Path

import pbf_amd

REPO = Path(__file__).resolve().parent.parent
HEADER = REPO / "include" / "pbf_amd.h"


def declared_functions() -> set[str]:
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    names = set(re.findall(r"\b(pbf_[a-z0-9_]+)\s*\(", text))
    return {n for n in names if not n.startswith("pbf_status")}


def test_header_declares_the_boundary():
    names = declared_functions()
    for required in ("pbf_solver_step", "pbf_grid_build", "pbf_collision_solve",
                     "pbf_k_find_neighbors", "pbf_k_compute_lambda",
                     "pbf_k_compute_delta_position", "pbf_k_apply_xsph_viscosity"):
        assert required in names
    assert len(names) >= 70


def test_library_loads_and_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(pbf_amd.LIB_PATH))
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def test_binding_covers_every_declared_symbol():
    assert set(pbf_amd.declared_symbols()) == declared_functions()


def test_exported_symbols_are_only_the_c_abi_and_cpp_api():
    out = subprocess.run(["nm", "-D", "--defined-only", str(pbf_amd.LIB_PATH)],
                         capture_output=True, text=True, check=True).stdout
    exported_c = {line.split()[-1] for line in out.splitlines()
                  if " T " in line and line.split()[-1].startswith("pbf_")}
    assert exported_c == declared_functions()


def test_binding_struct_layouts_match_the_header():
    assert ctypes.sizeof(pbf_amd.SimulationParams) == 72
    assert pbf_amd.SimulationParams.gravity.offset == 40
    assert pbf_amd.SimulationParams.scorrK.offset == 52
    assert pbf_amd.SimulationParams.vorticityStrength.offset == 68


def test_version_string_without_device():
    assert pbf_amd.lib().pbf_version().decode().startswith("pbf_amd")
