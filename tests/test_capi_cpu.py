"""CPU checks of the drop-in boundary: the HIP library loads without a GPU and exports every
entry point declared in include/nse2d_hip.h, and the Python binding declares all of them."""
import os
import re
import subprocess

import nse2d_hip as nh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "nse2d_hip.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^\s*(?!typedef\b)(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s+\**\s*([A-Za-z_][A-Za-z0-9_]*)\s*\(",
                       src, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "return")))


def test_header_declares_reference_boundary():
    names = declared_functions()
    for n in ("testapi_uv_from_psi_mapped", "testapi_omega_from_psi_interior_mapped",
              "testapi_thom_wall_vorticity_mapped", "testapi_arakawa_J_mapped",
              "testapi_compute_adv_from_uv_mapped", "testapi_compute_adv_flux_mapped",
              "testapi_arakawa_J_mapped_split", "poisson2d_solve", "launch_max_uv_from_psi_mapped"):
        assert n in names
    assert len(names) >= 60


def test_library_exports_every_declared_symbol():
    lib = nh.lib()  # loads without touching the GPU
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    missing = [n for n in declared_functions() if n not in nh.SIGNATURES]
    assert not missing, missing


def test_dynamic_symbol_table():
    out = subprocess.run(["nm", "-D", "--defined-only", nh.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert set(declared_functions()) <= exported


def test_executables_built():
    for p in (nh.EXE_PATH, nh.EXE_FREESLIP_PATH):
        assert os.access(p, os.X_OK), p
