"""Code-generation checks on the gfx950 code objects built from this repository's HIP sources.

A device body that the compiler leaves as a real call (s_swappc) reaches LDS through generic
pointers and loses most of its scheduling; a kernel that spills moves registers through scratch.
Both have cost the hot kernels 20-60 % of their time without changing a result, so the parity tests
cannot catch them. These checks read the objects of the in-tree build (nse2d-channel-gpu_amd/
_build) with the ROCm LLVM tools; they are skipped when either is missing.
"""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "nse2d-channel-gpu_amd", "_build")
LLVM = "/opt/rocm/llvm/bin"
OBJECTS = ["channel", "poisson", "diag", "ops", "closure"]


def _tools():
    need = [os.path.join(LLVM, t) for t in ("llvm-objcopy", "clang-offload-bundler", "llvm-objdump", "llvm-readelf")]
    return need if all(os.path.isfile(t) for t in need) else None


def _device_object(name, tmp_path):
    obj = os.path.join(BUILD, name + ".o")
    if not os.path.isfile(obj):
        pytest.skip(f"{obj} not built")
    tools = _tools()
    if tools is None:
        pytest.skip("ROCm LLVM tools not found")
    objcopy, bundler, _, _ = tools
    bundle, co = tmp_path / (name + ".bundle"), tmp_path / (name + ".co")
    subprocess.check_call([objcopy, f"--dump-section=.hip_fatbin={bundle}", obj])
    subprocess.check_call([bundler, "--type=o", "--unbundle", "--targets=hipv4-amdgcn-amd-amdhsa--gfx950",
                           f"--input={bundle}", f"--output={co}"])
    return str(co)


def _kernel_notes(co):
    txt = subprocess.run([os.path.join(LLVM, "llvm-readelf"), "--notes", co], capture_output=True, text=True,
                         check=True).stdout
    out, cur = {}, None
    for line in txt.splitlines():
        m = re.search(r"\.name:\s+(\S+)", line)
        if m:
            cur = m.group(1)
            out[cur] = {}
        m = re.search(r"\.(vgpr_spill_count|sgpr_spill_count|vgpr_count|private_segment_fixed_size):\s+(\d+)", line)
        if m and cur:
            out[cur][m.group(1)] = int(m.group(2))
    return out


@pytest.mark.parametrize("name", OBJECTS)
def test_kernels_make_no_device_calls(name, tmp_path):
    co = _device_object(name, tmp_path)
    dis = subprocess.run([os.path.join(LLVM, "llvm-objdump"), "-d", co], capture_output=True, text=True,
                         check=True).stdout
    calls = [ln for ln in dis.splitlines() if "s_swappc" in ln]
    assert not calls, f"{name}: {len(calls)} device function calls left in the code object"


@pytest.mark.parametrize("name", OBJECTS)
def test_kernels_do_not_spill(name, tmp_path):
    """No kernel moves registers through scratch memory. Scalar registers spilled into lanes of a
    vector register (sgpr_spill_count with no private segment) stay on chip and are allowed: the
    two-column stage kernels keep the row factors of their rows in scalar registers that way."""
    notes = _kernel_notes(_device_object(name, tmp_path))
    assert notes, "no kernel metadata found"
    bad = {k: v for k, v in notes.items()
           if v.get("vgpr_spill_count", 0) or v.get("private_segment_fixed_size", 0)}
    assert not bad, f"{name}: kernels that spill to scratch: {sorted(bad)[:5]}"

