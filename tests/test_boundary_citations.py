"""The drop-in boundary's citations: every launcher and test wrapper declared in
include/nse2d_hip.h keeps a reference function's name, and the file:line range cited in the
comment above its declaration must contain that name in the reference source. Runs where the
reference checkout is present (the build container); skipped elsewhere."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/src"


def _cited_declarations():
    src = open(os.path.join(ROOT, "include", "nse2d_hip.h")).read()
    pat = r"/\*((?:(?!\*/).)*?)\*/\s*\n\s*(?:int|void)\s+(\w+)\s*\("
    return [(name, re.findall(r"(\w+\.cuh?):(\d+)(?:-(\d+))?", com)) for com, name in re.findall(pat, src, re.S)]


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference sources not present")
def test_same_name_entry_points_cite_their_declaration():
    checked = 0
    for name, cites in _cited_declarations():
        if not (name.startswith("launch_") or name.startswith("testapi_")):
            continue
        assert cites, name
        found = False
        for f, a, b in cites:
            lines = open(os.path.join(REF, f)).read().split("\n")
            seg = "\n".join(lines[int(a) - 1:int(b or a)])
            found |= name in seg
        assert found, (name, cites)
        checked += 1
    assert checked >= 25
