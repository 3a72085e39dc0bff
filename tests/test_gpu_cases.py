"""The reference's verification suite on the HIP path: all 21 cases of tests/run_tests.py run by
the HIP driver (nse2d_main, in this process), compared row by row with the reference fixtures and
checked by the six exact/identity tests plus reference.json (harness.py restates run_tests.py)."""
import os

import numpy as np
import pytest

import harness
import nse2d_hip as nh

pytestmark = pytest.mark.gpu
ALL = [c[0] for c in harness.cases()]


@pytest.fixture(scope="module")
def outroot(tmp_path_factory):
    root = str(tmp_path_factory.mktemp("gpu_cases"))
    nh.require_gpu()
    for name in ALL:
        print(f"[cases] running {name}", flush=True)
        harness.run_case(nh.run_main, name, root)
    return root


@pytest.mark.parametrize("name", ALL)
def test_t0_row_matches_fixture(outroot, name):
    got = harness.read_diag(os.path.join(outroot, name))
    bad = harness.compare_rows(harness.golden_diag(name), got, rows=1, rel=5e-8)
    assert not bad, bad[:5]


SHORT = ["eig_space_Ny33_b0", "eig_space_Ny33_b2", "eig_space_Ny65_b0", "eig_space_Ny129_b2", "eig_time_dt0.64",
         "eig_time_dt0.32", "eig_time_dt0.16", "eig_time_dt0.08", "eig_time_ref", "mms_N64_b0", "mms_N64_b2",
         "mms_N128_b2", "cons_dt0.004", "cons_dt0.002", "poiseuille"]


@pytest.mark.parametrize("name", SHORT)
def test_all_rows_match_fixture(outroot, name):
    got, ref = harness.read_diag(os.path.join(outroot, name)), harness.golden_diag(name)
    assert len(got) == len(ref)
    np.testing.assert_array_equal(got["step"], ref["step"])
    bad = harness.compare_rows(ref, got, rel=5e-7)
    scale = np.max(np.abs(ref["eta"])) + np.max(np.abs(ref["Omega"]))
    bad = [b for b in bad if not (b[1] in ("dOmega_rhs",) and abs(b[2] - b[3]) < 1e-8 * scale)]
    assert not bad, bad[:5]


def test_channel_rows_up_to_t2(outroot):
    got, ref = harness.read_diag(os.path.join(outroot, "channel")), harness.golden_diag("channel")
    assert len(got) == len(ref)
    rows = int(np.sum(ref["t"] <= 2.0 + 1e-9))
    bad = harness.compare_rows(ref, got, rows=rows, rel=1e-6)
    assert not bad, bad[:5]


@pytest.mark.parametrize("test", [t[0] for t in harness.TESTS])
def test_reference_checks(outroot, test):
    ok, lines, metrics = harness.check(test, outroot)
    print("\n".join(lines))
    assert ok, lines
