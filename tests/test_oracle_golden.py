"""Pins the oracle to the reference's own fixtures (tests/reference/output/*, copied to
tests/golden/reference): row 1 (t = 0) of all 21 cases, complete runs of the short cases, and
the early part of the channel case including the --closure_diag columns."""
import os

import numpy as np
import pytest

import harness
import oracle

ALL_CASES = [c[0] for c in harness.cases()]


def _run_oracle(name, tmp_path, extra=()):
    return harness.run_case(oracle.run_main, name, str(tmp_path), extra)


def test_case_list_matches_fixture_commands():
    """harness.cases() reproduces the command line stored with every fixture."""
    for name, walls, args in harness.cases():
        cmd = open(os.path.join(harness.GOLDEN, "output", name, "command.txt")).read().split()
        exe = "nse2d_freeslip" if walls == "freeslip" else "nse2d"
        assert cmd[0] == exe
        assert cmd[1:cmd.index("--outdir")] == args, name


@pytest.mark.parametrize("name", ALL_CASES)
def test_t0_row(name, tmp_path):
    """Grid metrics, IC, initial Poisson solve + Thom closure, CFL maxima, every budget
    reduction and the host formulas, against the fixture's first row (9 significant digits)."""
    d = _run_oracle(name, tmp_path, ["--max_steps", "0", "--snap_save", "0"])
    got = harness.read_diag(d)
    ref = harness.golden_diag(name)
    bad = harness.compare_rows(ref, got, rows=1, rel=5e-8)
    assert not bad, bad[:5]


SHORT_RUNS = ["eig_space_Ny33_b0", "eig_space_Ny33_b2", "eig_time_dt0.64", "eig_time_dt0.32",
              "eig_time_dt0.08", "mms_N64_b0", "mms_N64_b2", "cons_dt0.004", "poiseuille"]


@pytest.mark.parametrize("name", SHORT_RUNS)
def test_full_run_rows(name, tmp_path):
    """Every diagnostics row of a complete run: SSPRK stages, dt rule, output-time clipping,
    snapping, and the row formulas at later times."""
    d = _run_oracle(name, tmp_path, ["--snap_save", "0"] if name != "cons_dt0.004" else [])
    got, ref = harness.read_diag(d), harness.golden_diag(name)
    assert len(got) == len(ref)
    bad = harness.compare_rows(ref, got, rel=2e-7)
    # <omega rhs> near a steady state is a difference of O(1) terms; compare it on the scale of
    # the enstrophy budget instead of its own (tiny) value.
    scale = np.max(np.abs(ref["eta"])) + np.max(np.abs(ref["Omega"]))
    bad = [b for b in bad if not (b[1] == "dOmega_rhs" and abs(b[2] - b[3]) < 1e-8 * scale)]
    assert not bad, bad[:5]
    np.testing.assert_array_equal(got["step"], ref["step"])


@pytest.mark.parametrize("name", ["eig_space_Ny33_b0", "poiseuille", "cons_dt0.004"])
def test_full_run_profiles(name, tmp_path):
    """profiles_timeavg.csv: row moments (clamped walls), time weighting, u_tau and the writer."""
    d = _run_oracle(name, tmp_path, ["--snap_save", "0"])
    got = np.genfromtxt(os.path.join(d, "profiles_timeavg.csv"), delimiter=",", names=True)
    ref = harness.golden_profiles(name)
    vel2 = np.max(ref["U"] ** 2) + np.max(ref["V"] ** 2)
    for c in ref.dtype.names:
        # second moments are differences <u^2> - U^2: judge them on the U^2 scale
        scale = vel2 if c in ("uu", "vv", "uv") else np.max(np.abs(ref[c]))
        if scale < 1e-12:
            continue
        assert np.max(np.abs(got[c] - ref[c])) <= 1e-6 * scale, c


def test_channel_early_rows_with_closure_diag(tmp_path):
    """Channel case (mix IC, forcing, drag, adaptive dt, closure_diag) up to t = 0.5: all 32
    columns including the wall bookkeeping and the scheme-consistent budget terms."""
    d = _run_oracle("channel", tmp_path, ["--tend", "0.52", "--snap_save", "0"])
    got, ref = harness.read_diag(d), harness.golden_diag("channel")
    assert len(got) == 11
    bad = harness.compare_rows(ref, got, rows=11, rel=5e-7)
    assert not bad, bad[:5]


def test_golden_check_detects_planted_mistakes(tmp_path):
    """The t = 0 comparison is not vacuous: the oracle run with a wrong metric (beta 2 -> 2.001),
    a wrong wall closure (free-slip instead of Thom) or a wrong viscosity fails it."""
    ref = harness.golden_diag("poiseuille")
    walls, args = harness.case_args("poiseuille")
    plants = [
        [a if a != "2" else "2.001" for a in args],
        args + ["--walls", "freeslip"],
        [a if a != "500.0" else "500.5" for a in args],
    ]
    for k, a in enumerate(plants):
        d = str(tmp_path / f"plant{k}")
        assert oracle.run_main(a + ["--max_steps", "0", "--snap_save", "0", "--outdir", d]) == 0
        assert harness.compare_rows(ref, harness.read_diag(d), rows=1, rel=5e-8), k
