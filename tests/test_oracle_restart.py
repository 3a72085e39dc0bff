"""Restart through a snapshot in the oracle driver (main.cu:129-147, 313-324; init.cu:215-279;
io.cu:173-178): N steps, a snapshot, and a resumed run of N more steps from it give the same state
and the same diagnostics rows as 2N steps in one run, bit for bit. The resumed run starts at the
snapshot's time, writes no t = 0 outputs, and (as in the reference) restores neither the step
counter nor dt: the step column restarts at 0 and dt at --dt."""
import math
import os

import numpy as np
import pytest

import harness
import oracle

BASE = ["--Nx", "64", "--Ny", "33", "--Lx", repr(2 * math.pi), "--h", "1", "--stretch", "tanh", "--beta", "2",
        "--Ub", "1", "--Reb", "2000", "--F0", "0.05", "--nforce", "4", "--drag", "1e-3", "--init", "lamband",
        "--amp", "0.05", "--band_kmax", "8", "--seed", "17", "--dt", "1e-3", "--no-adapt", "--diag_save", "0.01",
        "--statsEvery", "100000"]


@pytest.mark.parametrize("walls", [[], ["--walls", "freeslip"]], ids=["noslip", "freeslip"])
def test_oracle_resume_equals_continuous_run(tmp_path, walls):
    a, b = str(tmp_path / "whole"), str(tmp_path / "resumed")
    assert oracle.run_main(BASE + walls + ["--tend", "0.04", "--snap_save", "0.02", "--outdir", a]) == 0
    mid = os.path.join(a, "snaps", "snap_t0.020000")
    assert os.path.isfile(mid + ".bin")
    assert oracle.run_main(BASE + walls + ["--tend", "0.04", "--snap_save", "0.02", "--resume", "--init", "load",
                                           "--load_path", mid + ".json", "--outdir", b]) == 0
    sa, sb = harness.last_snap(a, 0.04), harness.last_snap(b, 0.04)
    assert np.array_equal(sa.omega, sb.omega) and np.array_equal(sa.psi, sb.psi)
    da, db = harness.read_diag(a), harness.read_diag(b)
    # no t = 0 row on resume; the rows after the snapshot are the whole run's rows
    assert db["t"][0] > 0.02 - 1e-12
    n = len(db)
    for col in ("t", "K", "Omega", "eta", "epsilon", "tauw_bot", "tauw_top"):
        np.testing.assert_array_equal(db[col], da[col][-n:])
    np.testing.assert_array_equal(db["step"], da["step"][-n:] - 20)
    # the resumed run wrote no snapshot at its start time
    assert not os.path.isfile(os.path.join(b, "snaps", "snap_t0.020000.bin"))


def test_oracle_resume_needs_a_snapshot(tmp_path):
    assert oracle.run_main(BASE + ["--tend", "0.01", "--resume", "--init", "load", "--load_path",
                                   str(tmp_path / "missing.json"), "--outdir", str(tmp_path / "o")]) != 0
