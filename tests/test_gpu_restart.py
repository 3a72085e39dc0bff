"""Snapshot and restart through the HIP driver (main.cu:129-147, init.cu:215-279, io.cu:203-259):
N fixed steps, a snapshot, and a resumed run of N more steps from it give the same state as 2N
steps in one run, bit for bit. The resumed run reads omega from the snapshot and recomputes psi
and the wall vorticity by the initial solve and closure (main.cu:211-213), which are the same
operations the stepper's last stage applied."""
import math
import os

import numpy as np
import pytest

import harness
import nse2d_hip as nh

pytestmark = pytest.mark.gpu

BASE = ["--Nx", "128", "--Ny", "65", "--Lx", repr(2 * math.pi), "--h", "1", "--stretch", "tanh", "--beta", "2",
        "--Ub", "1", "--Reb", "2000", "--F0", "0.05", "--nforce", "4", "--drag", "1e-3", "--init", "lamband",
        "--amp", "0.05", "--band_kmax", "8", "--seed", "17", "--dt", "1e-3", "--no-adapt", "--diag_save", "0.01",
        "--statsEvery", "100000"]


@pytest.mark.parametrize("free_slip", [False, True])
def test_resume_equals_continuous_run(tmp_path, free_slip):
    print(f"[restart] 128x65, 20 + 20 steps vs 40 (free_slip={free_slip})", flush=True)
    run = nh.run_main
    extra = ["--walls", "freeslip"] if free_slip else []
    a, b = str(tmp_path / "whole"), str(tmp_path / "resumed")
    os.makedirs(a)
    os.makedirs(b)
    assert run(BASE + extra + ["--tend", "0.04", "--snap_save", "0.02", "--outdir", a]) == 0
    mid = os.path.join(a, "snaps", "snap_t0.020000")
    assert os.path.isfile(mid + ".bin")
    assert run(BASE + extra + ["--tend", "0.04", "--snap_save", "0.02", "--resume", "--init", "load",
                               "--load_path", mid + ".json", "--outdir", b]) == 0
    # the driver sets its channel's wall closure in the channel's descriptor, not library-wide
    assert nh.lib().nse2d_get_wall_type() == 0
    sa, sb = harness.last_snap(a, 0.04), harness.last_snap(b, 0.04)
    da, db = harness.read_diag(a), harness.read_diag(b)
    print(f"[restart] max |omega| diff {np.max(np.abs(sa.omega - sb.omega)):.3e}, "
          f"max |psi| diff {np.max(np.abs(sa.psi - sb.psi)):.3e}", flush=True)
    assert np.array_equal(sa.omega, sb.omega)
    assert np.array_equal(sa.psi, sb.psi)
    # the resumed run's rows continue the whole run's rows at the same times and values
    assert int(db["step"][-1]) == int(da["step"][-1]) - 20
    np.testing.assert_array_equal(db["K"][-1], da["K"][-1])
