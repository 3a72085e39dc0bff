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
import oracle

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


@pytest.mark.parametrize("free_slip", [False, True])
def test_resume_matches_oracle_resume(tmp_path, free_slip):
    """The HIP driver and the oracle driver resumed from the same snapshot (main.cu:129-147,
    init.cu:215-279): the same rows at the CSV's 9 digits and the same final state within 1e-10.
    Neither restores the step counter or dt (the reference restores neither)."""
    print(f"[restart] HIP resume vs oracle resume (free_slip={free_slip})", flush=True)
    extra = ["--walls", "freeslip"] if free_slip else []
    a = str(tmp_path / "whole")
    os.makedirs(a)
    assert nh.run_main(BASE + extra + ["--tend", "0.02", "--snap_save", "0.02", "--outdir", a]) == 0
    mid = os.path.join(a, "snaps", "snap_t0.020000.json")
    out = {}
    for name, run in (("hip", nh.run_main), ("oracle", oracle.run_main)):
        d = str(tmp_path / name)
        os.makedirs(d)
        assert run(BASE + extra + ["--tend", "0.04", "--snap_save", "0.02", "--resume", "--init", "load",
                                   "--load_path", mid, "--outdir", d]) == 0
        out[name] = d
    dh, do = harness.read_diag(out["hip"]), harness.read_diag(out["oracle"])
    assert len(dh) == len(do) >= 2
    np.testing.assert_array_equal(dh["step"], do["step"])
    assert dh["step"][0] == 10 and dh["t"][0] > 0.02  # no t = 0 row, step counter from 0
    bad = harness.compare_rows(do, dh, rel=5e-8)
    assert not bad, bad[:5]
    sh, so = harness.last_snap(out["hip"], 0.04), harness.last_snap(out["oracle"], 0.04)
    e_w = np.linalg.norm(sh.omega - so.omega) / np.linalg.norm(so.omega)
    e_p = np.linalg.norm(sh.psi - so.psi) / np.linalg.norm(so.psi)
    print(f"[restart] resumed HIP vs resumed oracle: omega {e_w:.2e} psi {e_p:.2e}", flush=True)
    assert e_w <= 1e-10 and e_p <= 1e-10
