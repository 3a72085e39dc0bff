"""Ensembles of independent channels (nse2d_ensemble_*): every member advances exactly as the same
channel advanced alone (bitwise, same kernel shapes), graph replays equal direct launches, one
member against the oracle with the host dt rule, and the autotuner leaves the state untouched."""
import ctypes
import math
import os

import numpy as np
import pytest

import nse2d_hip as nh
import oracle
from gpu_util import Grid, rel_err, sync
from oracle import P
from test_gpu_step import OracleStepper

pytestmark = pytest.mark.gpu

# (Nx, Ny, Re_b, seed): two size classes of uneven membership
MEMBERS = [(512, 65, 2.0e3, 11), (512, 65, 5.0e3, 12), (512, 65, 1.0e4, 13), (1024, 33, 3.0e3, 14),
           (1024, 33, 8.0e3, 15), (2048, 129, 4.0e3, 16)]


def member_ic(Nx, Ny, seed):
    args = ["--Nx", str(Nx), "--Ny", str(Ny), "--Lx", repr(2 * math.pi), "--init", "lamband", "--amp", "0.05",
            "--band_kmax", "32", "--stretch", "tanh", "--beta", "2", "--Ub", "1", "--seed", str(seed)]
    return oracle.init_omega(args, Nx, Ny)


class Setup:
    def __init__(self, members=MEMBERS):
        self.grids = {}
        self.descs, self.ics, self.members = [], [], members
        for Nx, Ny, Reb, seed in members:
            if (Nx, Ny) not in self.grids:
                self.grids[(Nx, Ny)] = Grid(Nx, Ny, 2.0)
            g = self.grids[(Nx, Ny)]
            d = nh.ChannelDesc()
            d.Nx, d.Ny, d.Lx, d.h = Nx, Ny, g.Lx, 1.0
            d.nu, d.F0, d.lin_drag, d.nforce = 2.0 / Reb, 0.05, 1e-3, 12
            d.psi_bot, d.psi_top = 0.0, 2.0
            d.cfl, d.cvisc, d.dt_max = 0.5, 0.5, 1e-2
            d.dy_min = float(np.min(np.abs(g.m.dy_edge)))
            d.d_a_node, d.d_a_edge, d.d_y_node = g.dp("a_node"), g.dp("a_edge"), g.dp("y")
            d.d_dy_edge, d.d_w_node = g.dp("dy_edge"), g.dp("w_node")
            self.descs.append(d)
            self.ics.append(member_ic(Nx, Ny, seed))

    def ensemble(self, graph=True):
        arr = (nh.ChannelDesc * len(self.descs))(*self.descs)
        e = nh.lib().nse2d_ensemble_create(arr, len(self.descs), None)
        assert e, nh.last_error()
        for m, w in enumerate(self.ics):
            nh.upload_ptr(nh.lib().nse2d_ensemble_field(e, m, nh.FIELD_OMEGA), w)
        nh.lib().nse2d_ensemble_init(e)
        nh.lib().nse2d_ensemble_set_graph(e, 1 if graph else 0)
        return e

    def fields(self, e, m):
        Nx, Ny = self.members[m][:2]
        return (nh.download_ptr(nh.lib().nse2d_ensemble_field(e, m, nh.FIELD_OMEGA), Nx * Ny, (Ny, Nx)),
                nh.download_ptr(nh.lib().nse2d_ensemble_field(e, m, nh.FIELD_PSI), Nx * Ny, (Ny, Nx)))

    def scalars(self, e):
        return nh.download_ptr(nh.lib().nse2d_ensemble_scalars(e), 8 * len(self.members)).reshape(-1, 8)

    def single(self, m, steps, start=None):
        """Member m alone: from its IC through nse2d_channel_init, or from the given (omega, psi)."""
        d = self.descs[m]
        ch = nh.lib().nse2d_channel_create(ctypes.byref(d), None)
        assert ch, nh.last_error()
        nh.upload_ptr(nh.lib().nse2d_channel_field(ch, nh.FIELD_OMEGA), self.ics[m])
        nh.lib().nse2d_channel_init(ch)
        if start is not None:
            nh.upload_ptr(nh.lib().nse2d_channel_field(ch, nh.FIELD_OMEGA), start[0])
            nh.upload_ptr(nh.lib().nse2d_channel_field(ch, nh.FIELD_PSI), start[1])
        nh.lib().nse2d_channel_step_auto(ch, steps)
        sync()
        Nx, Ny = self.members[m][:2]
        w = nh.download_ptr(nh.lib().nse2d_channel_field(ch, nh.FIELD_OMEGA), Nx * Ny, (Ny, Nx))
        psi = nh.download_ptr(nh.lib().nse2d_channel_field(ch, nh.FIELD_PSI), Nx * Ny, (Ny, Nx))
        t = nh.download_ptr(nh.lib().nse2d_channel_time_ptr(ch), 1)[0]
        dt = nh.download_ptr(nh.lib().nse2d_channel_dt_ptr(ch), 1)[0]
        nh.lib().nse2d_channel_destroy(ch)
        return w, psi, t, dt


def test_members_equal_single_channels_bitwise():
    print("[ensemble] 6 members, 20 graph steps vs single channels", flush=True)
    S = Setup()
    e = S.ensemble()
    nh.lib().nse2d_ensemble_step_auto(e, 20)
    nh.lib().nse2d_ensemble_sync(e)
    sc = S.scalars(e)
    for m in range(len(S.members)):
        w_e, psi_e = S.fields(e, m)
        w_s, psi_s, t_s, dt_s = S.single(m, 20)
        assert np.array_equal(w_e, w_s) and np.array_equal(psi_e, psi_s), (m, np.max(np.abs(w_e - w_s)))
        assert sc[m, 0] == t_s and sc[m, 1] == dt_s, (m, sc[m, :2], t_s, dt_s)
    # members differ in Re_b, so their dt sequences differ
    assert len(set(np.round(sc[:, 1], 15))) > 1
    nh.lib().nse2d_ensemble_destroy(e)


def test_graph_replays_equal_direct_launches():
    """Graph replays interleaved with eager work (downloads, synchronisations) against the same
    ensemble launched directly, compared after every replay."""
    S = Setup(MEMBERS[:2] + MEMBERS[3:4])
    a, b = S.ensemble(graph=True), S.ensemble(graph=False)
    for k in range(4):
        nh.lib().nse2d_ensemble_step_auto(a, 1)
        nh.lib().nse2d_ensemble_step_auto(b, 1)
        nh.lib().nse2d_ensemble_sync(a)
        nh.lib().nse2d_ensemble_sync(b)
        for m in range(3):
            wa, pa = S.fields(a, m)
            wb, pb = S.fields(b, m)
            assert np.array_equal(wa, wb) and np.array_equal(pa, pb), (k, m)
        assert np.array_equal(S.scalars(a), S.scalars(b)), k
    nh.lib().nse2d_ensemble_destroy(a)
    nh.lib().nse2d_ensemble_destroy(b)


def test_member_against_oracle_with_host_dt():
    S = Setup()
    e = S.ensemble()
    m = 3
    Nx, Ny, Reb, _ = S.members[m]
    g = S.grids[(Nx, Ny)]
    nu, cfl, cvisc, dt_max = 2.0 / Reb, 0.5, 0.5, 1e-2
    o = OracleStepper(g, nu, 0.05, 12, 1e-3, 2.0)
    w_o = S.ics[m].copy()
    psi_o = o.solve(w_o)
    dy_min = float(np.min(np.abs(g.m.dy_edge)))
    t = 0.0
    for k in range(10):
        mx = np.zeros(3)
        oracle.lib().orc_max_uv_from_psi_mapped(P(psi_o), Nx, Ny, g.inv_dx, g.inv_deta, P(g.m.a_node), P(mx))
        dt_x = cfl * (g.dx / mx[0]) if mx[0] > 0 else 1e30
        dt_y = cfl * ((2.0 / (Ny - 1)) / mx[2]) if mx[2] > 0 else 1e30
        dt = min(dt_max, min(min(dt_x, dt_y), cvisc * 0.5 / (nu * (g.inv_dx2 + 1.0 / dy_min ** 2))))
        psi_o, w_o = o.step(psi_o, w_o, t, dt)
        t += dt
    nh.lib().nse2d_ensemble_step_auto(e, 10)
    nh.lib().nse2d_ensemble_sync(e)
    w_e, psi_e = S.fields(e, m)
    sc = S.scalars(e)
    ew, ep = rel_err(w_e, w_o), rel_err(psi_e, psi_o)
    print(f"[ensemble] member {m} ({Nx}x{Ny}) after 10 steps: omega {ew:.2e} psi {ep:.2e}", flush=True)
    assert abs(sc[m, 0] - t) <= 1e-14 and ew <= 1e-10 and ep <= 1e-10
    nh.lib().nse2d_ensemble_destroy(e)


def test_autotune_keeps_state_and_members_match_chosen_shapes(monkeypatch):
    S = Setup()
    e = S.ensemble()
    nh.lib().nse2d_ensemble_sync(e)
    before = [S.fields(e, m) for m in range(len(S.members))]
    sc0 = S.scalars(e)
    buf = ctypes.create_string_buffer(512)
    assert nh.lib().nse2d_ensemble_autotune(e, 2, buf, 512) == 0
    report = buf.value.decode()
    print("[ensemble] autotune:", report, flush=True)
    for m in range(len(S.members)):
        w, p = S.fields(e, m)
        assert np.array_equal(w, before[m][0]) and np.array_equal(p, before[m][1]), m
    assert np.array_equal(S.scalars(e), sc0)
    # the chosen shapes, applied to single channels through the tuning switches, reproduce the
    # members bit for bit
    parts = [x.strip() for x in report.split(";")]
    tile = parts[0].split()[1]
    cls = {}
    for part in parts[1:]:
        size, rest = part.split(":")
        Nx, Ny = (int(v) for v in size.split("x"))
        radix = int(rest.split("radix")[1].split(",")[0])
        kc = int(rest.split("KC")[1])
        cls[(Nx, Ny)] = (radix, kc)
    nh.lib().nse2d_ensemble_step_auto(e, 8)
    nh.lib().nse2d_ensemble_sync(e)
    monkeypatch.setenv("NSE2D_STAGE_TILE", tile)
    for m in (0, 3, 5):
        Nx, Ny = S.members[m][:2]
        radix, kc = cls[(Nx, Ny)]
        monkeypatch.setenv("NSE2D_FFT_RADIX", str(radix))
        monkeypatch.setenv("NSE2D_THOMAS_KC", str(kc))
        w_s, psi_s, _, _ = S.single(m, 8, start=before[m])  # the initial solve ran before tuning
        w_e, psi_e = S.fields(e, m)
        assert np.array_equal(w_e, w_s) and np.array_equal(psi_e, psi_s), m
    nh.lib().nse2d_ensemble_destroy(e)


def test_rejects_unsupported_members():
    S = Setup(MEMBERS[:1])
    d = S.descs[0]
    bad = nh.ChannelDesc.from_buffer_copy(d)
    bad.Nx = 256
    arr = (nh.ChannelDesc * 2)(d, bad)
    assert not nh.lib().nse2d_ensemble_create(arr, 2, None)
    assert "power of two" in nh.last_error()
    bad = nh.ChannelDesc.from_buffer_copy(d)
    bad.mms = 1
    arr = (nh.ChannelDesc * 1)(bad)
    assert not nh.lib().nse2d_ensemble_create(arr, 1, None)
    assert "manufactured" in nh.last_error()


@pytest.mark.parametrize("tile", ["p8x2x1", "p8x4x1"])
def test_persistent_stage_equals_one_tile_per_workgroup(tile, monkeypatch):
    """The persistent grouped stage kernel (a fixed grid walking every member's tiles; selected by
    NSE2D_STAGE_TILE=p... or by the autotuner) advances the ensemble exactly as the grouped kernel
    with one workgroup per tile."""
    print(f"[ensemble] {tile} vs w{tile[1:]}", flush=True)
    S = Setup()
    runs = []
    for t in (tile, "w" + tile[1:]):
        monkeypatch.setenv("NSE2D_STAGE_TILE", t)
        e = S.ensemble()
        nh.lib().nse2d_ensemble_step_auto(e, 10)
        nh.lib().nse2d_ensemble_sync(e)
        runs.append(([S.fields(e, m) for m in range(len(S.members))], S.scalars(e)))
        nh.lib().nse2d_ensemble_destroy(e)
    (fa, sa), (fb, sb) = runs
    for m in range(len(S.members)):
        assert np.array_equal(fa[m][0], fb[m][0]) and np.array_equal(fa[m][1], fb[m][1]), m
    assert np.array_equal(sa, sb)


def config4_members(seed=20241203):
    """The 64-member ensemble workload: 32 channels at 512x257, 24 at 1024x513 and 8 at
    2048x1025, Re_b log-uniform in [2e3, 1e4] (seeded), each with its own IC seed (the members of
    bench.py --workload ensemble)."""
    rng = np.random.default_rng(seed)
    out = []
    for Nx, Ny, count in ((512, 257, 32), (1024, 513, 24), (2048, 1025, 8)):
        for _ in range(count):
            out.append((Nx, Ny, float(np.exp(rng.uniform(np.log(2e3), np.log(1e4)))), 1000 + len(out)))
    return out


def test_config4_sixty_four_members(progress):
    """The full 64-member ensemble, three graph-replayed steps with autotuned shapes: every member
    equals its single-channel run under the chosen shapes bit for bit, and the first member of each
    size class follows the oracle stepping with the host dt rule (main.cu:451-458)."""
    members = config4_members()
    progress(f"[ensemble cfg4] {len(members)} members, 3 graph steps")
    S = Setup(members)
    e = S.ensemble()
    buf = ctypes.create_string_buffer(512)
    assert nh.lib().nse2d_ensemble_autotune(e, 1, buf, 512) == 0
    report = buf.value.decode()
    progress("[ensemble cfg4] autotune: " + report)
    nh.lib().nse2d_ensemble_sync(e)
    start = [S.fields(e, m) for m in range(len(members))]
    nh.lib().nse2d_ensemble_step_auto(e, 3)
    nh.lib().nse2d_ensemble_sync(e)
    sc = S.scalars(e)
    assert np.all(np.isfinite(sc[:, :2]))
    parts = [x.strip() for x in report.split(";")]
    tile = parts[0].split()[1]
    cls = {}
    for part in parts[1:]:
        size, rest = part.split(":")
        Nx, Ny = (int(v) for v in size.split("x"))
        cls[(Nx, Ny)] = (int(rest.split("radix")[1].split(",")[0]), int(rest.split("KC")[1]))
    env = {k: os.environ.get(k) for k in ("NSE2D_STAGE_TILE", "NSE2D_FFT_RADIX", "NSE2D_THOMAS_KC")}
    try:
        os.environ["NSE2D_STAGE_TILE"] = tile
        for m, (Nx, Ny, Reb, _) in enumerate(members):
            os.environ["NSE2D_FFT_RADIX"], os.environ["NSE2D_THOMAS_KC"] = (str(v) for v in cls[(Nx, Ny)])
            w_s, psi_s, t_s, dt_s = S.single(m, 3, start=start[m])
            w_e, psi_e = S.fields(e, m)
            assert np.array_equal(w_e, w_s) and np.array_equal(psi_e, psi_s), (m, np.max(np.abs(w_e - w_s)))
            assert sc[m, 1] == dt_s, (m, sc[m, 1], dt_s)
    finally:
        for k, v in env.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    progress("[ensemble cfg4] 64 members equal their single-channel runs bit for bit")
    for m in (0, 32, 56):
        Nx, Ny, Reb, _ = members[m]
        g = S.grids[(Nx, Ny)]
        nu, cfl, cvisc, dt_max = 2.0 / Reb, 0.5, 0.5, 1e-2
        dy_min = float(np.min(np.abs(g.m.dy_edge)))
        res = []
        for order in (0, 3):  # the oracle, and the oracle in the second FP64 order (the bar)
            oracle.lib().orc_set_eval_order(order)
            try:
                o = OracleStepper(g, nu, 0.05, 12, 1e-3, 2.0)
                w_o = S.ics[m].copy()
                psi_o = o.solve(w_o)
                t = 0.0
                for k in range(3):
                    mx = np.zeros(3)
                    oracle.lib().orc_max_uv_from_psi_mapped(P(psi_o), Nx, Ny, g.inv_dx, g.inv_deta, P(g.m.a_node),
                                                            P(mx))
                    dt_x = cfl * (g.dx / mx[0]) if mx[0] > 0 else 1e30
                    dt_y = cfl * ((2.0 / (Ny - 1)) / mx[2]) if mx[2] > 0 else 1e30
                    dt = min(dt_max, min(min(dt_x, dt_y), cvisc * 0.5 / (nu * (g.inv_dx2 + 1.0 / dy_min ** 2))))
                    psi_o, w_o = o.step(psi_o, w_o, t, dt)
                    t += dt
            finally:
                oracle.lib().orc_set_eval_order(0)
            res.append((w_o, psi_o, t))
        (w_o, psi_o, t), (w_q, psi_q, _) = res
        w_e, psi_e = S.fields(e, m)
        ew, ei, ep = rel_err(w_e, w_o), rel_err(w_e[1:-1], w_o[1:-1]), rel_err(psi_e, psi_o)
        fw, fi = rel_err(w_q, w_o), rel_err(w_q[1:-1], w_o[1:-1])
        progress(f"[ensemble cfg4] member {m} ({Nx}x{Ny}, Re_b {Reb:.0f}) vs oracle after 3 steps: omega {ew:.2e} "
                 f"(interior {ei:.2e}) psi {ep:.2e}; reordered oracle omega {fw:.2e} (interior {fi:.2e})")
        assert abs(sc[m, 0] - t) <= 1e-14 * t
        assert ep <= 1e-10
        assert ew <= max(1e-10, 10.0 * fw) and ei <= max(1e-10, 10.0 * fi)
    nh.lib().nse2d_ensemble_destroy(e)
