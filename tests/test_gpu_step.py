"""The fused SSPRK(3,3) step (nse2d_channel_step: stage kernels + Poisson + Thom) against the
oracle's unfused sequence of main.cu:482-538, for one step and after 1000 steps (bar 1e-10), and
the device-resident dt path against the host dt rule.

A step composes three stage updates, three Poisson solves and three Thom closures, each checked
on its own at 1e-13 (test_gpu_ops, test_gpu_poisson_diag). The closure multiplies the near-wall
streamfunction by A = 2(1/h1^2 + 1/(h1 h2)) ~ 1e4-1e5, so two FP64 solves that agree to rounding
give wall vorticities that differ by A times that rounding. The step comparison is therefore
split: psi and the interior rows of omega against the oracle, and the wall rows of omega against
the closure applied by the oracle to the GPU's own psi. The same conditioning makes a one-ulp
difference in a stage output grow to ~1e-12 in the interior after the next stage, so the tight
parity check is stagewise: each stage recomputed by the oracle from the GPU's previous stage."""
import ctypes
import math

import numpy as np
import pytest

import nse2d_hip as nh
import oracle
from gpu_util import Grid, rel_err, sync
from oracle import P

pytestmark = pytest.mark.gpu
L2PI = 2 * math.pi


class GpuPoisson:
    """The HIP Poisson solver (poisson2d_* C-ABI) as a host-array function."""

    def __init__(self, g):
        self.g = g
        self.p = nh.lib().poisson2d_create(g.Nx, g.Ny, g.Lx, None)
        assert self.p, nh.last_error()
        assert nh.lib().poisson2d_enable_mapped(self.p, g.dp("a_node"), g.dp("a_edge"), g.inv_deta, None) == 0
        self.dw, self.dpsi = nh.DeviceBuffer(g.Nx * g.Ny), nh.DeviceBuffer(g.Nx * g.Ny)

    def __call__(self, w, psi_top):
        self.dw.upload(w)
        nh.lib().poisson2d_solve(self.p, self.dw.ptr, self.dpsi.ptr, 0.0, 0.0, psi_top, None)
        return self.dpsi.download(w.shape)

    def close(self):
        nh.lib().poisson2d_destroy(self.p)


class OracleStepper:
    """The oracle's restatement of one SSPRK(3,3) step (main.cu:482-538). With `solver` given,
    the Poisson solves use it instead of the oracle's (composition checks)."""

    def __init__(self, g, nu, F0, nforce, drag, psi_top, free_slip=0, mms=None, solver=None):
        self.g, self.nu, self.F0, self.nforce, self.drag = g, nu, F0, nforce, drag
        self.psi_top, self.fs, self.mms, self.solver = psi_top, free_slip, mms, solver
        n = (g.Ny, g.Nx)
        self.J, self.lap, self.rhs = np.zeros(n), np.zeros(n), np.zeros(n)

    def solve(self, w):
        if self.solver is not None:
            psi = self.solver(w, self.psi_top)
        else:
            psi = oracle.poisson_solve(w, self.g.Lx, self.g.m, 0.0, self.psi_top)
        oracle.lib().orc_thom_wall_vorticity_mapped(P(w), P(psi), self.g.Nx, self.g.Ny, P(self.g.m.dy_edge), 0.0,
                                                    self.psi_top, self.fs)
        return psi

    def build_rhs(self, psi, w, t):
        g, L = self.g, oracle.lib()
        L.orc_arakawa_J_mapped(P(psi), P(w), P(self.J), g.Nx, g.Ny, g.inv_dx, g.inv_deta, P(g.m.a_node))
        L.orc_laplacian_fd_mapped(P(w), P(self.lap), g.Nx, g.Ny, g.inv_dx2, g.inv_deta2, P(g.m.a_node), P(g.m.a_edge))
        L.orc_rhs_mapped(P(self.J), P(self.lap), P(w), P(self.rhs), g.Nx, g.Ny, self.nu, self.F0, self.nforce, g.h,
                         self.drag, P(g.m.y))
        if self.mms:
            L.orc_mms_add_forcing(P(self.rhs), g.Nx, g.Ny, g.Lx, g.h, t, self.nu, self.drag, *self.mms, P(g.m.y))
        return self.rhs

    def step(self, psi, w, t, dt):
        N = w.size
        L = oracle.lib()
        w1, w2 = np.zeros_like(w), np.zeros_like(w)
        r = self.build_rhs(psi, w, t)
        L.orc_axpy(P(w1), P(w), P(r), dt, N)
        psi = self.solve(w1)
        r = self.build_rhs(psi, w1, t + dt)
        L.orc_ssprk2(P(w2), P(w), P(w1), P(r), dt, N)
        psi = self.solve(w2)
        r = self.build_rhs(psi, w2, t + 0.5 * dt)
        L.orc_ssprk3(P(w), P(w), P(w2), P(r), dt, N)
        psi = self.solve(w)
        return psi, w


def make_channel(g, nu, F0, nforce, drag, psi_top, mms=None, cfl=0.5, cvisc=0.5, dt_max=1e-2, wall_type=0):
    d = nh.ChannelDesc()
    d.Nx, d.Ny, d.Lx, d.h = g.Nx, g.Ny, g.Lx, g.h
    d.nu, d.F0, d.lin_drag, d.nforce = nu, F0, drag, nforce
    d.psi_bot, d.psi_top = 0.0, psi_top
    if mms:
        d.mms, (d.mms_A0, d.mms_eps, d.mms_Om, d.mms_kx) = 1, mms
    d.cfl, d.cvisc, d.dt_max = cfl, cvisc, dt_max
    d.dy_min = float(np.min(np.abs(g.m.dy_edge)))
    d.d_a_node, d.d_a_edge, d.d_y_node = g.dp("a_node"), g.dp("a_edge"), g.dp("y")
    d.d_dy_edge, d.d_w_node = g.dp("dy_edge"), g.dp("w_node")
    d.wall_type = wall_type
    ch = nh.lib().nse2d_channel_create(ctypes.byref(d), None)
    assert ch, nh.last_error()
    return ch


def field(ch, which, g):
    return nh.download_ptr(nh.lib().nse2d_channel_field(ch, which), g.Nx * g.Ny, (g.Ny, g.Nx))


def step_errors(g, w_gpu, psi_gpu, w_o, psi_o, psi_top, free_slip=0):
    """(psi, interior omega, wall omega vs closure of the GPU psi, wall omega vs oracle)."""
    w_cl = w_gpu.copy()
    oracle.lib().orc_thom_wall_vorticity_mapped(P(w_cl), P(psi_gpu), g.Nx, g.Ny, P(g.m.dy_edge), 0.0, psi_top,
                                                free_slip)
    return (rel_err(psi_gpu, psi_o), rel_err(w_gpu[1:-1], w_o[1:-1]),
            rel_err(w_gpu[[0, -1]], w_cl[[0, -1]]), rel_err(w_gpu[[0, -1]], w_o[[0, -1]]))


def channel_ic(Nx, Ny, Lx=L2PI):
    args = ["--Nx", str(Nx), "--Ny", str(Ny), "--Lx", repr(Lx), "--init", "lamband", "--amp", "0.05",
            "--band_kmax", str(min(32, Nx // 4)), "--stretch", "tanh", "--beta", "2", "--Ub", "1", "--seed", "7"]
    return oracle.init_omega(args, Nx, Ny)


@pytest.mark.parametrize("Nx,Ny,mms", [(64, 33, None), (256, 129, None), (128, 65, (1.0, 0.1, 1.0, 2))])
def test_one_step_matches_oracle(Nx, Ny, mms):
    print(f"[step] {Nx}x{Ny} mms={mms}", flush=True)
    g = Grid(Nx, Ny, 2.0)
    nu, F0, nf, drag = 1e-3, 0.05, 12, 1e-3
    w0 = channel_ic(Nx, Ny)
    o = OracleStepper(g, nu, F0, nf, drag, 2.0, mms=mms)
    w_o = w0.copy()
    psi_o = o.solve(w_o)
    ch = make_channel(g, nu, F0, nf, drag, 2.0, mms=mms)
    nh.upload_ptr(nh.lib().nse2d_channel_field(ch, nh.FIELD_OMEGA), w0)
    nh.lib().nse2d_channel_init(ch)
    sync()
    e0 = step_errors(g, field(ch, nh.FIELD_OMEGA, g), field(ch, nh.FIELD_PSI, g), w_o, psi_o, 2.0)
    assert e0[0] <= 1e-13 and e0[1] == 0.0 and e0[2] <= 1e-13
    t, dt = 0.3, 1e-3
    psi_o, w_o = o.step(psi_o, w_o, t, dt)
    nh.lib().nse2d_channel_step(ch, t, dt, nh.STEP_STORE_RHS)
    sync()
    e_psi, e_int, e_cl, e_wall = step_errors(g, field(ch, nh.FIELD_OMEGA, g), field(ch, nh.FIELD_PSI, g), w_o, psi_o, 2.0)
    e_rhs = rel_err(field(ch, nh.FIELD_RHS, g)[1:-1], o.rhs[1:-1])
    print(f"[step] {Nx}x{Ny} psi {e_psi:.2e} omega interior {e_int:.2e} rhs interior {e_rhs:.2e} "
          f"wall vs closure {e_cl:.2e} wall vs oracle {e_wall:.2e}", flush=True)
    assert e_psi <= 2e-13 and e_cl <= 1e-13
    assert e_int <= 5e-12 and e_rhs <= 5e-10 and e_wall <= 1e-10  # closure conditioning, see module doc
    nh.lib().nse2d_channel_destroy(ch)


def stagewise_errors(Nx, Ny, mms, beta=2.0, nu=1e-3, w0=None, free_slip=0):
    """One fused step checked stage by stage: the oracle computes each stage's update from the
    GPU's own previous stage (omega_s from the W1/W2 fields, psi_s = the HIP solve of it), so no
    difference is carried into the Thom closure. Returns relative errors of the interior rows of
    every stage output, of its wall rows against the closure of its own psi, of the final psi and
    of the stored stage-3 RHS."""
    g = Grid(Nx, Ny, beta)
    F0, nf, drag = 0.05, 12, 1e-3
    if w0 is None:
        w0 = channel_ic(Nx, Ny)
    solver = GpuPoisson(g)
    o = OracleStepper(g, nu, F0, nf, drag, 2.0, mms=mms, solver=solver)
    o.fs = free_slip
    ch = make_channel(g, nu, F0, nf, drag, 2.0, mms=mms, wall_type=2 if free_slip else 1)
    nh.upload_ptr(nh.lib().nse2d_channel_field(ch, nh.FIELD_OMEGA), w0)
    nh.lib().nse2d_channel_init(ch)
    t, dt = 0.3, 1e-3
    nh.lib().nse2d_channel_step(ch, t, dt, 0)  # move off the initial condition
    t += dt
    sync()
    w_n, psi_n = field(ch, nh.FIELD_OMEGA, g), field(ch, nh.FIELD_PSI, g)
    nh.lib().nse2d_channel_step(ch, t, dt, nh.STEP_STORE_RHS)
    sync()
    W1, W2, W3 = (field(ch, k, g) for k in (nh.FIELD_W1, nh.FIELD_W2, nh.FIELD_OMEGA))
    psi_g, rhs_g = field(ch, nh.FIELD_PSI, g), field(ch, nh.FIELD_RHS, g)
    nh.lib().nse2d_channel_destroy(ch)
    L, N = oracle.lib(), w_n.size
    errs = {}

    def walls(name, w_gpu):
        psi_s = solver(w_gpu, 2.0)
        w_cl = w_gpu.copy()
        L.orc_thom_wall_vorticity_mapped(P(w_cl), P(psi_s), g.Nx, g.Ny, P(g.m.dy_edge), 0.0, 2.0, free_slip)
        errs[name + " walls"] = rel_err(w_gpu[[0, -1]], w_cl[[0, -1]])
        return psi_s

    ref = np.zeros_like(w_n)
    r = o.build_rhs(psi_n, w_n, t)
    L.orc_axpy(P(ref), P(w_n), P(r), dt, N)
    errs["stage1 interior"] = rel_err(W1[1:-1], ref[1:-1])
    psi1 = walls("stage1", W1)
    r = o.build_rhs(psi1, W1, t + dt)
    L.orc_ssprk2(P(ref), P(w_n), P(W1), P(r), dt, N)
    errs["stage2 interior"] = rel_err(W2[1:-1], ref[1:-1])
    psi2 = walls("stage2", W2)
    r = o.build_rhs(psi2, W2, t + 0.5 * dt)
    errs["rhs"] = rel_err(rhs_g, r)
    L.orc_ssprk3(P(ref), P(w_n), P(W2), P(r), dt, N)
    errs["stage3 interior"] = rel_err(W3[1:-1], ref[1:-1])
    psi3 = walls("stage3", W3)
    errs["psi"] = rel_err(psi_g, psi3)
    solver.close()
    return errs


def _fmt(errs):
    return " ".join(f"{k} {v:.2e}" for k, v in errs.items())


def _check_stagewise(errs):
    # interior updates and closures from identical inputs: rounding only. The stored RHS sums
    # -J and nu lap(omega), which on the first interior rows are both far larger than the sum.
    for k, v in errs.items():
        assert v <= (1e-12 if k == "rhs" else 1e-14), (k, v)


@pytest.mark.parametrize("Nx,Ny,mms,fs", [(64, 33, None, 0), (256, 129, None, 0), (128, 65, (1.0, 0.1, 1.0, 2), 0),
                                          (512, 65, None, 0), (1024, 33, None, 0), (1024, 33, None, 1),
                                          (2048, 65, None, 0), (4096, 257, None, 0), (4096, 4, None, 0)])
def test_step_stagewise(Nx, Ny, mms, fs):
    """The fused step against the oracle's stage sequence (J, Laplacian, RHS, MMS, SSPRK
    combinations, Thom; main.cu:482-538), one stage at a time. Nx >= 512 exercises the wall
    closure fused into the inverse transform (no-slip and free-slip; Ny = 4, where rows 1 and
    Ny-2 are each other's partner rows); smaller rows use the separate closure kernel."""
    errs = stagewise_errors(Nx, Ny, mms, free_slip=fs)
    print(f"[stagewise] {Nx}x{Ny}: " + _fmt(errs), flush=True)
    _check_stagewise(errs)


def test_wall_type_is_per_channel():
    """A no-slip and a free-slip channel (descriptor wall_type 1 and 2) stepped in turn, with the
    library-wide default set to the other closure in between: each follows the oracle step with its
    own closure (ops.cu:108-156), so the choice lives in the channel, not in process state."""
    Nx, Ny = 512, 33
    g = Grid(Nx, Ny, 2.0)
    nu, F0, nf, drag, dt = 1e-3, 0.05, 12, 1e-3, 1e-3
    w0 = channel_ic(Nx, Ny)
    chans, orcs = [], []
    for fs in (0, 1):
        ch = make_channel(g, nu, F0, nf, drag, 2.0, wall_type=2 if fs else 1)
        nh.upload_ptr(nh.lib().nse2d_channel_field(ch, nh.FIELD_OMEGA), w0)
        nh.lib().nse2d_set_wall_type(1 - fs)
        nh.lib().nse2d_channel_init(ch)
        o = OracleStepper(g, nu, F0, nf, drag, 2.0, free_slip=fs)
        w_o = w0.copy()
        chans.append(ch)
        orcs.append([o, o.solve(w_o), w_o])
    try:
        for k in range(3):
            for fs in (0, 1):
                nh.lib().nse2d_set_wall_type(1 - fs)
                nh.lib().nse2d_channel_step(chans[fs], k * dt, dt, 0)
                o, psi_o, w_o = orcs[fs]
                orcs[fs][1], orcs[fs][2] = o.step(psi_o, w_o, k * dt, dt)
        sync()
        for fs in (0, 1):
            w_g, psi_g = field(chans[fs], nh.FIELD_OMEGA, g), field(chans[fs], nh.FIELD_PSI, g)
            e_psi, e_int, e_cl, _ = step_errors(g, w_g, psi_g, orcs[fs][2], orcs[fs][1], 2.0, free_slip=fs)
            print(f"[wall type per channel] free_slip={fs}: psi {e_psi:.2e}, interior {e_int:.2e}, "
                  f"walls vs own closure {e_cl:.2e}", flush=True)
            assert e_psi <= 1e-12 and e_int <= 1e-11 and e_cl <= 1e-13
            if fs:
                assert np.all(w_g[[0, -1]] == 0.0)
            else:
                assert np.min(np.abs(w_g[[0, -1]])) > 0
    finally:
        nh.lib().nse2d_set_wall_type(0)
        for ch in chans:
            nh.lib().nse2d_channel_destroy(ch)


def test_invalid_wall_type_is_rejected():
    g = Grid(64, 17, 2.0)
    d = nh.ChannelDesc()
    d.Nx, d.Ny, d.Lx, d.h, d.nu = g.Nx, g.Ny, g.Lx, g.h, 1e-3
    d.d_a_node, d.d_a_edge, d.d_y_node = g.dp("a_node"), g.dp("a_edge"), g.dp("y")
    d.d_dy_edge, d.d_w_node = g.dp("dy_edge"), g.dp("w_node")
    d.wall_type = 3
    assert not nh.lib().nse2d_channel_create(ctypes.byref(d), None)
    assert "wall_type" in nh.last_error()


STAGE_TILES = (["0"] + [f"{ry}x{sy}x{pf}" for ry, sy in ((8, 4), (16, 4), (16, 2), (32, 1)) for pf in (1, 2, 3)]
               + ["w8x4x1", "w8x4x2", "w4x4x1", "w16x2x1", "w8x2x1", "w4x1x1", "w4x2x1", "w4x1x2"])  # two columns per lane (Nx % 128 == 0)


@pytest.mark.parametrize("tile", STAGE_TILES)
@pytest.mark.parametrize("Nx,Ny,mms", [(256, 77, None), (64, 21, None)])
def test_stage_tile_shapes(tile, Nx, Ny, mms, monkeypatch):
    """Every stage-kernel shape (NSE2D_STAGE_TILE=RYxSYWxPF at channel creation; 0 = the generic
    kernel) on grids whose rows do not fill the last tile, stage by stage. The manufactured
    forcing always runs in the generic kernel (covered by test_step_stagewise)."""
    monkeypatch.setenv("NSE2D_STAGE_TILE", tile)
    errs = stagewise_errors(Nx, Ny, mms)
    print(f"[tile {tile}] {Nx}x{Ny}: " + _fmt(errs), flush=True)
    _check_stagewise(errs)


def test_thousand_steps_within_1e10():
    """1000 steps on a 256x129 stretched channel (forcing, drag, noise IC): GPU and oracle stay
    within 1e-10 normwise."""
    Nx, Ny = 256, 129
    g = Grid(Nx, Ny, 2.0)
    nu, F0, nf, drag, dt = 1e-4, 0.05, 12, 1e-3, 2e-3
    w0 = channel_ic(Nx, Ny)
    o = OracleStepper(g, nu, F0, nf, drag, 2.0)
    w_o = w0.copy()
    psi_o = o.solve(w_o)
    ch = make_channel(g, nu, F0, nf, drag, 2.0)
    nh.upload_ptr(nh.lib().nse2d_channel_field(ch, nh.FIELD_OMEGA), w0)
    nh.lib().nse2d_channel_init(ch)
    t = 0.0
    for k in range(1000):
        psi_o, w_o = o.step(psi_o, w_o, t, dt)
        nh.lib().nse2d_channel_step(ch, t, dt, 0)
        t += dt
        if k % 250 == 249:
            print(f"[1000 steps] step {k + 1}", flush=True)
    sync()
    w_g, psi_g = field(ch, nh.FIELD_OMEGA, g), field(ch, nh.FIELD_PSI, g)
    e_w, e_p = rel_err(w_g, w_o), rel_err(psi_g, psi_o)
    e_int = rel_err(w_g[1:-1], w_o[1:-1])
    print(f"[1000 steps] omega {e_w:.2e} (interior rows {e_int:.2e}) psi {e_p:.2e}", flush=True)
    assert e_w <= 1e-10 and e_p <= 1e-10
    nh.lib().nse2d_channel_destroy(ch)


def test_device_dt_matches_host_rule():
    """nse2d_channel_step_auto chooses dt on the device with main.cu:451-458's rule; the host rule
    on the oracle's maxima gives the same dt sequence and fields."""
    Nx, Ny = 128, 65
    g = Grid(Nx, Ny, 2.0)
    nu, F0, nf, drag = 2e-4, 0.05, 12, 1e-3
    cfl, cvisc, dt_max = 0.5, 0.5, 1e-2
    w0 = channel_ic(Nx, Ny)
    o = OracleStepper(g, nu, F0, nf, drag, 2.0)
    w_o = w0.copy()
    psi_o = o.solve(w_o)
    ch = make_channel(g, nu, F0, nf, drag, 2.0, cfl=cfl, cvisc=cvisc, dt_max=dt_max)
    nh.upload_ptr(nh.lib().nse2d_channel_field(ch, nh.FIELD_OMEGA), w0)
    nh.lib().nse2d_channel_init(ch)
    dy_min = float(np.min(np.abs(g.m.dy_edge)))
    deta = 2.0 / (Ny - 1)
    t = 0.0
    for k in range(20):
        m = np.zeros(3)
        oracle.lib().orc_max_uv_from_psi_mapped(P(psi_o), Nx, Ny, g.inv_dx, g.inv_deta, P(g.m.a_node), P(m))
        dt_x = cfl * (g.dx / m[0]) if m[0] > 0 else 1e30
        dt_y = cfl * (deta / m[2]) if m[2] > 0 else 1e30
        dt_vis = cvisc * 0.5 / (nu * (g.inv_dx2 + 1.0 / dy_min ** 2))
        dt = min(dt_max, min(min(dt_x, dt_y), dt_vis))
        psi_o, w_o = o.step(psi_o, w_o, t, dt)
        t += dt
        nh.lib().nse2d_channel_step_auto(ch, 1)
        sync()
        dt_dev = nh.download_ptr(nh.lib().nse2d_channel_dt_ptr(ch), 1)[0]
        assert abs(dt_dev - dt) <= 1e-15 * dt, (k, dt_dev, dt)
    t_dev = nh.download_ptr(nh.lib().nse2d_channel_time_ptr(ch), 1)[0]
    assert abs(t_dev - t) <= 1e-14
    assert rel_err(field(ch, nh.FIELD_OMEGA, g), w_o) <= 1e-10
    nh.lib().nse2d_channel_destroy(ch)


def test_kernel_class_timing_and_sampling():
    """nse2d_channel_profile brackets every launch of the chosen kernel classes with events;
    nse2d_channel_profile_every(k) times only steps 0, k, 2k, ... after the reset. Timing leaves the
    fields bit for bit as an untimed run makes them."""
    Nx, Ny = 256, 97
    g = Grid(Nx, Ny, 2.0)
    w0 = channel_ic(Nx, Ny)
    chans = []
    for _ in range(2):
        ch = make_channel(g, 2e-4, 0.05, 12, 1e-3, 2.0)
        nh.upload_ptr(nh.lib().nse2d_channel_field(ch, nh.FIELD_OMEGA), w0)
        nh.lib().nse2d_channel_init(ch)
        chans.append(ch)
    timed, plain = chans
    ncls = len(nh.PROF_CLASSES)
    L = nh.lib()

    def read(ch):
        ms, cnt = (ctypes.c_double * ncls)(), (ctypes.c_longlong * ncls)()
        L.nse2d_channel_profile_read(ch, ms, cnt)
        return np.array(ms[:]), np.array(cnt[:], dtype=np.int64)

    L.nse2d_channel_profile(timed, (1 << ncls) - 1)
    L.nse2d_channel_step_auto(timed, 4)
    ms, n = read(timed)
    per_step = n // 4
    print(f"[profile] launches per step {dict(zip(nh.PROF_CLASSES, per_step.tolist()))}", flush=True)
    assert np.array_equal(per_step * 4, n)
    for cls in ("stage", "fft_forward", "thomas", "fft_inverse"):
        k = nh.PROF_CLASSES.index(cls)
        assert per_step[k] == 3 and ms[k] > 0.0
    assert per_step[nh.PROF_CLASSES.index("cfl")] == 1
    L.nse2d_channel_profile(timed, (1 << ncls) - 1)
    L.nse2d_channel_profile_every(timed, 3)
    L.nse2d_channel_step_auto(timed, 7)  # steps 0, 3 and 6 of this segment are timed
    ms, n = read(timed)
    assert np.array_equal(n, per_step * 3), (n, per_step)
    L.nse2d_channel_profile(timed, 0)
    L.nse2d_channel_profile_every(timed, 1)
    L.nse2d_channel_step_auto(plain, 11)
    sync()
    assert np.array_equal(field(timed, nh.FIELD_OMEGA, g), field(plain, nh.FIELD_OMEGA, g))
    for ch in chans:
        L.nse2d_channel_destroy(ch)


@pytest.mark.parametrize("Nx,Ny", [(512, 257), (1024, 129), (2048, 97), (4096, 65), (8192, 35), (16384, 19)])
def test_fused_stage_transform_equals_separate_kernels(Nx, Ny, monkeypatch):
    """With NSE2D_STAGE_FFT=1 (at creation) production stages run as one kernel with the forward
    row transform where the row-FFT plan allows it; by default the stage kernel and the transform
    run separately. Both give the same fields, dt and time bit for bit over adaptive steps, on
    every instantiated plan (M = Nx/2 from 256 to 8192)."""
    print(f"[stage+fft] {Nx}x{Ny}", flush=True)
    g = Grid(Nx, Ny, 2.0)
    w0 = channel_ic(Nx, Ny)
    chans = []
    for flag in ("0", "1"):
        monkeypatch.setenv("NSE2D_STAGE_FFT", flag)
        ch = make_channel(g, 2e-4, 0.05, 12, 1e-3, 2.0)
        nh.upload_ptr(nh.lib().nse2d_channel_field(ch, nh.FIELD_OMEGA), w0)
        nh.lib().nse2d_channel_init(ch)
        chans.append(ch)
    try:
        for ch in chans:
            nh.lib().nse2d_channel_step_auto(ch, 6)
        sync()
        (a, pa), (b, pb) = ((field(ch, nh.FIELD_OMEGA, g), field(ch, nh.FIELD_PSI, g)) for ch in chans)
        ta, tb = (nh.download_ptr(nh.lib().nse2d_channel_time_ptr(ch), 2) for ch in chans)
        print(f"[stage+fft] {Nx}x{Ny}: max |omega| diff {np.max(np.abs(a - b)):.3e}, "
              f"max |psi| diff {np.max(np.abs(pa - pb)):.3e}", flush=True)
        assert np.all(np.isfinite(a))
        assert np.array_equal(a, b) and np.array_equal(pa, pb) and np.array_equal(ta, tb)
    finally:
        for ch in chans:
            nh.lib().nse2d_channel_destroy(ch)
