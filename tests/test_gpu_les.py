"""The learned-closure LES path (csrc/les.hip, nse2d_les_* and nse2d_channel_set_closure) on the GPU:

- the filter F (box in mapped y, sharp cut in x) against the numpy oracle (oracle/les_oracle.py),
  1e-13 normwise;
- the training pair (coarse state with Thom walls, closure target F(R_f) - R_c) against the same
  composition of the FP64 oracle's solve, closure, Jacobian, Laplacian and RHS, and the identity
  target = 0 when the coarse grid is the fine grid;
- the stage hook: one SSPRK(3,3) step with a fixed additive source against the oracle's step
  with the same source (the step bars of test_gpu_step);
- the closure term against the CNN oracle (bf16-emulating) applied to the normalised inputs;
- training on filtered DNS: the trained network's hold-out error against a zero closure (the gain
  of the learned part a priori), and coarse runs with and without it against the filtered DNS
  (the gain a posteriori).
"""
import ctypes
import math
import os
import sys

import numpy as np
import pytest

import nse2d_hip as nh
import oracle
from gpu_util import Grid, rel_err, smooth_random_field, sync
from oracle import P
from test_gpu_step import OracleStepper, field, make_channel, step_errors

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import cnn_oracle as co  # noqa: E402
import les_oracle as lo  # noqa: E402

import les_experiment as lx  # noqa: E402

pytestmark = pytest.mark.gpu
L2PI = 2 * math.pi


def cfg(Nx, Ny, Reb=5000.0, **kw):
    c = dict(Nx=Nx, Ny=Ny, Lx=L2PI, h=1.0, beta=2.0, Ub=1.0, Reb=Reb, F0=0.05, nforce=12, drag=1e-3, amp=0.05,
             band_kmax=8, seed=11, cfl=0.5, cvisc=0.5, dt_max=1e-2)
    c.update(kw)
    return c


def oracle_rhs(g, psi, w, nu, F0, nforce, drag):
    L = oracle.lib()
    J, lap, r = np.zeros_like(w), np.zeros_like(w), np.zeros_like(w)
    L.orc_arakawa_J_mapped(P(psi), P(w), P(J), g.Nx, g.Ny, g.inv_dx, g.inv_deta, P(g.m.a_node))
    L.orc_laplacian_fd_mapped(P(w), P(lap), g.Nx, g.Ny, g.inv_dx2, g.inv_deta2, P(g.m.a_node), P(g.m.a_edge))
    L.orc_rhs_mapped(P(J), P(lap), P(w), P(r), g.Nx, g.Ny, nu, F0, nforce, g.h, drag, P(g.m.y))
    return r


def fine_state(g, psi_top, seed):
    """A fine state as the solver holds it: psi = solve(omega), omega's wall rows by Thom."""
    w = smooth_random_field(g.Nx, g.Ny, seed, kmax=6, y=g.m.y) + 3.0 * g.m.y[:, None]
    psi = oracle.poisson_solve(w, g.Lx, g.m, 0.0, psi_top)
    oracle.lib().orc_thom_wall_vorticity_mapped(P(w), P(psi), g.Nx, g.Ny, P(g.m.dy_edge), 0.0, psi_top, 0)
    return w, psi


@pytest.mark.parametrize("fine,coarse", [((256, 129), (64, 33)), ((512, 257), (64, 33)), ((128, 65), (64, 65))])
def test_filter_matches_oracle(fine, coarse):
    print(f"[les] filter {fine} -> {coarse}", flush=True)
    F = lx.Channel(cfg(*fine), ic=False)
    C = lx.Channel(cfg(*coarse), ic=False)
    les = lx.Les(F, C, 1, 1)
    rx, ry = fine[0] // coarse[0], (fine[1] - 1) // (coarse[1] - 1)
    f = smooth_random_field(fine[0], fine[1], 5, kmax=12) + np.random.default_rng(1).normal(size=(fine[1], fine[0]))
    d_f, d_c = nh.DeviceBuffer.from_array(f), nh.DeviceBuffer(coarse[0] * coarse[1])
    assert nh.lib().nse2d_les_filter(les.h, d_f.ptr, d_c.ptr) == 0
    got = d_c.download((coarse[1], coarse[0]))
    ref = lo.filter_field(f, rx, ry)
    e = rel_err(got, ref)
    print(f"[les] filter rel err {e:.2e}", flush=True)
    les.destroy()
    F.destroy()
    C.destroy()
    assert e <= 1e-13


@pytest.mark.parametrize("fine,coarse", [((256, 129), (64, 33)), ((512, 257), (64, 33)),
                                         ((4096, 2049), (512, 257))],
                         ids=["256x129-64x33", "512x257-64x33", "4096x2049-512x257"])
def test_pair_matches_oracle(fine, coarse):
    """The training pair (coarse state with Thom walls, closure target F(R_f) - R_c) against the FP64
    oracle's composition, up to the learned-closure workload (4096x2049 DNS filtered 8x8 to the
    512x257 LES grid)."""
    print(f"[les] pair {fine} -> {coarse}", flush=True)
    cf, cc = cfg(*fine), cfg(*coarse)
    F = lx.Channel(cf, ic=False)
    C = lx.Channel(cc, ic=False)
    les = lx.Les(F, C, 2, 1)
    gf, gc = Grid(fine[0], fine[1], 2.0), Grid(coarse[0], coarse[1], 2.0)
    psi_top = 2.0
    wf, psif = fine_state(gf, psi_top, 9)
    nh.upload_ptr(F.field(nh.FIELD_OMEGA), wf)
    nh.upload_ptr(F.field(nh.FIELD_PSI), psif)
    les.pair(append=True)
    assert les.num_pairs() == 1
    wc, psic, tgt = (b.download((coarse[1], coarse[0])) for b in (les.wc, les.psic, les.tgt))
    # the oracle's composition
    rx, ry = fine[0] // coarse[0], (fine[1] - 1) // (coarse[1] - 1)
    nu = 2.0 / cf["Reb"]
    rf = oracle_rhs(gf, psif, wf, nu, 0.05, 12, 1e-3)
    wc_o = np.ascontiguousarray(lo.filter_field(wf, rx, ry))
    psic_o = oracle.poisson_solve(wc_o, gc.Lx, gc.m, 0.0, psi_top)
    oracle.lib().orc_thom_wall_vorticity_mapped(P(wc_o), P(psic_o), gc.Nx, gc.Ny, P(gc.m.dy_edge), 0.0, psi_top, 0)
    rc = oracle_rhs(gc, psic_o, wc_o, nu, 0.05, 12, 1e-3)
    frf = lo.filter_field(rf, rx, ry)
    tgt_o = np.zeros_like(frf)
    tgt_o[1:-1] = frf[1:-1] - rc[1:-1]
    e_w, e_psi = rel_err(wc[1:-1], wc_o[1:-1]), rel_err(psic, psic_o)
    # the target is a difference of two RHS: its bar is relative to the larger of them (with a
    # random fine state the coarse grid's own Thom walls make R_c near the walls far larger than
    # the filtered F(R_f))
    scale = max(np.max(np.abs(frf[1:-1])), np.max(np.abs(rc[1:-1])))
    e_t = float(np.max(np.abs(tgt - tgt_o)) / scale)
    print(f"[les] pair: omega_c {e_w:.2e}, psi_c {e_psi:.2e}, target {e_t:.2e} "
          f"(|target| / |F(R_f)| = {np.max(np.abs(tgt_o)) / np.max(np.abs(frf[1:-1])):.2e})", flush=True)
    les.destroy()
    F.destroy()
    C.destroy()
    assert e_w <= 1e-13 and e_psi <= 1e-12 and e_t <= 1e-11
    assert np.all(tgt[[0, -1]] == 0.0)


def test_target_vanishes_when_the_grids_coincide():
    """rx = ry = 1: F keeps every mode below the Nyquist one. With x-modes below Nx/4 only, the
    state and its RHS (quadratic in the state) have no Nyquist content, F is the identity on
    both, and the target is zero up to rounding."""
    c = cfg(64, 33)
    F = lx.Channel(c, ic=False)
    C = lx.Channel(c, ic=False)
    les = lx.Les(F, C, 1, 1)
    g = Grid(64, 33, 2.0)
    w = smooth_random_field(64, 33, 4, kmax=6, y=g.m.y) + 3.0 * g.m.y[:, None]
    Wk = np.fft.rfft(w, axis=1)
    Wk[:, 64 // 4:] = 0.0
    wf = np.ascontiguousarray(np.fft.irfft(Wk, n=64, axis=1))
    psif = oracle.poisson_solve(wf, g.Lx, g.m, 0.0, 2.0)
    oracle.lib().orc_thom_wall_vorticity_mapped(P(wf), P(psif), g.Nx, g.Ny, P(g.m.dy_edge), 0.0, 2.0, 0)
    nh.upload_ptr(F.field(nh.FIELD_OMEGA), wf)
    nh.upload_ptr(F.field(nh.FIELD_PSI), psif)
    les.pair(append=False)
    tgt = les.tgt.download((33, 64))
    rf = oracle_rhs(g, psif, wf, 2.0 / c["Reb"], 0.05, 12, 1e-3)
    r = float(np.max(np.abs(tgt)) / np.max(np.abs(rf[1:-1])))
    print(f"[les] identity target {r:.2e}", flush=True)
    les.destroy()
    F.destroy()
    C.destroy()
    assert r <= 1e-10


class SourceStepper(OracleStepper):
    """The oracle's step with a fixed additive source on the interior rows."""

    def __init__(self, *a, source=None, **kw):
        super().__init__(*a, **kw)
        self.source = source

    def build_rhs(self, psi, w, t):
        r = super().build_rhs(psi, w, t)
        r[1:-1] += self.source[1:-1]
        return r


CLOSURE_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                              ctypes.c_void_p)


@pytest.mark.parametrize("Nx,Ny", [(64, 33), (256, 77)])
def test_stage_hook_adds_the_source(Nx, Ny):
    print(f"[les] stage hook {Nx}x{Ny}", flush=True)
    g = Grid(Nx, Ny, 2.0)
    nu, F0, nforce, drag, psi_top, dt = 2e-3, 0.05, 12, 1e-3, 2.0, 2e-4
    src = (0.3 * smooth_random_field(Nx, Ny, 21, kmax=5, y=g.m.y)).astype(np.float32)
    d_src = nh.DeviceArray(src)
    calls = []

    def hook(ctx, w, psi, out, stream):
        calls.append(1)
        nh.lib().nse2d_memcpy_d2d(out, d_src.ptr, src.nbytes)

    cb = CLOSURE_FN(hook)
    w0 = smooth_random_field(Nx, Ny, 3, kmax=6, y=g.m.y) + 3.0 * g.m.y[:, None]
    o = SourceStepper(g, nu, F0, nforce, drag, psi_top, source=src.astype(np.float64))
    w_o = w0.copy()
    psi_o = o.solve(w_o)
    ch = make_channel(g, nu, F0, nforce, drag, psi_top)
    assert nh.lib().nse2d_channel_set_closure(ch, ctypes.cast(cb, ctypes.c_void_p), None) == 0
    nh.upload_ptr(nh.lib().nse2d_channel_field(ch, nh.FIELD_OMEGA), w0)
    nh.lib().nse2d_channel_init(ch)
    psi_o, w_o = o.step(psi_o, w_o, 0.0, dt)
    nh.lib().nse2d_channel_step(ch, 0.0, dt, 0)
    sync()
    e = step_errors(g, field(ch, nh.FIELD_OMEGA, g), field(ch, nh.FIELD_PSI, g), w_o, psi_o, psi_top)
    nh.lib().nse2d_channel_set_closure(ch, None, None)
    nh.lib().nse2d_channel_destroy(ch)
    print(f"[les] hook step: psi {e[0]:.2e}, interior {e[1]:.2e}, wall vs own closure {e[2]:.2e}", flush=True)
    assert len(calls) == 3
    assert e[0] <= 2e-13 and e[1] <= 2e-13 and e[2] <= 1e-13


def test_closure_term_matches_the_cnn_oracle():
    c = cfg(64, 33)
    F = lx.Channel(cfg(128, 65), ic=False)
    C = lx.Channel(c, ic=False)
    les = lx.Les(F, C, 1, 1, seed=3)
    g = Grid(64, 33, 2.0)
    w, psi = fine_state(g, 2.0, 6)
    mean = np.array([0.5, 1.0, 0.7, 0.0, 0.1])
    std = np.array([3.0, 0.6, 0.5, 0.2, 2.0])
    assert nh.lib().nse2d_les_set_scales(les.h, mean.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                         std.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == 0
    d_w, d_psi = nh.DeviceBuffer.from_array(w), nh.DeviceBuffer.from_array(psi)
    d_out = nh.DeviceArray(shape=(33, 64), dtype=np.float32)
    assert nh.lib().nse2d_les_closure(les.h, d_w.ptr, d_psi.ptr, d_out.ptr) == 0
    got = d_out.download().astype(np.float64)
    cnn = nh.lib().nse2d_les_cnn(les.h)
    p = np.zeros(nh.lib().nse2d_cnn_num_params(cnn), np.float32)
    nh.lib().nse2d_cnn_get_params(cnn, p.ctypes.data)
    # inputs as in csrc/les.hip: omega, psi, u, v (ops.cu:47-63), normalised, rounded to fp32
    u, v = np.zeros_like(w), np.zeros_like(w)
    oracle.lib().orc_uv_from_psi_mapped(P(psi), P(u), P(v), g.Nx, g.Ny, g.inv_dx, g.inv_deta, P(g.m.a_node))
    x = np.stack([(f - mean[k]) / std[k] for k, f in enumerate((w, psi, u, v))], axis=-1)
    x = x.astype(np.float32).astype(np.float64)[None]
    ref, _ = co.forward(p.astype(np.float64), x, emulate=True)
    ref = ref[0] * std[4] + mean[4]
    ref[[0, -1]] = 0.0
    e = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    print(f"[les] closure term vs CNN oracle {e:.2e}", flush=True)
    les.destroy()
    F.destroy()
    C.destroy()
    assert np.all(got[[0, -1]] == 0.0)
    assert e <= 5e-3


def test_closure_constraints():
    """nse2d_les_set_constraints: the mean-free term is the plain term minus its row averages, the
    dissipative one is the plain term with every point where omega * C > 0 set to 0, both together
    apply the row averages first; an unknown flag is refused."""
    F = lx.Channel(cfg(256, 129), ic=False)
    C = lx.Channel(cfg(128, 65), ic=False)
    les = lx.Les(F, C, 1, 1, seed=5)
    g = Grid(128, 65, 2.0)
    w, psi = fine_state(g, 2.0, 9)
    d_w, d_psi = nh.DeviceBuffer.from_array(w), nh.DeviceBuffer.from_array(psi)
    d_out = nh.DeviceArray(shape=(65, 128), dtype=np.float32)
    lib = nh.lib()

    def term(flags):
        assert lib.nse2d_les_set_constraints(les.h, flags) == 0
        assert lib.nse2d_les_closure(les.h, d_w.ptr, d_psi.ptr, d_out.ptr) == 0
        return d_out.download()

    base = term(0)
    mf = term(nh.CLOSURE_MEAN_FREE)
    dis = term(nh.CLOSURE_DISSIPATIVE)
    both = term(nh.CLOSURE_MEAN_FREE | nh.CLOSURE_DISSIPATIVE)
    bad = lib.nse2d_les_set_constraints(les.h, 4)
    again = term(0)
    les.destroy()
    F.destroy()
    C.destroy()
    means = base.astype(np.float64).mean(axis=1, keepdims=True)
    mf_ref = base - means.astype(np.float32)
    scale = float(np.max(np.abs(base)))
    e_mf = float(np.max(np.abs(mf - mf_ref))) / scale
    rows = float(np.max(np.abs(mf.astype(np.float64).sum(axis=1)) / np.abs(base).astype(np.float64).sum(axis=1).clip(1e-300)))
    print(f"[les] mean-free term {e_mf:.1e}, row sums {rows:.1e}; backscatter points "
          f"{int(np.sum(base * w > 0))} of {base.size}", flush=True)
    assert np.std(means) > 0.0  # the plain term has row means to remove
    assert e_mf <= 1e-6 and rows <= 1e-5
    assert np.array_equal(dis, np.where(base.astype(np.float64) * w > 0.0, np.float32(0.0), base))
    assert 0 < np.sum(base * w > 0) < base.size
    assert np.array_equal(both, np.where(mf.astype(np.float64) * w > 0.0, np.float32(0.0), mf))
    assert bad == -1 and "unknown flag" in nh.last_error()
    assert np.array_equal(again, base)


@pytest.mark.parametrize("fine,coarse", [((128, 65), (64, 33)), ((4096, 2049), (512, 257))],
                         ids=["128x65-64x33", "4096x2049-512x257"])
def test_once_per_step_closure_with_direct_calls_between(fine, coarse, progress):
    """Mode 2 (closure of each step's first stage state, held over the three stages) with direct
    nse2d_les_closure calls on other states between the steps: every direct call writes its
    output (the buffer is filled with NaN before each), repeated calls agree bit for bit, and each
    attached step equals the oracle step carrying the closure of that step's own first stage state
    as a fixed source (the stage-hook bars)."""
    progress(f"[les] mode 2 closure on the {coarse} coarse channel (fine {fine})")
    c = cfg(*coarse, Reb=2e4)
    F = lx.Channel(cfg(*fine, Reb=2e4), ic=False)
    C = lx.Channel(c, ic=False)
    les = lx.Les(F, C, 1, 1, seed=3)
    g = Grid(coarse[0], coarse[1], 2.0)
    Ny, Nx = coarse[1], coarse[0]
    mean = np.array([0.5, 1.0, 0.7, 0.0, 0.1])
    std = np.array([3.0, 0.6, 0.5, 0.2, 2.0])
    assert nh.lib().nse2d_les_set_scales(les.h, mean.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                         std.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == 0
    w0, _ = fine_state(g, 2.0, 6)
    C.set_omega(w0)
    les.attach(C, mode=2)
    other_w, other_psi = fine_state(g, 2.0, 9)
    d_ow, d_opsi = nh.DeviceBuffer.from_array(other_w), nh.DeviceBuffer.from_array(other_psi)
    d_out = nh.DeviceArray(shape=(Ny, Nx), dtype=np.float32)
    nan = np.full((Ny, Nx), np.nan, np.float32)

    def direct(wp, pp):
        d_out.upload(nan)
        assert nh.lib().nse2d_les_closure(les.h, wp, pp, d_out.ptr) == 0
        out = d_out.download()
        assert np.all(np.isfinite(out))
        return out

    dt, t = 2e-4, 0.0
    try:
        for k in range(3):
            w_n, psi_n = C.omega(), nh.download_ptr(C.field(nh.FIELD_PSI), g.Nx * g.Ny, (g.Ny, g.Nx))
            T = direct(C.field(nh.FIELD_OMEGA), C.field(nh.FIELD_PSI))
            X = [direct(d_ow.ptr, d_opsi.ptr) for _ in range(k + 1)]
            assert all(np.array_equal(X[0], x) for x in X)
            nh.lib().nse2d_channel_step(C.ch, t, dt, 0)
            sync()
            o = SourceStepper(g, C.nu, 0.05, 12, 1e-3, 2.0, source=T.astype(np.float64))
            psi_o, w_o = o.step(psi_n.copy(), w_n.copy(), t, dt)
            e = step_errors(g, C.omega(), nh.download_ptr(C.field(nh.FIELD_PSI), g.Nx * g.Ny, (g.Ny, g.Nx)),
                            w_o, psi_o, 2.0)
            progress(f"[les] mode-2 step {k} with {k + 1} direct calls before it: psi {e[0]:.2e}, interior "
                     f"{e[1]:.2e}, wall vs own closure {e[2]:.2e}")
            # whole-step bars of test_gpu_step: at 257 rows the Thom coefficient (~3e6) carries the
            # solvers' near-wall psi rounding into the wall vorticity and the next stage's interior
            bar_int = 2e-13 if coarse[1] <= 65 else 5e-12
            assert e[0] <= 2e-13 and e[1] <= bar_int and e[2] <= 1e-13
            t += dt
    finally:
        les.attach(C, mode=0)
        les.destroy()
        F.destroy()
        C.destroy()


def test_attach_rejects_a_channel_on_another_stream():
    """The stage kernel reads the closure buffer that the hook's inference writes on the LES's
    stream, so the channel must run on that stream."""
    c = cfg(64, 33)
    F = lx.Channel(cfg(128, 65), ic=False)
    C = lx.Channel(c, ic=False)
    les = lx.Les(F, C, 1, 1, seed=3)
    s = nh.lib().nse2d_stream_create()
    C2 = lx.Channel(c, stream=s, ic=False)
    try:
        assert nh.lib().nse2d_les_attach(les.h, C2.ch, 1) == -1
        assert "stream" in nh.last_error()
        assert nh.lib().nse2d_les_attach(les.h, C.ch, 1) == 0
        assert nh.lib().nse2d_les_attach(les.h, C.ch, 0) == 0
    finally:
        les.destroy()
        C2.destroy()
        C.destroy()
        F.destroy()
        nh.lib().nse2d_stream_destroy(s)


def test_training_on_filtered_dns_beats_the_zero_closure():
    """A small DNS (512x257, Re_b = 5000), 64 pairs filtered 8x8 to 64x33, 300 Adam steps: the
    trained closure must explain most of the hold-out target (normalised MSE below half that of
    a zero closure), coarse runs with and without it must stay finite over the window, and the
    closure must bring the coarse run closer to the filtered DNS."""
    print("[les] small experiment 512x257 -> 64x33", flush=True)
    r = lx.run(cfg(512, 257), cfg(64, 33), spinup_time=1.0, window_time=1.0, n_pairs=64, holdout=8,
               train_steps=300, batch=8, lr=2e-3, n_samples=16, log=lambda s: print("[les] " + s, flush=True))
    print(f"[les] result {r}", flush=True)
    assert r["holdout_mse_cnn"] < 0.5 * r["holdout_mse_zero"]
    assert r["runs"]["closure"]["finite"] and r["runs"]["no_closure"]["finite"]
    assert r["train_loss_first_last"][1] < r["train_loss_first_last"][0]
    # a posteriori: started from the same filtered state, the coarse run with the closure follows
    # the filtered DNS more closely than the run without it (measured: 0.2 % against 3.9 % at
    # the end of the window), and its Re_tau is closer to the filtered DNS's
    cl, nc, cs = r["runs"]["closure"], r["runs"]["no_closure"], r["runs"]["closure_per_step"]
    assert cl["traj_err_end"] < 0.5 * nc["traj_err_end"]
    assert cl["re_tau_rel_err"] < nc["re_tau_rel_err"]
    # the closure evaluated once per step (held over the three stages) keeps most of the gain
    assert cs["finite"] and cs["traj_err_end"] < 0.5 * nc["traj_err_end"]
