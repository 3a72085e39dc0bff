"""Parity of the HIP Poisson solver (LDS FFT + Thomas) and of the diagnostics reductions against
the oracle (poisson.cu / diag.cu restatements) on seeded inputs."""
import math

import numpy as np
import pytest

import nse2d_hip as nh
import oracle
from gpu_util import Grid, dev, rel_err, scalars, smooth_random_field, sync, zeros_dev
from oracle import P

pytestmark = pytest.mark.gpu

POISSON_GRIDS = [(4, 4, 2.0), (8, 9, 0.0), (64, 33, 2.0), (128, 129, 2.0), (256, 65, 0.0), (512, 257, 2.0),
                 (1024, 65, 2.0), (2048, 129, 2.0), (4096, 65, 1.0), (8192, 33, 2.0), (16384, 17, 2.5)]

# (Nx, radix) of every instantiated register row-FFT plan, and 0 = the LDS radix-4 kernels
FFT_PLANS = [(512, 4), (1024, 4), (1024, 8), (2048, 8), (2048, 16), (4096, 8), (4096, 16), (8192, 16),
             (16384, 16), (4096, 0), (512, 0)]


@pytest.mark.parametrize("Nx,radix", FFT_PLANS)
def test_poisson_every_fft_plan(Nx, radix, monkeypatch):
    """Each row-FFT plan (selected with NSE2D_FFT_RADIX at plan creation) on white-noise vorticity
    with homogeneous walls, so that psi is the noise response alone and every twiddle and
    butterfly carries weight, against the oracle."""
    monkeypatch.setenv("NSE2D_FFT_RADIX", str(radix))
    Ny = 33
    print(f"[fft plan] Nx={Nx} radix={radix}", flush=True)
    g = Grid(Nx, Ny, 2.0)
    w = np.random.default_rng(Nx + radix).normal(size=(Ny, Nx))
    ref = oracle.poisson_solve(w, g.Lx, g.m, 0.0, 0.0)
    p = nh.lib().poisson2d_create(Nx, Ny, g.Lx, None)
    assert p, nh.last_error()
    assert nh.lib().poisson2d_enable_mapped(p, g.dp("a_node"), g.dp("a_edge"), g.inv_deta, None) == 0
    dw, dpsi = dev(w), zeros_dev(Nx * Ny)
    nh.lib().poisson2d_solve(p, dw.ptr, dpsi.ptr, 0.0, 0.0, 0.0, None)
    got = dpsi.download((Ny, Nx))
    nh.lib().poisson2d_destroy(p)
    e = rel_err(got, ref)
    print(f"[fft plan] Nx={Nx} radix={radix} rel err {e:.2e}", flush=True)
    assert e <= 1e-13


@pytest.mark.parametrize("Nx,Ny,beta", POISSON_GRIDS)
def test_poisson_solve_matches_oracle(Nx, Ny, beta):
    print(f"[poisson] {Nx}x{Ny} beta={beta}", flush=True)
    g = Grid(Nx, Ny, beta)
    w = smooth_random_field(Nx, Ny, 11, kmax=min(8, Nx // 2), scale=5.0, y=g.m.y)
    for psi_bot, psi_top in ((0.0, 0.0), (0.0, 2.0)):
        ref = oracle.poisson_solve(w, g.Lx, g.m, psi_bot, psi_top)
        p = nh.lib().poisson2d_create(Nx, Ny, g.Lx, None)
        assert p, nh.last_error()
        assert nh.lib().poisson2d_enable_mapped(p, g.dp("a_node"), g.dp("a_edge"), g.inv_deta, None) == 0
        dw, dpsi = dev(w), zeros_dev(Nx * Ny)
        nh.lib().poisson2d_solve(p, dw.ptr, dpsi.ptr, 0.0, psi_bot, psi_top, None)
        got = dpsi.download((Ny, Nx))
        nh.lib().poisson2d_destroy(p)
        e = rel_err(got, ref)
        print(f"[poisson] {Nx}x{Ny} psi_top={psi_top} rel err {e:.2e}", flush=True)
        assert e <= 1e-13, (psi_bot, psi_top)
        np.testing.assert_allclose(got[0], psi_bot, atol=1e-15)
        np.testing.assert_allclose(got[-1], psi_top, atol=1e-14)


def test_poisson_full_size_residual():
    """4096x2049 (the benchmark grid): GPU solve vs oracle, and the discrete residual of
    -(spectral d_xx + mapped d_yy) psi = omega on the interior rows (a property independent of
    either implementation). At 2047 interior rows the serial FP64 elimination of the reference
    differs from an extended-precision solve by up to ~5e-13 for the kx ~ 0 modes
    (test_oracle_math.py::test_thomas_rounding_floor_at_2049_rows), so two FP64 solves are
    compared at 5e-13 here; below 1000 rows the bar is 1e-13."""
    Nx, Ny = 4096, 2049
    print(f"[poisson] full size {Nx}x{Ny}", flush=True)
    g = Grid(Nx, Ny, 2.0)
    w = smooth_random_field(Nx, Ny, 12, kmax=32, scale=5.0, y=g.m.y)
    ref = oracle.poisson_solve(w, g.Lx, g.m, 0.0, 2.0)
    p = nh.lib().poisson2d_create(Nx, Ny, g.Lx, None)
    nh.lib().poisson2d_enable_mapped(p, g.dp("a_node"), g.dp("a_edge"), g.inv_deta, None)
    dw, dpsi = dev(w), zeros_dev(Nx * Ny)
    nh.lib().poisson2d_solve(p, dw.ptr, dpsi.ptr, 0.0, 0.0, 2.0, None)
    got = dpsi.download((Ny, Nx))
    nh.lib().poisson2d_destroy(p)
    e = rel_err(got, ref)
    print(f"[poisson] full size rel err {e:.2e}", flush=True)
    assert e <= 5e-13
    k = np.fft.fftfreq(Nx, d=1.0 / Nx)
    k[Nx // 2] = Nx // 2
    kx2 = (2 * math.pi * k / g.Lx) ** 2
    dxx = np.real(np.fft.ifft(-kx2[None, :] * np.fft.fft(got, axis=1), axis=1))
    a, ae = g.m.a_node, g.m.a_edge
    dyy = a[1:-1, None] * (ae[1:, None] * (got[2:] - got[1:-1]) - ae[:-1, None] * (got[1:-1] - got[:-2])) * g.inv_deta2
    res = dxx[1:-1] + dyy + w[1:-1]
    assert np.max(np.abs(res)) <= 1e-8 * np.max(np.abs(w))


def test_poisson_rejects_bad_sizes():
    assert not nh.lib().poisson2d_create(48, 33, 1.0, None)
    assert "power of two" in nh.last_error()
    assert not nh.lib().poisson2d_create(64, 3, 1.0, None)


DIAG_GRIDS = [(64, 33, 2.0), (512, 257, 2.0), (256, 129, 0.0), (96, 37, 2.5)]


@pytest.fixture(scope="module", params=DIAG_GRIDS, ids=lambda g: f"{g[0]}x{g[1]}b{g[2]}")
def dcase(request):
    Nx, Ny, beta = request.param
    print(f"[diag] grid {Nx}x{Ny} beta={beta}", flush=True)
    g = Grid(Nx, Ny, beta)
    psi = smooth_random_field(Nx, Ny, 21, y=g.m.y)
    w = smooth_random_field(Nx, Ny, 22, scale=2.0, y=g.m.y)
    rhs = smooth_random_field(Nx, Ny, 23, y=g.m.y)
    return g, psi, w, rhs


def _s(buf, k=0):
    return float(buf.download()[k])


def _abs_scales(g, psi, w, rhs):
    """Sum of |terms| of the signed reductions (error bound scale of any summation order)."""
    Nx, Ny, L = g.Nx, g.Ny, oracle.lib()
    u, v = np.zeros_like(psi), np.zeros_like(psi)
    L.orc_uv_from_psi_mapped(P(psi), P(u), P(v), Nx, Ny, g.inv_dx, g.inv_deta, P(g.m.a_node))
    y, h = g.m.y, 1.0
    Fx = 0.05 * np.sin(12 * math.pi * ((y + h) / (2 * h)))
    fw = 0.05 * (12 * math.pi / (2 * h)) * np.abs(np.cos(12 * math.pi * ((y + h) / (2 * h))))
    lap = np.zeros_like(w)
    L.orc_laplacian_fd_mapped(P(w), P(lap), Nx, Ny, g.inv_dx2, g.inv_deta2, P(g.m.a_node), P(g.m.a_edge))
    wn = g.m.w_node
    return {
        "launch_power_input_mapped": float(np.sum(wn[:, None] * np.abs(u * Fx[:, None]))),
        "launch_enstrophy_input_mapped_full": float(np.sum(wn[:, None] * np.abs(w) * fw[:, None])),
        "launch_enstrophy_input_mapped": float(np.sum(wn[1:-1, None] * np.abs(w[1:-1]) * fw[1:-1, None])),
        "launch_omega_rhs_balance_mapped": float(np.sum(wn[:, None] * np.abs(w * rhs))),
        "launch_enstrophy_diss_balance_mapped": float(np.sum(wn[1:-1, None] * (w[1:-1] ** 2 + np.abs(w[1:-1] * lap[1:-1])))),
        "launch_mean_wall_omega": float(np.mean(np.abs(w[0])) + np.mean(np.abs(w[-1]))),
        "launch_mean_wall_omega_omegay_mapped": float(
            np.mean(np.abs(w[0]) * g.m.a_node[0] * (3 * np.abs(w[0]) + 4 * np.abs(w[1]) + np.abs(w[2])) * 0.5 * g.inv_deta)
            + np.mean(np.abs(w[-1]) * g.m.a_node[-1] * (3 * np.abs(w[-1]) + 4 * np.abs(w[-2]) + np.abs(w[-3])) * 0.5 * g.inv_deta)),
        "launch_enstrophy_sbp_mapped": float(
            np.sum(g.m.wt_node[:, None] * (np.roll(w, -1, axis=1) - w) ** 2) * g.inv_dx2
            + np.sum(g.m.a_edge[:, None] * np.diff(w, axis=0) ** 2) * g.inv_deta
            + (np.sum(np.abs(w[0] * (w[1] - w[0]))) * g.m.a_edge[0] + np.sum(np.abs(w[-1] * (w[-1] - w[-2]))) * g.m.a_edge[-1]) * g.inv_deta),
    }


def test_reductions_match_oracle(dcase):
    """Positive sums: relative error <= 1e-13. Signed sums: |error| <= 1e-13 * sum |terms|
    (the bound of a sum evaluated in any order)."""
    g, psi, w, rhs = dcase
    Nx, Ny, L = g.Nx, g.Ny, oracle.lib()
    dpsi, dw, drhs = dev(psi), dev(w), dev(rhs)
    out = scalars(2)
    tol = 1e-13
    scales = _abs_scales(g, psi, w, rhs)

    def ok(fn, got, ref):
        bound = tol * max(abs(ref), scales.get(fn, 0.0))
        assert abs(got - ref) <= bound or abs(got - ref) < 1e-300, (fn, got, ref, bound)

    def pair(orc_fn, hip_fn, orc_args, hip_args):
        ref = np.zeros(2)
        getattr(L, orc_fn)(*orc_args, P(ref))
        getattr(nh.lib(), hip_fn)(*hip_args, out.ptr, out.ptr + 8, None)
        for a, b in zip(out.download(), ref):
            ok(hip_fn, a, b)

    pair("orc_enstrophy_diss_mapped", "launch_enstrophy_diss_mapped",
         (P(w), Nx, Ny, g.inv_dx, g.inv_deta, P(g.m.a_node), P(g.m.w_node)),
         (dw.ptr, Nx, Ny, g.inv_dx, g.inv_deta, g.dp("a_node"), g.dp("w_node")))
    pair("orc_enstrophy_diss_mapped_full", "launch_enstrophy_diss_mapped_full",
         (P(w), Nx, Ny, g.inv_dx, g.inv_deta, P(g.m.a_node), P(g.m.w_node)),
         (dw.ptr, Nx, Ny, g.inv_dx, g.inv_deta, g.dp("a_node"), g.dp("w_node")))
    pair("orc_enstrophy_sbp_mapped", "launch_enstrophy_sbp_mapped",
         (P(w), Nx, Ny, g.inv_dx, g.inv_deta, P(g.m.a_edge), P(g.m.wt_node)),
         (dw.ptr, Nx, Ny, g.inv_dx, g.inv_deta, g.dp("a_edge"), g.dp("wt_node")))
    pair("orc_enstrophy_diss_balance_mapped", "launch_enstrophy_diss_balance_mapped",
         (P(w), Nx, Ny, g.inv_dx2, g.inv_deta2, P(g.m.a_node), P(g.m.a_edge), P(g.m.w_node)),
         (dw.ptr, Nx, Ny, g.inv_dx2, g.inv_deta2, g.dp("a_node"), g.dp("a_edge"), g.dp("w_node")))
    pair("orc_mean_wall_omega", "launch_mean_wall_omega", (P(w), Nx, Ny), (dw.ptr, Nx, Ny))
    pair("orc_mean_wall_omega_omegay_mapped", "launch_mean_wall_omega_omegay_mapped",
         (P(w), Nx, Ny, g.inv_deta, P(g.m.a_node)), (dw.ptr, Nx, Ny, g.inv_deta, g.dp("a_node")))

    def single(orc_fn, hip_fn, orc_args, hip_args):
        ref = getattr(L, orc_fn)(*orc_args)
        getattr(nh.lib(), hip_fn)(*hip_args, out.ptr, None)
        ok(hip_fn, _s(out), ref)

    single("orc_kinetic_energy_mapped", "launch_kinetic_energy_mapped",
           (P(psi), Nx, Ny, g.inv_dx, g.inv_deta, P(g.m.a_node), P(g.m.w_node)),
           (dpsi.ptr, Nx, Ny, g.inv_dx, g.inv_deta, g.dp("a_node"), g.dp("w_node")))
    single("orc_energy_dissipation_mapped", "launch_energy_dissipation_mapped",
           (P(psi), P(w), Nx, Ny, g.inv_dx, g.inv_deta, P(g.m.a_node), P(g.m.a_edge), P(g.m.w_node)),
           (dpsi.ptr, dw.ptr, Nx, Ny, g.inv_dx, g.inv_deta, g.dp("a_node"), g.dp("a_edge"), g.dp("w_node")))
    single("orc_power_input_mapped", "launch_power_input_mapped",
           (P(psi), Nx, Ny, g.inv_deta, 1.0, 0.05, 12, P(g.m.a_node), P(g.m.y), P(g.m.w_node)),
           (dpsi.ptr, Nx, Ny, g.inv_deta, 1.0, 0.05, 12, g.dp("a_node"), g.dp("y"), g.dp("w_node")))
    single("orc_enstrophy_input_mapped_full", "launch_enstrophy_input_mapped_full",
           (P(w), Nx, Ny, 1.0, 0.05, 12, P(g.m.y), P(g.m.w_node)),
           (dw.ptr, Nx, Ny, 1.0, 0.05, 12, g.dp("y"), g.dp("w_node")))
    single("orc_enstrophy_input_mapped", "launch_enstrophy_input_mapped",
           (P(w), Nx, Ny, 1.0, 0.05, 12, P(g.m.y), P(g.m.w_node)),
           (dw.ptr, Nx, Ny, 1.0, 0.05, 12, g.dp("y"), g.dp("w_node")))
    single("orc_omega_rhs_balance_mapped", "launch_omega_rhs_balance_mapped",
           (P(w), P(rhs), Nx, Ny, P(g.m.w_node)), (dw.ptr, drhs.ptr, Nx, Ny, g.dp("w_node")))


def test_max_uv_and_row_moments(dcase):
    g, psi, w, rhs = dcase
    Nx, Ny = g.Nx, g.Ny
    dpsi = dev(psi)
    ref = np.zeros(3)
    oracle.lib().orc_max_uv_from_psi_mapped(P(psi), Nx, Ny, g.inv_dx, g.inv_deta, P(g.m.a_node), P(ref))
    out = scalars(3)
    nh.lib().launch_max_uv_from_psi_mapped(dpsi.ptr, Nx, Ny, g.inv_dx, g.inv_deta, g.dp("a_node"), out.ptr,
                                           out.ptr + 8, out.ptr + 16, None)
    got = out.download()
    assert rel_err(got, ref) <= 1e-14
    refs = [np.zeros(Ny) for _ in range(5)]
    oracle.lib().orc_row_moments_from_psi_mapped(P(psi), Nx, Ny, g.inv_dx, g.inv_deta, P(g.m.a_node),
                                                 *[P(r) for r in refs])
    outs = [zeros_dev(Ny) for _ in range(5)]
    nh.lib().launch_row_moments_from_psi_mapped(dpsi.ptr, Nx, Ny, g.inv_dx, g.inv_deta, g.dp("a_node"),
                                                *[o.ptr for o in outs], None)
    sync()
    # per-row sums of signed terms: bound by the row sums of |terms| (u, v as the kernel forms them)
    u, v = np.zeros_like(psi), np.zeros_like(psi)
    oracle.lib().orc_uv_from_psi_mapped(P(psi), P(u), P(v), Nx, Ny, g.inv_dx, g.inv_deta, P(g.m.a_node))
    abs_rows = [np.sum(np.abs(t), axis=1) for t in (u, v, u * u, v * v, u * v)]
    for o, r, a in zip(outs, refs, abs_rows):
        assert np.all(np.abs(o.download() - r) <= 1e-13 * a + 1e-300)


def _launcher_row(g, dpsi, dw, drhs, h=1.0, F0=0.05, nforce=12):
    """The outputs of nse2d_diag_row from the separate launchers, in the NSE2D_DIAG_* order."""
    Nx, Ny, L = g.Nx, g.Ny, nh.lib()
    o = scalars(16)
    b = o.ptr
    L.launch_max_uv_from_psi_mapped(dpsi.ptr, Nx, Ny, g.inv_dx, g.inv_deta, g.dp("a_node"), b, b + 8, b + 16, None)
    L.launch_kinetic_energy_mapped(dpsi.ptr, Nx, Ny, g.inv_dx, g.inv_deta, g.dp("a_node"), g.dp("w_node"), b + 24, None)
    L.launch_enstrophy_diss_mapped_full(dw.ptr, Nx, Ny, g.inv_dx, g.inv_deta, g.dp("a_node"), g.dp("w_node"), b + 32,
                                        b + 40, None)
    L.launch_omega_rhs_balance_mapped(dw.ptr, drhs.ptr, Nx, Ny, g.dp("w_node"), b + 48, None)
    L.launch_power_input_mapped(dpsi.ptr, Nx, Ny, g.inv_deta, h, F0, nforce, g.dp("a_node"), g.dp("y"),
                                g.dp("w_node"), b + 56, None)
    L.launch_energy_dissipation_mapped(dpsi.ptr, dw.ptr, Nx, Ny, g.inv_dx, g.inv_deta, g.dp("a_node"), g.dp("a_edge"),
                                       g.dp("w_node"), b + 64, None)
    L.launch_enstrophy_input_mapped_full(dw.ptr, Nx, Ny, h, F0, nforce, g.dp("y"), g.dp("w_node"), b + 72, None)
    L.launch_mean_wall_omega_omegay_mapped(dw.ptr, Nx, Ny, g.inv_deta, g.dp("a_node"), b + 80, b + 88, None)
    L.launch_mean_wall_omega(dw.ptr, Nx, Ny, b + 96, b + 104, None)
    return o.download()[:14]


@pytest.mark.parametrize("Nx,Ny,beta", DIAG_GRIDS + [(4096, 2049, 2.0), (300, 5, 2.0)])
def test_fused_diag_row_equals_launchers(Nx, Ny, beta):
    """nse2d_diag_row (one pass over psi, omega and rhs) against the separate launchers of the
    same reductions, bit for bit, including the optional row moments; the launchers themselves
    are checked against the oracle above."""
    print(f"[diag row] {Nx}x{Ny} beta={beta}", flush=True)
    g = Grid(Nx, Ny, beta)
    psi = smooth_random_field(Nx, Ny, 31, y=g.m.y)
    w = smooth_random_field(Nx, Ny, 32, scale=2.0, y=g.m.y)
    rhs = smooth_random_field(Nx, Ny, 33, y=g.m.y)
    dpsi, dw, drhs = dev(psi), dev(w), dev(rhs)
    ref = _launcher_row(g, dpsi, dw, drhs)
    out, mom = scalars(14), zeros_dev(5 * Ny)
    for moments in (None, mom.ptr):
        assert nh.lib().nse2d_diag_row(dpsi.ptr, dw.ptr, drhs.ptr, Nx, Ny, g.inv_dx, g.inv_deta, 1.0, 0.05, 12,
                                       g.dp("a_node"), g.dp("a_edge"), g.dp("y"), g.dp("w_node"), out.ptr, moments,
                                       None) == 0, nh.last_error()
        got = out.download()
        bad = [k for k in range(14) if not got[k] == ref[k]]
        assert not bad, [(k, got[k], ref[k]) for k in bad]
    outs = [zeros_dev(Ny) for _ in range(5)]
    nh.lib().launch_row_moments_from_psi_mapped(dpsi.ptr, Nx, Ny, g.inv_dx, g.inv_deta, g.dp("a_node"),
                                                *[o.ptr for o in outs], None)
    got_m = mom.download((5, Ny))
    for k in range(5):
        ref_m = outs[k].download()
        diff = np.nonzero(got_m[k] != ref_m)[0]
        assert diff.size == 0, (k, diff[:5], got_m[k][diff[:3]], ref_m[diff[:3]])
    assert nh.lib().nse2d_diag_row(dpsi.ptr, dw.ptr, drhs.ptr, Nx, 3, g.inv_dx, g.inv_deta, 1.0, 0.05, 12,
                                   g.dp("a_node"), g.dp("a_edge"), g.dp("y"), g.dp("w_node"), out.ptr, None,
                                   None) == -1


@pytest.mark.parametrize("Nx,Ny", [(64, 17), (512, 33), (4096, 65)])
def test_spectra1d_matches_oracle(Nx, Ny):
    """nse2d_spectra1d (diag.cuh:123-191 on the row FFT) against the oracle's restatement, over
    two accumulated samples, on the LDS and the register row-FFT paths."""
    print(f"[spectra] {Nx}x{Ny}", flush=True)
    rng = np.random.default_rng(Nx)
    L = nh.lib()
    sp = L.nse2d_spectra1d_create(Nx, Ny, 2.5, None)
    assert sp, nh.last_error()
    nk = Nx // 2 + 1
    acc = np.zeros(2 * nk)
    for _ in range(2):
        u, v = rng.normal(size=(Ny, Nx)), rng.normal(size=(Ny, Nx))
        du, dv = dev(u), dev(v)
        L.nse2d_spectra1d_accumulate(sp, du.ptr, dv.ptr, None)
        oracle.lib().orc_spectra1d_accumulate(P(u), P(v), Nx, Ny, P(acc))
    kx, eu, ev = np.zeros(nk), np.zeros(nk), np.zeros(nk)
    assert L.nse2d_spectra1d_read(sp, kx.ctypes.data, eu.ctypes.data, ev.ctypes.data, None) == 2
    L.nse2d_spectra1d_destroy(sp)
    np.testing.assert_allclose(kx, 2 * math.pi * np.arange(nk) / 2.5, rtol=1e-15)
    for got, ref in ((eu, acc[:nk] / 2), (ev, acc[nk:] / 2)):
        e = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
        print(f"[spectra] {Nx}x{Ny} rel err {e:.2e}", flush=True)
        assert e <= 1e-13
