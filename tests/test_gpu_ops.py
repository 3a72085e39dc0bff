"""Per-operator parity of the HIP C-ABI (ops.cu counterparts, test_api.cu wrappers) against the
oracle on the same seeded inputs. Bar: normwise relative error <= 1e-13 (FP64; the two sides
differ only by FMA contraction and operation order)."""
import math

import numpy as np
import pytest

import nse2d_hip as nh
import oracle
from gpu_util import Grid, dev, rel_err, smooth_random_field, sync, zeros_dev
from oracle import P

pytestmark = pytest.mark.gpu
TOL = 1e-13
GRIDS = [(64, 33, 2.0), (256, 129, 2.0), (128, 65, 0.0), (96, 37, 2.5), (64, 4, 2.0)]


@pytest.fixture(scope="module", params=GRIDS, ids=lambda g: f"{g[0]}x{g[1]}b{g[2]}")
def case(request):
    Nx, Ny, beta = request.param
    print(f"[ops] grid {Nx}x{Ny} beta={beta}", flush=True)
    g = Grid(Nx, Ny, beta)
    psi = smooth_random_field(Nx, Ny, 1, y=g.m.y)
    w = smooth_random_field(Nx, Ny, 2, scale=3.0, y=g.m.y)
    return g, psi, w


def test_arakawa_and_split(case):
    g, psi, w = case
    Nx, Ny = g.Nx, g.Ny
    ref = np.zeros_like(w)
    oracle.lib().orc_arakawa_J_mapped(P(psi), P(w), P(ref), Nx, Ny, g.inv_dx, g.inv_deta, P(g.m.a_node))
    dpsi, dw, dJ = dev(psi), dev(w), zeros_dev(Nx * Ny)
    nh.lib().testapi_arakawa_J_mapped(dpsi.ptr, dw.ptr, dJ.ptr, Nx, Ny, g.inv_dx, g.inv_deta, g.dp("a_node"), None)
    assert rel_err(dJ.download((Ny, Nx)), ref) <= TOL
    refs = [np.zeros_like(w) for _ in range(4)]
    oracle.lib().orc_arakawa_J_mapped_split(P(psi), P(w), *[P(r) for r in refs], Nx, Ny, g.inv_dx, g.inv_deta,
                                           P(g.m.a_node))
    outs = [zeros_dev(Nx * Ny) for _ in range(4)]
    nh.lib().testapi_arakawa_J_mapped_split(dpsi.ptr, dw.ptr, *[o.ptr for o in outs], Nx, Ny, g.inv_dx, g.inv_deta,
                                            g.dp("a_node"), None)
    for o, r in zip(outs, refs):
        assert rel_err(o.download((Ny, Nx)), r) <= TOL
    # NULL outputs are allowed
    nh.lib().testapi_arakawa_J_mapped_split(dpsi.ptr, dw.ptr, None, outs[1].ptr, None, None, Nx, Ny, g.inv_dx,
                                            g.inv_deta, g.dp("a_node"), None)
    sync()
    assert rel_err(outs[1].download((Ny, Nx)), refs[1]) <= TOL


def test_laplacian_rhs_mms(case):
    g, psi, w = case
    Nx, Ny = g.Nx, g.Ny
    lap_ref, J_ref, rhs_ref = np.zeros_like(w), np.zeros_like(w), np.zeros_like(w)
    oracle.lib().orc_laplacian_fd_mapped(P(w), P(lap_ref), Nx, Ny, g.inv_dx2, g.inv_deta2, P(g.m.a_node),
                                         P(g.m.a_edge))
    oracle.lib().orc_arakawa_J_mapped(P(psi), P(w), P(J_ref), Nx, Ny, g.inv_dx, g.inv_deta, P(g.m.a_node))
    oracle.lib().orc_rhs_mapped(P(J_ref), P(lap_ref), P(w), P(rhs_ref), Nx, Ny, 1e-3, 0.05, 12, 1.0, 1e-3,
                                P(g.m.y))
    dw, dlap, dJ, drhs = dev(w), zeros_dev(Nx * Ny), dev(J_ref), zeros_dev(Nx * Ny)
    nh.lib().launch_laplacian_fd_mapped(dw.ptr, dlap.ptr, Nx, Ny, g.inv_dx2, g.inv_deta2, g.dp("a_node"),
                                        g.dp("a_edge"), None)
    assert rel_err(dlap.download((Ny, Nx)), lap_ref) <= TOL
    dlap.upload(lap_ref)
    nh.lib().launch_rhs_mapped(dJ.ptr, dlap.ptr, dw.ptr, drhs.ptr, Nx, Ny, 1e-3, 0.05, 12, 1.0, 1e-3, g.dp("y"), None)
    assert rel_err(drhs.download((Ny, Nx)), rhs_ref) <= TOL
    args = (Nx, Ny, g.Lx, 1.0, 0.37, 1e-2, 1e-3, 1.0, 0.1, 1.0, 2)
    oracle.lib().orc_mms_add_forcing(P(rhs_ref), *args, P(g.m.y))
    nh.lib().launch_mms_add_forcing(drhs.ptr, *args, g.dp("y"), None)
    assert rel_err(drhs.download((Ny, Nx)), rhs_ref) <= TOL


def test_velocities_omega_thom(case):
    g, psi, w = case
    Nx, Ny = g.Nx, g.Ny
    u_ref, v_ref = np.zeros_like(w), np.zeros_like(w)
    oracle.lib().orc_uv_from_psi_mapped(P(psi), P(u_ref), P(v_ref), Nx, Ny, g.inv_dx, g.inv_deta, P(g.m.a_node))
    dpsi, du, dv = dev(psi), zeros_dev(Nx * Ny), zeros_dev(Nx * Ny)
    nh.lib().testapi_uv_from_psi_mapped(dpsi.ptr, du.ptr, dv.ptr, Nx, Ny, g.inv_dx, g.inv_deta, g.dp("a_node"), None)
    assert rel_err(du.download((Ny, Nx)), u_ref) <= TOL
    assert rel_err(dv.download((Ny, Nx)), v_ref) <= TOL
    # omega from psi (interior only; wall rows untouched)
    w_ref = w.copy()
    oracle.lib().orc_omega_from_psi_interior_mapped(P(w_ref), P(psi), Nx, Ny, g.inv_dx2, g.inv_deta2,
                                                    P(g.m.a_node), P(g.m.a_edge))
    dw = dev(w)
    nh.lib().testapi_omega_from_psi_interior_mapped(dw.ptr, dpsi.ptr, Nx, Ny, g.inv_dx2, g.inv_deta2,
                                                    g.dp("a_node"), g.dp("a_edge"), None)
    got = dw.download((Ny, Nx))
    assert rel_err(got, w_ref) <= TOL
    np.testing.assert_array_equal(got[0], w[0])
    # Thom closure, no-slip and free-slip
    for fs in (0, 1):
        w_ref = w.copy()
        oracle.lib().orc_thom_wall_vorticity_mapped(P(w_ref), P(psi), Nx, Ny, P(g.m.dy_edge), 0.1, 2.0, fs)
        dw = dev(w)
        nh.lib().nse2d_set_wall_type(fs)
        nh.lib().testapi_thom_wall_vorticity_mapped(dw.ptr, dpsi.ptr, Nx, Ny, g.dp("dy_edge"), 0.1, 2.0, None)
        sync()
        nh.lib().nse2d_set_wall_type(0)
        got = dw.download((Ny, Nx))
        np.testing.assert_array_equal(got[1:-1], w[1:-1])
        if fs:
            assert np.all(got[[0, -1]] == 0.0)
        else:
            assert rel_err(got[[0, -1]], w_ref[[0, -1]]) <= TOL


def test_advection_forms(case):
    g, psi, w = case
    Nx, Ny = g.Nx, g.Ny
    u, v = smooth_random_field(Nx, Ny, 3), smooth_random_field(Nx, Ny, 4)
    du, dv, dw, dout = dev(u), dev(v), dev(w), zeros_dev(Nx * Ny)
    for o_fn, h_fn in (("orc_compute_adv_from_uv_mapped", "testapi_compute_adv_from_uv_mapped"),
                       ("orc_compute_adv_flux_mapped", "testapi_compute_adv_flux_mapped")):
        ref = np.zeros_like(w)
        getattr(oracle.lib(), o_fn)(P(u), P(v), P(w), P(ref), Nx, Ny, g.inv_dx, g.inv_deta, P(g.m.a_node))
        getattr(nh.lib(), h_fn)(du.ptr, dv.ptr, dw.ptr, dout.ptr, Nx, Ny, g.inv_dx, g.inv_deta, g.dp("a_node"), None)
        assert rel_err(dout.download((Ny, Nx)), ref) <= TOL


def test_ssprk_combinations(case):
    g, psi, w = case
    N = g.Nx * g.Ny
    rng = np.random.default_rng(5)
    a, b, r = rng.normal(size=N), rng.normal(size=N), rng.normal(size=N)
    da, db, dr, dout = dev(a), dev(b), dev(r), zeros_dev(N)
    dt = 3.7e-4
    ref = np.zeros(N)
    oracle.lib().orc_axpy(P(ref), P(a), P(r), dt, N)
    nh.lib().launch_axpy(dout.ptr, da.ptr, dr.ptr, dt, N, None)
    assert rel_err(dout.download(), ref) <= TOL
    oracle.lib().orc_ssprk2(P(ref), P(a), P(b), P(r), dt, N)
    nh.lib().launch_ssprk2(dout.ptr, da.ptr, db.ptr, dr.ptr, dt, N, None)
    assert rel_err(dout.download(), ref) <= TOL
    oracle.lib().orc_ssprk3(P(ref), P(a), P(b), P(r), dt, N)
    nh.lib().launch_ssprk3(da.ptr, da.ptr, db.ptr, dr.ptr, dt, N, None)  # in place, as main.cu:530
    assert rel_err(da.download(), ref) <= TOL
