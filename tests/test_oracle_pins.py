"""Closed-form and invariant pins of the oracle functions that no reference fixture exercises:
the test-API operators (ops.cu:47-101, :339-467) and the two reductions that the reference never
calls (diag.cu:103-156, :423-467). Each check takes the operator as a callable; the last test
feeds it copies of the operator with one planted mistake each (a dropped factor, a wrong sign, a
wrong index, a wrong wall rule) and requires every copy to fail, so the checks are not vacuous.

The closed forms use the facts that fix the discrete operators:
  - centred differences of sin(kx) on a periodic row: sin(k(x+dx)) - sin(k(x-dx)) = 2 cos(kx) sin(k dx);
  - a quantity linear in eta has the exact centred eta-difference 2 c deta, and eta^2 gives 4 eta deta;
  - psi built so that a_{j+1/2} (psi_{j+1} - psi_j) = eta_{j+1/2} deta makes the conservative mapped
    y-Laplacian exactly a_j (any metric, any stretching);
  - the trapezoid rule on a uniform grid is exact for trigonometric polynomials whose period divides
    the interval, so sum_j w_j cos^2(n pi (y_j + h) / 2h) = h;
  - the Arakawa terms J2 and J3 are in divergence form (zero weighted sums), J2 conserves enstrophy
    and J3 energy, and all three reduce to -U delta_x omega on a uniform flow psi = U y.
"""
import math

import numpy as np
import pytest

import oracle
from oracle import P

L = oracle.lib


def _grid(Ny, stretched):
    return oracle.Metrics(Ny, 1.0, 1, 2.0) if stretched else oracle.Metrics(Ny, 1.0, 0, 0.0)


def _eta(Ny):
    return -1.0 + 2.0 * np.arange(Ny) / (Ny - 1)


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ---------------------------------------------------------------------------------------------
# The oracle operators as callables (the functions under test).

def orc_uv(psi, m, inv_dx):
    Ny, Nx = psi.shape
    u, v = np.full_like(psi, np.nan), np.full_like(psi, np.nan)
    L().orc_uv_from_psi_mapped(P(_c(psi)), P(u), P(v), Nx, Ny, inv_dx, (Ny - 1) / 2.0, P(m.a_node))
    return u, v


def orc_omega_from_psi(w_init, psi, m, inv_dx):
    Ny, Nx = psi.shape
    w = _c(w_init.copy())
    L().orc_omega_from_psi_interior_mapped(P(w), P(_c(psi)), Nx, Ny, inv_dx ** 2, ((Ny - 1) / 2.0) ** 2,
                                           P(m.a_node), P(m.a_edge))
    return w


def orc_adv_uv(u, v, w, m, inv_dx):
    Ny, Nx = w.shape
    J = np.full_like(w, np.nan)
    L().orc_compute_adv_from_uv_mapped(P(_c(u)), P(_c(v)), P(_c(w)), P(J), Nx, Ny, inv_dx, (Ny - 1) / 2.0,
                                       P(m.a_node))
    return J


def orc_adv_flux(u, v, w, m, inv_dx):
    Ny, Nx = w.shape
    J = np.full_like(w, np.nan)
    L().orc_compute_adv_flux_mapped(P(_c(u)), P(_c(v)), P(_c(w)), P(J), Nx, Ny, inv_dx, (Ny - 1) / 2.0,
                                    P(m.a_node))
    return J


def orc_split(psi, w, m, inv_dx):
    Ny, Nx = w.shape
    out = [np.full_like(w, np.nan) for _ in range(4)]
    L().orc_arakawa_J_mapped_split(P(_c(psi)), P(_c(w)), *[P(o) for o in out], Nx, Ny, inv_dx, (Ny - 1) / 2.0,
                                   P(m.a_node))
    return out


def orc_balance(w, m, inv_dx):
    Ny, Nx = w.shape
    out = np.zeros(2)
    L().orc_enstrophy_diss_balance_mapped(P(_c(w)), Nx, Ny, inv_dx ** 2, ((Ny - 1) / 2.0) ** 2, P(m.a_node),
                                          P(m.a_edge), P(m.w_node), P(out))
    return out


def orc_input(w, m, h, F0, n):
    Ny, Nx = w.shape
    return L().orc_enstrophy_input_mapped(P(_c(w)), Nx, Ny, h, F0, n, P(m.y), P(m.w_node))


# ---------------------------------------------------------------------------------------------
# Checks (closed forms and invariants). Each raises AssertionError on a mismatch.

def check_uv(op):
    """ops.cu:47-63. psi linear in eta gives u = a_j c inside and a_j c / 2 on the wall rows (the
    clamped wall difference); psi = sin(kx) gives v = -cos(kx) sin(k dx) / dx and u = 0."""
    Nx, Ny, c, k = 32, 17, 0.7, 3
    dx = 2 * math.pi / Nx
    x = np.arange(Nx) * dx
    for stretched in (False, True):
        m = _grid(Ny, stretched)
        psi = np.tile((c * _eta(Ny))[:, None], (1, Nx))
        u, v = op(psi, m, 1 / dx)
        ref = (c * m.a_node)[:, None] * np.ones((1, Nx))
        ref[[0, -1]] *= 0.5
        np.testing.assert_allclose(u, ref, rtol=1e-14, atol=0)
        assert np.all(v == 0.0)
    m = _grid(Ny, True)
    psi = np.tile(np.sin(k * x), (Ny, 1))
    u, v = op(psi, m, 1 / dx)
    np.testing.assert_allclose(v, np.tile(-np.cos(k * x) * math.sin(k * dx) / dx, (Ny, 1)), rtol=0, atol=1e-13)
    assert np.all(u == 0.0)


def _psi_with_unit_ylap(m):
    """psi_j with a_{j+1/2} (psi_{j+1} - psi_j) = eta_{j+1/2} deta, so that the mapped y-Laplacian
    a_j [a_{j+1/2}(psi_{j+1}-psi_j) - a_{j-1/2}(psi_j-psi_{j-1})] / deta^2 equals a_j."""
    Ny = len(m.a_node)
    deta = 2.0 / (Ny - 1)
    eta_half = -1.0 + deta * (np.arange(Ny - 1) + 0.5)
    steps = eta_half * deta / m.a_edge
    return np.concatenate([[0.0], np.cumsum(steps)])


def check_omega_from_psi(op):
    """ops.cu:72-93. Uniform grid: psi = sin(kx) + y^2 gives omega = (4/dx^2) sin^2(k dx/2) sin(kx) - 2.
    Stretched grid: psi with a unit mapped y-Laplacian gives omega = -a_j. Wall rows untouched."""
    Nx, Ny, k = 32, 17, 3
    dx = 2 * math.pi / Nx
    x = np.arange(Nx) * dx
    m = _grid(Ny, False)
    psi = np.tile(np.sin(k * x), (Ny, 1)) + (m.y ** 2)[:, None]
    sentinel = np.full((Ny, Nx), 7.25)
    w = op(sentinel, psi, m, 1 / dx)
    ref = (4 / dx ** 2) * math.sin(k * dx / 2) ** 2 * np.sin(k * x)[None, :] - 2.0
    np.testing.assert_allclose(w[1:-1], np.broadcast_to(ref, (Ny - 2, Nx)), rtol=0, atol=1e-11)
    assert np.all(w[[0, -1]] == 7.25)
    m = _grid(Ny, True)
    psi = np.tile(_psi_with_unit_ylap(m)[:, None], (1, Nx))
    w = op(sentinel, psi, m, 1 / dx)
    np.testing.assert_allclose(w[1:-1], np.broadcast_to(-m.a_node[1:-1, None], (Ny - 2, Nx)), rtol=1e-11, atol=0)
    assert np.all(w[[0, -1]] == 7.25)


def check_adv_from_uv(op):
    """ops.cu:339-371. u = U(y) (per row), v = V, omega = sin(kx) + c eta:
    J = U(y_j) cos(kx) sin(k dx)/dx + V a_j c inside, 0 on the wall rows."""
    Nx, Ny, k, c, V = 32, 17, 2, 0.4, -0.9
    dx = 2 * math.pi / Nx
    x = np.arange(Nx) * dx
    for stretched in (False, True):
        m = _grid(Ny, stretched)
        Uy = 0.3 + m.y ** 2
        u = np.tile(Uy[:, None], (1, Nx))
        v = np.full((Ny, Nx), V)
        w = np.sin(k * x)[None, :] + (c * _eta(Ny))[:, None]
        J = op(u, v, w, m, 1 / dx)
        ref = Uy[:, None] * (np.cos(k * x) * math.sin(k * dx) / dx)[None, :] + (V * m.a_node * c)[:, None]
        np.testing.assert_allclose(J[1:-1], ref[1:-1], rtol=0, atol=1e-13)
        assert np.all(J[[0, -1]] == 0.0)


def check_adv_flux(op):
    """ops.cu:372-409. Constant u = U, v = V reduce the face averages to the centred form
    U delta_x omega + V a delta_eta omega; constant omega = W with u = sin(kx), v = eta^2 gives the
    centred divergence W cos(kx) sin(k dx)/dx + 2 a_j W eta_j (an advective form would give 0)."""
    Nx, Ny, k, U, V, W = 32, 17, 2, 0.6, -0.9, 1.3
    dx = 2 * math.pi / Nx
    x = np.arange(Nx) * dx
    for stretched in (False, True):
        m = _grid(Ny, stretched)
        eta = _eta(Ny)
        w = np.sin(k * x)[None, :] + (0.4 * eta)[:, None]
        J = op(np.full((Ny, Nx), U), np.full((Ny, Nx), V), w, m, 1 / dx)
        ref = U * (np.cos(k * x) * math.sin(k * dx) / dx)[None, :] + (V * m.a_node * 0.4)[:, None]
        np.testing.assert_allclose(J[1:-1], ref[1:-1], rtol=0, atol=1e-13)
        assert np.all(J[[0, -1]] == 0.0)
        u = np.tile(np.sin(k * x), (Ny, 1))
        v = np.tile((eta ** 2)[:, None], (1, Nx))
        J = op(u, v, np.full((Ny, Nx), W), m, 1 / dx)
        ref = W * (np.cos(k * x) * math.sin(k * dx) / dx)[None, :] + (2 * m.a_node * W * eta)[:, None]
        np.testing.assert_allclose(J[1:-1], ref[1:-1], rtol=0, atol=1e-12)
        assert np.all(J[[0, -1]] == 0.0)


def check_split(op):
    """ops.cu:415-467. (1) Uniform flow psi = c eta (u = a_j c): J1 = J2 = J3 = -a_j c delta_x omega.
    (2) On random fields that vanish on the two rows next to each wall, with wt_j = deta / a_j:
    sum wt J2 = sum wt J3 = 0 (divergence forms), sum wt omega J2 = 0 (enstrophy),
    sum wt psi J3 = 0 (energy), while J1 has neither property. (3) Jadv = -(J1+J2+J3)/3 equals the
    fused operator (ops.cu:283-326), and every output is zero on the wall rows."""
    Nx, Ny, c = 24, 13, 0.8
    dx = 2 * math.pi / Nx
    x = np.arange(Nx) * dx
    for stretched in (False, True):
        m = _grid(Ny, stretched)
        psi = np.tile((c * _eta(Ny))[:, None], (1, Nx))
        w = np.tile(np.sin(2 * x) + 0.3 * np.cos(5 * x), (Ny, 1))
        J1, J2, J3, Ja = op(psi, w, m, 1 / dx)
        dxw = (np.roll(w, -1, axis=1) - np.roll(w, 1, axis=1)) / (2 * dx)
        ref = -(c * m.a_node)[:, None] * dxw
        for Jk in (J1, J2, J3):
            np.testing.assert_allclose(Jk[1:-1], ref[1:-1], rtol=0, atol=1e-13)
        np.testing.assert_allclose(Ja[1:-1], -ref[1:-1], rtol=0, atol=1e-13)
    Nx, Ny = 40, 33
    m = _grid(Ny, True)
    rng = np.random.default_rng(11)
    psi, w = rng.normal(size=(Ny, Nx)), rng.normal(size=(Ny, Nx))
    for f in (psi, w):
        f[[0, 1, -2, -1]] = 0.0
    J1, J2, J3, Ja = op(psi, w, m, Nx / (2 * math.pi))
    wt = m.wt_node[:, None]

    def zero(s, terms):
        return abs(s) <= 1e-13 * terms

    assert zero(np.sum(wt * J2), np.sum(wt * np.abs(J2)))
    assert zero(np.sum(wt * J3), np.sum(wt * np.abs(J3)))
    assert zero(np.sum(wt * w * J2), np.sum(wt * np.abs(w * J2)))
    assert zero(np.sum(wt * psi * J3), np.sum(wt * np.abs(psi * J3)))
    assert not zero(np.sum(wt * psi * J1), np.sum(wt * np.abs(psi * J1)))
    assert not zero(np.sum(wt * w * J1), np.sum(wt * np.abs(w * J1)))
    assert not zero(np.sum(wt * w * J3), np.sum(wt * np.abs(w * J3)))
    Jf = np.zeros_like(w)
    L().orc_arakawa_J_mapped(P(_c(psi)), P(_c(w)), P(Jf), Nx, Ny, Nx / (2 * math.pi), (Ny - 1) / 2.0, P(m.a_node))
    np.testing.assert_allclose(Ja, Jf, rtol=0, atol=1e-14 * np.max(np.abs(Jf)))
    for Jk in (J1, J2, J3, Ja):
        assert np.all(Jk[[0, -1]] == 0.0)


def check_balance(op):
    """diag.cu:103-156 + 586-625: (sum over interior rows of w_j omega^2, sum of w_j omega lap omega)
    with the solver's Laplacian. Uniform grid, omega = cos(kx): (Nx/2) S and -(4/dx^2) sin^2(k dx/2)
    (Nx/2) S with S = sum_{interior} w_j = 2h - dy. Stretched grid, omega from the unit mapped
    y-Laplacian construction: sum w omega^2 and sum w omega a_j over interior rows."""
    Nx, Ny, k = 32, 17, 3
    dx = 2 * math.pi / Nx
    x = np.arange(Nx) * dx
    m = _grid(Ny, False)
    S = 2.0 - m.dy_edge[0]
    w = np.tile(np.cos(k * x), (Ny, 1))
    got = op(w, m, 1 / dx)
    lam = (4 / dx ** 2) * math.sin(k * dx / 2) ** 2
    np.testing.assert_allclose(got, [Nx / 2 * S, -lam * Nx / 2 * S], rtol=1e-13)
    m = _grid(Ny, True)
    q = _psi_with_unit_ylap(m)
    w = np.tile(q[:, None], (1, Nx))
    got = op(w, m, 1 / dx)
    jj = slice(1, Ny - 1)
    ref = [Nx * np.sum(m.w_node[jj] * q[jj] ** 2), Nx * np.sum(m.w_node[jj] * q[jj] * m.a_node[jj])]
    np.testing.assert_allclose(got, ref, rtol=1e-12)


def check_enstrophy_input(op):
    """diag.cu:423-467 (interior rows only). Uniform grid, omega = f_omega(y) =
    -F0 (n pi / 2h) cos(n pi (y+h)/2h): the trapezoid sum of cos^2 over [-h, h] is h exactly, and the
    two wall rows (cos^2 = 1, weight dy/2 each) are left out, so P = F0^2 (n pi/2h)^2 Nx (h - dy).
    omega = cos(kx) g(y) has zero x-mean, so P = 0."""
    Nx, Ny, h, F0, n = 16, 65, 1.3, 0.05, 4
    m = oracle.Metrics(Ny, h, 0, 0.0)
    fw = -F0 * (n * math.pi / (2 * h)) * np.cos(n * math.pi * (m.y + h) / (2 * h))
    w = np.tile(fw[:, None], (1, Nx))
    got = op(w, m, h, F0, n)
    ref = F0 ** 2 * (n * math.pi / (2 * h)) ** 2 * Nx * (h - m.dy_edge[0])
    assert abs(got - ref) <= 1e-13 * abs(ref), (got, ref)
    x = np.arange(Nx) * 2 * math.pi / Nx
    w = np.cos(3 * x)[None, :] * (1 + m.y ** 2)[:, None]
    assert abs(op(w, m, h, F0, n)) <= 1e-15


CHECKS = {
    "uv_from_psi_mapped": (check_uv, orc_uv),
    "omega_from_psi_interior_mapped": (check_omega_from_psi, orc_omega_from_psi),
    "compute_adv_from_uv_mapped": (check_adv_from_uv, orc_adv_uv),
    "compute_adv_flux_mapped": (check_adv_flux, orc_adv_flux),
    "arakawa_J_mapped_split": (check_split, orc_split),
    "enstrophy_diss_balance_mapped": (check_balance, orc_balance),
    "enstrophy_input_mapped": (check_enstrophy_input, orc_input),
}


@pytest.mark.parametrize("name", list(CHECKS))
def test_oracle_function_pinned(name):
    check, op = CHECKS[name]
    check(op)


# ---------------------------------------------------------------------------------------------
# Planted mistakes: copies of an operator with one plausible error each. Every copy must fail
# its check. The copies wrap the oracle and alter what it computes or what it is given.

def _shift_x(f, s):
    return np.roll(f, -s, axis=1)


def _mut_uv_no_clamp(psi, m, inv_dx):
    u, v = orc_uv(psi, m, inv_dx)
    u[0], u[-1] = 2 * u[0], 2 * u[-1]          # full one-sided difference at the walls
    return u, v


def _mut_uv_sign(psi, m, inv_dx):
    u, v = orc_uv(psi, m, inv_dx)
    return u, -v                                # v = +psi_x


def _mut_uv_no_metric(psi, m, inv_dx):
    u, v = orc_uv(psi, m, inv_dx)
    return u / m.a_node[:, None], v             # a_j dropped


def _mut_omega_edge_index(w_init, psi, m, inv_dx):
    class M:                                    # a_{j+1/2} and a_{j-1/2} exchanged
        pass
    mm = M()
    mm.a_node = m.a_node
    mm.a_edge = _c(np.concatenate([m.a_edge[1:], m.a_edge[-1:]]))
    return orc_omega_from_psi(w_init, psi, mm, inv_dx)


def _mut_omega_writes_walls(w_init, psi, m, inv_dx):
    w = orc_omega_from_psi(w_init, psi, m, inv_dx)
    w[0], w[-1] = 0.0, 0.0                      # wall rows overwritten
    return w


def _mut_omega_sign(w_init, psi, m, inv_dx):
    w = orc_omega_from_psi(w_init, psi, m, inv_dx)
    w[1:-1] *= -1                               # omega = +lap psi
    return w


def _mut_adv_uv_neighbour(u, v, w, m, inv_dx):
    return orc_adv_uv(np.roll(u, 1, axis=0), v, w, m, inv_dx)   # u taken one row off


def _mut_adv_uv_no_metric(u, v, w, m, inv_dx):
    J = orc_adv_uv(u, 0 * v, w, m, inv_dx)
    Jv = orc_adv_uv(0 * u, v, w, m, inv_dx)
    return J + Jv / np.where(m.a_node[:, None] == 0, 1, m.a_node[:, None])


def _mut_flux_as_advective(u, v, w, m, inv_dx):
    return orc_adv_uv(u, v, w, m, inv_dx)       # advective instead of flux form


def _mut_flux_half(u, v, w, m, inv_dx):
    J = orc_adv_flux(u, v, w, m, inv_dx)
    Jy = orc_adv_flux(0 * u, v, w, m, inv_dx)
    return J - 0.5 * Jy                         # the y-flux derivative halved (0.5 inv_deta)


def _mut_split_swap23(psi, w, m, inv_dx):
    J1, J2, J3, Ja = orc_split(psi, w, m, inv_dx)
    return J1, J3, J2, Ja                       # J2 and J3 exchanged


def _mut_split_j2_sign(psi, w, m, inv_dx):
    J1, J2, J3, Ja = orc_split(psi, w, m, inv_dx)
    J2 = J2.copy()
    J2[1::2] *= -1                              # a sign error on alternate rows of J2
    return J1, J2, J3, Ja


def _mut_split_jadv_sign(psi, w, m, inv_dx):
    J1, J2, J3, Ja = orc_split(psi, w, m, inv_dx)
    return J1, J2, J3, -Ja


def _mut_split_wrong_metric_row(psi, w, m, inv_dx):
    class M:
        pass
    mm = M()
    mm.a_node = _c(np.roll(m.a_node, 1))       # a_{j-1} in place of a_j
    mm.wt_node = m.wt_node
    return orc_split(psi, w, mm, inv_dx)


def _mut_balance_all_rows(w, m, inv_dx):
    out = orc_balance(w, m, inv_dx)             # wall rows included in the enstrophy sum
    return out + np.array([m.w_node[0] * np.sum(w[0] ** 2) + m.w_node[-1] * np.sum(w[-1] ** 2), 0.0])


def _mut_balance_sign(w, m, inv_dx):
    out = orc_balance(w, m, inv_dx)
    return np.array([out[0], -out[1]])          # eta = -<omega lap omega>


def _mut_balance_no_weight(w, m, inv_dx):
    class M:
        pass
    mm = M()
    mm.a_node, mm.a_edge = m.a_node, m.a_edge
    mm.w_node = _c(np.ones_like(m.w_node))
    return orc_balance(w, mm, inv_dx)


def _mut_input_full_rows(w, m, h, F0, n):
    Ny, Nx = w.shape
    return L().orc_enstrophy_input_mapped_full(P(_c(w)), Nx, Ny, h, F0, n, P(m.y), P(m.w_node))


def _mut_input_mode_off_by_one(w, m, h, F0, n):
    return orc_input(w, m, h, F0, n + 1)        # forcing mode off by one


PLANTS = [
    ("uv_from_psi_mapped", _mut_uv_no_clamp),
    ("uv_from_psi_mapped", _mut_uv_sign),
    ("uv_from_psi_mapped", _mut_uv_no_metric),
    ("omega_from_psi_interior_mapped", _mut_omega_edge_index),
    ("omega_from_psi_interior_mapped", _mut_omega_writes_walls),
    ("omega_from_psi_interior_mapped", _mut_omega_sign),
    ("compute_adv_from_uv_mapped", _mut_adv_uv_neighbour),
    ("compute_adv_from_uv_mapped", _mut_adv_uv_no_metric),
    ("compute_adv_flux_mapped", _mut_flux_as_advective),
    ("compute_adv_flux_mapped", _mut_flux_half),
    ("arakawa_J_mapped_split", _mut_split_swap23),
    ("arakawa_J_mapped_split", _mut_split_j2_sign),
    ("arakawa_J_mapped_split", _mut_split_jadv_sign),
    ("arakawa_J_mapped_split", _mut_split_wrong_metric_row),
    ("enstrophy_diss_balance_mapped", _mut_balance_all_rows),
    ("enstrophy_diss_balance_mapped", _mut_balance_sign),
    ("enstrophy_diss_balance_mapped", _mut_balance_no_weight),
    ("enstrophy_input_mapped", _mut_input_full_rows),
    ("enstrophy_input_mapped", _mut_input_mode_off_by_one),
]


@pytest.mark.parametrize("name,mutant", PLANTS, ids=[f"{n}-{f.__name__}" for n, f in PLANTS])
def test_check_detects_planted_mistake(name, mutant):
    check, _ = CHECKS[name]
    with pytest.raises(AssertionError):
        check(mutant)


# Where each exported oracle function is pinned: test file and test function. The fixture tests
# run orc_run_main (the main.cu restatement) over the 21 reference cases; the functions listed
# under them are the ones that driver calls, so their values appear in the compared rows.
PINNED_BY = {
    "orc_build_y_metrics": [("test_oracle_math.py", "test_grid_metrics_closed_forms"),
                            ("test_oracle_golden.py", "test_t0_row")],
    "orc_table_y_metrics": [("test_tabulated_grid.py", "")],
    "orc_uv_from_psi_mapped": [("test_oracle_pins.py", "check_uv")],
    "orc_omega_from_psi_interior_mapped": [("test_oracle_pins.py", "check_omega_from_psi")],
    "orc_thom_wall_vorticity_mapped": [("test_oracle_math.py", "test_thom_uniform_grid_formula")],
    "orc_arakawa_J_mapped": [("test_oracle_math.py", "test_arakawa_uniform_flow_is_centred_advection"),
                             ("test_oracle_math.py", "test_arakawa_conserves_energy_and_enstrophy_on_mapped_grid")],
    "orc_arakawa_J_mapped_split": [("test_oracle_pins.py", "check_split")],
    "orc_compute_adv_from_uv_mapped": [("test_oracle_pins.py", "check_adv_from_uv")],
    "orc_compute_adv_flux_mapped": [("test_oracle_pins.py", "check_adv_flux")],
    "orc_laplacian_fd_mapped": [("test_oracle_math.py", "test_uniform_laplacian_closed_forms")],
    "orc_fft_z2z": [("test_oracle_math.py", "test_fft_matches_numpy_and_naive_dft")],
    "orc_dft_naive": [("test_oracle_math.py", "test_fft_matches_numpy_and_naive_dft")],
    "orc_set_eval_order": [("test_oracle_math.py", "test_fft_matches_numpy_and_naive_dft"),
                           ("test_oracle_math.py", "test_fft_orders_round_differently"),
                           ("test_oracle_math.py",
                            "test_reversed_elimination_matches_dense_solve_and_rounds_differently")],
    "orc_fill_kx2": [("test_oracle_math.py", "test_kx2_convention")],
    "orc_precompute_cp_mapped": [("test_oracle_math.py", "test_thomas_rounding_floor_at_2049_rows")],
    "orc_poisson_tridiag_mapped": [("test_oracle_math.py", "test_thomas_rounding_floor_at_2049_rows")],
    "orc_poisson_solve": [("test_oracle_math.py", "test_poisson_matches_dense_solve")],
    "orc_poisson_solve_ld": [("test_oracle_math.py", "test_extended_precision_solve_matches_dense_and_fp64")],
    "orc_spectra1d_accumulate": [("test_oracle_math.py", "test_spectra_parseval_and_single_mode")],
    "orc_enstrophy_diss_balance_mapped": [("test_oracle_pins.py", "check_balance")],
    "orc_enstrophy_input_mapped": [("test_oracle_pins.py", "check_enstrophy_input")],
}
VIA_FIXTURE_DRIVER = [
    "orc_run_main", "orc_init_omega", "orc_rhs_mapped", "orc_mms_add_forcing", "orc_axpy", "orc_ssprk2",
    "orc_ssprk3", "orc_enstrophy_diss_mapped", "orc_enstrophy_diss_mapped_full", "orc_enstrophy_sbp_mapped",
    "orc_kinetic_energy_mapped", "orc_energy_dissipation_mapped", "orc_power_input_mapped",
    "orc_enstrophy_input_mapped_full", "orc_row_moments_from_psi_mapped", "orc_mean_wall_omega",
    "orc_mean_wall_omega_omegay_mapped", "orc_omega_rhs_balance_mapped", "orc_max_uv_from_psi_mapped",
]
TOOLS = ["orc_time_steps", "orc_num_threads", "orc_get_eval_order"]   # timing helpers of the CPU baseline, no arithmetic


def test_every_exported_oracle_function_is_pinned():
    import os
    here = os.path.dirname(os.path.abspath(__file__))
    src = open(os.path.join(os.path.dirname(here), "oracle", "nse2d_oracle.cpp")).read()
    for name in oracle.SIG:
        assert (name in PINNED_BY) + (name in VIA_FIXTURE_DRIVER) + (name in TOOLS) == 1, name
        for f, test in PINNED_BY.get(name, []):
            text = open(os.path.join(here, f)).read()
            assert test == "" or f"def {test}(" in text, (name, f, test)
    # the driver reaches each function of the fixture group (called by name inside the oracle)
    for name in VIA_FIXTURE_DRIVER:
        assert src.count(name + "(") >= 2 or name in ("orc_run_main", "orc_init_omega"), name
