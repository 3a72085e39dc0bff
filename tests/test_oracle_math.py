"""Pins of the oracle that do not come from the reference's fixtures: textbook routines (DFT,
dense linear solve), closed forms (uniform-grid Laplacian, Thom coefficients, Arakawa on a
uniform flow), the discrete Arakawa invariants, and the rounding floor of the serial FP64
elimination at the benchmark's 2049 rows."""
import math

import numpy as np
import pytest

import oracle
from oracle import P

L = oracle.lib


@pytest.mark.parametrize("order", [0, 1], ids=["dit", "dif"])
@pytest.mark.parametrize("N", [2, 4, 8, 16, 64, 256, 1024, 12, 30])
def test_fft_matches_numpy_and_naive_dft(N, order):
    """Both evaluation orders of the oracle's FFT (decimation in time, the default, and in
    frequency, a measurement tool) against numpy and the O(N^2) DFT."""
    rng = np.random.default_rng(N)
    z = rng.normal(size=N) + 1j * rng.normal(size=N)
    for inv in (0, 1):
        buf = np.ascontiguousarray(np.stack([z.real, z.imag], axis=1).ravel())
        L().orc_set_eval_order(order)
        try:
            L().orc_fft_z2z(P(buf), N, 1, inv)
        finally:
            L().orc_set_eval_order(0)
        got = buf[0::2] + 1j * buf[1::2]
        ref = np.fft.ifft(z) * N if inv else np.fft.fft(z)
        assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))
        src = np.ascontiguousarray(np.stack([z.real, z.imag], axis=1).ravel())
        out = np.zeros_like(src)
        L().orc_dft_naive(P(src), P(out), N, inv)
        naive = out[0::2] + 1j * out[1::2]
        assert np.max(np.abs(naive - ref)) <= 1e-13 * np.max(np.abs(ref))


def test_fft_orders_round_differently():
    """The two orders are different FP64 evaluations of one transform (else they could not
    measure how closely two FP64 solvers agree): equal to rounding, not bit for bit."""
    N = 4096
    z = np.random.default_rng(1).normal(size=2 * N)
    a, b = z.copy(), z.copy()
    L().orc_fft_z2z(P(a), N, 1, 0)
    L().orc_set_eval_order(1)
    try:
        L().orc_fft_z2z(P(b), N, 1, 0)
    finally:
        L().orc_set_eval_order(0)
    d = np.max(np.abs(a - b)) / np.max(np.abs(a))
    assert 0 < d < 1e-14, d


def test_reversed_elimination_matches_dense_solve_and_rounds_differently():
    """The elimination from the top wall down (evaluation order flag 2, a measurement tool) solves
    the same operator: against the dense solve on a small grid, and at the benchmark's 2049 rows
    it differs from the reference's bottom-up elimination by rounding only (the FP64 spread of the
    y solve)."""
    Nx, Ny, Lx = 8, 9, 5.0
    m = oracle.Metrics(Ny, 1.0, 1, 2.0)
    w = np.random.default_rng(5).normal(size=(Ny, Nx))
    a = oracle.poisson_solve(w, Lx, m, 0.0, 2.0)
    L().orc_set_eval_order(2)
    try:
        b = oracle.poisson_solve(w, Lx, m, 0.0, 2.0)
    finally:
        L().orc_set_eval_order(0)
    np.testing.assert_allclose(b, a, rtol=0, atol=1e-14 * np.max(np.abs(a)))
    Nx, Ny = 16, 2049
    m = oracle.Metrics(Ny, 1.0, 1, 2.0)
    w = np.random.default_rng(6).normal(size=(Ny, Nx)) + 3 * m.y[:, None]
    a = oracle.poisson_solve(w, 2 * math.pi, m, 0.0, 2.0)
    L().orc_set_eval_order(3)
    try:
        b = oracle.poisson_solve(w, 2 * math.pi, m, 0.0, 2.0)
    finally:
        L().orc_set_eval_order(0)
    d = np.max(np.abs(a - b)) / np.max(np.abs(a))
    print(f"bottom-up vs top-down elimination at 2049 rows: {d:.2e}")
    assert 0 < d < 2e-12


def test_kx2_convention():
    """poisson.cu:19: k <= Nx/2 -> +k (the Nyquist is +Nx/2), k > Nx/2 -> k - Nx."""
    Nx, Lx = 8, 3.0
    kx2 = np.zeros(Nx)
    L().orc_fill_kx2(P(kx2), Nx, Lx)
    kphys = np.array([0, 1, 2, 3, 4, -3, -2, -1]) * 2 * math.pi / Lx
    np.testing.assert_array_equal(kx2, kphys ** 2)


def test_grid_metrics_closed_forms():
    Ny, beta = 65, 2.0
    m = oracle.Metrics(Ny, 1.0, 1, beta)
    eta = -1 + 2 * np.arange(Ny) / (Ny - 1)
    np.testing.assert_allclose(m.y, np.tanh(beta * eta) / np.tanh(beta), rtol=0, atol=1e-15)
    y_eta = beta / np.cosh(beta * eta) ** 2 / np.tanh(beta)
    np.testing.assert_allclose(m.a_node, 1 / y_eta, rtol=1e-14)
    np.testing.assert_allclose(m.a_edge, 0.5 * (m.a_node[1:] + m.a_node[:-1]), rtol=0, atol=0)
    assert abs(np.sum(m.w_node) - 2.0) < 1e-14
    u = oracle.Metrics(33, 1.0, 0, 0.0)
    np.testing.assert_allclose(u.dy_edge, 2.0 / 32, rtol=1e-15)
    np.testing.assert_allclose(u.a_node, 1.0, rtol=1e-15)  # y_eta = dy / deta = 1 for h = 1
    assert u.w_node[0] == 0.5 * u.w_node[1]


def test_uniform_laplacian_closed_forms():
    """delta_xx sin(kx) = -(4/dx^2) sin^2(k dx/2) sin(kx); mapped delta_yy of y^2 on a uniform
    grid = 2; wall rows are zero (ops.cu:483)."""
    Nx, Ny = 32, 17
    m = oracle.Metrics(Ny, 1.0, 0, 0.0)
    dx = 2 * math.pi / Nx
    x = np.arange(Nx) * dx
    k = 3
    q = np.tile(np.sin(k * x), (Ny, 1)) + (m.y ** 2)[:, None]
    out = np.zeros_like(q)
    inv_deta = (Ny - 1) / 2.0
    L().orc_laplacian_fd_mapped(P(q), P(out), Nx, Ny, 1 / dx ** 2, inv_deta ** 2, P(m.a_node), P(m.a_edge))
    ref = -(4 / dx ** 2) * math.sin(k * dx / 2) ** 2 * np.sin(k * x)[None, :] + 2.0
    np.testing.assert_allclose(out[1:-1], np.broadcast_to(ref, (Ny - 2, Nx)), rtol=0, atol=1e-11)
    assert np.all(out[[0, -1]] == 0.0)


def test_arakawa_uniform_flow_is_centred_advection():
    """For psi = U y (u = U, v = 0) every Arakawa term reduces to U delta_x omega, so
    J_adv = U (omega_{i+1} - omega_{i-1}) / (2 dx); this fixes the signs of J1, J2, J3."""
    Nx, Ny, U = 16, 9, 0.7
    m = oracle.Metrics(Ny, 1.0, 0, 0.0)
    dx = 2 * math.pi / Nx
    x = np.arange(Nx) * dx
    psi = np.ascontiguousarray(np.tile((U * m.y)[:, None], (1, Nx)))
    w = np.ascontiguousarray(np.tile(np.sin(2 * x) + 0.3 * np.cos(5 * x), (Ny, 1)))
    J = np.zeros_like(w)
    L().orc_arakawa_J_mapped(P(psi), P(w), P(J), Nx, Ny, 1 / dx, (Ny - 1) / 2.0, P(m.a_node))
    ref = U * (np.roll(w, -1, axis=1) - np.roll(w, 1, axis=1)) / (2 * dx)
    np.testing.assert_allclose(J[1:-1], ref[1:-1], rtol=0, atol=1e-13)
    assert np.all(J[[0, -1]] == 0.0)


def test_arakawa_conserves_energy_and_enstrophy_on_mapped_grid():
    """With the operator weights wt_j = deta / a_j the mapped Arakawa Jacobian is exactly
    antisymmetric: sum wt psi J = sum wt omega J = 0 for fields that vanish near the walls."""
    Nx, Ny = 48, 41
    m = oracle.Metrics(Ny, 1.0, 1, 2.5)
    rng = np.random.default_rng(3)
    psi, w = rng.normal(size=(Ny, Nx)), rng.normal(size=(Ny, Nx))
    for f in (psi, w):
        f[[0, 1, -2, -1]] = 0.0
    J = np.zeros_like(w)
    L().orc_arakawa_J_mapped(P(psi), P(w), P(J), Nx, Ny, Nx / (2 * math.pi), (Ny - 1) / 2.0, P(m.a_node))
    wt = m.wt_node[:, None]
    scale = np.sum(wt * np.abs(w * J))
    assert abs(np.sum(wt * w * J)) <= 1e-13 * scale
    assert abs(np.sum(wt * psi * J)) <= 1e-13 * np.sum(wt * np.abs(psi * J))


def test_thom_uniform_grid_formula():
    """On a uniform grid the second-order closure is omega_0 = -(8 psi_1 - 7 psi_0 - psi_2) / (2 dy^2)."""
    Nx, Ny = 8, 11
    m = oracle.Metrics(Ny, 1.0, 0, 0.0)
    rng = np.random.default_rng(4)
    psi = rng.normal(size=(Ny, Nx))
    w = np.zeros_like(psi)
    pb, pt = 0.25, 1.5
    L().orc_thom_wall_vorticity_mapped(P(w), P(psi), Nx, Ny, P(m.dy_edge), pb, pt, 0)
    dy = m.dy_edge[0]
    np.testing.assert_allclose(w[0], -(8 * psi[1] - 7 * pb - psi[2]) / (2 * dy * dy), rtol=1e-12)
    np.testing.assert_allclose(w[-1], -(8 * psi[-2] - 7 * pt - psi[-3]) / (2 * dy * dy), rtol=1e-12)
    L().orc_thom_wall_vorticity_mapped(P(w), P(psi), Nx, Ny, P(m.dy_edge), pb, pt, 1)
    assert np.all(w[[0, -1]] == 0.0)


def test_poisson_matches_dense_solve():
    """The spectral-x / FD-y operator built as a dense matrix and solved with numpy."""
    Nx, Ny, Lx = 8, 9, 5.0
    m = oracle.Metrics(Ny, 1.0, 1, 2.0)
    rng = np.random.default_rng(5)
    w = rng.normal(size=(Ny, Nx))
    pb, pt = 0.0, 2.0
    psi = oracle.poisson_solve(w, Lx, m, pb, pt)
    k = np.array([j if j <= Nx // 2 else j - Nx for j in range(Nx)]) * 2 * math.pi / Lx
    F = np.fft.fft(np.eye(Nx), axis=0)
    Dxx = np.real(np.linalg.inv(F) @ np.diag(-k ** 2) @ F)
    inv_deta2 = ((Ny - 1) / 2.0) ** 2
    n = (Ny - 2) * Nx
    A = np.zeros((n, n))
    b = np.zeros(n)
    for r in range(Ny - 2):
        j = r + 1
        asub = m.a_node[j] * m.a_edge[j - 1] * inv_deta2
        csup = m.a_node[j] * m.a_edge[j] * inv_deta2
        sl = slice(r * Nx, (r + 1) * Nx)
        A[sl, sl] = Dxx - (asub + csup) * np.eye(Nx)
        b[sl] = -w[j]
        if r > 0:
            A[sl, (r - 1) * Nx:r * Nx] = asub * np.eye(Nx)
        else:
            b[sl] -= asub * pb
        if r < Ny - 3:
            A[sl, (r + 1) * Nx:(r + 2) * Nx] = csup * np.eye(Nx)
        else:
            b[sl] -= csup * pt
    dense = np.linalg.solve(A, b).reshape(Ny - 2, Nx)
    np.testing.assert_allclose(psi[1:-1], dense, rtol=0, atol=1e-13 * np.max(np.abs(dense)))
    np.testing.assert_allclose(psi[0], pb, atol=1e-15)
    np.testing.assert_allclose(psi[-1], pt, atol=1e-15)


def _thomas_longdouble(a, c, kk, f):
    """The elimination of poisson.cu:70-133 in 80-bit extended precision (measurement only)."""
    a, c = a.astype(np.longdouble), c.astype(np.longdouble)
    fr, fi = f.real.astype(np.longdouble), f.imag.astype(np.longdouble)
    n = len(f)
    b = -(a + c) - np.longdouble(kk)
    cp = np.zeros(n, np.longdouble)
    dr, di = np.zeros(n, np.longdouble), np.zeros(n, np.longdouble)
    cp[0], dr[0], di[0] = c[0] / b[0], fr[0] / b[0], fi[0] / b[0]
    for r in range(1, n):
        den = b[r] - a[r] * cp[r - 1]
        cp[r] = c[r] / den
        dr[r], di[r] = (fr[r] - a[r] * dr[r - 1]) / den, (fi[r] - a[r] * di[r - 1]) / den
    xr, xi = dr.copy(), di.copy()
    for r in range(n - 2, -1, -1):
        xr[r], xi[r] = dr[r] - cp[r] * xr[r + 1], di[r] - cp[r] * xi[r + 1]
    return xr.astype(float) + 1j * xi.astype(float)


def test_thomas_rounding_floor_at_2049_rows():
    """The serial FP64 elimination (the reference's, restated by the oracle) against an
    extended-precision run of the same elimination on the tanh beta=2 grid with 2049 rows: the
    FP64 rounding floor for the kx ~ 0 modes is of order 1e-13, which bounds how closely any two
    FP64 solvers can agree at this size."""
    Ny, Nx, Lx = 2049, 4, 2 * math.pi
    m = oracle.Metrics(Ny, 1.0, 1, 2.0)
    inv_deta2 = ((Ny - 1) / 2.0) ** 2
    a = m.a_node[1:-1] * m.a_edge[:-1] * inv_deta2
    c = m.a_node[1:-1] * m.a_edge[1:] * inv_deta2
    kx2 = np.zeros(Nx)
    L().orc_fill_kx2(P(kx2), Nx, Lx)
    cp = np.zeros(Nx * (Ny - 2))
    L().orc_precompute_cp_mapped(P(cp), P(m.a_node), P(m.a_edge), P(kx2), Nx, Ny, inv_deta2)
    rng = np.random.default_rng(6)
    yy = m.y[1:-1]
    worst = 0.0
    for k in (0, 1):
        f = (np.cos(3 * yy) + 0.3 * rng.normal(size=Ny - 2)) * (1 + 0.5j)
        hat = np.zeros((Ny, Nx), complex)
        hat[1:-1, k] = -f  # the tridiagonal kernel solves with -omega_hat on the right
        buf = np.ascontiguousarray(np.stack([hat.real, hat.imag], axis=-1).ravel())
        L().orc_poisson_tridiag_mapped(P(buf), P(cp), P(m.a_node), P(m.a_edge), P(kx2), Nx, Ny, inv_deta2, 0.0, 0.0)
        got = buf.reshape(Ny, Nx, 2)[1:-1, k]
        got = got[:, 0] + 1j * got[:, 1]
        exact = _thomas_longdouble(a, c, kx2[k], f)
        err = np.max(np.abs(got - exact)) / np.max(np.abs(exact))
        worst = max(worst, err)
    print(f"FP64 elimination vs extended precision at 2049 rows: {worst:.2e}")
    assert 1e-15 < worst < 2e-12


def test_extended_precision_solve_matches_dense_and_fp64():
    """orc_poisson_solve_ld (80-bit; used to measure the FP64 floor of whole runs) solves the same
    operator: against the dense solve and against the FP64 solve on a small grid."""
    Nx, Ny, Lx = 8, 9, 5.0
    m = oracle.Metrics(Ny, 1.0, 1, 2.0)
    w = np.random.default_rng(5).normal(size=(Ny, Nx))
    a = oracle.poisson_solve(w, Lx, m, 0.0, 2.0)
    b = oracle.poisson_solve(w, Lx, m, 0.0, 2.0, extended=True)
    np.testing.assert_allclose(b, a, rtol=0, atol=1e-14 * np.max(np.abs(a)))
    Nx, Ny = 256, 129
    m = oracle.Metrics(Ny, 1.0, 1, 2.0)
    w = np.random.default_rng(6).normal(size=(Ny, Nx))
    a = oracle.poisson_solve(w, 2 * math.pi, m, 0.0, 2.0)
    b = oracle.poisson_solve(w, 2 * math.pi, m, 0.0, 2.0, extended=True)
    assert np.max(np.abs(a - b)) <= 1e-13 * np.max(np.abs(a))


def test_spectra_parseval_and_single_mode():
    """Spectra1D (diag.cuh:123-191): sum_k E_u(k) is the mean of u^2 over the grid points
    (Parseval with the half-spectrum weights), and u = cos(2 pi m x / Lx) puts 1/2 at k = m."""
    Nx, Ny = 64, 9
    rng = np.random.default_rng(7)
    u, v = rng.normal(size=(Ny, Nx)), rng.normal(size=(Ny, Nx))
    acc = np.zeros(2 * (Nx // 2 + 1))
    L().orc_spectra1d_accumulate(P(u), P(v), Nx, Ny, P(acc))
    Eu, Ev = acc[:Nx // 2 + 1], acc[Nx // 2 + 1:]
    assert abs(Eu.sum() - np.mean(u * u)) <= 1e-14 * np.mean(u * u)
    assert abs(Ev.sum() - np.mean(v * v)) <= 1e-14 * np.mean(v * v)
    m = 5
    x = np.arange(Nx) * (3.0 / Nx)
    c = np.ascontiguousarray(np.tile(np.cos(2 * math.pi * m * x / 3.0), (Ny, 1)))
    acc[:] = 0
    L().orc_spectra1d_accumulate(P(c), P(c), Nx, Ny, P(acc))
    Eu = acc[:Nx // 2 + 1]
    assert abs(Eu[m] - 0.5) <= 1e-14 and np.max(np.abs(np.delete(Eu, m))) <= 1e-14
