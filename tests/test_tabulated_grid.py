"""Tabulated y-grids (--stretch table, grid.cu:46-78). The reference reads y from an HDF5 dataset
/y; this build reads the same Ny values as text (no HDF5 library in the image). The oracle's
restatement is pinned by closed forms, and the library's host code must give the oracle's
metrics bit for bit."""
import ctypes
import os

import numpy as np
import pytest

import nse2d_hip as nh
import oracle


def eta(Ny):
    return -1.0 + 2.0 * np.arange(Ny) / (Ny - 1)


def test_uniform_table_gives_the_uniform_metrics():
    """A uniform table: every difference quotient is dy/deta, so the metrics equal the uniform
    branch's (grid.cu:10-27) up to the rounding of the tabulated values."""
    Ny, h = 33, 1.0
    u = oracle.Metrics(Ny, h, stretch=0)
    y, a_node, a_edge, dy_edge, w_node = oracle.table_metrics(u.y)
    np.testing.assert_array_equal(y, u.y)
    np.testing.assert_allclose(a_node, u.a_node, rtol=1e-13)
    np.testing.assert_allclose(a_edge, u.a_edge, rtol=1e-13)
    np.testing.assert_allclose(dy_edge, u.dy_edge, rtol=1e-13)
    np.testing.assert_allclose(w_node, u.w_node, rtol=1e-13)


@pytest.mark.parametrize("c", [0.2, -0.35])
def test_quadratic_table_closed_forms(c):
    """y = h (eta + c (eta^2 - 1)) (monotone for |c| < 1/2, y(-+1) = -+h): the centred difference
    of a quadratic is exact, a_node = 1 / (h (1 + 2 c eta)) inside; the one-sided ends give
    1 / (h (1 + c (eta_0 + eta_1))) and 1 / (h (1 + c (eta_{Ny-2} + eta_{Ny-1}))). The edge values
    are the node means, dy_edge the differences, and the trapezoid weights sum to 2h."""
    Ny, h = 41, 1.3
    e = eta(Ny)
    yt = h * (e + c * (e * e - 1.0))
    y, a_node, a_edge, dy_edge, w_node = oracle.table_metrics(yt)
    inner = 1.0 / (h * (1.0 + 2.0 * c * e[1:-1]))
    np.testing.assert_allclose(a_node[1:-1], inner, rtol=1e-12)
    np.testing.assert_allclose(a_node[0], 1.0 / (h * (1.0 + c * (e[0] + e[1]))), rtol=1e-12)
    np.testing.assert_allclose(a_node[-1], 1.0 / (h * (1.0 + c * (e[-2] + e[-1]))), rtol=1e-12)
    np.testing.assert_allclose(a_edge, 0.5 * (a_node[:-1] + a_node[1:]), rtol=0, atol=0)
    np.testing.assert_array_equal(dy_edge, np.diff(yt))
    assert abs(w_node.sum() - 2.0 * h) <= 1e-13
    assert w_node[0] == 0.5 * dy_edge[0] and w_node[-1] == 0.5 * dy_edge[-1]


def _table_file(tmp_path, y):
    path = os.path.join(str(tmp_path), "ytable.txt")
    with open(path, "w") as f:
        f.write("\n".join(repr(float(v)) for v in y) + "\n")
    return path


def test_library_host_metrics_match_oracle(tmp_path):
    """nse2d_host_metrics(Ny, h, "table:<file>", ...) reads the text table and gives the oracle's
    metrics of the same values bit for bit (the same differences, in the same order)."""
    Ny, h = 65, 1.0
    e = eta(Ny)
    yt = h * (e + 0.3 * np.sin(np.pi * e) / np.pi)  # y' = h (1 + 0.3 cos(pi eta)) > 0
    path = _table_file(tmp_path, yt)
    got = [np.zeros(Ny), np.zeros(Ny), np.zeros(Ny - 1), np.zeros(Ny - 1), np.zeros(Ny)]
    rc = nh.lib().nse2d_host_metrics(Ny, h, ("table:" + path).encode(), 0.0, *[a.ctypes.data for a in got])
    assert rc == 0, nh.last_error()
    want = oracle.table_metrics(yt)
    for name, g, w in zip(("y", "a_node", "a_edge", "dy_edge", "w_node"), got, want):
        assert np.array_equal(g, w), name


def test_library_host_metrics_table_errors(tmp_path):
    """A missing file or a table with fewer than Ny values is an error with a message."""
    Ny = 17
    out = [np.zeros(Ny), np.zeros(Ny), np.zeros(Ny - 1), np.zeros(Ny - 1), np.zeros(Ny)]
    rc = nh.lib().nse2d_host_metrics(Ny, 1.0, b"table:/nonexistent/ytable.txt", 0.0, *[a.ctypes.data for a in out])
    assert rc != 0 and "ytable" in nh.last_error()
    path = _table_file(tmp_path, np.linspace(-1.0, 1.0, Ny - 3))
    rc = nh.lib().nse2d_host_metrics(Ny, 1.0, ("table:" + path).encode(), 0.0, *[a.ctypes.data for a in out])
    assert rc != 0 and "fewer than Ny" in nh.last_error()
