"""The allocation-free CFL maxima (launch_max_uv_from_psi_mapped_noalloc) against the allocating
launcher (launch_max_uv_from_psi_mapped, diag.cu:702-866) through the C-ABI: the same kernels, so
the three maxima must agree bit for bit, on two calls in a row (the second on a smaller grid, so
the fixed array holds the first call's partials when it starts), at the largest supported row
count, and with Ny < 3 (zeros, diag.cu:826-832). Ny beyond the fixed array is rejected on the
host with nothing launched."""
import numpy as np
import pytest

import nse2d_hip as nh
import oracle
from gpu_util import Grid, dev, scalars, smooth_random_field, sync
from oracle import P

pytestmark = pytest.mark.gpu

MAX_ROWS = 16385  # NSE2D_NOALLOC_MAX_ROWS


def _maxima(Nx, Ny, seed, beta=2.0):
    """(allocating launcher, allocation-free launcher, oracle) maxima of one seeded psi."""
    print(f"[noalloc] {Nx}x{Ny} seed {seed}", flush=True)
    g = Grid(Nx, Ny, beta)
    psi = smooth_random_field(Nx, Ny, seed, y=g.m.y) + 2.0 * g.m.y[:, None]
    d = dev(psi)
    a, b = scalars(3), scalars(3)
    L = nh.lib()
    L.launch_max_uv_from_psi_mapped(d.ptr, Nx, Ny, g.inv_dx, g.inv_deta, g.dp("a_node"), a.ptr, a.ptr + 8,
                                    a.ptr + 16, None)
    rc = L.launch_max_uv_from_psi_mapped_noalloc(d.ptr, Nx, Ny, g.inv_dx, g.inv_deta, g.dp("a_node"), b.ptr,
                                                 b.ptr + 8, b.ptr + 16, None)
    assert rc == 0, nh.last_error()
    sync()
    ref = np.zeros(3)
    oracle.lib().orc_max_uv_from_psi_mapped(P(np.ascontiguousarray(psi)), Nx, Ny, g.inv_dx, g.inv_deta,
                                            P(g.m.a_node), P(ref))
    return a.download(), b.download(), ref


def test_noalloc_equals_allocating_launcher_on_two_calls():
    for Nx, Ny, seed in ((1024, 513, 1), (256, 129, 2)):
        got_a, got_b, ref = _maxima(Nx, Ny, seed)
        assert np.array_equal(got_a, got_b), (got_a, got_b)
        assert np.all(np.abs(got_b - ref) <= 1e-14 * np.abs(ref)), (got_b, ref)


def test_noalloc_at_the_largest_row_count():
    got_a, got_b, ref = _maxima(64, MAX_ROWS, 3)
    assert np.array_equal(got_a, got_b)
    assert np.all(np.abs(got_b - ref) <= 1e-14 * np.abs(ref))


def test_noalloc_degenerate_grid_gives_zeros():
    g = Grid(64, 3, 2.0)
    d = dev(np.ones((2, 64)))
    b = scalars(3)
    assert nh.lib().launch_max_uv_from_psi_mapped_noalloc(d.ptr, 64, 2, g.inv_dx, g.inv_deta, g.dp("a_node"),
                                                          b.ptr, b.ptr + 8, b.ptr + 16, None) == 0
    sync()
    assert np.all(b.download() == 0.0)


def test_noalloc_rejects_rows_beyond_its_array():
    b = scalars(3)
    rc = nh.lib().launch_max_uv_from_psi_mapped_noalloc(None, 64, MAX_ROWS + 1, 1.0, 1.0, None, b.ptr, b.ptr + 8,
                                                        b.ptr + 16, None)
    assert rc == -1
    assert "exceeds" in nh.last_error()
    sync()
    assert np.all(np.isnan(b.download()))  # nothing was launched
