"""Pins of the LES filter oracle (oracle/les_oracle.py) by closed forms: no reference fixture
exists for it (the reference has no LES)."""
import math
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import les_oracle as lo  # noqa: E402


def eta(n):
    return -1.0 + 2.0 * np.arange(n) / (n - 1)


@pytest.mark.parametrize("rx,ry", [(4, 4), (8, 8), (2, 1), (1, 2)])
def test_constant_is_kept(rx, ry):
    f = np.full((8 * 4 + 1, 64), 3.25)
    out = lo.filter_field(f, rx, ry)
    assert out.shape == ((f.shape[0] - 1) // ry + 1, 64 // rx)
    assert np.max(np.abs(out - 3.25)) < 1e-14


@pytest.mark.parametrize("rx", [2, 4, 8])
def test_modes_below_the_cut_pass_and_above_vanish(rx):
    nxf, ny = 128, 9
    nxc = nxf // rx
    x = np.arange(nxf) * 2 * math.pi / nxf
    for k in range(0, nxf // 2 + 1):
        f = np.tile(np.cos(k * x + 0.3), (ny, 1))
        out = lo.xsharp(f, rx)
        want = np.tile(np.cos(k * x[::rx] + 0.3), (ny, 1)) if k < nxc // 2 else np.zeros((ny, nxc))
        assert np.max(np.abs(out - want)) < 1e-13, (rx, k)


def test_cut_matches_the_brute_force_sum():
    rng = np.random.default_rng(3)
    f = rng.normal(size=(5, 64))
    for rx in (2, 4, 8):
        assert np.max(np.abs(lo.xsharp(f, rx) - lo.xsharp_dft(f, rx))) < 1e-13


@pytest.mark.parametrize("ry", [2, 4, 8])
def test_box_is_exact_on_linear_profiles_and_averages_curvature(ry):
    nyc = 9
    nyf = ry * (nyc - 1) + 1
    e = eta(nyf)
    f = np.tile((2.0 - 3.0 * e)[:, None], (1, 4))
    out = lo.ybox(f, ry)
    assert np.max(np.abs(out - (2.0 - 3.0 * eta(nyc))[:, None])) < 1e-14
    # e^2: trapezoid average over [eta_J - D/2, eta_J + D/2] with D = ry d_eta is eta_J^2 plus the
    # box average of (eta - eta_J)^2 (D^2/12) plus the trapezoid error (d_eta^2 / 6)
    g = np.tile((e * e)[:, None], (1, 4))
    out = lo.ybox(g, ry)
    de = 2.0 / (nyf - 1)
    D = ry * de
    want = eta(nyc) ** 2 + D * D / 12.0 + de * de / 6.0
    want[0], want[-1] = 1.0, 1.0
    assert np.max(np.abs(out[:, 0] - want)) < 1e-14


def test_a_planted_mistake_fails_the_pins():
    """Dropping the trapezoid half weights changes the linear-profile result: the pins see it."""
    ry, nyc = 4, 9
    nyf = ry * (nyc - 1) + 1
    e = eta(nyf)
    f = np.tile((2.0 - 3.0 * e)[:, None], (1, 4))
    bad = np.empty((nyc, 4))
    bad[0], bad[-1] = f[0], f[-1]
    for J in range(1, nyc - 1):
        bad[J] = f[ry * J - 2: ry * J + 3].sum(axis=0) / ry
    assert np.max(np.abs(bad - (2.0 - 3.0 * eta(nyc))[:, None])) > 1e-3
