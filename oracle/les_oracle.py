"""CPU oracle of the LES filter (TEST INFRASTRUCTURE ONLY: the checker of the filter in
nse2d-channel-gpu_amd/csrc/les.hip; never used by the product path). No reference counterpart:
the reference solver has no LES. Parity of this restatement is pinned by closed forms in
tests/test_les_oracle.py (constants, single Fourier modes on both sides of the cut, linear
profiles in eta, a brute-force DFT of the cut).

F = sharp spectral cut in x after a box filter in mapped y:
  ybox:   coarse row J = (1/ry) [f(ry J - ry/2)/2 + sum_{|m| < ry/2} f(ry J + m) + f(ry J + ry/2)/2]
          (trapezoid rule over ry fine cells of eta centred on the coarse node), wall rows kept;
  xsharp: keep the Fourier modes |k| < Nx_c/2 of each row (Nx_c = Nx_f/rx) and sample the
          result at every rx-th fine node.
The GPU takes another route for the x step (a truncated half spectrum inverted at Nx_c); here it
is numpy's rfft/irfft at the fine size.
"""
import numpy as np


def ybox(f, ry):
    """[Ny_f][Nx] -> [Ny_c][Nx], Ny_f - 1 = ry (Ny_c - 1), ry = 1 or even."""
    f = np.asarray(f, dtype=np.float64)
    if ry == 1:
        return f.copy()
    assert ry % 2 == 0 and (f.shape[0] - 1) % ry == 0
    nyc = (f.shape[0] - 1) // ry + 1
    h = ry // 2
    w = np.ones(ry + 1)
    w[0] = w[-1] = 0.5
    out = np.empty((nyc, f.shape[1]))
    out[0], out[-1] = f[0], f[-1]
    for J in range(1, nyc - 1):
        out[J] = (w @ f[ry * J - h: ry * J + h + 1]) / ry
    return out


def xsharp(f, rx):
    """[rows][Nx_f] -> [rows][Nx_f / rx]: modes |k| < Nx_c / 2 kept, every rx-th node."""
    f = np.asarray(f, dtype=np.float64)
    nxf = f.shape[1]
    nxc = nxf // rx
    F = np.fft.rfft(f, axis=1)
    F[:, nxc // 2:] = 0.0
    return np.fft.irfft(F, n=nxf, axis=1)[:, ::rx]


def xsharp_dft(f, rx):
    """xsharp by the explicit sum over |k| < Nx_c/2 (O(N^2); tiny sizes only)."""
    f = np.asarray(f, dtype=np.float64)
    nxf = f.shape[1]
    nxc = nxf // rx
    n = np.arange(nxf)
    out = np.zeros((f.shape[0], nxc))
    xs = np.arange(nxc) * rx
    for k in range(-(nxc // 2) + 1, nxc // 2):
        Xk = f @ np.exp(-2j * np.pi * k * n / nxf)
        out += np.real(np.outer(Xk, np.exp(2j * np.pi * k * xs / nxf))) / nxf
    return out


def filter_field(f, rx, ry):
    return xsharp(ybox(f, ry), rx)
