"""Helpers of the GPU parity tests: seeded inputs, device round trips through the C-ABI and
error norms. GPU calls go through nse2d_hip (libnse2d_hip.so); nothing here falls back to the CPU."""
import math

import numpy as np

import nse2d_hip as nh
import oracle


def rel_err(got, ref):
    """Normwise relative error max(||g-r||_2/||r||_2, ||g-r||_inf/||r||_inf)."""
    got, ref = np.asarray(got, dtype=float), np.asarray(ref, dtype=float)
    d = got - ref
    n2, ni = np.linalg.norm(ref.ravel()), np.max(np.abs(ref))
    if n2 == 0:
        return float(np.max(np.abs(d)))
    return max(float(np.linalg.norm(d.ravel()) / n2), float(np.max(np.abs(d)) / ni))


def smooth_random_field(Nx, Ny, seed, kmax=6, scale=1.0, y=None):
    """Band-limited random field (sum of a few Fourier x sine-y modes plus white noise)."""
    rng = np.random.default_rng(seed)
    x = np.arange(Nx) * 2 * math.pi / Nx
    yy = np.linspace(-1, 1, Ny) if y is None else y
    f = np.zeros((Ny, Nx))
    for _ in range(8):
        kx, ky = rng.integers(0, kmax + 1), rng.integers(1, kmax + 1)
        f += rng.normal() * np.outer(np.sin(0.5 * ky * math.pi * (yy + 1) + rng.uniform(0, 1)),
                                     np.cos(kx * x + rng.uniform(0, 2 * math.pi)))
    f += 1e-2 * rng.normal(size=(Ny, Nx))
    return scale * f


class Grid:
    """Metrics (oracle-built, identical to the driver's) on host and device."""

    def __init__(self, Nx, Ny, beta=2.0, Lx=2 * math.pi, h=1.0):
        self.Nx, self.Ny, self.Lx, self.h = Nx, Ny, Lx, h
        self.m = oracle.Metrics(Ny, h, 1 if beta > 0 else 0, beta)
        self.dx = Lx / Nx
        self.inv_dx = 1.0 / self.dx
        self.inv_dx2 = self.inv_dx ** 2
        self.inv_deta = (Ny - 1) / 2.0
        self.inv_deta2 = self.inv_deta ** 2
        self.d = {k: nh.DeviceBuffer.from_array(getattr(self.m, k))
                  for k in ("y", "a_node", "a_edge", "dy_edge", "w_node", "wt_node")}

    def dp(self, k):
        return self.d[k].ptr


def dev(a):
    return nh.DeviceBuffer.from_array(np.ascontiguousarray(a, dtype=np.float64))


def zeros_dev(n):
    return nh.DeviceBuffer(n, np.zeros(n))


def scalars(n):
    return nh.DeviceBuffer(n, np.full(n, np.nan))


def sync():
    nh.lib().nse2d_device_sync()
