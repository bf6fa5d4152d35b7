"""Seeded synthetic inputs of the benchmark configurations.

Each function returns ``(syst, out, in_)`` in the reference's calling
convention (cal_smatrix_RGF.m:12-50 / cal_smatrix_RGF_3D.m:12-47).
"""
from __future__ import annotations

import math

import numpy as np


def slab_2d_small(seed: int = 1):
    """2D N=64, 200 slices, eps ~ U[1,4], eps_L = eps_R = 1, lambda = 15 px, periodic, full S."""
    rng = np.random.default_rng(seed)
    eps = 1 + 3 * rng.random((64, 200))
    syst = dict(epsilon=eps, epsilon_L=1.0, epsilon_R=1.0, k0dx=2 * math.pi / 15, yBC="periodic")
    return syst, ["L", "R"], ["L", "R"]


def slab_2d_large(seed: int = 2, nx: int = 2000):
    """2D N=1024, 2000 slices, eps = 1 + 1.25 Bernoulli(0.3) + 0.01i, eps_L = 1, eps_R = 2.25,
    lambda = 10 px, Dirichlet, full S."""
    rng = np.random.default_rng(seed)
    eps = 1 + 1.25 * (rng.random((1024, nx)) < 0.3) + 0.01j
    syst = dict(epsilon=eps, epsilon_L=1.0, epsilon_R=2.25, k0dx=2 * math.pi / 10, yBC="Dirichlet")
    return syst, ["L", "R"], ["L", "R"]


def spheres_3d(seed: int = 3, nx: int = 64, ny: int = 64, nz: int = 128, n_spheres: int = 40):
    """3D N = nx*ny = 4096, nz = 128 slices along z; background eps 1 with 40 spheres of
    radius U[4,10] px and eps 2.1 (centres uniform, wrapped transversally, as the
    transverse BC is Bloch-periodic); Bloch phases U[0, 2pi); lambda = 12 px; full S."""
    rng = np.random.default_rng(seed)
    eps = np.ones((nx, ny, nz))
    X, Y, Z = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    for _ in range(n_spheres):
        c = rng.random(3) * np.array([nx, ny, nz])
        r = 4 + 6 * rng.random()
        dx = np.abs(X - c[0])
        dx = np.minimum(dx, nx - dx)
        dy = np.abs(Y - c[1])
        dy = np.minimum(dy, ny - dy)
        dz = Z - c[2]
        eps[dx * dx + dy * dy + dz * dz < r * r] = 2.1
    phases = 2 * math.pi * rng.random(2)
    syst = dict(epsilon=eps, epsilon_L=1.0, epsilon_R=1.0, k0dx=2 * math.pi / 12, xyBC=[phases[0], phases[1]])
    return syst, ["L", "R"], ["L", "R"]


def batched_sweep(seed: int = 4, count: int = 256, nx: int = 500):
    """256 2D solves, widths in {128,192,256,384,512}, 500 slices, eps ~ U[1,3],
    lambda ~ U[8,20] px with Im(k0dx)/Re(k0dx) ~ U[0,1e-3]; periodic, eps_L = eps_R = 1."""
    rng = np.random.default_rng(seed)
    widths = np.array([128, 192, 256, 384, 512])
    out = []
    for _ in range(count):
        ny = int(rng.choice(widths))
        lam = 8 + 12 * rng.random()
        k0dx = (2 * math.pi / lam) * (1 + 1j * 1e-3 * rng.random())
        eps = 1 + 2 * rng.random((ny, nx))
        out.append((dict(epsilon=eps, epsilon_L=1.0, epsilon_R=1.0, k0dx=k0dx, yBC="periodic"), ["L", "R"], ["L", "R"]))
    return out


def adjoint_2d(seed: int = 5, ny: int = 512, nx: int = 1024):
    """2D N=512, 1024 slices, eps ~ U[1,4], lambda = 15 px, periodic, eps_L = eps_R = 1;
    objective T = sum |t_ij|^2 (inputs L, outputs R)."""
    rng = np.random.default_rng(seed)
    eps = 1 + 3 * rng.random((ny, nx))
    syst = dict(epsilon=eps, epsilon_L=1.0, epsilon_R=1.0, k0dx=2 * math.pi / 15, yBC="periodic")
    return syst, ["R"], ["L"]
