"""GPU parity of the adjoint transmission gradient (rgf_transmission_gradient).

Checks, for the FP64 adjoint and the mixed-precision one (FP32 slice
inverses, forward fields and adjoint rows refined in FP64): against the
oracle's sparse-LU adjoint (independent of the slice recursion) at small
sizes, checkpoint-interval invariance, central finite differences (h = 1e-6)
at 20 seeded pixels on the 2D N=512, 1024-slice workload with the cost
bound (gradient <= 3 forward solves), and the torch.autograd.Function.
"""
import time

import numpy as np
import pytest

import oracle.rgf_oracle as o
import rgf_amd

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "fp32_refined"])
@pytest.mark.parametrize("bc", ["periodic", "Dirichlet", "Neumann", 0.8])
@pytest.mark.parametrize("direction", ["LR", "RL"])
def test_gradient_vs_oracle_2d(bc, direction, precision):
    rng = np.random.default_rng(len(str(bc)) + (0 if direction == "LR" else 100))
    ny, nx = 21, 13
    eps = 1 + 2 * rng.random((ny, nx)) + 0.03j * rng.random((ny, nx))
    syst = dict(epsilon=eps, epsilon_L=1.0, epsilon_R=1.7, k0dx=1.05 + 0.01j, yBC=bc)
    if direction == "LR":
        T, gre, gim = rgf_amd.transmission_gradient(syst, ["R"], ["L"], checkpoint_every=4, precision=precision)
        To, gro, gio = o.transmission_gradient(syst)
    else:
        T, gre, gim = rgf_amd.transmission_gradient(syst, ["L"], ["R"], checkpoint_every=4, precision=precision)
        m = dict(syst, epsilon=eps[:, ::-1], epsilon_L=1.7, epsilon_R=1.0)
        To, gro, gio = o.transmission_gradient(m)
        gro, gio = gro[:, ::-1], gio[:, ::-1]
    scale = max(np.abs(gro).max(), np.abs(gio).max())
    assert abs(T - To) < 1e-10 * max(1.0, To)
    assert np.abs(gre - gro).max() < 1e-9 * scale
    assert np.abs(gim - gio).max() < 1e-9 * scale


@pytest.mark.parametrize("precision", ["fp64", "fp32_refined"])
def test_gradient_vs_oracle_3d(precision):
    # a Bloch slab: H is not symmetric, the adjoint needs the transposed stencil
    rng = np.random.default_rng(7)
    eps = 1 + rng.random((4, 5, 6))
    syst = dict(epsilon=eps, epsilon_L=1.0, epsilon_R=1.2, k0dx=1.2, xyBC=[0.4, 1.9])
    T, gre, gim = rgf_amd.transmission_gradient(syst, ["R"], ["L"], three_d=True, checkpoint_every=2,
                                                precision=precision)
    To, gro, gio = o.transmission_gradient(syst, three_d=True)
    assert abs(T - To) < 1e-10 * max(1, To)
    assert np.abs(gre - gro).max() < 1e-9 * np.abs(gro).max()


def test_channel_subset_and_checkpoint_invariance():
    rng = np.random.default_rng(8)
    eps = 1 + 3 * rng.random((40, 29))
    syst = dict(epsilon=eps, epsilon_L=1.0, epsilon_R=1.0, k0dx=2 * np.pi / 9, yBC="periodic")
    sel_in, sel_out = {"ind_in_L": [2, 1, 4]}, {"ind_out_R": [3, 5]}
    ref = None
    for every in (0, 1, 3, 7, 32):
        T, gre, _ = rgf_amd.transmission_gradient(syst, sel_out, sel_in, checkpoint_every=every)
        if ref is None:
            ref = (T, gre)
            S, _ = rgf_amd.cal_smatrix_RGF(syst, sel_out, sel_in)
            assert abs(T - np.sum(np.abs(S) ** 2)) < 1e-11
        else:
            assert abs(T - ref[0]) < 1e-12 and np.abs(gre - ref[1]).max() < 1e-12


@pytest.fixture(scope="module")
def cfg5_fd():
    """The adjoint workload (2D N=512, 1024 slices, eps~U[1,4], lambda=15,
    periodic) and central differences (h = 1e-6, FP64 solves) of T at 20
    seeded pixels, computed once for both precisions."""
    from rgf_amd.workloads import adjoint_2d
    syst, out, inn = adjoint_2d()
    rng = np.random.default_rng(20)
    ny, nx = syst["epsilon"].shape
    pix = [(int(rng.integers(ny)), int(rng.integers(nx))) for _ in range(20)]
    h = 1e-6
    fd = []
    for (m, n) in pix:
        vals = []
        for sgn in (1, -1):
            e = syst["epsilon"].copy()
            e[m, n] += sgn * h
            Sp, _ = rgf_amd.cal_smatrix_RGF(dict(syst, epsilon=e), out, inn)
            vals.append(np.sum(np.abs(Sp) ** 2))
        fd.append((vals[0] - vals[1]) / (2 * h))
    T64, g64, _ = rgf_amd.transmission_gradient(syst, out, inn, checkpoint_every=32)
    return syst, out, inn, pix, np.array(fd), T64, g64


@pytest.mark.parametrize("precision", ["fp64", "fp32_refined"])
def test_adjoint_workload_fd_20_pixels_and_cost(cfg5_fd, precision):
    """Adjoint vs central FD at 20 pixels, the mixed-precision gradient vs the
    FP64 one, and the cost: gradient <= 3 forward solves of the same precision."""
    syst, out, inn, pix, fd, T64, g64 = cfg5_fd
    rgf_amd.cal_smatrix_RGF(syst, out, inn, precision=precision)  # warm-up (workspace)
    rgf_amd.transmission_gradient(syst, out, inn, checkpoint_every=32, precision=precision)
    t0 = time.perf_counter()
    S, _ = rgf_amd.cal_smatrix_RGF(syst, out, inn, precision=precision)
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    T, gre, _ = rgf_amd.transmission_gradient(syst, out, inn, checkpoint_every=32, precision=precision)
    t_grad = time.perf_counter() - t0
    assert abs(T - np.sum(np.abs(S) ** 2)) < 1e-9 * T
    assert abs(T - T64) < 1e-9 * T
    for k, (m, n) in enumerate(pix):
        # FD truncation O(h^2) plus cancellation ~ 1e-16 * T / h
        assert abs(fd[k] - gre[m, n]) < 1e-7 * max(1.0, T) + 1e-5 * abs(gre[m, n]), (m, n, fd[k], gre[m, n])
    assert np.abs(gre - g64).max() < 1e-9 * np.abs(g64).max()
    print(f"{precision}: forward {t_fwd:.3f} s, gradient {t_grad:.3f} s, ratio {t_grad / t_fwd:.2f}", flush=True)
    assert t_grad <= 3 * t_fwd


def test_autograd_forward_without_gradient_is_one_solve():
    torch = pytest.importorskip("torch")
    from rgf_amd.autograd import transmission
    rng = np.random.default_rng(10)
    e0 = 1 + 2 * rng.random((24, 15))
    eps = torch.tensor(e0, dtype=torch.float64, device="cuda")  # requires_grad False
    T = transmission(eps, 1.0, 1.0, 1.1, "periodic")
    S, _ = rgf_amd.cal_smatrix_RGF(dict(epsilon=e0, epsilon_L=1.0, epsilon_R=1.0, k0dx=1.1, yBC="periodic"),
                                   ["R"], ["L"])
    assert abs(T.item() - np.sum(np.abs(S) ** 2)) < 1e-12
    epsg = eps.clone().requires_grad_(True)
    T2 = transmission(epsg, 1.0, 1.0, 1.1, "periodic", precision="fp32_refined")
    T2.backward()
    assert abs(T2.item() - T.item()) < 1e-10 and epsg.grad is not None


def test_torch_autograd_function():
    torch = pytest.importorskip("torch")
    from rgf_amd.autograd import transmission
    rng = np.random.default_rng(9)
    e0 = 1 + 2 * rng.random((24, 15)) + 0.02j * rng.random((24, 15))
    eps = torch.tensor(e0, dtype=torch.complex128, device="cuda", requires_grad=True)
    T = transmission(eps, 1.0, 1.3, 1.1, "Dirichlet")
    T.backward()
    To, gro, gio = o.transmission_gradient(dict(epsilon=e0, epsilon_L=1.0, epsilon_R=1.3, k0dx=1.1, yBC="Dirichlet"))
    assert abs(T.item() - To) < 1e-10
    g = eps.grad.cpu().numpy()
    assert np.abs(g.real - gro).max() < 1e-9 * np.abs(gro).max()
    assert np.abs(g.imag - gio).max() < 1e-9 * np.abs(gro).max()
    # real eps on the GPU, a 3D slab
    e3 = torch.tensor(1 + rng.random((4, 4, 5)), dtype=torch.float64, device="cuda", requires_grad=True)
    T3 = transmission(e3, 1.0, 1.0, 1.3, "periodic", three_d=True)
    T3.backward()
    To3, gr3, _ = o.transmission_gradient(dict(epsilon=e3.detach().cpu().numpy(), epsilon_L=1.0, epsilon_R=1.0,
                                               k0dx=1.3, xyBC="periodic"), three_d=True)
    assert np.abs(e3.grad.cpu().numpy() - gr3).max() < 1e-9 * np.abs(gr3).max()
