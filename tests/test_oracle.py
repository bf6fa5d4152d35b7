"""Pins the FP64 oracle (oracle/rgf_oracle.py) without running the MATLAB reference.

Every check here is independent of the oracle's own formulas: brute-force
full-system solves, closed forms, invariants, high-precision arithmetic and
textbook stencils.  Each is chosen so that a plausible slip in the restatement
(a dropped term, a wrong sign or branch, a transposed operand, a wrong loop
direction) fails at least one of them.
"""
import math

import numpy as np
import pytest

import oracle.rgf_oracle as o

BCS_2D = ["periodic", "Dirichlet", "Neumann", "DirichletNeumann", "NeumannDirichlet", 0.7, -2.1]


def _rand_eps(rng, shape, cplx=True):
    e = 1 + 2 * rng.random(shape)
    if cplx:
        e = e + 0.05j * rng.random(shape)
    return e


# ----------------------------------------------------------------------------- Laplacian (a8)
@pytest.mark.parametrize("n", [1, 2, 3, 7])
def test_laplacian_textbook(n):
    D = o.build_laplacian_1d(n, "Dirichlet")
    ref = -2 * np.eye(n) + np.eye(n, k=1) + np.eye(n, k=-1)
    assert np.allclose(D, ref)
    P = o.build_laplacian_1d(n, "periodic")
    if n >= 3:
        ref = -2 * np.eye(n) + np.eye(n, k=1) + np.eye(n, k=-1)
        ref[0, -1] += 1
        ref[-1, 0] += 1
        assert np.allclose(P, ref)
    ka = 0.9
    Bl = o.build_laplacian_1d(n, ka)
    if n >= 3:
        assert np.isclose(Bl[0, -1], np.exp(-1j * ka)) and np.isclose(Bl[-1, 0], np.exp(1j * ka))
    if n == 1:
        assert np.isclose(Bl[0, 0], -4 * math.sin(ka / 2) ** 2)
    Nm = o.build_laplacian_1d(n, "Neumann")
    if n >= 2:
        assert np.isclose(Nm[0, 0], -1) and np.isclose(Nm[-1, -1], -1)
    assert np.allclose(o.build_laplacian_1d(n, "PEC"), D)


@pytest.mark.parametrize("bc", BCS_2D)
@pytest.mark.parametrize("ny", [1, 2, 5, 8])
def test_laplacian_eigenmodes_and_unitary_phi(bc, ny):
    """L phi_a = -4 sin^2(ky_a/2) phi_a ties setup_channels to build_laplacian_1d."""
    ch = o.setup_channels(ny, bc, 0.5, 1.0, 1.0)
    phi = ch["phi"]
    Lap = o.build_laplacian_1d(ny, bc)
    lam = -4 * np.sin(ch["kydx"] / 2) ** 2
    assert np.allclose(Lap @ phi, phi * lam[None, :], atol=1e-12)
    assert np.allclose(phi.conj().T @ phi, np.eye(ny), atol=1e-12)
    assert np.all(np.diff(ch["kydx"]) > 0)


def test_bloch_string_rejected():
    with pytest.raises(ValueError):
        o.setup_channels(4, "Bloch", 0.5, 1.0)
    with pytest.raises(ValueError):
        o.setup_channels(4, "nonsense", 0.5, 1.0)
    with pytest.raises(ValueError):
        o.setup_channels(4, 0.3 + 0.1j, 0.5, 1.0)


# ----------------------------------------------------------------------------- channels (a4, a5)
@pytest.mark.parametrize("k2eps", [0.3, 2.0, 3.9, 4.5, 7.0, 2.0 + 0.3j, 2.0 - 0.3j, 5 - 0.2j, 5 + 0.2j,
                                   -0.5 + 0.1j, -0.5 - 0.1j])
def test_dispersion_and_branch(k2eps):
    ny = 9
    for bc in ["periodic", "Dirichlet", "Neumann", 1.3]:
        ch = o.setup_channels(ny, bc, 1.0, k2eps)
        kx = ch["L"]["kxdx"]
        lhs = 4 * np.sin(kx / 2) ** 2 + 4 * np.sin(ch["kydx"] / 2) ** 2
        assert np.allclose(lhs, k2eps, atol=1e-12)
        z = ch["L"]["z"]
        prop = ch["L"]["ind_prop"] - 1
        assert np.array_equal(prop, np.flatnonzero((z.real > 0) & (z.real < 1)))
        if np.imag(k2eps) == 0:
            # documented choice (setup_channels.m:54-56)
            assert np.all((kx[prop].real > 0) & (kx[prop].real < np.pi) & (kx[prop].imag == 0))
            ev = np.setdiff1d(np.arange(ny), prop)
            assert np.all(kx[ev].imag >= 0)
            assert np.all(np.isclose(np.mod(kx[ev].real, np.pi), 0) | np.isclose(np.mod(kx[ev].real, np.pi), np.pi))
            assert np.all(np.abs(np.exp(1j * kx)) <= 1 + 1e-15)


@pytest.mark.parametrize("x", [-0.7, 0.2, 0.9, 1.6])
def test_branch_vertical_continuity(x):
    """Continuation along a vertical line from the real axis (setup_channels.m:57-60)."""
    k0 = o.kx_from_z(complex(x, 0.0))
    for y in (-1e-10, 1e-10):
        k1 = o.kx_from_z(complex(x, y))
        assert abs(np.exp(1j * k1) - np.exp(1j * k0)) < 1e-8
        assert abs(np.sin(k1) - np.sin(k0)) < 1e-8


def test_matlab_representative_above_cutoff():
    """real z > 1 gives Re kx = -pi (the representative MATLAB produces after its flip)."""
    k = o.kx_from_z(4.0)
    assert np.isclose(k.real, -np.pi) and np.isclose(k.imag, 2 * math.acosh(2.0))


def test_ind_prop_conj_periodic():
    for ny in (8, 9, 12, 13):
        for k0dx in (0.4, 1.1, 2.0):
            ch = o.setup_channels(ny, "periodic", k0dx, 1.0, 1.0)
            kp = ch["kydx"][ch["L"]["ind_prop"] - 1]
            c = ch["L"]["ind_prop_conj"] - 1
            phi = o.mode_matrix_2d(ny, "Bloch", kp)
            phic = o.mode_matrix_2d(ny, "Bloch", kp[c])
            assert np.allclose(phic, phi.conj(), atol=1e-12)


def test_channels_3d_column_major_and_phi():
    nx, ny = 4, 5
    ch = o.setup_channels_3D(nx, ny, [0.3, -1.1], 1.2, 1.0, 2.0)
    assert ch["kxdx"].shape == (nx, ny)
    assert np.allclose(ch["kxdx"][:, 0], ch["kxdx"][:, 3])
    phi = ch["phi"]
    assert np.allclose(phi.conj().T @ phi, np.eye(nx * ny), atol=1e-12)
    A0 = o.slice_operator_3d(nx, ny, [0.3, -1.1])
    lam = 2 + 4 * (np.sin(ch["kxdx"] / 2) ** 2 + np.sin(ch["kydx"] / 2) ** 2).ravel(order="F")
    assert np.allclose(A0 @ phi, phi * lam[None, :], atol=1e-12)
    # column a = vec(fun_phi(kx(a), ky(a))) with x fast (setup_channels_3D.m:118)
    a = 7
    ix, iy = a % nx, a // nx
    f = np.exp(1j * np.arange(1, nx + 1)[:, None] * ch["kxdx"][ix, iy]) * \
        np.exp(1j * np.arange(1, ny + 1)[None, :] * ch["kydx"][ix, iy]) / math.sqrt(nx * ny)
    assert np.allclose(phi[:, a], f.ravel(order="F"))
    with pytest.raises(ValueError):
        o.setup_channels_3D(3, 3, "Neumann", 1.0, 1.0)


# ----------------------------------------------------------------------------- lead G0 (a7)
@pytest.mark.parametrize("bc", BCS_2D)
def test_lead_self_consistency_and_retarded(bc):
    ny = 11
    k0dx, eL = 0.9, 1.7
    ch = o.setup_channels(ny, bc, k0dx, eL, eL)
    G0 = o.lead_greens_function(ch["phi"], ch["L"]["kxdx"])
    Alead = o.slice_operator_2d(ny, bc) - k0dx ** 2 * eL * np.eye(ny)
    assert np.allclose(G0, np.linalg.inv(Alead - G0), atol=1e-12)
    # retarded: with absorption the lead recursion g <- (A_lead - g)^-1 from g=0
    # converges to the decaying solution; the oracle's complex-eps branch must
    # reproduce it, and its eta -> 0+ limit must be the real-eps G0.
    for eta in (0.3, 0.05):
        e = eL + 1j * eta / k0dx ** 2
        che = o.setup_channels(ny, bc, k0dx, e, e)
        G0e = o.lead_greens_function(che["phi"], che["L"]["kxdx"])
        Ae = o.slice_operator_2d(ny, bc) - k0dx ** 2 * e * np.eye(ny)
        g = np.zeros((ny, ny), complex)
        for _ in range(4000):
            g = np.linalg.inv(Ae - g)
        assert np.abs(g - G0e).max() < 1e-9
    che = o.setup_channels(ny, bc, k0dx, eL + 1e-9j, eL + 1e-9j)
    assert np.abs(o.lead_greens_function(che["phi"], che["L"]["kxdx"]) - G0).max() < 1e-6


# ----------------------------------------------------------------------------- end-to-end (a1-a13)
def _full_2d(rng, ny, nx, bc, eL, eR, k0dx, cplx=True):
    eps = _rand_eps(rng, (ny, nx), cplx)
    syst = dict(epsilon=eps, epsilon_L=eL, epsilon_R=eR, k0dx=k0dx, yBC=bc)
    return syst


@pytest.mark.parametrize("bc", BCS_2D)
@pytest.mark.parametrize("nx", [0, 1, 6])
def test_recursion_vs_brute_force_2d(bc, nx):
    rng = np.random.default_rng(abs(hash((str(bc), nx))) % 2 ** 31)
    syst = _full_2d(rng, 9, nx, bc, 1.0, 2.3, 2 * np.pi / 5.5)
    S, ch = o.cal_smatrix_RGF(syst, ["L", "R"], ["L", "R"])
    iL = np.arange(1, ch["L"]["N_prop"] + 1)
    iR = np.arange(1, ch["R"]["N_prop"] + 1)
    A0 = o.slice_operator_2d(9, bc)
    Sb = o.brute_force_smatrix([syst["epsilon"][:, n] for n in range(nx)], A0, syst["k0dx"] ** 2,
                               ch["phi"], ch["L"], ch["R"], 1.0, 2.3, iL, iR, iL, iR)
    assert np.abs(S - Sb).max() < 1e-12


def test_brute_force_fields_satisfy_slice_equations():
    """The field profiles returned by the brute-force solve: interior slices
    obey (A0 - k^2 eps_n) psi_n = psi_{n-1} + psi_{n+1}, the lead slices carry
    the source and self-energy, and S follows from the boundary slices."""
    rng = np.random.default_rng(8)
    ny, nx, k0dx = 9, 5, 2 * np.pi / 5.5
    syst = _full_2d(rng, ny, nx, "periodic", 1.0, 2.3, k0dx)
    S, ch = o.cal_smatrix_RGF(syst, ["L", "R"], ["L", "R"])
    iL = np.arange(1, ch["L"]["N_prop"] + 1)
    iR = np.arange(1, ch["R"]["N_prop"] + 1)
    A0 = o.slice_operator_2d(ny, "periodic")
    eps = syst["epsilon"]
    Sb, F = o.brute_force_smatrix([eps[:, n] for n in range(nx)], A0, k0dx ** 2, ch["phi"], ch["L"], ch["R"],
                                  1.0, 2.3, iL, iR, iL, iR, return_fields=True)
    assert F.shape == (ny, nx + 2, len(iL) + len(iR))
    assert np.abs(S - Sb).max() < 1e-12
    for n in range(1, nx + 1):
        lhs = (A0 - np.diag(k0dx ** 2 * eps[:, n - 1])) @ F[:, n]
        assert np.abs(lhs - F[:, n - 1] - F[:, n + 1]).max() < 1e-10 * np.abs(F).max()
    # left lead slice: (A_lead - G0_L) psi_0 - psi_1 = B_L (left inputs), 0 (right)
    G0L = o.lead_greens_function(ch["phi"], ch["L"]["kxdx"])
    BL = ch["phi"][:, ch["L"]["ind_prop"][iL - 1] - 1] * ch["L"]["sqrt_mu"][iL - 1][None, :]
    r0 = (A0 - k0dx ** 2 * 1.0 * np.eye(ny) - G0L) @ F[:, 0] - F[:, 1]
    assert np.abs(r0[:, :len(iL)] - BL).max() < 1e-10 * np.abs(F).max()
    assert np.abs(r0[:, len(iL):]).max() < 1e-10 * np.abs(F).max()


def test_five_schemes_agree_with_general():
    rng = np.random.default_rng(3)
    syst = _full_2d(rng, 10, 7, "periodic", 1.2, 2.0, 2 * np.pi / 6)
    S, ch = o.cal_smatrix_RGF(syst, ["L", "R"], ["L", "R"])
    nL, nR = ch["L"]["N_prop"], ch["R"]["N_prop"]
    blocks = {
        (("L",), ("L",)): S[:nL, :nL],
        (("R",), ("R",)): S[nL:, nL:],
        (("L", "R"), ("L",)): S[:, :nL],
        (("L", "R"), ("R",)): S[:, nL:],
        (("R",), ("L",)): S[nL:, :nL],
        (("L",), ("R",)): S[:nL, nL:],
    }
    for (out, inn), ref in blocks.items():
        Sx, _ = o.cal_smatrix_RGF(syst, list(out), list(inn))
        assert np.abs(Sx - ref).max() < 1e-12


def test_struct_selection_order_duplicates_and_delta():
    rng = np.random.default_rng(5)
    syst = _full_2d(rng, 12, 5, "Dirichlet", 1.0, 1.0, 2 * np.pi / 6)
    S, ch = o.cal_smatrix_RGF(syst, ["L", "R"], ["L", "R"])
    nL = ch["L"]["N_prop"]
    out = {"ind_out_L": [3, 1, 3], "ind_out_R": [2]}
    inn = {"ind_in_L": [1, 3], "ind_in_R": [2, 1]}
    Sx, _ = o.cal_smatrix_RGF(syst, out, inn)
    rows = [2, 0, 2, nL + 1]
    cols = [0, 2, nL + 1, nL + 0]
    ref = S[np.ix_(rows, cols)].copy()
    # the general S already has delta subtracted once per matching index; with
    # duplicate rows, intersect() subtracts it only at the first occurrence
    ref[2, 1] += 1.0
    assert np.abs(Sx - ref).max() < 1e-12


def test_empty_request_returns_real_zeros():
    rng = np.random.default_rng(6)
    syst = _full_2d(rng, 6, 3, "periodic", 1.0, 1.0, 1.0)
    S, _ = o.cal_smatrix_RGF(syst, {"ind_out_L": []}, ["L"])
    assert S.shape == (0, o.setup_channels(6, "periodic", 1.0, 1.0)["L"]["N_prop"])
    assert not np.iscomplexobj(S)


def test_index_validation():
    rng = np.random.default_rng(7)
    syst = _full_2d(rng, 6, 3, "periodic", 1.0, 1.0, 1.0)
    with pytest.raises(ValueError):
        o.cal_smatrix_RGF(syst, ["L"], {"ind_in_L": [99]})
    with pytest.raises(ValueError):
        o.cal_smatrix_RGF(syst, ["L"], {"ind_in_L": [1.5]})
    with pytest.raises(ValueError):
        o.cal_smatrix_RGF(syst, "L", ["L"])
    with pytest.raises(ValueError):
        o.cal_smatrix_RGF({k: v for k, v in syst.items() if k != "k0dx"}, ["L"], ["L"])


def test_uniform_slab_closed_form():
    ny, nx = 16, 9
    for bc in ["periodic", "Dirichlet", 0.4]:
        syst = dict(epsilon=1.7 * np.ones((ny, nx)), epsilon_L=1.7, epsilon_R=1.7, k0dx=0.8, yBC=bc)
        S, ch = o.cal_smatrix_RGF(syst, ["L", "R"], ["L", "R"])
        n = ch["L"]["N_prop"]
        kx = ch["L"]["kxdx"][ch["L"]["ind_prop"] - 1]
        assert np.abs(S[:n, :n]).max() < 1e-12
        assert np.abs(S[n:, :n] - np.diag(np.exp(1j * kx * (nx + 1)))).max() < 1e-12


def test_single_interface_closed_form():
    ny = 10
    eL, eR, k0dx = 1.0, 2.5, 1.0
    syst = dict(epsilon=np.zeros((ny, 0)), epsilon_L=eL, epsilon_R=eR, k0dx=k0dx, yBC="Dirichlet")
    S, ch = o.cal_smatrix_RGF(syst, ["L", "R"], ["L", "R"])
    nL = ch["L"]["N_prop"]
    for a in range(nL):
        m = ch["L"]["ind_prop"][a] - 1
        b = int(np.flatnonzero(ch["R"]["ind_prop"] - 1 == m)[0])
        gL, gR = np.exp(1j * ch["L"]["kxdx"][m]), np.exp(1j * ch["R"]["kxdx"][m])
        muL, muR = np.sin(ch["L"]["kxdx"][m]), np.sin(ch["R"]["kxdx"][m])
        r = -2j * muL * gL / (1 - gL * gR) - 1
        t = -2j * np.sqrt(muL * muR) * gL * gR / (1 - gL * gR)
        assert abs(S[a, a] - r) < 1e-12
        assert abs(S[nL + b, a] - t) < 1e-12


@pytest.mark.parametrize("bc", ["periodic", "Dirichlet", "Neumann"])
def test_unitarity_and_reciprocity(bc):
    rng = np.random.default_rng(11)
    syst = _full_2d(rng, 20, 12, bc, 1.0, 1.5, 2 * np.pi / 7, cplx=False)
    S, ch = o.cal_smatrix_RGF(syst, ["L", "R"], ["L", "R"])
    assert np.abs(S.conj().T @ S - np.eye(S.shape[0])).max() < 1e-12   # example_script.m:27
    perm = np.concatenate([ch["L"]["ind_prop_conj"], ch["L"]["N_prop"] + ch["R"]["ind_prop_conj"]]) - 1
    St = S[perm, :]
    assert np.abs(St.T - St).max() < 1e-12                              # example_script.m:32-34


def test_absorption_is_contractive():
    rng = np.random.default_rng(12)
    syst = _full_2d(rng, 14, 10, "periodic", 1.0, 1.0, 2 * np.pi / 7)
    S, _ = o.cal_smatrix_RGF(syst, ["L", "R"], ["L", "R"])
    assert np.linalg.eigvalsh(S.conj().T @ S).max() <= 1 + 1e-12


def test_mirror_symmetry():
    rng = np.random.default_rng(13)
    syst = _full_2d(rng, 9, 6, "Dirichlet", 1.0, 2.0, 2 * np.pi / 6)
    S, ch = o.cal_smatrix_RGF(syst, ["L", "R"], ["L", "R"])
    m = dict(syst, epsilon=syst["epsilon"][:, ::-1], epsilon_L=2.0, epsilon_R=1.0)
    Sm, chm = o.cal_smatrix_RGF(m, ["L", "R"], ["L", "R"])
    nL = ch["L"]["N_prop"]
    nR = ch["R"]["N_prop"]
    assert np.abs(S[:nL, :nL] - Sm[nR:, nR:]).max() < 1e-12
    assert np.abs(S[nL:, :nL] - Sm[:nR, nR:]).max() < 1e-12


def test_bloch_zero_equals_periodic():
    rng = np.random.default_rng(14)
    syst = _full_2d(rng, 8, 5, "periodic", 1.0, 1.0, 1.1)
    S1, _ = o.cal_smatrix_RGF(syst, ["L", "R"], ["L", "R"])
    S2, _ = o.cal_smatrix_RGF(dict(syst, yBC=0.0), ["L", "R"], ["L", "R"])
    assert np.abs(S1 - S2).max() < 1e-13


def test_3d_vs_brute_force_and_unitarity():
    rng = np.random.default_rng(15)
    nx, ny, nz = 4, 3, 4
    for bc in ["periodic", "Dirichlet", [0.5, 2.0]]:
        eps = 1 + rng.random((nx, ny, nz))
        syst = dict(epsilon=eps, epsilon_L=1.0, epsilon_R=1.4, k0dx=1.3, xyBC=bc)
        S, ch = o.cal_smatrix_RGF_3D(syst, ["L", "R"], ["L", "R"])
        assert np.abs(S.conj().T @ S - np.eye(S.shape[0])).max() < 1e-12
        iL = np.arange(1, ch["L"]["N_prop"] + 1)
        iR = np.arange(1, ch["R"]["N_prop"] + 1)
        A0 = o.slice_operator_3d(nx, ny, bc)
        Sb = o.brute_force_smatrix([eps[:, :, n].ravel(order="F") for n in range(nz)], A0, 1.3 ** 2,
                                   ch["phi"], ch["L"], ch["R"], 1.0, 1.4, iL, iR, iL, iR)
        assert np.abs(S - Sb).max() < 1e-12


def test_3d_y_invariant_reduces_to_2d():
    """eps independent of y: each ky block is a 2D solve with k0dx^2 eps shifted."""
    rng = np.random.default_rng(16)
    nx, ny, nz = 5, 4, 3
    kxa, kya = 0.7, 1.9
    prof = 1 + rng.random((nx, nz))
    eps = np.repeat(prof[:, None, :], ny, axis=1)
    k0dx = 1.4
    S3, ch3 = o.cal_smatrix_RGF_3D(dict(epsilon=eps, epsilon_L=1.0, epsilon_R=1.0, k0dx=k0dx,
                                        xyBC=[kxa, kya]), ["L", "R"], ["L", "R"])
    nL = ch3["L"]["N_prop"]
    prop3 = ch3["L"]["ind_prop"] - 1
    for iy in range(ny):
        ky = ch3["kydx"][0, iy]
        shift = 4 * math.sin(ky / 2) ** 2 / k0dx ** 2
        S2, ch2 = o.cal_smatrix_RGF(dict(epsilon=prof - shift, epsilon_L=1.0 - shift, epsilon_R=1.0 - shift,
                                         k0dx=k0dx, yBC=kxa), ["L", "R"], ["L", "R"])
        # 3D channels a with iy(a) == iy, in order; 2D channel order is ascending kx index
        sel = [a for a in range(nL) if prop3[a] // nx == iy]
        if not sel:
            continue
        n2 = ch2["L"]["N_prop"]
        assert n2 == len(sel)
        idx3 = sel + [nL + a for a in sel]
        assert np.abs(S3[np.ix_(idx3, idx3)] - S2).max() < 1e-12


def test_high_precision_tiny():
    """mpmath 40-digit full-system solve bounds the oracle's own FP64 error."""
    mp = pytest.importorskip("mpmath")
    mp.mp.dps = 40
    ny, nx = 3, 2
    rng = np.random.default_rng(17)
    eps = _rand_eps(rng, (ny, nx))
    k0dx, eL, eR = mp.mpf(1.1), mp.mpf(1), mp.mpf(2)
    syst = dict(epsilon=eps, epsilon_L=1.0, epsilon_R=2.0, k0dx=1.1, yBC="Dirichlet")
    S, ch = o.cal_smatrix_RGF(syst, ["L", "R"], ["L", "R"])
    # high-precision channels from the closed forms
    ky = [mp.mpf(m) * mp.pi / (ny + 1) for m in range(1, ny + 1)]
    phi = mp.matrix(ny, ny)
    for m in range(ny):
        for a in range(ny):
            phi[m, a] = mp.sin((m + 1) * ky[a]) * mp.sqrt(mp.mpf(2) / (ny + 1))

    def side(e):
        kx = []
        for a in range(ny):
            z = k0dx ** 2 * e / 4 - mp.sin(ky[a] / 2) ** 2
            kx.append(2 * mp.asin(mp.sqrt(z)) if 0 < z < 1 else 2j * mp.asinh(mp.sqrt(-z)))
        return kx

    kL, kR = side(eL), side(eR)

    def G0(kx):
        D = mp.diag([mp.exp(1j * k) for k in kx])
        return phi * D * phi.T

    A0 = mp.matrix(ny, ny)
    for m in range(ny):
        A0[m, m] = 4
        if m > 0:
            A0[m, m - 1] = -1
            A0[m - 1, m] = -1
    Ntot = ny * (nx + 2)
    H = mp.matrix(Ntot, Ntot)
    blocks = [A0 - k0dx ** 2 * eL * mp.eye(ny) - G0(kL)]
    for n in range(nx):
        blocks.append(A0 - k0dx ** 2 * mp.diag([mp.mpc(complex(v)) for v in eps[:, n]]))
    blocks.append(A0 - k0dx ** 2 * eR * mp.eye(ny) - G0(kR))
    for b, blk in enumerate(blocks):
        for i in range(ny):
            for j in range(ny):
                H[b * ny + i, b * ny + j] = blk[i, j]
        if b + 1 < len(blocks):
            for i in range(ny):
                H[b * ny + i, (b + 1) * ny + i] = -1
                H[(b + 1) * ny + i, b * ny + i] = -1
    Hi = H ** -1
    pL = [a for a in range(ny) if 0 < float(mp.re(k0dx ** 2 * eL / 4 - mp.sin(ky[a] / 2) ** 2)) < 1]
    pR = [a for a in range(ny) if 0 < float(mp.re(k0dx ** 2 * eR / 4 - mp.sin(ky[a] / 2) ** 2)) < 1]
    chans = [(0, a, kL) for a in pL] + [(nx + 1, a, kR) for a in pR]
    for r, (sr, ar, kr) in enumerate(chans):
        for c, (sc, ac, kc) in enumerate(chans):
            acc = mp.mpc(0)
            for i in range(ny):
                for j in range(ny):
                    acc += phi[i, ar] * Hi[sr * ny + i, sc * ny + j] * phi[j, ac]
            val = -2j * mp.sqrt(mp.sin(kr[ar])) * mp.sqrt(mp.sin(kc[ac])) * acc - (1 if r == c else 0)
            assert abs(complex(val) - S[r, c]) < 1e-13


# ----------------------------------------------------------------------------- adjoint gradient
# transmission_gradient takes the forward fields and the adjoint rows from one
# sparse LU of the whole system; the finite differences below difference the
# S matrix of the slice recursion (cal_smatrix_RGF, itself pinned by the
# brute-force tests above).  Both parts are checked: a wrong sign or a dropped
# conjugate in dT/dIm eps, or a transposed Psi/Phi index, fails here.
@pytest.mark.parametrize("bc", BCS_2D)
@pytest.mark.parametrize("cplx_k", [False, True])
def test_transmission_gradient_vs_central_differences_2d(bc, cplx_k):
    rng = np.random.default_rng(17 + BCS_2D.index(bc) + 100 * cplx_k)
    ny, nx = 7, 5
    eps = _rand_eps(rng, (ny, nx), cplx=True)
    k0dx = 1.3 * (1 + 0.01j) if cplx_k else 1.3
    syst = dict(epsilon=eps, epsilon_L=1.0, epsilon_R=1.4, k0dx=k0dx, yBC=bc)
    T, gre, gim = o.transmission_gradient(syst)
    S, _ = o.cal_smatrix_RGF(syst, ["R"], ["L"])
    assert abs(T - float(np.sum(np.abs(S) ** 2))) < 1e-12 * max(1.0, T)
    pixels = [(i, j) for i in range(ny) for j in range(nx)]
    fd_re = o.transmission_gradient_fd(syst, pixels, h=1e-5, part="re")
    fd_im = o.transmission_gradient_fd(syst, pixels, h=1e-5, part="im")
    scale = max(np.abs(fd_re).max(), np.abs(fd_im).max())
    assert scale > 1e-3
    assert np.abs(gre[tuple(np.array(pixels).T)] - fd_re).max() < 1e-7 * scale
    assert np.abs(gim[tuple(np.array(pixels).T)] - fd_im).max() < 1e-7 * scale


def test_transmission_gradient_vs_central_differences_real_eps():
    # real eps: dT/dRe eps from FD of a real perturbation, dT/dIm eps from a
    # perturbation that makes the slab lossy
    rng = np.random.default_rng(5)
    eps = 1 + 2 * rng.random((6, 4))
    syst = dict(epsilon=eps, epsilon_L=1.0, epsilon_R=1.0, k0dx=1.1, yBC="periodic")
    T, gre, gim = o.transmission_gradient(syst)
    pixels = [(i, j) for i in range(6) for j in range(4)]
    fd_re = o.transmission_gradient_fd(syst, pixels, h=1e-5, part="re")
    fd_im = o.transmission_gradient_fd(syst, pixels, h=1e-5, part="im")
    idx = tuple(np.array(pixels).T)
    scale = max(np.abs(fd_re).max(), np.abs(fd_im).max())
    assert np.abs(gre[idx] - fd_re).max() < 1e-7 * scale
    assert np.abs(gim[idx] - fd_im).max() < 1e-7 * scale
    # a passive real slab: T is reduced by any absorption, so dT/dIm eps <= 0
    assert np.all(gim <= 1e-12)


@pytest.mark.parametrize("bc", [[0.4, 2.1], "Dirichlet"])
def test_transmission_gradient_vs_central_differences_3d(bc):
    rng = np.random.default_rng(23)
    nx, ny, nz = 3, 4, 3
    eps = _rand_eps(rng, (nx, ny, nz), cplx=True)
    syst = dict(epsilon=eps, epsilon_L=1.0, epsilon_R=1.2, k0dx=1.2 + 0.005j, xyBC=bc)
    T, gre, gim = o.transmission_gradient(syst, three_d=True)
    pixels = [(i, j, k) for i in range(nx) for j in range(ny) for k in range(nz)]
    fd_re = o.transmission_gradient_fd(syst, pixels, h=1e-5, three_d=True, part="re")
    fd_im = o.transmission_gradient_fd(syst, pixels, h=1e-5, three_d=True, part="im")
    idx = tuple(np.array(pixels).T)
    scale = max(np.abs(fd_re).max(), np.abs(fd_im).max())
    assert np.abs(gre[idx] - fd_re).max() < 1e-7 * scale
    assert np.abs(gim[idx] - fd_im).max() < 1e-7 * scale


def test_transmission_gradient_fd_catches_planted_sign_error():
    # the FD check above must fail for a gradient with a flipped Im part
    rng = np.random.default_rng(9)
    eps = _rand_eps(rng, (5, 3), cplx=True)
    syst = dict(epsilon=eps, epsilon_L=1.0, epsilon_R=1.0, k0dx=1.2, yBC="Dirichlet")
    _, _, gim = o.transmission_gradient(syst)
    pixels = [(i, j) for i in range(5) for j in range(3)]
    fd_im = o.transmission_gradient_fd(syst, pixels, h=1e-5, part="im")
    idx = tuple(np.array(pixels).T)
    assert np.abs(-gim[idx] - fd_im).max() > 1e-3 * np.abs(fd_im).max()
