"""FP64 CPU restatement of the RGF S-matrix solve -- TEST INFRASTRUCTURE ONLY.

This module is the oracle that the HIP implementation in ``rgf_amd`` is checked
against.  Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` may import it, and only as a checker or as the timed CPU
baseline; the product path never calls it.

It restates, in numpy/scipy (LAPACK getrf/getri through ``numpy.linalg.inv``,
gesv through ``numpy.linalg.solve``, zgemm through ``@``), the algorithm of the
MATLAB reference chiaweihsu/RGF:

* ``build_laplacian_1d``   -- src/build_laplacian_1d.m:1-77
* ``setup_channels``       -- src/setup_channels.m:1-238
* ``setup_channels_3D``    -- src/setup_channels_3D.m:1-188
* ``cal_smatrix_RGF``      -- src/cal_smatrix_RGF.m:1-291 (all five schemes)
* ``cal_smatrix_RGF_3D``   -- src/cal_smatrix_RGF_3D.m:1-306
* ``brute_force_smatrix``  -- independent check: a direct sparse solve of the
  whole block-tridiagonal system (no recursion), used to pin the restatement.

The MATLAB reference cannot be executed here (no MATLAB/Octave), so parity is
pinned by: the brute-force full-system solve, closed forms (uniform slab,
single interface, Laplacian eigenvectors, lead self-consistency), invariants
(unitarity, reciprocity, mirror symmetry, scheme agreement, Bloch ka=0 vs
periodic, 3D y-invariant vs 2D) and an mpmath high-precision solve at tiny
sizes -- see tests/test_oracle.py.

Conventions that differ from a literal transcription (documented, not bugs):

* Branch of kxdx.  The reference computes ``2*asin(sqrt(z))`` and then flips
  signs to compensate for MATLAB's real-argument ``asin(x>1)`` having a
  negative imaginary part (setup_channels.m:211-215).  C99/numpy ``asin`` has
  the opposite sign on that cut, so a verbatim copy would be wrong.  The
  physical rule the reference documents (setup_channels.m:52-60) is
  implemented directly in ``kx_from_z``: real z in (0,1) -> real kx in (0,pi);
  real z<=0 -> kx = 2i*asinh(sqrt(-z)); real z>1 -> kx = -pi + 2i*acosh(sqrt(z))
  (the representative MATLAB returns); complex z -> principal branch, negated
  where Im z<0 and (Re z<0 or Re z>1).
* Bloch BC with n=1 in build_laplacian_1d: spdiags receives diagonal 0 twice
  (build_laplacian_1d.m:51); MATLAB's sparse() sums duplicates, giving
  -4 sin^2(ka/2), which is what is implemented.
"""
from __future__ import annotations

import cmath
import math
from typing import Any

import numpy as np

__all__ = [
    "build_laplacian_1d", "parse_bc", "kx_from_z", "setup_channels",
    "setup_channels_3D", "cal_smatrix_RGF", "cal_smatrix_RGF_3D",
    "brute_force_smatrix", "select_channels", "mode_matrix_2d",
    "mode_matrix_3d", "lead_greens_function", "slice_operator_2d",
    "slice_operator_3d", "transmission_gradient_fd",
]

_BC_ALIASES = {
    "dirichlet": "Dirichlet", "pec": "Dirichlet",
    "neumann": "Neumann", "pmc": "Neumann",
    "dirichletneumann": "DirichletNeumann", "pecpmc": "DirichletNeumann",
    "neumanndirichlet": "NeumannDirichlet", "pmcpec": "NeumannDirichlet",
}


def parse_bc(BC) -> tuple[str, Any]:
    """Normalise a BC argument to ('Bloch', ka) or (name, None).

    Follows setup_channels.m:102-115 / build_laplacian_1d.m:31-44: a numeric
    BC means Bloch with ka=BC; 'periodic' means Bloch with ka=0; 'Bloch' as a
    string is an error; other strings are case-insensitive names.
    """
    if isinstance(BC, (int, float, complex, np.number)) and not isinstance(BC, bool):
        return "Bloch", BC
    if not isinstance(BC, str):
        raise ValueError("Boundary condition BC has to be a character vector or string, "
                         "or numeric scalar (for Bloch periodic BC)")
    low = BC.lower()
    if low == "bloch":
        raise ValueError("To use Bloch periodic boundary condition, set BC to k*a where k is "
                         "the Bloch wave number and a is the periodicity")
    if low == "periodic":
        return "Bloch", 0.0
    if low in _BC_ALIASES:
        return _BC_ALIASES[low], None
    raise ValueError(f"Boundary condition BC = {BC} is not an eligible option")


def build_laplacian_1d(n: int, BC) -> np.ndarray:
    """Dense (d/dx)^2 dx^2 as -grad1^H grad1 (build_laplacian_1d.m:46-75)."""
    if not (isinstance(n, (int, np.integer)) and n >= 0):
        raise ValueError("The number of sites n must be a non-negative integer scalar")
    kind, ka = parse_bc(BC)
    if kind == "Bloch":
        # rows m: df(m-0.5) = f(m) - f(m-1), with f(0) = f(n) exp(-i ka) (:48-51)
        g = np.zeros((n, n), dtype=complex)
        for m in range(n):
            g[m, m] += 1.0
            if m >= 1:
                g[m, m - 1] += -1.0
        if n >= 1:
            g[0, n - 1] += -cmath.exp(-1j * ka)   # duplicates add, as sparse() does
    elif kind == "Dirichlet":       # (:52-55), (n+1) x n
        g = np.zeros((n + 1, n), dtype=complex)
        for m in range(n + 1):
            if m < n:
                g[m, m] = 1.0
            if m >= 1:
                g[m, m - 1] = -1.0
    elif kind == "Neumann":         # (:56-59), (n-1) x n
        g = np.zeros((max(n - 1, 0), n), dtype=complex)
        for m in range(n - 1):
            g[m, m + 1] = 1.0
            g[m, m] = -1.0
    elif kind == "DirichletNeumann":  # (:60-63), n x n
        g = np.zeros((n, n), dtype=complex)
        for m in range(n):
            g[m, m] = 1.0
            if m >= 1:
                g[m, m - 1] = -1.0
    else:                           # NeumannDirichlet (:64-67), n x n
        g = np.zeros((n, n), dtype=complex)
        for m in range(n):
            g[m, m] = -1.0
            if m + 1 < n:
                g[m, m + 1] = 1.0
    return -(g.conj().T @ g)


def kx_from_z(z: complex) -> complex:
    """Longitudinal wavenumber 2*asin(sqrt(z)) on the physical branch.

    Restates setup_channels.m:185-215 (see the module docstring for why the
    rule is stated by cases instead of copying the MATLAB sign flip).
    """
    x, y = float(np.real(z)), float(np.imag(z))
    if y == 0.0:
        if 0.0 <= x <= 1.0:
            return complex(2.0 * math.asin(math.sqrt(x)), 0.0)
        if x < 0.0:
            return complex(0.0, 2.0 * math.asinh(math.sqrt(-x)))
        return complex(-math.pi, 2.0 * math.acosh(math.sqrt(x)))
    kx = 2.0 * cmath.asin(cmath.sqrt(complex(x, y)))
    if y < 0.0 and (x < 0.0 or x > 1.0):
        kx = -kx
    return kx


def _longitudinal(k0dx2_eps: complex, sin2_t: np.ndarray) -> dict:
    """setup_longitudinal (setup_channels.m:179-219; setup_channels_3D.m:146-188).

    ``sin2_t`` is sum of sin^2(k_t/2) over the transverse directions, in the
    channel order (column-major in 3D).  Returns 1-based ``ind_prop``.
    """
    z = 0.25 * k0dx2_eps - sin2_t
    zc = np.asarray(z, dtype=complex)
    kx = np.array([kx_from_z(v) for v in zc.ravel()], dtype=complex)
    prop = np.flatnonzero((zc.real > 0) & (zc.real < 1))
    sqrt_mu = np.sqrt(np.sin(kx[prop]).astype(complex))
    return {"kxdx": kx, "ind_prop": prop + 1, "N_prop": int(prop.size),
            "sqrt_mu": sqrt_mu, "z": zc}


def setup_channels(ny: int, BC, k0dx, epsilon_L, epsilon_R=None) -> dict:
    """2D lead channels (setup_channels.m:84-176).

    Returns a dict with ``kydx`` (length ny), ``phi`` (ny x ny mode matrix,
    the value of ``fun_phi(kydx)``), ``bc`` ('Bloch' or a name), ``ka`` and
    sides ``L``/``R`` with ``kxdx``, ``ind_prop`` (1-based), ``N_prop``,
    ``sqrt_mu`` and, where defined, ``ind_prop_conj`` (1-based).
    """
    if not (isinstance(ny, (int, np.integer)) and ny > 0):
        raise ValueError("ny must be a positive integer scalar")
    kind, ka = parse_bc(BC)
    if kind == "Bloch" and np.iscomplexobj(ka) and np.imag(ka) != 0:
        raise ValueError(f"k*a = {np.real(ka):g} + 1i*{np.imag(ka):g} is a complex number; has to "
                         "be real for a complete orthonormal transverse basis.")
    m = np.arange(1, ny + 1)
    ind_zero_ky = None
    if kind == "Bloch":
        ka = float(np.real(ka))
        ind_zero_ky = (ny + 1) // 2 if ny % 2 == 1 else ny // 2          # (:125-129)
        kydx = ka / ny + (m - ind_zero_ky) * (2 * math.pi / ny)           # (:130)
    elif kind == "Dirichlet":
        kydx = m * (math.pi / (ny + 1))                                    # (:136)
    elif kind == "Neumann":
        kydx = (m - 1) * (math.pi / ny)                                    # (:141)
    else:                                                                  # DN / ND
        kydx = (m - 0.5) * (math.pi / (ny + 0.5))                          # (:149,:154)
    ch = {"kydx": kydx, "bc": kind, "ka": ka, "ny": ny, "ind_zero_ky": ind_zero_ky}
    ch["phi"] = mode_matrix_2d(ny, kind, kydx)
    k0dx2 = complex(k0dx) ** 2
    sin2 = np.sin(kydx / 2) ** 2
    ch["L"] = _longitudinal(k0dx2 * epsilon_L, sin2)
    _add_conj(ch["L"], kind, ka, ind_zero_ky)
    if epsilon_R is not None and not (np.isscalar(epsilon_R) and np.isnan(epsilon_R)):
        if epsilon_R == epsilon_L:                                         # (:166-167)
            ch["R"] = ch["L"]
        else:
            ch["R"] = _longitudinal(k0dx2 * epsilon_R, sin2)
            _add_conj(ch["R"], kind, ka, ind_zero_ky)
    return ch


def _add_conj(side: dict, kind: str, ka, ind_zero_ky) -> None:
    """ind_prop_conj (setup_channels.m:221-236)."""
    n = side["N_prop"]
    if kind != "Bloch":
        side["ind_prop_conj"] = np.arange(1, n + 1)
    elif ka == 0:
        if (ind_zero_ky in set(side["ind_prop"].tolist())) or n % 2 == 0:
            side["ind_prop_conj"] = np.arange(n, 0, -1)
        else:
            side["ind_prop_conj"] = np.concatenate([np.arange(n - 1, 0, -1), [n]])


def mode_matrix_2d(ny: int, kind: str, kydx: np.ndarray) -> np.ndarray:
    """phi = fun_phi(kydx): column a is transverse mode a (setup_channels.m:133-156)."""
    m = np.arange(1, ny + 1)[:, None].astype(float)
    k = np.asarray(kydx, dtype=float)[None, :]
    if kind == "Bloch":
        return np.exp(1j * m * k) / math.sqrt(ny)
    if kind == "Dirichlet":
        return (np.sin(m * k) * math.sqrt(2 / (ny + 1))).astype(complex)
    if kind == "Neumann":
        return ((np.cos((m - 0.5) * k) - (k == 0) * (1 - math.sqrt(0.5))) * math.sqrt(2 / ny)).astype(complex)
    if kind == "DirichletNeumann":
        return (np.sin(m * k) * math.sqrt(2 / (ny + 0.5))).astype(complex)
    return (np.cos((m - 0.5) * k) * math.sqrt(2 / (ny + 0.5))).astype(complex)


def setup_channels_3D(nx: int, ny: int, BC, k0dx, epsilon_L, epsilon_R=None) -> dict:
    """3D lead channels (setup_channels_3D.m:76-140).

    ``kxdx``/``kydx`` are nx x ny arrays; channel index a is the column-major
    linear index (x fast).  ``phi`` is the nxy x nxy mode matrix whose column a
    is vec(fun_phi(kxdx(a), kydx(a))) (cal_smatrix_RGF_3D.m:185-188).
    """
    if not (isinstance(nx, (int, np.integer)) and nx > 0):
        raise ValueError("nx must be a positive integer scalar")
    if not (isinstance(ny, (int, np.integer)) and ny > 0):
        raise ValueError("ny must be a positive integer scalar")
    if isinstance(BC, str):
        low = BC.lower()
        if low == "bloch":
            raise ValueError("To use Bloch periodic boundary condition, set BC to [kx*ax, ky*ay]")
        if low == "periodic":
            kind, kxa, kya = "Bloch", 0.0, 0.0
        elif low in ("dirichlet", "pec"):
            kind, kxa, kya = "Dirichlet", None, None
        else:
            raise ValueError(f"Boundary condition BC = {BC} is not an eligible option")
    else:
        v = np.asarray(BC).ravel()
        if v.size != 2:
            raise ValueError("Boundary condition BC has to be a character vector or string, or "
                             "numeric vector with 2 elements (for Bloch periodic BC)")
        for lab, val in (("kx", v[0]), ("ky", v[1])):
            if np.imag(val) != 0:
                raise ValueError(f"{lab}*a is a complex number; has to be real")
        kind, kxa, kya = "Bloch", float(np.real(v[0])), float(np.real(v[1]))
    ix = np.arange(1, nx + 1)
    iy = np.arange(1, ny + 1)
    if kind == "Bloch":
        zx = (nx + 1) // 2 if nx % 2 == 1 else nx // 2
        zy = (ny + 1) // 2 if ny % 2 == 1 else ny // 2
        kx1 = kxa / nx + (ix - zx) * (2 * math.pi / nx)                   # (:115)
        ky1 = kya / ny + (iy - zy) * (2 * math.pi / ny)                   # (:116)
    else:
        kx1 = ix * (math.pi / (nx + 1))                                    # (:120)
        ky1 = iy * (math.pi / (ny + 1))                                    # (:121)
    kxdx = np.repeat(kx1[:, None], ny, axis=1)
    kydx = np.repeat(ky1[None, :], nx, axis=0)
    ch = {"kxdx": kxdx, "kydx": kydx, "bc": kind, "kxa": kxa, "kya": kya, "nx": nx, "ny": ny}
    ch["phi"] = mode_matrix_3d(nx, ny, kind, kx1, ky1)
    k0dx2 = complex(k0dx) ** 2
    sin2 = (np.sin(kxdx / 2) ** 2 + np.sin(kydx / 2) ** 2).ravel(order="F")
    ch["L"] = _longitudinal(k0dx2 * epsilon_L, sin2)
    ch["L"]["kzdx"] = ch["L"]["kxdx"].reshape((nx, ny), order="F")
    if epsilon_R is not None and not (np.isscalar(epsilon_R) and np.isnan(epsilon_R)):
        if epsilon_R == epsilon_L:
            ch["R"] = ch["L"]
        else:
            ch["R"] = _longitudinal(k0dx2 * epsilon_R, sin2)
            ch["R"]["kzdx"] = ch["R"]["kxdx"].reshape((nx, ny), order="F")
    return ch


def mode_matrix_3d(nx, ny, kind, kx1, ky1) -> np.ndarray:
    """nxy x nxy mode matrix; row (ix + nx*iy), column (a + nx*b) (setup_channels_3D.m:118,123)."""
    mx = np.arange(1, nx + 1)[:, None]
    my = np.arange(1, ny + 1)[:, None]
    if kind == "Bloch":
        fx = np.exp(1j * mx * kx1[None, :]) / math.sqrt(nx)
        fy = np.exp(1j * my * ky1[None, :]) / math.sqrt(ny)
    else:
        fx = np.sin(mx * kx1[None, :]) * math.sqrt(2 / (nx + 1))
        fy = np.sin(my * ky1[None, :]) * math.sqrt(2 / (ny + 1))
    # column-major vec with x fast: phi[(ix,iy),(a,b)] = fx[ix,a] fy[iy,b] -> kron(fy, fx)
    return np.kron(fy, fx).astype(complex)


def lead_greens_function(phi: np.ndarray, kxdx: np.ndarray) -> np.ndarray:
    """G0 = phi diag(exp(i kx)) phi^H (cal_smatrix_RGF.m:198-203)."""
    return (phi * np.exp(1j * kxdx)[None, :]) @ phi.conj().T


def select_channels(spec, side_counts: tuple[int, int], kind: str) -> tuple[np.ndarray, np.ndarray]:
    """Parse ``in``/``out`` (cal_smatrix_RGF.m:103-175).

    ``spec`` is a list of strings (MATLAB cell) or a dict (MATLAB struct) with
    ``ind_in_L``/``ind_in_R`` (kind='in') or ``ind_out_L``/``ind_out_R``.
    Returns 1-based index arrays (L, R), user order preserved.
    """
    nL, nR = side_counts
    iL = np.zeros(0, dtype=np.int64)
    iR = np.zeros(0, dtype=np.int64)
    if isinstance(spec, dict):
        for side, n in (("L", nL), ("R", nR)):
            key = f"ind_{kind}_{side}"
            if key in spec:
                v = np.asarray(spec[key]).ravel()
                if v.size and not (np.issubdtype(v.dtype, np.number) and np.all(np.round(v) == v)
                                   and v.min() > 0 and v.max() <= n):
                    raise ValueError(f"{kind}.{key}, when specified, must be an array of positive "
                                     f"integers not exceeding N_prop_{side}={n}")
                v = v.astype(np.int64)
                if side == "L":
                    iL = v
                else:
                    iR = v
    elif isinstance(spec, (list, tuple)):
        for el in spec:
            if el in ("left", "L"):
                iL = np.arange(1, nL + 1)
            elif el in ("right", "R"):
                iR = np.arange(1, nR + 1)
            # anything else: the reference warns and ignores (:124-134)
    else:
        raise ValueError(f"input argument '{kind}' must be a structure or a cell array")
    return iL, iR


def slice_operator_2d(ny: int, BC) -> np.ndarray:
    """A0 = -L_y + 2I (cal_smatrix_RGF.m:209)."""
    return -build_laplacian_1d(ny, BC) + 2 * np.eye(ny)


def slice_operator_3d(nx: int, ny: int, xyBC) -> np.ndarray:
    """A0 = -kron(I,Lx) - kron(Ly,I) + 2I (cal_smatrix_RGF_3D.m:209-224)."""
    if isinstance(xyBC, str):
        xBC = yBC = xyBC
    else:
        xBC, yBC = np.asarray(xyBC).ravel()[:2]
        xBC, yBC = float(np.real(xBC)), float(np.real(yBC))
    Lx = build_laplacian_1d(nx, xBC)
    Ly = build_laplacian_1d(ny, yBC)
    return -np.kron(np.eye(ny), Lx) - np.kron(Ly, np.eye(nx)) + 2 * np.eye(nx * ny)


def _delta(ind_out_L, ind_in_L, ind_out_R, ind_in_R, shape):
    """Kronecker-delta subtraction via intersect (cal_smatrix_RGF.m:286-289)."""
    D = np.zeros(shape)
    nOL, nIL = len(ind_out_L), len(ind_in_L)
    for (o, i, ro, ci) in ((ind_out_L, ind_in_L, 0, 0), (ind_out_R, ind_in_R, nOL, nIL)):
        common = np.intersect1d(o, i)
        for c in common:
            a = int(np.flatnonzero(o == c)[0])    # first occurrence
            b = int(np.flatnonzero(i == c)[0])
            D[ro + a, ci + b] = 1.0
    return D


def _rgf_core(slices, A0, k0dx2, G0_L, G0_R, B_L, B_R, C_L, C_R, nin, nout):
    """The five slice-loop schemes (cal_smatrix_RGF.m:215-280).

    ``slices`` is a list of per-slice epsilon vectors (length N each), in
    increasing n.
    """
    N = A0.shape[0]
    I = np.eye(N)
    N_in_L, N_in_R = nin
    N_out_L, N_out_R = nout
    nx = len(slices)

    def A_nn(n):
        return A0 - np.diag(k0dx2 * slices[n])

    if N_in_R == 0 and N_out_R == 0:
        G_LL = G0_R
        for n in range(nx - 1, -1, -1):
            G_LL = np.linalg.inv(A_nn(n) - G_LL)
        G_LL = np.linalg.solve(I - G0_L @ G_LL, G0_L)
        return -2j * (C_L @ (G_LL @ B_L))
    if N_in_R == 0:
        G_LL = G0_R
        G_RL = G0_R
        for n in range(nx - 1, -1, -1):
            G_LL = np.linalg.inv(A_nn(n) - G_LL)
            G_RL = G_RL @ G_LL
        G_LL = np.linalg.solve(I - G0_L @ G_LL, G0_L)
        G_RL = G_RL @ G_LL
        return -2j * np.vstack([(C_L @ G_LL) @ B_L, C_R @ (G_RL @ B_L)])
    if N_in_L == 0 and N_out_L == 0:
        G_RR = G0_L
        for n in range(nx):
            G_RR = np.linalg.inv(A_nn(n) - G_RR)
        G_RR = np.linalg.solve(I - G0_R @ G_RR, G0_R)
        return -2j * (C_R @ (G_RR @ B_R))
    if N_in_L == 0:
        G_RR = G0_L
        G_LR = G0_L
        for n in range(nx):
            G_RR = np.linalg.inv(A_nn(n) - G_RR)
            G_LR = G_LR @ G_RR
        G_RR = np.linalg.solve(I - G0_R @ G_RR, G0_R)
        G_LR = G_LR @ G_RR
        return -2j * np.vstack([C_L @ (G_LR @ B_R), (C_R @ G_RR) @ B_R])
    G_RR = G0_L
    G_RL = G0_L
    G_LL = G0_L
    G_LR = G0_L
    for n in range(nx):
        G_RR = np.linalg.inv(A_nn(n) - G_RR)
        G_RL = G_RR @ G_RL
        G_LL = G_LL + G_LR @ G_RL
        G_LR = G_LR @ G_RR
    G_RR = np.linalg.solve(I - G0_R @ G_RR, G0_R)
    G_RL = G_RR @ G_RL
    G_LL = G_LL + G_LR @ G_RL
    G_LR = G_LR @ G_RR
    top = np.hstack([C_L @ (G_LL @ B_L), C_L @ (G_LR @ B_R)])
    bot = np.hstack([C_R @ (G_RL @ B_L), C_R @ (G_RR @ B_R)])
    return -2j * np.vstack([top, bot])


def _validate_syst(syst, bc_key):
    if not isinstance(syst, dict):
        raise ValueError("syst must be a structure")
    for key in ("epsilon", "epsilon_L", "epsilon_R", "k0dx", bc_key):
        if key not in syst:
            raise ValueError(f"syst.{key} must be provided")


def _bc_matrices(phi, ch, ind_in_L, ind_in_R, ind_out_L, ind_out_R):
    """B and C (cal_smatrix_RGF.m:189-195): note sqrt_mu is not conjugated in C."""
    L, R = ch["L"], ch["R"]
    B_L = phi[:, L["ind_prop"][ind_in_L - 1] - 1] * L["sqrt_mu"][ind_in_L - 1][None, :]
    B_R = phi[:, R["ind_prop"][ind_in_R - 1] - 1] * R["sqrt_mu"][ind_in_R - 1][None, :]
    C_L = L["sqrt_mu"][ind_out_L - 1][:, None] * phi[:, L["ind_prop"][ind_out_L - 1] - 1].conj().T
    C_R = R["sqrt_mu"][ind_out_R - 1][:, None] * phi[:, R["ind_prop"][ind_out_R - 1] - 1].conj().T
    return B_L, B_R, C_L, C_R


def _smatrix_common(slices, A0, phi, ch, syst, out, in_):
    ind_in_L, ind_in_R = select_channels(in_, (ch["L"]["N_prop"], ch["R"]["N_prop"]), "in")
    ind_out_L, ind_out_R = select_channels(out, (ch["L"]["N_prop"], ch["R"]["N_prop"]), "out")
    nin = (len(ind_in_L), len(ind_in_R))
    nout = (len(ind_out_L), len(ind_out_R))
    if sum(nin) == 0 or sum(nout) == 0:
        return np.zeros((sum(nout), sum(nin)))
    B_L, B_R, C_L, C_R = _bc_matrices(phi, ch, ind_in_L, ind_in_R, ind_out_L, ind_out_R)
    kL = ch["L"]["kxdx"]
    kR = ch["R"]["kxdx"]
    G0_L = lead_greens_function(phi, kL)
    G0_R = G0_L if syst["epsilon_R"] == syst["epsilon_L"] else lead_greens_function(phi, kR)
    k0dx2 = complex(syst["k0dx"]) ** 2
    S = _rgf_core(slices, A0, k0dx2, G0_L, G0_R, B_L, B_R, C_L, C_R, nin, nout)
    return S - _delta(ind_out_L, ind_in_L, ind_out_R, ind_in_R, S.shape)


def cal_smatrix_RGF(syst: dict, out, in_):
    """2D S matrix (cal_smatrix_RGF.m:76-291). Returns (S, channels).

    Accepts a numeric ``yBC`` (Bloch), which the reference's validation at
    cal_smatrix_RGF.m:89 rejects even though its help text (:32-34) and
    setup_channels accept it.
    """
    _validate_syst(syst, "yBC")
    eps = np.asarray(syst["epsilon"])
    if eps.ndim == 1:
        eps = eps[:, None]
    ny, nx = eps.shape
    if ny == 0:
        raise ValueError("need at least one site along the transverse (y) direction")
    ch = setup_channels(ny, syst["yBC"], syst["k0dx"], syst["epsilon_L"], syst["epsilon_R"])
    A0 = slice_operator_2d(ny, syst["yBC"])
    slices = [eps[:, n] for n in range(nx)]
    S = _smatrix_common(slices, A0, ch["phi"], ch, syst, out, in_)
    return S, ch


def cal_smatrix_RGF_3D(syst: dict, out, in_):
    """3D S matrix (cal_smatrix_RGF_3D.m:74-306). ``epsilon`` has shape (nx, ny, nz)."""
    _validate_syst(syst, "xyBC")
    eps = np.asarray(syst["epsilon"])
    if eps.ndim == 2:
        eps = eps[:, :, None]
    nx, ny, nz = eps.shape
    if nx * ny == 0:
        raise ValueError("need at least one site along the transverse directions (x,y)")
    ch = setup_channels_3D(nx, ny, syst["xyBC"], syst["k0dx"], syst["epsilon_L"], syst["epsilon_R"])
    A0 = slice_operator_3d(nx, ny, syst["xyBC"])
    slices = [eps[:, :, n].ravel(order="F") for n in range(nz)]
    S = _smatrix_common(slices, A0, ch["phi"], ch, syst, out, in_)
    return S, ch


def brute_force_smatrix(slices, A0, k0dx2, phi, chL, chR, epsL, epsR,
                        ind_in_L, ind_in_R, ind_out_L, ind_out_R, G0_L=None, G0_R=None, return_fields=False):
    """Full S from one sparse solve of the whole block-tridiagonal system.

    With ``return_fields`` also returns the solution itself, the field
    profiles: an N x (L+2) x n_in array (slice n of input a is
    ``Psi[:, n, a]``), inputs L then R.

    Slices 0..L+1; slice 0 and L+1 are the lead surfaces with diagonal blocks
    A_lead - Sigma where Sigma = G0 is the lead self-energy (G0 satisfies
    G0 = (A_lead - G0)^-1), off-diagonal blocks -I.  No recursion is used, so
    this pins the recursion order, all five schemes and the final attach.
    """
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    N = A0.shape[0]
    L = len(slices)
    if G0_L is None:
        G0_L = lead_greens_function(phi, chL["kxdx"])
    if G0_R is None:
        G0_R = lead_greens_function(phi, chR["kxdx"])
    blocks = [[None] * (L + 2) for _ in range(L + 2)]
    I = sp.identity(N, dtype=complex, format="csr")
    blocks[0][0] = sp.csr_matrix(A0 - k0dx2 * epsL * np.eye(N) - G0_L)
    for n in range(L):
        blocks[n + 1][n + 1] = sp.csr_matrix(A0 - np.diag(k0dx2 * slices[n]))
    blocks[L + 1][L + 1] = sp.csr_matrix(A0 - k0dx2 * epsR * np.eye(N) - G0_R)
    for n in range(L + 1):
        blocks[n][n + 1] = -I
        blocks[n + 1][n] = -I
    H = sp.bmat(blocks, format="csc")
    B_L = phi[:, chL["ind_prop"][ind_in_L - 1] - 1] * chL["sqrt_mu"][ind_in_L - 1][None, :]
    B_R = phi[:, chR["ind_prop"][ind_in_R - 1] - 1] * chR["sqrt_mu"][ind_in_R - 1][None, :]
    C_L = chL["sqrt_mu"][ind_out_L - 1][:, None] * phi[:, chL["ind_prop"][ind_out_L - 1] - 1].conj().T
    C_R = chR["sqrt_mu"][ind_out_R - 1][:, None] * phi[:, chR["ind_prop"][ind_out_R - 1] - 1].conj().T
    m = B_L.shape[1] + B_R.shape[1]
    rhs = np.zeros((N * (L + 2), m), dtype=complex)
    rhs[:N, :B_L.shape[1]] = B_L
    rhs[N * (L + 1):, B_L.shape[1]:] = B_R
    Psi = spla.splu(H).solve(rhs)
    P0 = Psi[:N]
    P1 = Psi[N * (L + 1):]
    S = -2j * np.vstack([C_L @ P0, C_R @ P1])
    S = S - _delta(ind_out_L, ind_in_L, ind_out_R, ind_in_R, S.shape)
    if return_fields:
        return S, Psi.reshape(L + 2, N, m).transpose(1, 0, 2)
    return S


def transmission_gradient_fd(syst: dict, pixels, h: float = 1e-6, three_d: bool = False, part: str = "re"):
    """Central finite differences of T = sum |t_ij|^2 (in L, out R) w.r.t. the
    real (part="re") or imaginary (part="im") part of eps pixels."""
    solve = cal_smatrix_RGF_3D if three_d else cal_smatrix_RGF
    step = {"re": 1.0, "im": 1j}[part]
    cplx = np.iscomplexobj(syst["epsilon"]) or part == "im"
    eps0 = np.array(syst["epsilon"], dtype=complex if cplx else float)
    out = []
    for idx in pixels:
        vals = []
        for s in (+1, -1):
            e = eps0.copy()
            e[idx] = e[idx] + s * h * step
            S, _ = solve({**syst, "epsilon": e}, ["R"], ["L"])
            vals.append(float(np.sum(np.abs(S) ** 2)))
        out.append((vals[0] - vals[1]) / (2 * h))
    return np.array(out)


def transmission_gradient(syst: dict, three_d: bool = False):
    """T = sum |t_ij|^2 (inputs L, outputs R) and dT/dRe(eps), dT/dIm(eps), FP64.

    Independent of the slice recursion: the forward fields Psi = H^-1 [B_L at
    slice 0] and the adjoint rows Phi = C_R H^-1 (row of slice L+1) come from
    one sparse LU of the whole block-tridiagonal system H (slices 0..L+1, leads
    folded into the end blocks).  With t = -2i C_R G_{L+1,0} B_L and
    dH_nn = -k0dx^2 d(eps) on the diagonal,
        dt_ij / d eps_{m,n} = -2i k0dx^2 Phi_n[i,m] Psi_n[m,j]
    and for g = sum_ij conj(t_ij) dt_ij/d eps:  dT/dRe eps = 2 Re g,
    dT/dIm eps = -2 Im g.
    """
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    eps = np.asarray(syst["epsilon"])
    if three_d:
        nx, ny, L = eps.shape
        ch = setup_channels_3D(nx, ny, syst["xyBC"], syst["k0dx"], syst["epsilon_L"], syst["epsilon_R"])
        A0 = slice_operator_3d(nx, ny, syst["xyBC"])
        slices = [eps[:, :, n].ravel(order="F") for n in range(L)]
    else:
        ny, L = eps.shape
        ch = setup_channels(ny, syst["yBC"], syst["k0dx"], syst["epsilon_L"], syst["epsilon_R"])
        A0 = slice_operator_2d(ny, syst["yBC"])
        slices = [eps[:, n] for n in range(L)]
    N = A0.shape[0]
    k2 = complex(syst["k0dx"]) ** 2
    phi = ch["phi"]
    G0L = lead_greens_function(phi, ch["L"]["kxdx"])
    G0R = lead_greens_function(phi, ch["R"]["kxdx"])
    blocks = [[None] * (L + 2) for _ in range(L + 2)]
    I = sp.identity(N, dtype=complex, format="csr")
    blocks[0][0] = sp.csr_matrix(A0 - k2 * syst["epsilon_L"] * np.eye(N) - G0L)
    for n in range(L):
        blocks[n + 1][n + 1] = sp.csr_matrix(A0 - np.diag(k2 * slices[n]))
    blocks[L + 1][L + 1] = sp.csr_matrix(A0 - k2 * syst["epsilon_R"] * np.eye(N) - G0R)
    for n in range(L + 1):
        blocks[n][n + 1] = -I
        blocks[n + 1][n] = -I
    H = sp.bmat(blocks, format="csc")
    L_, R_ = ch["L"], ch["R"]
    B = phi[:, L_["ind_prop"] - 1] * L_["sqrt_mu"][None, :]
    C = R_["sqrt_mu"][:, None] * phi[:, R_["ind_prop"] - 1].conj().T
    lu = spla.splu(H)
    rhs = np.zeros((N * (L + 2), B.shape[1]), complex)
    rhs[:N] = B
    Psi = lu.solve(rhs)
    rhs2 = np.zeros((N * (L + 2), C.shape[0]), complex)
    rhs2[N * (L + 1):] = C.T
    PhiT = lu.solve(rhs2, trans="T")          # (C H^-1)^T = H^-T C^T
    t = -2j * (C @ Psi[N * (L + 1):])
    T = float(np.sum(np.abs(t) ** 2))
    gre = np.zeros((N, L))
    gim = np.zeros((N, L))
    for n in range(L):
        Ps = Psi[N * (n + 1):N * (n + 2)]          # N x nin
        Ph = PhiT[N * (n + 1):N * (n + 2)].T       # nout x N
        g = -2j * k2 * np.einsum("mj,ij,im->m", Ps, t.conj(), Ph)
        gre[:, n] = 2 * g.real
        gim[:, n] = -2 * g.imag
    if three_d:
        gre = gre.reshape((nx, ny, L), order="F")
        gim = gim.reshape((nx, ny, L), order="F")
    return T, gre, gim
