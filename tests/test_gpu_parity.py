"""GPU parity of librgf against the FP64 oracle and the golden fixtures.

Tolerances: complex FP64 path <= 1e-10 max-entry error against the oracle's
S (entries are O(1)); FP32+refinement path <= 1e-9.  Kernel-level checks
compare the MFMA GEMM and the slice-inverse kernel with numpy.
"""
import ast
import os

import numpy as np
import pytest

import oracle.rgf_oracle as o
import rgf_amd

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "rgf_golden.npz")
TOL64 = 1e-10
TOL_MIXED = 1e-9


def _crand(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-13), ("fp32_refined", 2e-6)])
@pytest.mark.parametrize("mnk", [(1, 1, 1), (17, 33, 5), (64, 64, 64), (130, 70, 200), (300, 257, 129),
                                 (513, 1100, 96),
                                 # thin products with a long inner dimension: split along K over the
                                 # streamlined tiles (ragged M, N a multiple of 4) and, with N = 70
                                 # (rows not 16-byte aligned), over the general tiles; both reduced
                                 # by the split-K reduction kernel
                                 (1000, 92, 2048), (4096, 184, 1024), (300, 70, 1500)])
def test_mfma_complex_gemm(precision, tol, mnk):
    m, n, k = mnk
    rng = np.random.default_rng(m * 7 + n * 3 + k)
    A, B = _crand(rng, m, k), _crand(rng, k, n)
    Cm = rgf_amd.dense_gemm(A, B, precision=precision)
    ref = A @ B
    scale = np.abs(A).max() * np.abs(B).max() * k
    assert np.abs(Cm - ref).max() / scale < tol


def _well_posed(rng, n):
    return _crand(rng, n, n) / np.sqrt(n) + 4 * np.eye(n)


@pytest.mark.parametrize("n,b", [(1, 0), (5, 0), (64, 0), (100, 16), (200, 32), (333, 64), (700, 0), (129, 64),
                                 (130, 64), (192, 64), (1000, 32)])
def test_slice_inverse_fp64(n, b):
    rng = np.random.default_rng(n + b)
    A = _well_posed(rng, n)
    X = rgf_amd.dense_inverse(A, "fp64", b)
    assert np.abs(X @ A - np.eye(n)).max() < 1e-12


@pytest.mark.parametrize("n,b", [(7, 0), (128, 0), (300, 64), (1000, 128), (200, 128), (301, 64), (520, 128),
                                 (384, 128)])
def test_slice_inverse_fp32(n, b):
    rng = np.random.default_rng(n + b + 1)
    A = _well_posed(rng, n)
    X = rgf_amd.dense_inverse(A, "fp32_refined", b)
    assert np.abs(X @ A - np.eye(n)).max() < 5e-5


@pytest.mark.parametrize("n,b", [(128, 0), (520, 128), (1000, 128), (301, 64)])
def test_slice_inverse_fp32_3m(n, b):
    """The FP32 inverse with three real products per complex product (the
    default of the mixed path) at the accuracy of the four-product kernels."""
    rng = np.random.default_rng(n + b + 7)
    A = _well_posed(rng, n)
    X = rgf_amd.dense_inverse(A, "fp32_refined", b, complex_3m=True)
    assert np.abs(X @ A - np.eye(n)).max() < 5e-5


def test_dense_inverse_3m_flag_needs_fp32():
    # invalid arguments come back as ValueError (RGF_ERR_INVALID_ARGUMENT)
    with pytest.raises(ValueError, match="FP32 kernels"):
        rgf_amd.dense_inverse(np.eye(8, dtype=complex), "fp64", complex_3m=True)


def test_slice_inverse_pivots_inside_block():
    # zero diagonal: needs row interchanges within the diagonal block
    n = 48
    rng = np.random.default_rng(3)
    P = np.eye(n)[rng.permutation(n)]
    A = P @ (np.eye(n) * 3 + 0.1 * _crand(rng, n, n))
    A[np.arange(n), np.arange(n)] = 0.0
    X = rgf_amd.dense_inverse(A, "fp64", 64)
    assert np.abs(X - np.linalg.inv(A)).max() < 1e-12


@pytest.mark.parametrize("precision,n,b,tol", [("fp64", 96, 64, 1e-12), ("fp64", 200, 0, 1e-12),
                                               ("fp32_refined", 256, 128, 2e-5), ("fp32_refined", 150, 64, 2e-5)])
@pytest.mark.parametrize("where", ["first", "later", "small"])
def test_slice_inverse_fallback(precision, n, b, tol, where):
    # The unpivoted register path must hand the block to the pivoting path:
    # a zero pivot in the first panel, a zero pivot in a later panel, or a
    # tiny pivot whose multipliers exceed the growth bound.
    rng = np.random.default_rng(n + b + len(where))
    A = 3 * np.eye(n) + 0.2 * _crand(rng, n, n) / np.sqrt(n)
    k = {"first": 0, "later": 37, "small": 21}[where]
    A[k, k] = 0.0 if where != "small" else 1e-7
    A[k, (k + 5) % n] = 2.0
    A[(k + 5) % n, k] = 1.5
    X = rgf_amd.dense_inverse(A, precision, b)
    ref = np.linalg.inv(A)
    assert np.abs(X - ref).max() / np.abs(ref).max() < tol


def _second_block_nearly_singular(rng, n=256, b=64, scale=1e-9):
    """A well-conditioned matrix whose Schur complement block at rows and
    columns [b, 2b) is nearly singular: a block-row swap of a diagonally
    dominant matrix whose rows [2b, 3b) are tiny in the columns [0, 2b)."""
    B = 3 * np.eye(n) + 0.2 * _crand(rng, n, n) / np.sqrt(n)
    B[2 * b:3 * b, :2 * b] *= scale
    P = np.arange(n)
    P[b:2 * b], P[2 * b:3 * b] = np.arange(2 * b, 3 * b), np.arange(b, 2 * b)
    return B[P]


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32_refined", 2e-5)])
def test_slice_inverse_ill_conditioned_block_pivots_over_all_rows(precision, tol):
    # The blocked kernels pivot inside the diagonal block only: this
    # well-conditioned matrix, whose second 64-wide Schur-complement block is
    # nearly singular, is refused by their growth check (not inverted silently
    # with the growth of its multipliers) and inverted again with partial
    # pivoting over all rows, as MATLAB's inv does (cal_smatrix_RGF.m:265).
    rng = np.random.default_rng(11)
    A = _second_block_nearly_singular(rng)
    assert np.linalg.cond(A) < 10
    X = rgf_amd.dense_inverse(A, precision, 64)
    ref = np.linalg.inv(A)
    assert np.abs(X - ref).max() / np.abs(ref).max() < tol


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32_refined", 2e-5)])
@pytest.mark.parametrize("n", [5, 200, 333])
def test_pivoted_slice_inverse(precision, tol, n):
    # the pivoted inverse on its own (block_size < 0), including a matrix
    # whose leading entries are zero (row interchanges across the whole matrix)
    rng = np.random.default_rng(n + 17)
    A = _well_posed(rng, n)
    A[[0, n - 1]] = A[[n - 1, 0]]
    X = rgf_amd.dense_inverse(A, precision, -1)
    ref = np.linalg.inv(A)
    assert np.abs(X - ref).max() / np.abs(ref).max() < tol


def test_pivoted_slice_inverse_singular_is_refused():
    A = np.ones((40, 40), dtype=complex)
    with pytest.raises(rgf_amd.RGFError, match="zero or non-finite pivot"):
        rgf_amd.dense_inverse(A, "fp64", -1)


@pytest.mark.parametrize("precision,tol", [("fp64", TOL64), ("fp32_refined", TOL_MIXED)])
def test_high_index_slab_vs_oracle(precision, tol):
    # eps up to 12 at 6 px per vacuum wavelength: k0dx^2 eps runs through the
    # band edge 4 (slice matrices far from diagonal dominance)
    rng = np.random.default_rng(12)
    eps = 1 + 11 * rng.random((256, 40))
    syst = dict(epsilon=eps, epsilon_L=1.0, epsilon_R=2.0, k0dx=2 * np.pi / 6, yBC="periodic")
    S, _ = rgf_amd.cal_smatrix_RGF(syst, ["L", "R"], ["L", "R"], precision=precision)
    Sref, _ = o.cal_smatrix_RGF(syst, ["L", "R"], ["L", "R"])
    err = float(np.abs(S - Sref).max())
    print(f"high-index slab {precision}: {err:.2e}", flush=True)
    assert err < tol


def _cells(s):
    return [c for c in s]


def _load_golden():
    z = np.load(GOLDEN, allow_pickle=False)
    return z, [str(x) for x in z["names"]]


def _golden_case(z, name):
    eps = z[f"{name}/eps"]
    k0dx, eL, eR = z[f"{name}/params"]
    bc = str(z[f"{name}/bc"])
    out, inn = str(z[f"{name}/sel"]).split("|")
    S = z[f"{name}/S"]
    k0dx = k0dx.real if k0dx.imag == 0 else k0dx
    eL = eL.real if eL.imag == 0 else eL
    eR = eR.real if eR.imag == 0 else eR
    return eps, k0dx, eL, eR, bc, out, inn, S


@pytest.mark.parametrize("precision,tol", [("fp64", TOL64), ("fp32_refined", TOL_MIXED)])
def test_golden_fixtures(precision, tol):
    z, names = _load_golden()
    for name in names:
        eps, k0dx, eL, eR, bc, out, inn, Sref = _golden_case(z, name)
        if eps.ndim == 3:
            xy = bc if bc in ("periodic", "Dirichlet") else ast.literal_eval(bc)
            syst = dict(epsilon=eps, epsilon_L=eL, epsilon_R=eR, k0dx=k0dx, xyBC=xy)
            S, _ = rgf_amd.cal_smatrix_RGF_3D(syst, _cells(out), _cells(inn), precision=precision)
        else:
            try:
                yb = float(bc)
            except ValueError:
                yb = bc
            syst = dict(epsilon=eps, epsilon_L=eL, epsilon_R=eR, k0dx=k0dx, yBC=yb)
            S, _ = rgf_amd.cal_smatrix_RGF(syst, _cells(out), _cells(inn), precision=precision)
        assert S.shape == Sref.shape, name
        err = np.abs(S - Sref).max()
        print(f"{name} {precision} err={err:.2e}", flush=True)
        assert err < tol, (name, err)


SELECTIONS = [
    (["L", "R"], ["L", "R"]), (["L"], ["L"]), (["R"], ["R"]), (["R"], ["L"]), (["L"], ["R"]),
    (["L", "R"], ["L"]), (["L", "R"], ["R"]), (["R", "L"], ["L"]),
]


@pytest.mark.parametrize("bc", ["periodic", "Dirichlet", "Neumann", "DirichletNeumann", "NeumannDirichlet", 1.1])
@pytest.mark.parametrize("sel", range(len(SELECTIONS)))
@pytest.mark.parametrize("precision,tol", [("fp64", TOL64), ("fp32_refined", TOL_MIXED)])
def test_smatrix_2d_schemes(bc, sel, precision, tol):
    out, inn = SELECTIONS[sel]
    rng = np.random.default_rng(sel * 31 + len(str(bc)))
    ny, nx = 37, 23
    eps = 1 + 2 * rng.random((ny, nx)) + 0.02j * rng.random((ny, nx))
    syst = dict(epsilon=eps, epsilon_L=1.0, epsilon_R=1.69, k0dx=2 * np.pi / 6.5, yBC=bc)
    S, _ = rgf_amd.cal_smatrix_RGF(syst, out, inn, precision=precision)
    Sref, _ = o.cal_smatrix_RGF(syst, out, inn)
    assert np.abs(S - Sref).max() < tol


def test_struct_selection_and_delta():
    rng = np.random.default_rng(9)
    syst = dict(epsilon=1 + rng.random((30, 10)), epsilon_L=1.0, epsilon_R=1.0, k0dx=1.0, yBC="periodic")
    out = {"ind_out_L": [3, 1, 3], "ind_out_R": [2, 5]}
    inn = {"ind_in_L": [1, 3], "ind_in_R": [5, 2, 1]}
    S, _ = rgf_amd.cal_smatrix_RGF(syst, out, inn)
    Sref, _ = o.cal_smatrix_RGF(syst, out, inn)
    assert np.abs(S - Sref).max() < TOL64


@pytest.mark.parametrize("precision,tol", [("fp64", TOL64), ("fp32_refined", TOL_MIXED)])
@pytest.mark.parametrize("where", ["cuda", "cpu"])
def test_epsilon_as_tensor(precision, tol, where):
    """syst.epsilon given as a torch tensor (device pointers for a CUDA
    tensor, host pointers for a CPU one) gives the S of the numpy input:
    2D complex with the sweep mirrored or not, 3D from a non-contiguous view."""
    import torch
    rng = np.random.default_rng(77)
    eps2 = 1 + 2 * rng.random((37, 23)) + 0.02j * rng.random((37, 23))
    syst = dict(epsilon=eps2, epsilon_L=1.0, epsilon_R=1.69, k0dx=2 * np.pi / 6.5, yBC="Neumann")
    tsyst = dict(syst, epsilon=torch.as_tensor(eps2, device=where))
    for out, inn in ((["L", "R"], ["L"]), (["L"], ["R"])):
        S, _ = rgf_amd.cal_smatrix_RGF(syst, out, inn, precision=precision)
        St, _ = rgf_amd.cal_smatrix_RGF(tsyst, out, inn, precision=precision)
        assert np.abs(St - S).max() <= 1e-13
        Sref, _ = o.cal_smatrix_RGF(syst, out, inn)
        assert np.abs(St - Sref).max() < tol
    e3 = 1 + 1.1 * rng.random((7, 5, 6))
    t3 = torch.as_tensor(e3, device=where).permute(2, 1, 0)        # (6, 5, 7), not contiguous
    syst3 = dict(epsilon=np.ascontiguousarray(e3.transpose(2, 1, 0)), epsilon_L=1.0, epsilon_R=1.3, k0dx=1.2,
                 xyBC=[0.7, 2.9])
    S, _ = rgf_amd.cal_smatrix_RGF_3D(syst3, ["L", "R"], ["L", "R"], precision=precision)
    St, _ = rgf_amd.cal_smatrix_RGF_3D(dict(syst3, epsilon=t3), ["L", "R"], ["L", "R"], precision=precision)
    assert np.abs(St - S).max() <= 1e-13
    Sref, _ = o.cal_smatrix_RGF_3D(syst3, ["L", "R"], ["L", "R"])
    assert np.abs(St - Sref).max() < tol


def test_epsilon_tensor_device_mismatch():
    import torch
    syst = dict(epsilon=torch.ones((8, 5), dtype=torch.float64, device="cuda:0"), epsilon_L=1.0, epsilon_R=1.0,
                k0dx=1.0, yBC="periodic")
    with pytest.raises(ValueError, match="is on device 0"):
        rgf_amd.cal_smatrix_RGF(syst, ["R"], ["L"], device=1)


@pytest.mark.parametrize("precision,tol", [("fp64", TOL64), ("fp32_refined", TOL_MIXED)])
def test_smatrix_3d(precision, tol):
    rng = np.random.default_rng(21)
    for bc in ["periodic", "Dirichlet", [0.7, 2.9]]:
        nx, ny, nz = 6, 5, 7
        eps = 1 + 1.1 * rng.random((nx, ny, nz))
        syst = dict(epsilon=eps, epsilon_L=1.0, epsilon_R=1.3, k0dx=1.2, xyBC=bc)
        for out, inn in (["L", "R"], ["L", "R"]), (["R"], ["L"]):
            S, _ = rgf_amd.cal_smatrix_RGF_3D(syst, out, inn, precision=precision)
            Sref, _ = o.cal_smatrix_RGF_3D(syst, out, inn)
            assert np.abs(S - Sref).max() < tol


@pytest.mark.parametrize("shape", [(1, 6, 4), (6, 1, 4), (1, 1, 3), (9, 2, 3), (2, 9, 3)])
@pytest.mark.parametrize("bc", ["periodic", "Dirichlet", [1.3, -0.4]])
def test_smatrix_3d_thin_transverse(shape, bc):
    """Transverse shapes where one direction has one or two sites: the lead
    Green's function goes through the separable two-contraction form when
    ny > 1 and through the dense product when ny = 1."""
    rng = np.random.default_rng(sum(shape) + len(str(bc)))
    eps = 1 + 1.5 * rng.random(shape) + 0.01j * rng.random(shape)
    syst = dict(epsilon=eps, epsilon_L=1.0, epsilon_R=1.44, k0dx=2.2, xyBC=bc)
    S, _ = rgf_amd.cal_smatrix_RGF_3D(syst, ["L", "R"], ["L", "R"])
    Sref, _ = o.cal_smatrix_RGF_3D(syst, ["L", "R"], ["L", "R"])
    assert S.shape == Sref.shape
    if S.size:
        assert np.abs(S - Sref).max() < TOL64


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_mixed_3m_modes(mode):
    """Three-product complex GEMMs in the slice inverses (bit 0) and/or the
    refinement sweeps (bit 1): the refined S meets the mixed tolerance against
    the oracle and FP64 on a 3D problem with several slice-inverse blocks."""
    rng = np.random.default_rng(31)
    eps = 1 + 1.5 * rng.random((16, 18, 12))
    syst = dict(epsilon=eps, epsilon_L=1.0, epsilon_R=1.2, k0dx=1.0, xyBC="periodic")
    S, _, info = rgf_amd.cal_smatrix_RGF_3D(syst, ["L", "R"], ["L", "R"], precision="fp32_refined",
                                            options={"complex_3m": mode, "block_size": 64}, return_info=True)
    S64, _ = rgf_amd.cal_smatrix_RGF_3D(syst, ["L", "R"], ["L", "R"], precision="fp64")
    Sref, _ = o.cal_smatrix_RGF_3D(syst, ["L", "R"], ["L", "R"])
    assert np.abs(S64 - Sref).max() < TOL64
    assert np.abs(S - Sref).max() < TOL_MIXED
    assert info["refine_residual"] <= 1e-13


def _fields_ref_2d(syst, out, inn):
    """Oracle fields by the brute-force full-system solve (2D)."""
    eps = np.asarray(syst["epsilon"])
    ny, nx = eps.shape
    S, ch = o.cal_smatrix_RGF(syst, out, inn)
    iL = np.arange(1, ch["L"]["N_prop"] + 1)
    iR = np.arange(1, ch["R"]["N_prop"] + 1)
    sel_in = {"L": iL, "R": iR}
    inL = sel_in["L"] if "L" in inn else np.zeros(0, dtype=int)
    inR = sel_in["R"] if "R" in inn else np.zeros(0, dtype=int)
    A0 = o.slice_operator_2d(ny, syst["yBC"])
    Sb, F = o.brute_force_smatrix([eps[:, n] for n in range(nx)], A0, syst["k0dx"] ** 2, ch["phi"], ch["L"],
                                  ch["R"], syst["epsilon_L"], syst["epsilon_R"], inL, inR, iL, iR,
                                  return_fields=True)
    return F


@pytest.mark.parametrize("bc,inn", [("periodic", ["L", "R"]), ("Dirichlet", ["R"]), (0.9, ["L"])])
def test_field_profiles_2d(bc, inn):
    """Fields of the refined solve against the brute-force full-system solve
    (with inputs ["L"] only, the sweep runs mirrored: more left selections)."""
    rng = np.random.default_rng(41)
    eps = 1 + 2 * rng.random((14, 11))
    syst = dict(epsilon=eps, epsilon_L=1.0, epsilon_R=1.5, k0dx=1.1, yBC=bc)
    S, ch, info, F = rgf_amd.cal_smatrix_RGF(syst, ["L", "R"], inn, precision="fp32_refined", return_info=True,
                                             return_fields=True)
    Fref = _fields_ref_2d(syst, ["L", "R"], inn)
    assert F.shape == Fref.shape == (14, 13, S.shape[1])
    assert np.abs(F - Fref).max() <= 1e-9 * np.abs(Fref).max()
    Sref, _ = o.cal_smatrix_RGF(syst, ["L", "R"], inn)
    assert np.abs(S - Sref).max() < TOL_MIXED
    assert info["swept_right_to_left"] == (inn == ["L"])


def test_field_profiles_3d():
    rng = np.random.default_rng(42)
    nx, ny, nz = 5, 4, 6
    eps = 1 + rng.random((nx, ny, nz))
    syst = dict(epsilon=eps, epsilon_L=1.0, epsilon_R=1.3, k0dx=1.2, xyBC=[0.4, 2.2])
    S, ch, F = rgf_amd.cal_smatrix_RGF_3D(syst, ["L", "R"], ["L", "R"], precision="fp32_refined",
                                          return_fields=True)
    Sref, chr_ = o.cal_smatrix_RGF_3D(syst, ["L", "R"], ["L", "R"])
    iL = np.arange(1, chr_["L"]["N_prop"] + 1)
    iR = np.arange(1, chr_["R"]["N_prop"] + 1)
    A0 = o.slice_operator_3d(nx, ny, [0.4, 2.2])
    _, Fref = o.brute_force_smatrix([eps[:, :, n].ravel(order="F") for n in range(nz)], A0, 1.2 ** 2,
                                    chr_["phi"], chr_["L"], chr_["R"], 1.0, 1.3, iL, iR, iL, iR, return_fields=True)
    assert F.shape == (nx, ny, nz + 2, S.shape[1])
    Fr = Fref.reshape((nx, ny, nz + 2, S.shape[1]), order="F")
    assert np.abs(F - Fr).max() <= 1e-9 * np.abs(Fr).max()
    assert np.abs(S - Sref).max() < TOL_MIXED


def test_field_profiles_need_refined_precision():
    syst = dict(epsilon=np.ones((6, 4)), epsilon_L=1.0, epsilon_R=1.0, k0dx=1.0, yBC="periodic")
    with pytest.raises(ValueError, match="refined"):
        rgf_amd.cal_smatrix_RGF(syst, ["L", "R"], ["L", "R"], precision="fp64", return_fields=True)


def test_scan_matches_single_solves_and_oracle(tmp_path):
    """Frequency x Bloch-phase scan through the batched entry, with inputs
    picked by incident angle, against single solves and the oracle; the
    results round-trip through .npz and .mat."""
    from rgf_amd.sweep import load_scan, scan
    rng = np.random.default_rng(51)
    eps = 1 + 2 * rng.random((40, 12))
    syst = dict(epsilon=eps, epsilon_L=1.0, epsilon_R=1.0, k0dx=0.5, yBC="periodic")
    ks = [2 * np.pi / 12, 2 * np.pi / 10]
    bcs = [0.3, 1.7]
    res = scan(syst, k0dx=ks, bc=bcs, out=["L", "R"], angles_deg=[-30, 0, 30], chunk=3,
               path=str(tmp_path / "s.npz"))
    assert len(res["S"]) == 4
    for q, (k, b) in enumerate([(k, b) for k in ks for b in bcs]):
        s = dict(syst, k0dx=k, yBC=b)
        inn = {"ind_in_L": res["in_index"][q]}
        S1, _ = rgf_amd.cal_smatrix_RGF(s, ["L", "R"], inn)
        assert np.abs(res["S"][q] - S1).max() <= 1e-12
        Sref, _ = o.cal_smatrix_RGF(s, ["L", "R"], inn)
        assert np.abs(res["S"][q] - Sref).max() < TOL64
    back = load_scan(str(tmp_path / "s.npz"))
    assert all(np.array_equal(a, b) for a, b in zip(back["S"], res["S"]))
    from rgf_amd.sweep import save_scan
    save_scan(res, str(tmp_path / "s.mat"))
    backm = load_scan(str(tmp_path / "s.mat"))
    assert all(np.array_equal(a, b) for a, b in zip(backm["S"], res["S"]))


def _cfg1(seed=1):
    rng = np.random.default_rng(seed)
    eps = 1 + 3 * rng.random((64, 200))
    return dict(epsilon=eps, epsilon_L=1.0, epsilon_R=1.0, k0dx=2 * np.pi / 15, yBC="periodic")


@pytest.mark.parametrize("precision,tol", [("fp64", TOL64), ("fp32_refined", TOL_MIXED)])
def test_cfg1_full_size(precision, tol):
    """2D N=64, 200 slices, periodic, eps~U[1,4], lambda=15 px, full S (18 x 18)."""
    syst = _cfg1()
    S, ch, info = rgf_amd.cal_smatrix_RGF(syst, ["L", "R"], ["L", "R"], precision=precision, return_info=True)
    Sref, _ = o.cal_smatrix_RGF(syst, ["L", "R"], ["L", "R"])
    assert S.shape == (18, 18)
    assert np.abs(S - Sref).max() < tol
    assert np.abs(S.conj().T @ S - np.eye(18)).max() < 1e-9
    perm = np.concatenate([ch["L"]["ind_prop_conj"], ch["L"]["N_prop"] + ch["R"]["ind_prop_conj"]]) - 1
    St = S[perm]
    assert np.abs(St.T - St).max() < 1e-9
    print(info, flush=True)


def test_wide_slab_mixed_vs_fp64():
    """2D N=1024 Dirichlet slab with absorption and eps_R != eps_L (cfg-2 shape, 40 slices)."""
    rng = np.random.default_rng(2)
    eps = 1 + 1.25 * (rng.random((1024, 40)) < 0.3) + 0.01j
    syst = dict(epsilon=eps, epsilon_L=1.0, epsilon_R=2.25, k0dx=2 * np.pi / 10, yBC="Dirichlet")
    S64, ch = rgf_amd.cal_smatrix_RGF(syst, ["L", "R"], ["L", "R"])
    S32, _ = rgf_amd.cal_smatrix_RGF(syst, ["L", "R"], ["L", "R"], precision="fp32_refined")
    assert S64.shape == (208 + 320, 208 + 320)
    assert np.abs(S32 - S64).max() < TOL_MIXED
    assert np.linalg.eigvalsh(S64.conj().T @ S64).max() <= 1 + 1e-9


def test_batched_ragged_vs_single_and_oracle():
    """rgf_solve_batched on a ragged batch (2D widths 24..64, a 3D slab, mixed
    channel requests, both precisions): every S equals the single-solve S of
    the same problem and matches the oracle."""
    rng = np.random.default_rng(11)
    items = []
    for k, ny in enumerate([24, 40, 64, 32, 48, 24, 56, 40, 64]):
        nx = int(rng.integers(10, 40))
        eps = 1 + 2 * rng.random((ny, nx)) + 0.01j * rng.random((ny, nx))
        lam = 8 + 12 * rng.random()
        syst = dict(epsilon=eps, epsilon_L=1.0, epsilon_R=1.0 + 0.5 * (k % 2),
                    k0dx=(2 * np.pi / lam) * (1 + 1e-3j * rng.random()), yBC=["periodic", "Dirichlet", 0.7][k % 3])
        sel = (["L", "R"], ["L", "R"]) if k % 3 else (["R"], {"ind_in_L": [1, 2]})
        items.append((syst, *sel))
    eps3 = 1 + rng.random((5, 6, 7))
    items.append((dict(epsilon=eps3, epsilon_L=1.0, epsilon_R=1.2, k0dx=1.1, xyBC=[0.3, 1.7]), ["L", "R"], ["L"]))
    # fp32_refined with batch_engine 2: every problem on the worker contexts
    # (the lockstep engine has tests of its own)
    for precision, tol in (("fp64", 1e-10), ("fp32_refined", 1e-9)):
        Sb = rgf_amd.cal_smatrix_RGF_batched(items, precision=precision,
                                             options={"batch_concurrency": 4, "batch_engine": 2})
        for (syst, out, inn), S in zip(items, Sb):
            three_d = np.ndim(syst["epsilon"]) == 3
            single = rgf_amd.cal_smatrix_RGF_3D if three_d else rgf_amd.cal_smatrix_RGF
            S1, _ = single(syst, out, inn, precision=precision)
            assert S.shape == S1.shape
            assert np.abs(S - S1).max() <= 1e-12
            So = (o.cal_smatrix_RGF_3D if three_d else o.cal_smatrix_RGF)(syst, out, inn)[0]
            assert np.abs(S - So).max() < tol


def test_batched_sweep_workload_timing():
    """The 256-solve sweep workload shape (widths 128..512, complex k0dx), cut
    to 32 solves of 100 slices: batched results equal single solves, and the
    batched call is timed against the sequential loop."""
    import time
    from rgf_amd.workloads import batched_sweep
    items = [(dict(s, epsilon=s["epsilon"][:, :100]), out, inn) for s, out, inn in batched_sweep(count=32)]
    t0 = time.perf_counter()
    Sb = rgf_amd.cal_smatrix_RGF_batched(items)
    tb = time.perf_counter() - t0
    t0 = time.perf_counter()
    Ss = [rgf_amd.cal_smatrix_RGF(s, out, inn)[0] for s, out, inn in items]
    ts = time.perf_counter() - t0
    for a, b in zip(Sb, Ss):
        assert np.abs(a - b).max() <= 1e-12
        assert np.linalg.eigvalsh(a.conj().T @ a).max() <= 1 + 1e-9  # absorbing (Im k0dx > 0): passive
    print(f"batched {tb:.3f} s, sequential {ts:.3f} s, speed-up {ts / tb:.2f}", flush=True)


def test_cfg2_full_size_mixed_vs_fp64():
    """2D N=1024 Dirichlet, 2000 slices, eps = 1 + 1.25 Bernoulli(0.3) + 0.01i,
    eps_L = 1, eps_R = 2.25, k0dx = 2 pi / 10, full S (528 x 528): the mixed
    path against the FP64 path, passivity of both."""
    rng = np.random.default_rng(12)
    eps = 1 + 1.25 * (rng.random((1024, 2000)) < 0.3) + 0.01j
    syst = dict(epsilon=eps, epsilon_L=1.0, epsilon_R=2.25, k0dx=2 * np.pi / 10, yBC="Dirichlet")
    S64, _ = rgf_amd.cal_smatrix_RGF(syst, ["L", "R"], ["L", "R"])
    S32, _, info = rgf_amd.cal_smatrix_RGF(syst, ["L", "R"], ["L", "R"], precision="fp32_refined", return_info=True)
    assert S64.shape == (528, 528)
    assert np.abs(S32 - S64).max() < TOL_MIXED
    for S in (S64, S32):
        assert np.linalg.eigvalsh(S.conj().T @ S).max() <= 1 + 1e-9
    print(f"cfg2: refine iterations {info['refine_iters']}, residual {info['refine_residual']:.2e}", flush=True)


def test_cfg3_full_size_mixed_vs_fp64():
    """3D N = 64 x 64 = 4096, 128 slices, 40 spheres eps 2.1, Bloch random
    phases, lambda 12 px, full S (182 x 182): the mixed path at the library's
    tolerance and at the benchmark's (relative residual 1e-9) against the FP64
    path; all three unitary (real eps, equal leads)."""
    from rgf_amd.workloads import spheres_3d
    syst, out, inn = spheres_3d()
    S64, _ = rgf_amd.cal_smatrix_RGF_3D(syst, out, inn)
    S32, _, info = rgf_amd.cal_smatrix_RGF_3D(syst, out, inn, precision="fp32_refined", return_info=True)
    Sb, _, ib = rgf_amd.cal_smatrix_RGF_3D(syst, out, inn, precision="fp32_refined", options={"refine_tol": 1e-9},
                                           return_info=True)
    assert S64.shape == (182, 182)
    assert np.abs(S32 - S64).max() < 1e-11
    assert np.abs(Sb - S64).max() < TOL_MIXED
    assert ib["refine_residual"] <= 1e-9 and ib["refine_iters"] <= info["refine_iters"]
    for S in (S64, S32, Sb):
        assert np.abs(S.conj().T @ S - np.eye(182)).max() < TOL_MIXED
    Ss, _, is_ = rgf_amd.cal_smatrix_RGF_3D(syst, out, inn, precision="fp32_refined",
                                            options={"refine_tol": 1e-16, "refine_s_tol": 1e-9}, return_info=True)
    err_s = np.abs(Ss - S64).max()
    assert err_s <= is_["refine_s_err_estimate"] <= 1e-9
    print(f"cfg3: refine {info['refine_iters']} / {ib['refine_iters']} / {is_['refine_iters']} sweeps, "
          f"S err {np.abs(S32 - S64).max():.2e} / {np.abs(Sb - S64).max():.2e} / {err_s:.2e} "
          f"(estimate {is_['refine_s_err_estimate']:.2e})", flush=True)


def test_refinement_stops_on_s_error_estimate():
    """refine_s_tol: the refinement stops once q/(1-q) max|S_k - S_(k-1)| (x2)
    is below the tolerance; the S error against the FP64 path stays below the
    estimate (2D slabs of several kinds, the residual criterion switched off)."""
    rng = np.random.default_rng(31)
    cases = [dict(epsilon=1 + 3 * rng.random((64, 120)), epsilon_L=1.0, epsilon_R=1.0, k0dx=2 * np.pi / 15,
                  yBC="periodic"),
             dict(epsilon=1 + 1.25 * (rng.random((96, 80)) < 0.3) + 0.01j, epsilon_L=1.0, epsilon_R=2.25,
                  k0dx=2 * np.pi / 10, yBC="Dirichlet"),
             dict(epsilon=1 + rng.random((48, 200)), epsilon_L=1.3, epsilon_R=1.0, k0dx=0.9, yBC=1.1)]
    for syst in cases:
        S64, _ = rgf_amd.cal_smatrix_RGF(syst, ["L", "R"], ["L", "R"])
        for stol in (1e-6, 1e-9):
            S, _, info = rgf_amd.cal_smatrix_RGF(syst, ["L", "R"], ["L", "R"], precision="fp32_refined",
                                                 options={"refine_s_tol": stol, "refine_tol": 1e-16},
                                                 return_info=True)
            err = np.abs(S - S64).max()
            est = info["refine_s_err_estimate"]
            assert 0 < est <= stol, (est, stol)
            assert err <= est + 1e-13, (err, est)


def test_batched_s_error_stop():
    """The S-error stop through the batched entry: every S within 1e-9 of the
    FP64 path with the residual criterion switched off."""
    rng = np.random.default_rng(41)
    items = []
    for k, ny in enumerate([32, 48, 64, 40]):
        eps = 1 + 2 * rng.random((ny, 60)) + 0.005j * rng.random((ny, 60))
        syst = dict(epsilon=eps, epsilon_L=1.0, epsilon_R=1.0 + 0.3 * (k % 2), k0dx=2 * np.pi / (9 + 2 * k),
                    yBC=["periodic", "Dirichlet"][k % 2])
        items.append((syst, ["L", "R"], ["L", "R"]))
    S64 = rgf_amd.cal_smatrix_RGF_batched(items)
    S32 = rgf_amd.cal_smatrix_RGF_batched(items, precision="fp32_refined",
                                          options={"refine_s_tol": 1e-9, "refine_tol": 1e-16})
    for a, b in zip(S32, S64):
        assert np.abs(a - b).max() <= 1e-9
