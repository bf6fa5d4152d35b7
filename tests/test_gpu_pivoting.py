"""Solves with the slice inverses pivoted over all rows (rgf_options.pivoting):
the fallback the solve drivers take when the blocked Gauss-Jordan kernels
refuse a slice, forced from the start (pivoting = 1) and compared with the
default path on problems both can solve."""
import numpy as np
import pytest

import rgf_amd

pytestmark = pytest.mark.gpu


def _slab(seed, n=96, L=30, yBC="periodic"):
    rng = np.random.default_rng(seed)
    eps = 1 + 3 * rng.random((n, L)) + 0.01j * rng.random((n, L))
    return dict(epsilon=eps, epsilon_L=1.0, epsilon_R=2.25, k0dx=2 * np.pi / 11, yBC=yBC)


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-11), ("fp32_refined", 1e-9)])
@pytest.mark.parametrize("yBC", ["periodic", "Dirichlet"])
def test_pivoted_solve_matches_default(precision, tol, yBC):
    syst = _slab(3, yBC=yBC)
    S0, _, i0 = rgf_amd.cal_smatrix_RGF(syst, ["L", "R"], ["L", "R"], precision=precision, return_info=True)
    S1, _, i1 = rgf_amd.cal_smatrix_RGF(syst, ["L", "R"], ["L", "R"], precision=precision,
                                        options=dict(pivoting=1), return_info=True)
    assert i0["pivoted"] == 0 and i1["pivoted"] == 1
    err = float(np.abs(S1 - S0).max())
    print(f"pivoted vs blocked ({precision}, {yBC}): {err:.2e}", flush=True)
    assert err < tol


def test_pivoted_solve_3d_and_no_fallback_option():
    rng = np.random.default_rng(5)
    eps = 1 + 2 * rng.random((8, 9, 12))
    syst = dict(epsilon=eps, epsilon_L=1.0, epsilon_R=1.5, k0dx=2 * np.pi / 7, xyBC=[0.7, 2.9])
    S0, _ = rgf_amd.cal_smatrix_RGF_3D(syst, ["L", "R"], ["L", "R"], precision="fp64")
    S1, _ = rgf_amd.cal_smatrix_RGF_3D(syst, ["L", "R"], ["L", "R"], precision="fp64", options=dict(pivoting=1))
    S2, _ = rgf_amd.cal_smatrix_RGF_3D(syst, ["L", "R"], ["L", "R"], precision="fp64", options=dict(pivoting=2))
    assert np.abs(S1 - S0).max() < 1e-11
    assert np.array_equal(S2, S0)  # nothing refused: the same kernels, the same numbers


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-9), ("fp32_refined", 1e-7)])
def test_pivoted_gradient_matches_default(precision, tol):
    syst = _slab(7, n=64, L=20)
    T0, g0r, g0i = rgf_amd.transmission_gradient(syst, ["R"], ["L"], precision=precision)
    T1, g1r, g1i = rgf_amd.transmission_gradient(syst, ["R"], ["L"], precision=precision,
                                                 options=dict(pivoting=1))
    assert abs(T1 - T0) <= 1e-10 * abs(T0)
    scale = max(np.abs(g0r).max(), np.abs(g0i).max())
    assert np.abs(g1r - g0r).max() / scale < tol
    assert np.abs(g1i - g0i).max() / scale < tol


def test_pivoted_batched_solves_match_single_solves():
    # pivoting = 1 keeps the problems out of the lockstep engine (its kernels
    # pivot inside each diagonal block only): worker contexts, pivoted inverses
    items = [(_slab(11 + k, n=n, L=L), ["L", "R"], ["L", "R"]) for k, (n, L) in enumerate([(40, 9), (96, 12), (130, 7)])]
    Sb, infos = rgf_amd.cal_smatrix_RGF_batched(items, precision="fp32_refined", options=dict(pivoting=1),
                                                return_info=True)
    for (syst, out, inn), S, info in zip(items, Sb, infos):
        S1, _ = rgf_amd.cal_smatrix_RGF(syst, out, inn, precision="fp32_refined")
        assert info["pivoted"] == 1
        assert np.abs(S - S1).max() < 1e-9
