"""GPU parity at the benchmark workloads' full transverse sizes against the
FP64 oracle's S (tests/golden/rgf_golden_full.npz, written by
tests/golden/make_golden.py --full):

  * 2D N = 1024 (the large-slab workload's first 32 slices: Dirichlet,
    eps_L = 1, eps_R = 2.25, absorbing pixels), full 528 x 528 S;
  * 3D N = 4096 (the 3D workload's first 6 slices: 64 x 64 transverse,
    spheres, random Bloch phases), full 182 x 182 S.

These pin the FP64 GPU path, which the benchmark's S error is measured
against, and the mixed path at the headline N.  Tolerances: 1e-10 (FP64)
and 1e-9 (FP32 + refinement) on the max-entry error; |S| entries are O(1).
"""

import numpy as np
import pytest

import rgf_amd
from golden_full import load_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32_refined", 1e-9)])
def test_2d_n1024_full_s_vs_oracle(precision, tol):
    syst, S_ref = load_case("2d_n1024")
    S, _ = rgf_amd.cal_smatrix_RGF(syst, ["L", "R"], ["L", "R"], precision=precision)
    assert S.shape == S_ref.shape == (528, 528)
    err = float(np.abs(S - S_ref).max())
    print(f"2d_n1024 {precision}: max|S - S_oracle| = {err:.2e}", flush=True)
    assert err < tol


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32_refined", 1e-9)])
def test_3d_n4096_full_s_vs_oracle(precision, tol):
    syst, S_ref = load_case("3d_n4096")
    S, _ = rgf_amd.cal_smatrix_RGF_3D(syst, ["L", "R"], ["L", "R"], precision=precision)
    assert S.shape == S_ref.shape
    err = float(np.abs(S - S_ref).max())
    print(f"3d_n4096 {precision}: max|S - S_oracle| = {err:.2e}", flush=True)
    assert err < tol
