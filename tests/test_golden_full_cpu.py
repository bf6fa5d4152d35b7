"""The full-size fixture holds the workloads it names: its rebuilt inputs are
the seeded generators' first slices, and its S passes the invariants the
physics fixes (passivity with absorption in 2D, unitarity in 3D)."""
import numpy as np

from rgf_amd.workloads import slab_2d_large, spheres_3d
from golden_full import load_case


def test_2d_n1024_inputs_are_the_workload():
    syst, S = load_case("2d_n1024")
    ref, _, _ = slab_2d_large(nx=32)
    assert np.array_equal(syst["epsilon"], ref["epsilon"])
    for k in ("epsilon_L", "epsilon_R", "k0dx", "yBC"):
        assert syst[k] == ref[k]
    # absorbing slab: S^H S <= I (all singular values at most 1)
    assert np.linalg.norm(S, 2) <= 1 + 1e-12


def test_3d_n4096_inputs_are_the_workload():
    syst, S = load_case("3d_n4096")
    ref, _, _ = spheres_3d()
    assert np.array_equal(syst["epsilon"], ref["epsilon"][:, :, :6])
    assert np.allclose(syst["xyBC"], ref["xyBC"], rtol=0, atol=0)
    assert syst["k0dx"] == ref["k0dx"]
    # real eps: unitary to rounding
    assert np.abs(S.conj().T @ S - np.eye(S.shape[1])).max() < 1e-10
