"""GPU error paths of the C ABI: refinement that does not converge, a factor
store that does not fit in device memory, and errors inside a batched call."""
import numpy as np
import pytest

import rgf_amd
from rgf_amd._lib import RGF_ERR_NOT_CONVERGED, RGFError

pytestmark = pytest.mark.gpu


def _slab(seed=0, ny=32, nx=20):
    rng = np.random.default_rng(seed)
    eps = 1 + 2 * rng.random((ny, nx))
    return dict(epsilon=eps, epsilon_L=1.0, epsilon_R=1.0, k0dx=0.9, yBC="periodic")


def test_refinement_not_converged_is_reported():
    opts = {"max_refine_iter": 1, "refine_tol": 1e-16}
    with pytest.raises(RGFError) as e:
        rgf_amd.cal_smatrix_RGF(_slab(), ["L", "R"], ["L", "R"], precision="fp32_refined", options=opts)
    assert e.value.status == RGF_ERR_NOT_CONVERGED
    assert "refinement stopped" in str(e.value)
    # the handle stays usable
    S, _ = rgf_amd.cal_smatrix_RGF(_slab(), ["L", "R"], ["L", "R"], precision="fp32_refined")
    assert np.abs(S.conj().T @ S - np.eye(S.shape[1])).max() < 1e-10


def test_factor_store_larger_than_device_memory():
    # 3D 64 x 64, 20000 slices: the FP32 factor store alone is 2.7 TB
    eps = np.ones((64, 64, 20000), dtype=np.float64)
    syst = dict(epsilon=eps, epsilon_L=1.0, epsilon_R=1.0, k0dx=2 * np.pi / 12, xyBC="periodic")
    with pytest.raises(MemoryError, match="factor store"):
        rgf_amd.cal_smatrix_RGF_3D(syst, ["R"], ["L"], precision="fp32_refined")


def test_batched_error_names_the_problem():
    items = [(_slab(k), ["L", "R"], ["L", "R"]) for k in range(3)]
    with pytest.raises(RGFError) as e:
        rgf_amd.cal_smatrix_RGF_batched(items, precision="fp32_refined",
                                        options={"max_refine_iter": 1, "refine_tol": 1e-16})
    assert e.value.status == RGF_ERR_NOT_CONVERGED
    assert "batched solve, problem 0" in str(e.value)
    # a good batch afterwards on the same handle
    Sb = rgf_amd.cal_smatrix_RGF_batched(items)
    for (syst, out, inn), S in zip(items, Sb):
        assert np.abs(S - rgf_amd.cal_smatrix_RGF(syst, out, inn)[0]).max() <= 1e-12
