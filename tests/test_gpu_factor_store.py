"""The mixed path with fewer device slices than slice inverses
(rgf_options.factor_slots, and free memory): the first inverses stay
resident, the rest are rebuilt in segments by every refinement sweep.  The
rebuilt inverses are the first pass's numbers, so S and the gradient are
bitwise those of the all-resident solve; a 3D N=4096 slab of 2560 slices,
whose factor store exceeds the GPU's memory, is solved against the FP64
path."""
import time

import numpy as np
import pytest

import oracle.rgf_oracle as o
import rgf_amd
from rgf_amd.workloads import spheres_3d

pytestmark = pytest.mark.gpu


def _slab(ny=160, nx=40, seed=12):
    rng = np.random.default_rng(seed)
    eps = 1 + 2 * rng.random((ny, nx)) + 0.01j * rng.random((ny, nx))
    return dict(epsilon=eps, epsilon_L=1.0, epsilon_R=1.7, k0dx=2 * np.pi / 11, yBC=0.4)


@pytest.mark.parametrize("slots", [30, 17, 13])
def test_rebuilt_factors_give_the_resident_S(slots):
    syst = _slab()
    S_all, _, i_all = rgf_amd.cal_smatrix_RGF(syst, ["L", "R"], ["L", "R"], precision="fp32_refined",
                                              return_info=True)
    S_cut, _, i_cut = rgf_amd.cal_smatrix_RGF(syst, ["L", "R"], ["L", "R"], precision="fp32_refined",
                                              return_info=True, options={"factor_slots": slots})
    print(f"slots {slots}: resident {i_cut['factor_slices_resident']}, rebuilt {i_cut['factor_slices_rebuilt']}, "
          f"iters {i_cut['refine_iters']}", flush=True)
    assert i_all["factor_slices_resident"] == 42 and i_all["factor_slices_rebuilt"] == 0
    assert 1 <= i_cut["factor_slices_resident"] < 42 and i_cut["factor_slices_rebuilt"] > 0
    assert i_cut["refine_iters"] == i_all["refine_iters"]
    assert np.array_equal(S_cut, S_all)
    So = o.cal_smatrix_RGF(syst, ["L", "R"], ["L", "R"])[0]
    assert np.abs(S_cut - So).max() < 1e-9


def test_rebuilt_factors_in_the_mixed_gradient():
    syst = _slab(ny=96, nx=24)
    T1, gr1, gi1 = rgf_amd.transmission_gradient(syst, precision="fp32_refined")
    T2, gr2, gi2 = rgf_amd.transmission_gradient(syst, precision="fp32_refined", options={"factor_slots": 11})
    assert T1 == T2 and np.array_equal(gr1, gr2) and np.array_equal(gi1, gi2)


def test_too_few_slots_is_refused():
    with pytest.raises(MemoryError, match="factor store"):
        rgf_amd.cal_smatrix_RGF(_slab(), ["L"], ["R"], precision="fp32_refined", options={"factor_slots": 3})


def test_long_3d_slab_beyond_device_memory():
    """64 x 64 transverse (N = 4096), 2560 slices (the 128-slice spheres slab
    repeated 20 times): 2562 FP32 slice inverses are 344 GB, more than the
    GPU holds.  Inputs on the left, every output."""
    syst, _, _ = spheres_3d()
    syst = dict(syst, epsilon=np.tile(syst["epsilon"], (1, 1, 20)))
    out, inn = ["L", "R"], ["L"]
    t0 = time.perf_counter()
    S32, _, i32 = rgf_amd.cal_smatrix_RGF_3D(syst, out, inn, precision="fp32_refined", return_info=True)
    t32 = time.perf_counter() - t0
    t0 = time.perf_counter()
    S64, _ = rgf_amd.cal_smatrix_RGF_3D(syst, out, inn)
    t64 = time.perf_counter() - t0
    err = float(np.abs(S32 - S64).max())
    print(f"mixed {t32:.1f} s (resident {i32['factor_slices_resident']}, rebuilt {i32['factor_slices_rebuilt']}, "
          f"iters {i32['refine_iters']}), fp64 {t64:.1f} s, max |dS| {err:.2e}", flush=True)
    assert i32["factor_slices_resident"] < 2562 and i32["factor_slices_rebuilt"] > 0
    assert err < 1e-9
