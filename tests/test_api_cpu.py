"""Host-side marshalling of the Python mirror: epsilon given as a numpy
array or as a torch tensor reaches the library as the same column-major
FP64 planes (cal_smatrix_RGF.m / _3D.m take MATLAB column-major arrays)."""
import ctypes as C

import numpy as np
import pytest
import torch

from rgf_amd import api


def planes(prob, n):
    re = np.ctypeslib.as_array(C.cast(prob.p.eps_re, C.POINTER(C.c_double)), shape=(n,)).copy()
    im = (np.ctypeslib.as_array(C.cast(prob.p.eps_im, C.POINTER(C.c_double)), shape=(n,)).copy()
          if prob.p.eps_im else None)
    return re, im


@pytest.mark.parametrize("three_d", [False, True])
def test_tensor_and_array_give_the_same_planes(three_d):
    rng = np.random.default_rng(4)
    shape = (5, 4, 3) if three_d else (6, 7)
    eps = 1 + rng.random(shape) + 0.1j * rng.random(shape)
    key = "xyBC" if three_d else "yBC"
    syst = dict(epsilon=eps, epsilon_L=1.0, epsilon_R=2.0, k0dx=0.9, **{key: "Dirichlet"})
    build = api._problem_3d if three_d else api._problem_2d
    pa = build(syst)
    view = torch.as_tensor(eps)
    if three_d:
        view = view.permute(2, 1, 0).contiguous().permute(2, 1, 0)   # same values, another memory order
    pt = build(dict(syst, epsilon=view))
    assert list(pa.p.dims) == list(pt.p.dims) and pt.p.eps_on_device == 0 and pt.device is None
    ra, ia = planes(pa, eps.size)
    rt, it = planes(pt, eps.size)
    assert np.array_equal(ra, eps.real.ravel(order="F")) and np.array_equal(ia, eps.imag.ravel(order="F"))
    assert np.array_equal(ra, rt) and np.array_equal(ia, it)


def test_tensor_dtypes():
    p = api._problem_3d(dict(epsilon=torch.ones((3, 4, 2), dtype=torch.int32), epsilon_L=1.0, epsilon_R=1.0,
                             k0dx=1.0, xyBC="periodic"))
    assert list(p.p.dims) == [3, 4, 2] and not p.p.eps_im
    with pytest.raises(ValueError, match="numeric"):
        api._problem_2d(dict(epsilon=torch.ones((3, 4), dtype=torch.bool), epsilon_L=1.0, epsilon_R=1.0, k0dx=1.0,
                             yBC="periodic"))
