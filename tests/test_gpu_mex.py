"""The MEX gateway end to end on the GPU: S and dT/d eps through
cal_smatrix_RGF_mex (MATLAB's column-major, interleaved-complex arrays)
equal what the Python mirror returns for the same problem, and match the
oracle."""
import os

import numpy as np
import pytest

import oracle.rgf_oracle as o
import rgf_amd
from mexgw import LIB, Gateway

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(LIB), reason="MEX harness not built (python -m rgf_amd.build)")]


@pytest.fixture(scope="module")
def gw():
    g = Gateway()
    yield g
    g.lib.h_exit()   # mexAtExit: releases the gateway's handle


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32_refined", 1e-9)])
def test_mex_2d_matches_python_and_oracle(gw, precision, tol):
    rng = np.random.default_rng(5)
    eps = 1 + 2 * rng.random((29, 17)) + 0.02j * rng.random((29, 17))
    syst = dict(epsilon=eps, epsilon_L=1.0, epsilon_R=1.69, k0dx=2 * np.pi / 6.5, yBC="DirichletNeumann")
    for out, inn in ((["L", "R"], ["L", "R"]), ({"ind_out_R": [2, 1, 2]}, {"ind_in_L": [3, 1]})):
        S, info = gw.call(2, syst, out, inn, {"precision": precision})
        Sp, _ = rgf_amd.cal_smatrix_RGF(syst, out, inn, precision=precision)
        assert S.shape == Sp.shape
        assert np.abs(S - Sp).max() <= 1e-13
        Sref, _ = o.cal_smatrix_RGF(syst, out, inn)
        assert np.abs(S - Sref).max() < tol
        assert info["n_prop_L"] > 0 and info["seconds_total"] > 0


def test_mex_3d_matches_python(gw):
    rng = np.random.default_rng(6)
    eps = 1 + 1.1 * rng.random((6, 5, 7))
    syst = dict(epsilon=eps, epsilon_L=1.0, epsilon_R=1.3, k0dx=1.2, xyBC=[0.7, 2.9])
    S, _ = gw.call(2, syst, ["L", "R"], ["L"])
    Sp, _ = rgf_amd.cal_smatrix_RGF_3D(syst, ["L", "R"], ["L"])
    assert np.abs(S - Sp).max() <= 1e-13
    Sref, _ = o.cal_smatrix_RGF_3D(syst, ["L", "R"], ["L"])
    assert np.abs(S - Sref).max() < 1e-10


def test_mex_gradient_matches_python(gw):
    rng = np.random.default_rng(7)
    eps = 1 + 2 * rng.random((21, 9)) + 0.01j * rng.random((21, 9))
    syst = dict(epsilon=eps, epsilon_L=1.0, epsilon_R=1.0, k0dx=2 * np.pi / 7, yBC="periodic")
    T, g = gw.call(2, "gradient", syst, ["R"], ["L"], {"checkpoint_every": 4})
    Tp, gre, gim = rgf_amd.transmission_gradient(syst, ("R",), ("L",), checkpoint_every=4)
    assert abs(float(T.ravel()[0]) - Tp) <= 1e-12 * max(1.0, abs(Tp))
    assert g.shape == eps.shape and np.iscomplexobj(g)
    assert np.abs(g - (gre + 1j * gim)).max() <= 1e-12 * max(1.0, np.abs(gre).max())
