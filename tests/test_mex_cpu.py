"""The MEX gateway (rgf_amd/mex/cal_smatrix_RGF_mex.c) on the host: argument
handling of cal_smatrix_RGF.m:79-184 (errors, ignored cell elements, the
empty S of an empty selection), all reached before any device is used."""
import os

import numpy as np
import pytest

from mexgw import LIB, Gateway, MexError

pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="MEX harness not built (python -m rgf_amd.build)")


@pytest.fixture(scope="module")
def gw():
    return Gateway()


def syst2d(**kw):
    rng = np.random.default_rng(3)
    s = dict(epsilon=1 + rng.random((12, 5)), epsilon_L=1.0, epsilon_R=1.0, k0dx=1.0, yBC="periodic")
    s.update(kw)
    return s


def err(gw, *args):
    with pytest.raises(MexError) as e:
        gw.call(2, *args)
    return e.value


def test_usage_and_struct_checks(gw):
    assert "usage" in err(gw, syst2d(), ["L"]).msg
    assert err(gw, 1.0, ["L"], ["R"]).msg == "syst must be a structure"
    s = syst2d()
    del s["k0dx"]
    assert err(gw, s, ["L"], ["R"]).msg == "syst.k0dx must be provided"
    s = syst2d()
    del s["yBC"]
    assert err(gw, s, ["L"], ["R"]).msg == "syst.yBC must be provided"
    assert err(gw, syst2d(epsilon_L="a"), ["L"], ["R"]).msg == "syst.epsilon_L must be a numeric scalar"
    assert err(gw, "grad", syst2d(), ["R"], ["L"]).msg == "the only mode argument is 'gradient'"
    e = err(gw, syst2d(), 3.0, ["R"])
    assert e.ident == "RGF:input" and e.msg == "input argument 'out' must be a structure or a cell array"


def test_boundary_conditions(gw):
    assert err(gw, syst2d(yBC="foo"), ["L"], ["R"]).msg == "Boundary condition BC = foo is not an eligible option"
    assert "set BC to k*a" in err(gw, syst2d(yBC="Bloch"), ["L"], ["R"]).msg
    assert err(gw, syst2d(yBC=1 + 1j), ["L"], ["R"]).msg == "Bloch phase k*a has to be real"
    e3 = dict(epsilon=np.ones((3, 4, 2)), epsilon_L=1.0, epsilon_R=1.0, k0dx=1.0, xyBC="Neumann")
    assert err(gw, e3, ["L"], ["R"]).msg == "Boundary condition BC = Neumann is not an eligible option"
    e3["xyBC"] = [0.1, 0.2, 0.3]
    assert "2 elements" in err(gw, e3, ["L"], ["R"]).msg


def test_selection_checked_by_the_library(gw):
    # N_prop of this lead: k0dx = 1 on 12 periodic sites
    e = err(gw, syst2d(), {"ind_out_L": [1, 99]}, ["L"])
    assert e.ident == "RGF:error" and "N_prop_L=" in e.msg
    e = err(gw, syst2d(), {"ind_out_R": [1.5]}, ["L"])            # non-integer: rejected like out of range
    assert "positive integers" in e.msg
    e = err(gw, syst2d(), {"ind_out_R": [3e10]}, ["L"])           # beyond int32
    assert "positive integers" in e.msg


def test_empty_selection_gives_real_empty_S(gw):
    """cal_smatrix_RGF.m:181-184: an empty selection returns zeros(n_out, n_in)
    (real) without a solve; the ignored cell element is warned about."""
    S, info = gw.call(2, syst2d(), ["X"], ["L", "R"])
    assert "out{1} = X ignored" in gw.warnings
    assert S.shape[0] == 0 and S.shape[1] == info["n_prop_L"] + info["n_prop_R"] > 0
    assert not np.iscomplexobj(S)
    S, _ = gw.call(2, syst2d(), [], ["L"])
    assert S.shape == (0, info["n_prop_L"])
