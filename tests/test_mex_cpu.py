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
    assert err(gw, "grad", syst2d(), ["R"], ["L"]).msg == "the mode argument must be 'gradient', 'channels' or 'phi'"
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


# ---- the `channels` output: 'channels' and 'phi' modes (no GPU) ----

import oracle.rgf_oracle as o  # noqa: E402


def _same(a, b, tol=1e-13):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and (a.size == 0 or np.abs(a - b).max() <= tol * max(1.0, np.abs(b).max()))


def _check_side(got, want, kl, three):
    n = want["N_prop"]
    assert got["N_prop"] == n
    assert _same(got["ind_prop"], np.asarray(want["ind_prop"], float).reshape(1, n))
    assert _same(got["sqrt_mu"], np.asarray(want["sqrt_mu"]).reshape(1, n))
    assert _same(got[kl], want[kl] if three else np.asarray(want["kxdx"]).reshape(1, -1))
    # MATLAB stores an array with no imaginary part as real
    assert np.iscomplexobj(got[kl]) == bool(np.any(np.imag(want[kl]) != 0))
    if three:
        assert list(got) == ["kzdx", "ind_prop", "N_prop", "sqrt_mu"]
    elif "ind_prop_conj" in want:
        assert list(got) == ["kxdx", "ind_prop", "N_prop", "sqrt_mu", "ind_prop_conj"]
        assert _same(got["ind_prop_conj"], np.asarray(want["ind_prop_conj"], float).reshape(1, n))
    else:
        assert "ind_prop_conj" not in got   # Bloch ka != 0 (setup_channels.m:226-235)


@pytest.mark.parametrize("bc", ["periodic", "Dirichlet", "PMC", "DirichletNeumann", "pmcpec", 0.7, -2.1])
@pytest.mark.parametrize("k0dx,epsR", [(1.1, 1.0), (0.9 + 0.02j, 2.25 + 0.1j), (2.6, 1.7)])
def test_channels_mode_2d_vs_oracle(gw, bc, k0dx, epsR):
    """cal_smatrix_RGF_mex('channels', syst) field by field against the
    oracle's setup_channels (setup_channels.m:30-82, 160-176)."""
    ny = 15
    s = syst2d(epsilon=np.ones((ny, 3)), k0dx=k0dx, epsilon_L=1.3, epsilon_R=epsR, yBC=bc)
    (ch,) = gw.call(1, "channels", s)
    want = o.setup_channels(ny, bc, k0dx, 1.3, epsR)
    assert list(ch) == ["kydx", "fun_phi", "L", "R"]
    assert ch["fun_phi"].size == 0        # filled by the .m wrapper
    assert _same(ch["kydx"], want["kydx"].reshape(1, ny))
    for side in ("L", "R"):
        _check_side(ch[side], want[side], "kxdx", False)


@pytest.mark.parametrize("bc", ["periodic", "Dirichlet", [0.4, -1.3]])
@pytest.mark.parametrize("k0dx,epsR", [(1.0, 1.0), (1.2 + 0.01j, 2.0)])
def test_channels_mode_3d_vs_oracle(gw, bc, k0dx, epsR):
    nx, ny = 5, 4
    s = dict(epsilon=np.ones((nx, ny, 2)), epsilon_L=1.0, epsilon_R=epsR, k0dx=k0dx, xyBC=bc)
    (ch,) = gw.call(1, "channels", s)
    want = o.setup_channels_3D(nx, ny, bc, k0dx, 1.0, epsR)
    assert list(ch) == ["kxdx", "kydx", "fun_phi", "L", "R"]
    assert _same(ch["kxdx"], want["kxdx"]) and _same(ch["kydx"], want["kydx"])
    for side in ("L", "R"):
        w = dict(want[side])
        w["kzdx"] = w["kzdx"]
        _check_side(ch[side], w, "kzdx", True)


@pytest.mark.parametrize("bc", ["periodic", "Dirichlet", "Neumann", "PECPMC", "NeumannDirichlet", 1.3])
def test_phi_mode_2d(gw, bc):
    """fun_phi(kydx) over every channel: the oracle's mode matrix, unitary
    (setup_channels.m:36-41), and Neumann's kydx = 0 column constant."""
    ny = 11
    lead = dict(epsilon=np.zeros((ny, 0)), epsilon_L=1.0, epsilon_R=1.0, k0dx=1.0, yBC=bc)
    want = o.setup_channels(ny, bc, 1.0, 1.0)
    (phi,) = gw.call(1, "phi", lead, want["kydx"])
    assert phi.shape == (ny, ny)
    assert np.iscomplexobj(phi) == (not isinstance(bc, str) or bc == "periodic")
    assert _same(phi, want["phi"])
    assert np.abs(phi.conj().T @ phi - np.eye(ny)).max() < 1e-13
    if bc == "Neumann":
        assert np.allclose(phi[:, 0], 1 / np.sqrt(ny))
    # one wavenumber off the grid, and none
    (one,) = gw.call(1, "phi", lead, 0.37)
    assert _same(one, o.mode_matrix_2d(ny, o.parse_bc(bc)[0], np.array([0.37])))
    (none,) = gw.call(1, "phi", lead, np.zeros((1, 0)))
    assert none.shape == (ny, 0)


@pytest.mark.parametrize("bc", ["periodic", "Dirichlet", [0.4, -1.3]])
def test_phi_mode_3d(gw, bc):
    """fun_phi(kxdx(a), kydx(a)) is the nx x ny profile of channel a: column a
    of the oracle's mode matrix (setup_channels_3D.m:118,123)."""
    nx, ny = 4, 3
    lead = dict(epsilon=np.zeros((nx, ny, 0)), epsilon_L=1.0, epsilon_R=1.0, k0dx=1.0, xyBC=bc)
    want = o.setup_channels_3D(nx, ny, bc, 1.0, 1.0)
    kx, ky = want["kxdx"].ravel(order="F"), want["kydx"].ravel(order="F")
    for a in (0, 5, nx * ny - 1):
        (phi,) = gw.call(1, "phi", lead, kx[a], ky[a])
        assert phi.shape == (nx, ny)
        assert _same(phi.ravel(order="F"), want["phi"][:, a])
    (allp,) = gw.call(1, "phi", lead, kx, ky)
    assert allp.shape == (nx, ny, nx * ny)
    assert _same(allp.reshape(nx * ny, -1, order="F"), want["phi"])
    assert "same number" in err(gw, "phi", lead, kx, ky[:3]).msg
    assert "usage" in err(gw, "phi", lead, kx).msg
