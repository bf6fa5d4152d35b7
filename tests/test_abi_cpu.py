"""CPU checks of librgf: exports, and the host-side (FP64) parts of the C ABI.

No GPU is needed: the channel setup, Laplacian, selection validation and
sizing run on the host inside librgf and are compared with the oracle.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle.rgf_oracle as o
import rgf_amd
from rgf_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    lib = L.load_library()
    header = open(os.path.join(ROOT, "include", "rgf.h")).read()
    declared = set(re.findall(r"\b(rgf_[a-z0-9_]+)\s*\(", header))
    assert declared, "no declarations parsed"
    missing = [s for s in sorted(declared) if not hasattr(lib, s)]
    assert not missing, missing
    assert declared == set(L.SIGNATURES), declared ^ set(L.SIGNATURES)
    assert lib.rgf_abi_version() == 5


def test_status_strings_and_default_options():
    lib = L.load_library()
    assert lib.rgf_status_string(0) == b"ok"
    o_ = L.Options()
    lib.rgf_default_options(C.byref(o_))
    assert o_.precision == 0 and o_.refine_tol == 1e-13 and o_.max_refine_iter == 6
    assert o_.refine_s_tol == 0 and o_.complex_3m == 3 and o_.autotune == 1


def test_struct_layouts_match_the_header(tmp_path):
    """The ctypes mirrors of rgf_options / rgf_info / rgf_problem /
    rgf_selection / rgf_channels have the C header's size and field offsets
    (a C program built with gcc prints them)."""
    import shutil
    import subprocess
    if not shutil.which("gcc"):
        pytest.skip("gcc not available")
    structs = {"rgf_options": L.Options, "rgf_info": L.Info, "rgf_problem": L.Problem,
               "rgf_selection": L.Selection, "rgf_channels": L.Channels}
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "rgf.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'    printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'    printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["    return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = dict(ln.split() for ln in subprocess.run([str(exe)], check=True, capture_output=True,
                                                   text=True).stdout.splitlines())
    for cname, py in structs.items():
        assert int(got[cname]) == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert int(got[f"{cname}.{f}"]) == getattr(py, f).offset, (cname, f)


BCS = ["periodic", "Dirichlet", "Neumann", "DirichletNeumann", "NeumannDirichlet", "PEC", "pmc", 0.9, -2.5]


def _cmp_side(a, b, conj=True):
    assert a["N_prop"] == b["N_prop"]
    assert np.array_equal(a["ind_prop"], b["ind_prop"])
    assert np.allclose(a["kxdx"], b["kxdx"], atol=1e-13, rtol=0)
    assert np.allclose(a["sqrt_mu"], b["sqrt_mu"], atol=1e-13, rtol=0)
    if conj and "ind_prop_conj" in b:
        assert np.array_equal(a["ind_prop_conj"], b["ind_prop_conj"])
    elif conj:
        assert "ind_prop_conj" not in a


@pytest.mark.parametrize("bc", BCS)
@pytest.mark.parametrize("ny", [1, 2, 7, 16, 33])
@pytest.mark.parametrize("k", [(0.6, 1.0, 2.25), (1.3, 1.0 + 0.2j, 1.5 - 0.1j), (2.2, 1.2, 1.2), (0.4 + 0.01j, 1.0, 3.0)])
def test_channels_2d_match_oracle(bc, ny, k):
    k0dx, eL, eR = k
    a = rgf_amd.setup_channels(ny, bc, k0dx, eL, eR)
    b = o.setup_channels(ny, bc, k0dx, eL, eR)
    assert np.allclose(a["kydx"], b["kydx"], atol=1e-14)
    _cmp_side(a["L"], b["L"])
    _cmp_side(a["R"], b["R"])
    assert np.allclose(a["fun_phi"](a["kydx"]), b["phi"], atol=1e-13)


@pytest.mark.parametrize("bc", ["periodic", "Dirichlet", [0.3, 2.2], (np.pi, -1.0)])
@pytest.mark.parametrize("dims", [(1, 1), (3, 4), (8, 8), (5, 9)])
def test_channels_3d_match_oracle(bc, dims):
    nx, ny = dims
    a = rgf_amd.setup_channels_3D(nx, ny, bc, 1.1, 1.0, 2.0)
    b = o.setup_channels_3D(nx, ny, bc, 1.1, 1.0, 2.0)
    assert np.allclose(a["kxdx"], b["kxdx"]) and np.allclose(a["kydx"], b["kydx"])
    for s in ("L", "R"):
        assert a[s]["N_prop"] == b[s]["N_prop"]
        assert np.array_equal(a[s]["ind_prop"], b[s]["ind_prop"])
        assert np.allclose(a[s]["kzdx"], b[s]["kzdx"], atol=1e-13)
        assert np.allclose(a[s]["sqrt_mu"], b[s]["sqrt_mu"], atol=1e-13)
    col = 2 % (nx * ny)
    ix, iy = col % nx, col // nx
    assert np.allclose(a["fun_phi"](a["kxdx"][ix, iy], a["kydx"][ix, iy]).ravel(order="F"), b["phi"][:, col])


@pytest.mark.parametrize("bc", BCS)
@pytest.mark.parametrize("n", [0, 1, 2, 3, 10])
def test_laplacian_matches_oracle(bc, n):
    A = rgf_amd.build_laplacian_1d(n, bc).toarray()
    assert np.allclose(A, o.build_laplacian_1d(n, bc), atol=1e-15)


def test_validation_messages():
    with pytest.raises(ValueError, match="Bloch periodic"):
        rgf_amd.setup_channels(4, "Bloch", 1.0, 1.0)
    with pytest.raises(ValueError, match="not an eligible option"):
        rgf_amd.setup_channels(4, "foo", 1.0, 1.0)
    with pytest.raises(ValueError, match="complex number"):
        rgf_amd.setup_channels(4, 1 + 1j, 1.0, 1.0)
    with pytest.raises(ValueError, match="not an eligible option"):
        rgf_amd.setup_channels_3D(3, 3, "Neumann", 1.0, 1.0)
    syst = dict(epsilon=np.ones((6, 3)), epsilon_L=1.0, epsilon_R=1.0, k0dx=1.0, yBC="periodic")
    with pytest.raises(ValueError, match="N_prop_L"):
        rgf_amd.cal_smatrix_RGF(syst, ["L"], {"ind_in_L": [99]})
    with pytest.raises(ValueError, match="N_prop_R"):
        rgf_amd.cal_smatrix_RGF(syst, {"ind_out_R": [0]}, ["L"])
    with pytest.raises(ValueError, match="structure or a cell array"):
        rgf_amd.cal_smatrix_RGF(syst, "L", ["L"])
    with pytest.raises(ValueError, match="syst.k0dx must be provided"):
        rgf_amd.cal_smatrix_RGF({k: v for k, v in syst.items() if k != "k0dx"}, ["L"], ["L"])
    with pytest.raises(ValueError, match="transverse"):
        rgf_amd.cal_smatrix_RGF(dict(syst, epsilon=np.zeros((0, 3))), ["L"], ["L"])


def test_empty_selection_needs_no_device():
    syst = dict(epsilon=np.ones((6, 3)), epsilon_L=1.0, epsilon_R=1.0, k0dx=1.0, yBC="periodic")
    S, ch = rgf_amd.cal_smatrix_RGF(syst, {"ind_out_L": []}, ["L"])
    assert S.shape == (0, ch["L"]["N_prop"]) and not np.iscomplexobj(S)
    with pytest.warns(UserWarning):
        S, _ = rgf_amd.cal_smatrix_RGF(syst, ["L", "bogus"], [])
    assert S.shape == (ch["L"]["N_prop"], 0)


def test_query_sizes_match_oracle():
    lib = L.load_library()
    from rgf_amd.api import _problem_2d, _Sel
    syst = dict(epsilon=np.ones((40, 5)), epsilon_L=1.0, epsilon_R=2.25, k0dx=2 * np.pi / 10, yBC="Dirichlet")
    prob = _problem_2d(syst)
    ch = o.setup_channels(40, "Dirichlet", 2 * np.pi / 10, 1.0, 2.25)
    sel = _Sel([L.RGF_ALL, np.array([1, 2], np.int32)], [L.RGF_ALL, L.RGF_ALL])
    nL, nR, no, ni = (C.c_int32() for _ in range(4))
    assert lib.rgf_query(C.byref(prob.p), C.byref(sel.s), C.byref(nL), C.byref(nR), C.byref(no), C.byref(ni)) == 0
    assert (nL.value, nR.value) == (ch["L"]["N_prop"], ch["R"]["N_prop"])
    assert no.value == ch["L"]["N_prop"] + 2 and ni.value == ch["L"]["N_prop"] + ch["R"]["N_prop"]


def test_handle_creation_without_gpu_fails_cleanly():
    from conftest import gpu_available
    if gpu_available():
        pytest.skip("a GPU is present")
    lib = L.load_library()
    h = C.c_void_p()
    st = lib.rgf_create(C.byref(h), 0)
    assert st == L.RGF_ERR_DEVICE and not h.value
    assert b"device" in lib.rgf_last_error(None).lower()
