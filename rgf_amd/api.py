"""The reference's MATLAB interface over librgf.

``cal_smatrix_RGF``, ``cal_smatrix_RGF_3D``, ``setup_channels``,
``setup_channels_3D`` and ``build_laplacian_1d`` take the same arguments and
return the same things as the MATLAB functions of chiaweihsu/RGF
(src/cal_smatrix_RGF.m:1, src/cal_smatrix_RGF_3D.m:1, src/setup_channels.m:1,
src/setup_channels_3D.m:1, src/build_laplacian_1d.m:1), with Python types:
``syst`` is a dict, a MATLAB cell is a list of strings, a MATLAB struct is a
dict, matrices are numpy arrays indexed like MATLAB (ny x nx in 2D,
nx x ny x nz in 3D) and channel indices stay 1-based.  ``in`` is a Python
keyword, so that argument is spelled ``in_``.

All numerical work runs in librgf on the GPU; this module only marshals.
"""
from __future__ import annotations

import ctypes as C
import warnings

import numpy as np

from . import _lib as L

_BC_NAMES = {
    "dirichlet": L.RGF_BC_DIRICHLET, "pec": L.RGF_BC_DIRICHLET,
    "neumann": L.RGF_BC_NEUMANN, "pmc": L.RGF_BC_NEUMANN,
    "dirichletneumann": L.RGF_BC_DIRICHLET_NEUMANN, "pecpmc": L.RGF_BC_DIRICHLET_NEUMANN,
    "neumanndirichlet": L.RGF_BC_NEUMANN_DIRICHLET, "pmcpec": L.RGF_BC_NEUMANN_DIRICHLET,
}
_PRECISIONS = {"fp64": L.RGF_PRECISION_FP64, "fp32_refined": L.RGF_PRECISION_FP32_REFINED,
               "mixed": L.RGF_PRECISION_FP32_REFINED}


def _is_num(x) -> bool:
    return isinstance(x, (int, float, complex, np.number)) and not isinstance(x, bool)


def _bc_1d(BC) -> tuple[int, float]:
    """setup_channels.m:102-115: numeric -> Bloch ka; 'periodic' -> Bloch 0."""
    if _is_num(BC) or (isinstance(BC, np.ndarray) and BC.size == 1):
        ka = complex(np.asarray(BC).ravel()[0]) if isinstance(BC, np.ndarray) else complex(BC)
        if ka.imag != 0:
            raise ValueError(f"k*a = {ka.real:g} + 1i*{ka.imag:g} is a complex number; has to be real for a "
                             "complete orthonormal transverse basis.")
        return L.RGF_BC_BLOCH, ka.real
    if not isinstance(BC, str):
        raise ValueError("Boundary condition BC has to be a character vector or string, or numeric scalar "
                         "(for Bloch periodic BC)")
    low = BC.lower()
    if low == "bloch":
        raise ValueError("To use Bloch periodic boundary condition, set BC to k*a where k is the Bloch wave "
                         "number and a is the periodicity")
    if low == "periodic":
        return L.RGF_BC_BLOCH, 0.0
    if low in _BC_NAMES:
        return _BC_NAMES[low], 0.0
    raise ValueError(f"Boundary condition BC = {BC} is not an eligible option")


def _bc_xy(BC) -> tuple[list[int], list[float]]:
    """setup_channels_3D.m:81-99: a string for both directions or [kxa, kya]."""
    if isinstance(BC, str):
        low = BC.lower()
        if low == "bloch":
            raise ValueError("To use Bloch periodic boundary condition, set BC to [kx*ax, ky*ay] where (kx,ky) is "
                             "the Bloch wave vector and (ax,ay) is the periodicity")
        if low == "periodic":
            return [L.RGF_BC_BLOCH] * 2, [0.0, 0.0]
        if low in ("dirichlet", "pec"):
            return [L.RGF_BC_DIRICHLET] * 2, [0.0, 0.0]
        raise ValueError(f"Boundary condition BC = {BC} is not an eligible option")
    v = np.asarray(BC).ravel()
    if v.size != 2 or not np.issubdtype(v.dtype, np.number):
        raise ValueError("Boundary condition BC has to be a character vector or string, or numeric vector with 2 "
                         "elements (for Bloch periodic BC)")
    for lab, x in (("kx", v[0]), ("ky", v[1])):
        if np.imag(x) != 0:
            raise ValueError(f"{lab}*a = {np.real(x):g} + 1i*{np.imag(x):g} is a complex number; has to be real "
                             "for a complete orthonormal transverse basis.")
    return [L.RGF_BC_BLOCH] * 2, [float(np.real(v[0])), float(np.real(v[1]))]


def _scalar(x, name) -> complex:
    if not (_is_num(x) or (isinstance(x, np.ndarray) and x.size == 1)):
        raise ValueError(f"{name} must be a numeric scalar")
    return complex(np.asarray(x).ravel()[0])


class _Prob:
    """Owns the numpy buffers behind an rgf_problem."""

    def __init__(self, ndim, dims, eps, epsL, epsR, k0dx, bcs, kas):
        self.keep = []
        self.device = None  # device index of device-resident eps
        p = L.Problem()
        p.ndim = ndim
        p.dims[:] = list(dims)
        if eps is not None and eps.size:
            # column-major flattening: slices (columns in 2D, z-planes in 3D) are contiguous
            re = np.ascontiguousarray(np.real(eps).ravel(order="F"), dtype=np.float64)
            self.keep.append(re)
            p.eps_re = re.ctypes.data
            if np.iscomplexobj(eps):
                im = np.ascontiguousarray(np.imag(eps).ravel(order="F"), dtype=np.float64)
                self.keep.append(im)
                p.eps_im = im.ctypes.data
        for fld, v in (("eps_L", epsL), ("eps_R", epsR), ("k0dx", k0dx)):
            getattr(p, fld)[:] = [v.real, v.imag]
        p.bc[:] = list(bcs) + [0] * (2 - len(bcs))
        p.bloch_ka[:] = list(kas) + [0.0] * (2 - len(kas))
        p.eps_on_device = 0
        self.p = p


def _is_tensor(x) -> bool:
    return type(x).__module__.split(".")[0] == "torch"


def _problem_from_tensor(eps, epsilon_L, epsilon_R, k0dx, bc, three_d: bool) -> _Prob:
    """rgf_problem over a torch tensor in MATLAB index order (ny x nx, or
    nx x ny x nz).  A CUDA tensor is handed over as device pointers
    (rgf_problem.eps_on_device = 1): the column-major FP64 planes are formed
    on the device and torch's current stream is synchronized before the
    library reads them from its own stream."""
    import torch
    if eps.dtype == torch.bool:
        raise ValueError("syst.epsilon must be a numeric array")
    if three_d:
        if eps.dim() == 2:
            eps = eps[:, :, None]
        if eps.dim() != 3:
            raise ValueError("syst.epsilon must be a 3D array")
        nx, ny, nz = eps.shape
        if nx * ny == 0:
            raise ValueError("need at least one site along the transverse directions (x,y)")
        dims = (nx, ny, nz)
        bcs, kas = _bc_xy(bc)
        cm = eps.detach().permute(2, 1, 0)            # column-major view: slices contiguous
    else:
        if eps.dim() == 1:
            eps = eps[:, None]
        if eps.dim() != 2:
            raise ValueError("syst.epsilon must be a numeric matrix")
        ny, nx = eps.shape
        if ny == 0:
            raise ValueError("need at least one site along the transverse (y) direction")
        dims = (ny, nx, 1)
        b, ka = _bc_1d(bc)
        bcs, kas = [b], [ka]
        cm = eps.detach().t()
    prob = _Prob(3 if three_d else 2, dims, None, _scalar(epsilon_L, "epsilon_L"), _scalar(epsilon_R, "epsilon_R"),
                 _scalar(k0dx, "k0dx"), bcs, kas)
    re = cm.real.to(torch.float64).contiguous() if cm.is_complex() else cm.to(torch.float64).contiguous()
    prob.keep.append(re)
    prob.p.eps_re = re.data_ptr()
    if cm.is_complex():
        im = cm.imag.to(torch.float64).contiguous()
        prob.keep.append(im)
        prob.p.eps_im = im.data_ptr()
    prob.p.eps_on_device = 1 if eps.is_cuda else 0
    if eps.is_cuda:
        prob.device = eps.device.index
        torch.cuda.current_stream(eps.device).synchronize()
    return prob


def _problem_2d(syst, need_eps=True) -> _Prob:
    if not isinstance(syst, dict):
        raise ValueError("syst must be a structure")
    for key in ("epsilon", "epsilon_L", "epsilon_R", "k0dx", "yBC"):
        if key not in syst:
            raise ValueError(f"syst.{key} must be provided")
    if _is_tensor(syst["epsilon"]):
        return _problem_from_tensor(syst["epsilon"], syst["epsilon_L"], syst["epsilon_R"], syst["k0dx"], syst["yBC"],
                                    False)
    eps = np.asarray(syst["epsilon"])
    if eps.ndim == 1:
        eps = eps[:, None]
    if eps.ndim != 2 or not np.issubdtype(eps.dtype, np.number):
        raise ValueError("syst.epsilon must be a numeric matrix")
    ny, nx = eps.shape
    if ny == 0:
        raise ValueError("need at least one site along the transverse (y) direction")
    bc, ka = _bc_1d(syst["yBC"])
    return _Prob(2, (ny, nx, 1), eps, _scalar(syst["epsilon_L"], "syst.epsilon_L"),
                 _scalar(syst["epsilon_R"], "syst.epsilon_R"), _scalar(syst["k0dx"], "syst.k0dx"), [bc], [ka])


def _problem_3d(syst) -> _Prob:
    if not isinstance(syst, dict):
        raise ValueError("syst must be a structure")
    for key in ("epsilon", "epsilon_L", "epsilon_R", "k0dx", "xyBC"):
        if key not in syst:
            raise ValueError(f"syst.{key} must be provided")
    if _is_tensor(syst["epsilon"]):
        return _problem_from_tensor(syst["epsilon"], syst["epsilon_L"], syst["epsilon_R"], syst["k0dx"],
                                    syst["xyBC"], True)
    eps = np.asarray(syst["epsilon"])
    if eps.ndim == 2:
        eps = eps[:, :, None]
    if eps.ndim != 3 or not np.issubdtype(eps.dtype, np.number):
        raise ValueError("syst.epsilon must be a 3D array")
    nx, ny, nz = eps.shape
    if nx * ny == 0:
        raise ValueError("need at least one site along the transverse directions (x,y)")
    bcs, kas = _bc_xy(syst["xyBC"])
    return _Prob(3, (nx, ny, nz), eps, _scalar(syst["epsilon_L"], "syst.epsilon_L"),
                 _scalar(syst["epsilon_R"], "syst.epsilon_R"), _scalar(syst["k0dx"], "syst.k0dx"), bcs, kas)


def _count(sides, nprop) -> int:
    """Channels a parsed selection asks for (rgf_query's n_out / n_in)."""
    return sum(nprop[k] if isinstance(v, int) else len(v) for k, v in enumerate(sides))


def _query(prob: _Prob) -> tuple[int, int]:
    lib = L.load_library()
    nL, nR = C.c_int32(), C.c_int32()
    L.check(lib.rgf_query(C.byref(prob.p), None, C.byref(nL), C.byref(nR), None, None))
    return nL.value, nR.value


def _parse_selection(spec, kind: str, nprop: tuple[int, int]):
    """cal_smatrix_RGF.m:103-175 -> (L, R) where each is RGF_ALL or an int32 array."""
    sides = [np.zeros(0, np.int32), np.zeros(0, np.int32)]
    if isinstance(spec, dict):
        for s, side in enumerate(("L", "R")):
            key = f"ind_{kind}_{side}"
            if key in spec:
                v = np.asarray(spec[key]).ravel()
                ok = v.size == 0 or (np.issubdtype(v.dtype, np.number) and np.all(np.isreal(v))
                                     and np.all(np.round(np.real(v)) == np.real(v))
                                     and np.real(v).min() > 0 and np.real(v).max() <= nprop[s])
                if not ok:
                    raise ValueError(f"{kind}.{key}, when specified, must be an array of positive integers not "
                                     f"exceeding N_prop_{side}={nprop[s]}")
                sides[s] = np.real(v).astype(np.int32)
    elif isinstance(spec, (list, tuple)):
        for k, el in enumerate(spec, start=1):
            if not isinstance(el, str):
                warnings.warn(f"element {kind}{{{k}}} has wrong format and is ignored; should be one of "
                              "{'left', 'right', 'L', 'R'}")
                continue
            if el in ("left", "L"):
                sides[0] = L.RGF_ALL
            elif el in ("right", "R"):
                sides[1] = L.RGF_ALL
            else:
                warnings.warn(f"element {kind}{{{k}}} = {el} ignored; should be one of {{'left', 'right', 'L', 'R'}}")
    else:
        raise ValueError(f"input argument '{kind}' must be a structure or a cell array")
    return sides


class _Sel:
    def __init__(self, out_sides, in_sides):
        self.keep = []
        s = L.Selection()
        for fld, v in (("in_L", in_sides[0]), ("in_R", in_sides[1]), ("out_L", out_sides[0]),
                       ("out_R", out_sides[1])):
            if isinstance(v, int):
                setattr(s, "ind_" + fld, None)
                setattr(s, "n_" + fld, v)
            else:
                a = np.ascontiguousarray(v, dtype=np.int32)
                self.keep.append(a)
                setattr(s, "ind_" + fld, L.iptr(a) if a.size else None)
                setattr(s, "n_" + fld, int(a.size))
        self.s = s


def _options(precision="fp64", **kw) -> L.Options:
    lib = L.load_library()
    o = L.Options()
    lib.rgf_default_options(C.byref(o))
    if precision not in _PRECISIONS:
        raise ValueError(f"precision must be one of {sorted(_PRECISIONS)}")
    o.precision = _PRECISIONS[precision]
    known = {f[0] for f in L.Options._fields_} - {"precision"}
    for k, v in kw.items():
        if k not in known:
            raise ValueError(f"unknown option {k!r}; options are {sorted(known)}")
        if v is not None:
            setattr(o, k, v)
    return o


def _solve(prob: _Prob, out, in_, precision, device, opts, fields: bool = False):
    """S (n_out x n_in) and the info dict; with ``fields``, also the field
    profiles as an N x (L+2) x n_in array (rgf_solve_fields)."""
    lib = L.load_library()
    nprop = _query(prob)
    sel = _Sel(_parse_selection(out, "out", nprop), _parse_selection(in_, "in", nprop))
    nout, nin = C.c_int32(), C.c_int32()
    L.check(lib.rgf_query(C.byref(prob.p), C.byref(sel.s), None, None, C.byref(nout), C.byref(nin)))
    N = int(np.prod(prob.p.dims[:prob.p.ndim - 1]))
    nslab = int(prob.p.dims[prob.p.ndim - 1]) + 2
    if nout.value == 0 or nin.value == 0:
        S0 = np.zeros((nout.value, nin.value))                # cal_smatrix_RGF.m:181-184 (real zeros)
        if fields:
            return S0, None, np.zeros((N, nslab, nin.value), dtype=np.complex128, order="F")
        return S0, None
    if prob.device is not None:
        if device is not None and device >= 0 and device != prob.device:
            raise ValueError(f"syst.epsilon is on device {prob.device}, the solve was asked for device {device}")
        device = prob.device
    S = np.zeros((nout.value, nin.value), dtype=np.complex128, order="F")
    info = L.Info()
    o = _options(precision, **(opts or {}))
    with L.using(device) as h:
        if fields:
            F = np.zeros((N, nslab, nin.value), dtype=np.complex128, order="F")
            st = lib.rgf_solve_fields(h, C.byref(prob.p), C.byref(sel.s), C.byref(o), S.ctypes.data_as(L._dp),
                                      F.ctypes.data_as(L._dp), C.byref(info))
            L.check(st, h)
            return S, info.as_dict(), F
        st = lib.rgf_solve(h, C.byref(prob.p), C.byref(sel.s), C.byref(o), S.ctypes.data_as(L._dp),
                           C.byref(info))
        L.check(st, h)
    return S, info.as_dict()


def cal_smatrix_RGF(syst: dict, out, in_, *, precision: str = "fp64", device: int = -1, options: dict | None = None,
                    return_info: bool = False, return_fields: bool = False):
    """[smatrix, channels] = cal_smatrix_RGF(syst, out, in) (cal_smatrix_RGF.m:1-291).

    Unlike the reference's validation at cal_smatrix_RGF.m:89, a numeric
    ``yBC`` (Bloch phase ky*a, as its help text :32-34 documents) is accepted.
    ``precision`` is 'fp64' (complex FP64 throughout) or 'fp32_refined'.
    ``return_fields`` (precision 'fp32_refined' only) appends the field
    profiles, an ``ny x (nx + 2) x n_in`` array: column ``n`` of slice
    ``1..nx`` lines up with ``epsilon[:, n - 1]``, slices 0 and nx+1 are the
    lead slices next to the region, inputs in S's column order.
    """
    prob = _problem_2d(syst)
    res = _solve(prob, out, in_, precision, device, options, fields=return_fields)
    S, info = res[0], res[1]
    ch = setup_channels(prob.p.dims[0], syst["yBC"], syst["k0dx"], syst["epsilon_L"], syst["epsilon_R"])
    ret = (S, ch, info) if return_info else (S, ch)
    return ret + (res[2],) if return_fields else ret


def cal_smatrix_RGF_3D(syst: dict, out, in_, *, precision: str = "fp64", device: int = -1,
                       options: dict | None = None, return_info: bool = False, return_fields: bool = False):
    """[smatrix, channels] = cal_smatrix_RGF_3D(syst, out, in) (cal_smatrix_RGF_3D.m:1-306).

    ``return_fields`` (precision 'fp32_refined' only) appends the field
    profiles, an ``nx x ny x (nz + 2) x n_in`` array aligned with
    ``epsilon`` in slices 1..nz (slices 0 and nz+1 are the lead slices).
    """
    prob = _problem_3d(syst)
    res = _solve(prob, out, in_, precision, device, options, fields=return_fields)
    S, info = res[0], res[1]
    nx, ny = prob.p.dims[0], prob.p.dims[1]
    ch = setup_channels_3D(nx, ny, syst["xyBC"], syst["k0dx"], syst["epsilon_L"], syst["epsilon_R"])
    ret = (S, ch, info) if return_info else (S, ch)
    if not return_fields:
        return ret
    F = res[2]
    return ret + (F.reshape((nx, ny, F.shape[1], F.shape[2]), order="F"),)


def cal_smatrix_RGF_batched(items, *, precision: str = "fp64", device: int = -1, options: dict | None = None,
                            return_info: bool = False):
    """Many independent solves in one call (rgf_solve_batched).

    ``items`` is a sequence of ``(syst, out, in_)``; a syst whose ``epsilon``
    has three dimensions is a 3D problem (cal_smatrix_RGF_3D), otherwise 2D
    (cal_smatrix_RGF).  Up to ``options['batch_concurrency']`` (default 8)
    solves run at once.  Returns the list of S matrices (and the infos).
    """
    lib = L.load_library()
    items = list(items)
    probs, sels, Ss, keep = [], [], [], []
    for syst, out, in_ in items:
        prob = _problem_3d(syst) if np.ndim(syst["epsilon"]) == 3 else _problem_2d(syst)
        nprop = _query(prob)
        so, si = _parse_selection(out, "out", nprop), _parse_selection(in_, "in", nprop)
        probs.append(prob)
        sels.append(_Sel(so, si))
        Ss.append(np.zeros((_count(so, nprop), _count(si, nprop)), dtype=np.complex128, order="F"))
    n = len(items)
    if n == 0:
        return ([], []) if return_info else []
    P = (L.Problem * n)(*[pr.p for pr in probs])
    Sel = (L.Selection * n)(*[se.s for se in sels])
    # an empty request still needs a valid pointer; its S stays as returned by rgf_query (real zeros)
    dummy = np.zeros(2)
    outs = (L._dp * n)(*[(S if S.size else dummy).ctypes.data_as(L._dp) for S in Ss])
    infos = (L.Info * n)()
    o = _options(precision, **(options or {}))
    with L.using(device) as h:
        L.check(lib.rgf_solve_batched(h, n, P, Sel, C.byref(o), outs, infos), h)
    res = [S if S.size else np.zeros(S.shape) for S in Ss]
    return (res, [i.as_dict() for i in infos]) if return_info else res


def _channels_raw(prob: _Prob, N: int):
    lib = L.load_library()
    c = L.Channels()
    arrs = {
        "kt1": np.zeros(N), "kt2": np.zeros(N), "kl_L": np.zeros(2 * N), "kl_R": np.zeros(2 * N),
        "sqrt_mu_L": np.zeros(2 * N), "sqrt_mu_R": np.zeros(2 * N),
    }
    iarrs = {k: np.zeros(N, np.int32) for k in ("ind_prop_L", "ind_prop_R", "ind_prop_conj_L", "ind_prop_conj_R")}
    for k, a in arrs.items():
        setattr(c, k, L.dptr(a))
    for k, a in iarrs.items():
        setattr(c, k, L.iptr(a))
    L.check(lib.rgf_setup_channels(C.byref(prob.p), C.byref(c)))
    return c, arrs, iarrs


def _side(c, arrs, iarrs, s, conj: bool):
    n = c.n_prop_L if s == "L" else c.n_prop_R
    kl = arrs[f"kl_{s}"][0::2] + 1j * arrs[f"kl_{s}"][1::2]
    mu = arrs[f"sqrt_mu_{s}"][0:2 * n:2] + 1j * arrs[f"sqrt_mu_{s}"][1:2 * n:2]
    d = {"N_prop": n, "kxdx": kl, "ind_prop": iarrs[f"ind_prop_{s}"][:n].astype(np.int64),
         "sqrt_mu": mu if np.any(mu.imag) else mu.real}
    if conj and c.has_conj:
        d["ind_prop_conj"] = iarrs[f"ind_prop_conj_{s}"][:n].astype(np.int64)
    return d


def _profile_fn(prob: _Prob, N: int, ndim: int, bloch: bool, shape=None):
    """channels.fun_phi through rgf_transverse_profiles (setup_channels.m:133-156,
    setup_channels_3D.m:118,123): 2D fun_phi(kydx) -> N x len(kydx); 3D
    fun_phi(kxdx, kydx) -> nx x ny for scalars (nx x ny x K for K pairs)."""
    lib = L.load_library()

    def run(k1, k2=None):
        k1 = np.ascontiguousarray(np.atleast_1d(k1), dtype=np.float64).ravel()
        k2 = None if k2 is None else np.ascontiguousarray(np.atleast_1d(k2), dtype=np.float64).ravel()
        if k2 is not None and k2.size != k1.size:
            raise ValueError("kxdx and kydx must have the same number of elements")
        out = np.zeros(2 * N * max(k1.size, 1))
        L.check(lib.rgf_transverse_profiles(C.byref(prob.p), k1.size, L.dptr(k1),
                                            L.dptr(k2) if k2 is not None else None, L.dptr(out)))
        z = out[0::2] + 1j * out[1::2]
        z = z[:N * k1.size].reshape((N, k1.size), order="F")
        return z if bloch else z.real

    if ndim == 2:
        return lambda kydx: run(kydx)

    def fun_phi(kxdx, kydx):
        z = run(kxdx, kydx)
        K = z.shape[1]
        return z.reshape(shape, order="F") if K == 1 else z.reshape((*shape, K), order="F")
    return fun_phi


def setup_channels(ny: int, BC, k0dx, epsilon_L, epsilon_R=None) -> dict:
    """channels = setup_channels(ny, BC, k0dx, epsilon_L, epsilon_R) (setup_channels.m:1-176).

    ``fun_phi`` is a callable returning the ny x len(kydx) mode matrix.
    """
    if not (isinstance(ny, (int, np.integer)) and ny > 0):
        raise ValueError("ny must be a positive integer scalar")
    bc, ka = _bc_1d(BC)
    two_sided = epsilon_R is not None and not (_is_num(epsilon_R) and np.isnan(epsilon_R))
    eR = _scalar(epsilon_R, "epsilon_R") if two_sided else _scalar(epsilon_L, "epsilon_L")
    prob = _Prob(2, (ny, 0, 1), None, _scalar(epsilon_L, "epsilon_L"), eR, _scalar(k0dx, "k0dx"), [bc], [ka])
    c, arrs, iarrs = _channels_raw(prob, ny)
    fun_phi = _profile_fn(prob, ny, 2, bc == L.RGF_BC_BLOCH)
    ch = {"kydx": arrs["kt1"].copy(), "fun_phi": fun_phi, "L": _side(c, arrs, iarrs, "L", True)}
    if two_sided:
        ch["R"] = ch["L"] if complex(epsilon_R) == complex(epsilon_L) else _side(c, arrs, iarrs, "R", True)
    else:
        ch["R"] = {"N_prop": float("nan")}
    return ch


def setup_channels_3D(nx: int, ny: int, BC, k0dx, epsilon_L, epsilon_R=None) -> dict:
    """channels = setup_channels_3D(nx, ny, BC, k0dx, epsilon_L, epsilon_R) (setup_channels_3D.m:1-188)."""
    for n, lab in ((nx, "nx"), (ny, "ny")):
        if not (isinstance(n, (int, np.integer)) and n > 0):
            raise ValueError(f"{lab} must be a positive integer scalar")
    bcs, kas = _bc_xy(BC)
    two_sided = epsilon_R is not None and not (_is_num(epsilon_R) and np.isnan(epsilon_R))
    eR = _scalar(epsilon_R, "epsilon_R") if two_sided else _scalar(epsilon_L, "epsilon_L")
    prob = _Prob(3, (nx, ny, 0), None, _scalar(epsilon_L, "epsilon_L"), eR, _scalar(k0dx, "k0dx"), bcs, kas)
    N = nx * ny
    c, arrs, iarrs = _channels_raw(prob, N)
    fun_phi = _profile_fn(prob, N, 3, bcs[0] == L.RGF_BC_BLOCH, (nx, ny))
    ch = {"kxdx": arrs["kt1"].reshape((nx, ny), order="F").copy(),
          "kydx": arrs["kt2"].reshape((nx, ny), order="F").copy(), "fun_phi": fun_phi}
    for s in ("L", "R"):
        if s == "R" and not two_sided:
            ch["R"] = {"N_prop": float("nan")}
            continue
        d = _side(c, arrs, iarrs, s, False)
        d["kzdx"] = d.pop("kxdx").reshape((nx, ny), order="F")
        ch[s] = d
    if two_sided and complex(epsilon_R) == complex(epsilon_L):
        ch["R"] = ch["L"]
    return ch


def build_laplacian_1d(n: int, BC):
    """A = build_laplacian_1d(n, BC) as a scipy.sparse CSR matrix (build_laplacian_1d.m:1-77)."""
    import scipy.sparse as sp
    if not (isinstance(n, (int, np.integer)) and n >= 0):
        raise ValueError("The number of sites n must be a non-negative integer scalar")
    if _is_num(BC) and np.imag(BC) != 0:
        warnings.warn(f"k*a = {np.real(BC):g} + 1i*{np.imag(BC):g} is a complex number.")
        raise ValueError("complex Bloch phases are not supported by build_laplacian_1d")
    bc, ka = _bc_1d(BC)
    A = np.zeros(2 * n * n)
    L.check(L.load_library().rgf_build_laplacian_1d(n, bc, ka, L.dptr(A)))
    M = (A[0::2] + 1j * A[1::2]).reshape((n, n), order="F")
    return sp.csr_matrix(M if np.any(M.imag) else M.real)


RGF_DENSE_3M = 0x100


def dense_inverse(A: np.ndarray, precision: str = "fp64", block_size: int = 0, device: int = -1,
                  complex_3m: bool = False) -> np.ndarray:
    """In-place-style inverse through the slice-inverse kernel (kernel-level test entry);
    complex_3m: the FP32 kernels with three real products per complex product."""
    lib = L.load_library()
    M = np.asfortranarray(A, dtype=np.complex128).copy(order="F")
    prec = _PRECISIONS[precision] | (RGF_DENSE_3M if complex_3m else 0)
    with L.using(device) as h:
        L.check(lib.rgf_dense_inverse(h, prec, M.shape[0], M.ctypes.data_as(L._dp), block_size), h)
    return M


def dense_gemm(A: np.ndarray, B: np.ndarray, precision: str = "fp64", device: int = -1) -> np.ndarray:
    lib = L.load_library()
    A = np.asfortranarray(A, dtype=np.complex128)
    B = np.asfortranarray(B, dtype=np.complex128)
    Cm = np.zeros((A.shape[0], B.shape[1]), dtype=np.complex128, order="F")
    with L.using(device) as h:
        L.check(lib.rgf_dense_gemm(h, _PRECISIONS[precision], A.shape[0], B.shape[1], A.shape[1],
                                   A.ctypes.data_as(L._dp), B.ctypes.data_as(L._dp), Cm.ctypes.data_as(L._dp)), h)
    return Cm


def transmission_gradient(syst: dict, out=("R",), in_=("L",), *, three_d: bool = False, checkpoint_every: int = 32,
                          device: int = -1, return_info: bool = False, precision: str = "fp64",
                          options: dict | None = None):
    """T = sum |t_ij|^2 over the selected transmission block and dT/d eps.

    ``out``/``in_`` follow cal_smatrix_RGF (cells or structs) but must select
    inputs on one side and outputs on the other.  Returns ``(T, dT/dRe(eps),
    dT/dIm(eps))`` with the gradients shaped like ``syst['epsilon']``.
    ``precision`` 'fp64': the adjoint keeps every ``checkpoint_every``-th
    slice inverse and recomputes the rest segment by segment (0 keeps all of
    them).  'fp32_refined': FP32 slice inverses (all kept) with the forward
    fields and the adjoint rows each refined to FP64 accuracy.
    """
    prob = _problem_3d(syst) if three_d else _problem_2d(syst)
    lib = L.load_library()
    nprop = _query(prob)
    sel = _Sel(_parse_selection(list(out) if isinstance(out, tuple) else out, "out", nprop),
               _parse_selection(list(in_) if isinstance(in_, tuple) else in_, "in", nprop))
    shape = tuple(syst["epsilon"].shape)
    n = int(np.prod(shape))
    gre = np.zeros(n)
    gim = np.zeros(n)
    T = C.c_double()
    info = L.Info()
    o = _options(precision, **(options or {}))
    if prob.device is not None:
        device = prob.device
    with L.using(device) as h:
        st = lib.rgf_transmission_gradient(h, C.byref(prob.p), C.byref(sel.s), C.byref(o), int(checkpoint_every),
                                           C.byref(T), L.dptr(gre), L.dptr(gim), C.byref(info))
        L.check(st, h)
    gre = gre.reshape(shape, order="F")
    gim = gim.reshape(shape, order="F")
    if return_info:
        return T.value, gre, gim, info.as_dict()
    return T.value, gre, gim
