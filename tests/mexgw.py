"""ctypes driver of the MEX gateway (rgf_amd/mex/cal_smatrix_RGF_mex.c) built
against the host MEX API harness in tests/mex_harness: MATLAB values are
built from Python ones (dict -> struct, list of str -> cell, str -> char,
numbers and arrays -> double, column-major), mexFunction is called, and
errors come back as MexError with MATLAB's identifier."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mex_harness", "libmexgw.so")
_P = C.c_void_p


class MexError(Exception):
    def __init__(self, ident: str, msg: str):
        super().__init__(f"{ident}: {msg}")
        self.ident = ident
        self.msg = msg


class Gateway:
    def __init__(self, path: str = LIB):
        lib = C.CDLL(path)
        sig = {
            "h_double": (_P, [C.c_int, C.POINTER(C.c_int64), C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
            "h_char": (_P, [C.c_char_p]), "h_struct": (_P, []), "h_set": (None, [_P, C.c_char_p, _P]),
            "h_cell": (_P, [C.c_int]), "h_cell_set": (None, [_P, C.c_int, _P]),
            "h_call": (C.c_int, [C.c_int, C.POINTER(_P), C.c_int, C.POINTER(_P)]),
            "h_error_id": (C.c_char_p, []), "h_error": (C.c_char_p, []), "h_warnings": (C.c_char_p, []),
            "h_exit": (None, []), "h_class": (C.c_int, [_P]), "h_ndim": (C.c_int, [_P]),
            "h_dim": (C.c_int64, [_P, C.c_int]), "h_is_complex": (C.c_int, [_P]),
            "h_field": (_P, [_P, C.c_char_p]), "h_read": (None, [_P, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
            "h_nfields": (C.c_int, [_P]), "h_fieldname": (C.c_char_p, [_P, C.c_int]),
        }
        for name, (res, args) in sig.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        self.lib = lib
        self.warnings = ""

    def to_mx(self, x):
        lib = self.lib
        if isinstance(x, str):
            return lib.h_char(x.encode())
        if isinstance(x, dict):
            s = lib.h_struct()
            for k, v in x.items():
                lib.h_set(s, k.encode(), self.to_mx(v))
            return s
        if isinstance(x, (list, tuple)) and all(isinstance(e, str) for e in x):   # cell ({} when empty)
            c = lib.h_cell(len(x))
            for k, e in enumerate(x):
                lib.h_cell_set(c, k, self.to_mx(e))
            return c
        a = np.asarray(x)
        if a.ndim == 0:
            a = a.reshape(1, 1)
        elif a.ndim == 1:
            a = a.reshape(1, -1)                       # MATLAB row vector
        dims = (C.c_int64 * a.ndim)(*a.shape)
        cplx = np.iscomplexobj(a)
        re = np.ascontiguousarray(np.real(a).ravel(order="F"), dtype=np.float64)
        im = np.ascontiguousarray(np.imag(a).ravel(order="F"), dtype=np.float64) if cplx else None
        dp = C.POINTER(C.c_double)
        return lib.h_double(a.ndim, dims, int(cplx), re.ctypes.data_as(dp), im.ctypes.data_as(dp) if cplx else None)

    def from_mx(self, p):
        lib = self.lib
        if lib.h_class(p) == 2:                        # struct
            return {lib.h_fieldname(p, k).decode(): self.from_mx(lib.h_field(p, lib.h_fieldname(p, k)))
                    for k in range(lib.h_nfields(p))}
        shape = tuple(lib.h_dim(p, k) for k in range(lib.h_ndim(p)))
        n = int(np.prod(shape))
        re = np.zeros(n)
        im = np.zeros(n)
        dp = C.POINTER(C.c_double)
        lib.h_read(p, re.ctypes.data_as(dp), im.ctypes.data_as(dp))
        v = re + 1j * im if lib.h_is_complex(p) else re
        return v.reshape(shape, order="F")

    def call(self, nlhs: int, *args):
        lib = self.lib
        prhs = (_P * max(len(args), 1))(*[self.to_mx(a) for a in args])
        plhs = (_P * max(nlhs, 1))()
        rc = lib.h_call(nlhs, plhs, len(args), prhs)
        self.warnings = lib.h_warnings().decode()
        if rc:
            raise MexError(lib.h_error_id().decode(), lib.h_error().decode())
        return [self.from_mx(plhs[k]) if plhs[k] else None for k in range(nlhs)]
