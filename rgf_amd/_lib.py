"""ctypes binding of librgf.so (the C ABI declared in include/rgf.h)."""
from __future__ import annotations

import contextlib
import ctypes as C
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librgf.so")

RGF_OK = 0
RGF_ERR_INVALID_ARGUMENT = 1
RGF_ERR_BOUNDARY_CONDITION = 2
RGF_ERR_CHANNEL_INDEX = 3
RGF_ERR_OUT_OF_MEMORY = 4
RGF_ERR_DEVICE = 5
RGF_ERR_NUMERICAL = 6
RGF_ERR_NOT_CONVERGED = 7
RGF_ERR_BUFFER_TOO_SMALL = 8

RGF_ALL = -1
RGF_BC_BLOCH, RGF_BC_DIRICHLET, RGF_BC_NEUMANN, RGF_BC_DIRICHLET_NEUMANN, RGF_BC_NEUMANN_DIRICHLET = range(5)
RGF_PRECISION_FP64 = 0
RGF_PRECISION_FP32_REFINED = 1

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class Problem(C.Structure):
    _fields_ = [
        ("ndim", C.c_int32), ("dims", C.c_int64 * 3),
        ("eps_re", C.c_void_p), ("eps_im", C.c_void_p),
        ("eps_L", C.c_double * 2), ("eps_R", C.c_double * 2), ("k0dx", C.c_double * 2),
        ("bc", C.c_int32 * 2), ("bloch_ka", C.c_double * 2), ("eps_on_device", C.c_int32),
    ]


class Selection(C.Structure):
    _fields_ = [
        ("ind_in_L", _ip), ("n_in_L", C.c_int32),
        ("ind_in_R", _ip), ("n_in_R", C.c_int32),
        ("ind_out_L", _ip), ("n_out_L", C.c_int32),
        ("ind_out_R", _ip), ("n_out_R", C.c_int32),
    ]


class Options(C.Structure):
    _fields_ = [
        ("precision", C.c_int32), ("max_refine_iter", C.c_int32), ("refine_tol", C.c_double),
        ("block_size", C.c_int32), ("complex_3m", C.c_int32), ("profile", C.c_int32),
        ("single_stream", C.c_int32), ("batch_concurrency", C.c_int32), ("batch_engine", C.c_int32),
        ("batch_graph_rounds", C.c_int32), ("autotune", C.c_int32),
        ("refine_s_tol", C.c_double), ("factor_slots", C.c_int32), ("pivoting", C.c_int32),
    ]


class Info(C.Structure):
    _fields_ = [
        ("n_transverse", C.c_int64), ("n_slices", C.c_int64),
        ("n_prop_L", C.c_int32), ("n_prop_R", C.c_int32), ("n_out", C.c_int32), ("n_in", C.c_int32),
        ("swept_right_to_left", C.c_int32), ("refine_iters", C.c_int32),
        ("refine_residual", C.c_double), ("seconds_total", C.c_double), ("seconds_setup", C.c_double),
        ("seconds_slices", C.c_double), ("seconds_refine", C.c_double), ("flops_slices", C.c_double),
        ("kernel_seconds", C.c_double), ("kernel_flops", C.c_double), ("kernel_launches", C.c_int64),
        ("refine_s_err_estimate", C.c_double), ("max_multiplier", C.c_double),
        ("block_size", C.c_int32), ("autotuned", C.c_int32),
        ("factor_slices_resident", C.c_int64), ("factor_slices_rebuilt", C.c_int64),
        ("pivoted", C.c_int32),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class Channels(C.Structure):
    _fields_ = [
        ("kt1", _dp), ("kt2", _dp), ("kl_L", _dp), ("kl_R", _dp),
        ("ind_prop_L", _ip), ("ind_prop_R", _ip), ("sqrt_mu_L", _dp), ("sqrt_mu_R", _dp),
        ("ind_prop_conj_L", _ip), ("ind_prop_conj_R", _ip),
        ("n_prop_L", C.c_int32), ("n_prop_R", C.c_int32), ("has_conj", C.c_int32),
    ]


# symbol -> (restype, argtypes): every entry point of include/rgf.h
SIGNATURES = {
    "rgf_abi_version": (C.c_int32, []),
    "rgf_status_string": (C.c_char_p, [C.c_int]),
    "rgf_default_options": (None, [C.POINTER(Options)]),
    "rgf_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int32]),
    "rgf_destroy": (C.c_int, [C.c_void_p]),
    "rgf_last_error": (C.c_char_p, [C.c_void_p]),
    "rgf_set_stream": (C.c_int, [C.c_void_p, C.c_void_p]),
    "rgf_trim": (C.c_int, [C.c_void_p]),
    "rgf_query": (C.c_int, [C.POINTER(Problem), C.POINTER(Selection), _ip, _ip, _ip, _ip]),
    "rgf_setup_channels": (C.c_int, [C.POINTER(Problem), C.POINTER(Channels)]),
    "rgf_transverse_profiles": (C.c_int, [C.POINTER(Problem), C.c_int64, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                          C.POINTER(C.c_double)]),
    "rgf_build_laplacian_1d": (C.c_int, [C.c_int64, C.c_int32, C.c_double, _dp]),
    "rgf_solve": (C.c_int, [C.c_void_p, C.POINTER(Problem), C.POINTER(Selection), C.POINTER(Options), _dp,
                            C.POINTER(Info)]),
    "rgf_solve_fields": (C.c_int, [C.c_void_p, C.POINTER(Problem), C.POINTER(Selection), C.POINTER(Options), _dp,
                                   _dp, C.POINTER(Info)]),
    "rgf_solve_batched": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(Problem), C.POINTER(Selection),
                                    C.POINTER(Options), C.POINTER(_dp), C.POINTER(Info)]),
    "rgf_transmission_gradient": (C.c_int, [C.c_void_p, C.POINTER(Problem), C.POINTER(Selection),
                                            C.POINTER(Options), C.c_int32, _dp, _dp, _dp, C.POINTER(Info)]),
    "rgf_dense_inverse": (C.c_int, [C.c_void_p, C.c_int32, C.c_int64, _dp, C.c_int32]),
    "rgf_dense_inverse_noalloc": (C.c_int, [C.c_void_p, C.c_int32, C.c_int64, C.c_void_p, C.c_void_p, C.c_int64,
                                            C.c_int32, C.c_void_p, C.c_void_p]),
    "rgf_dense_gemm": (C.c_int, [C.c_void_p, C.c_int32, C.c_int64, C.c_int64, C.c_int64, _dp, _dp, _dp]),
}

_lib = None
_lock = threading.Lock()
_handles: dict[int, int] = {}
# A handle's cached device buffers and stream are not safe for concurrent use
# (ctypes releases the GIL during the call): every call that takes a handle
# holds that handle's lock, including set_stream / solve / reset sequences.
_handle_locks: dict[int, threading.RLock] = {}


class RGFError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(message)
        self.status = status


def load_library():
    """Load librgf.so (built in-tree by rgf_amd.build); raises if it is missing."""
    global _lib
    with _lock:
        if _lib is None:
            # torch ships its own libamdhip64.so.7 / libhsa-runtime64.so.1 with
            # the same SONAMEs as the system ROCm; whichever is loaded first
            # serves the whole process, and torch cannot start on a runtime
            # another library initialised first.  Let torch initialise first.
            try:
                import torch
                torch.cuda.is_available()
            except ImportError:
                pass
            if not os.path.exists(LIB_PATH):
                raise ImportError(f"{LIB_PATH} is not built; run `python -m rgf_amd.build`")
            lib = C.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            _lib = lib
    return _lib


def check(status: int, handle=None) -> None:
    if status == RGF_OK:
        return
    lib = load_library()
    msg = lib.rgf_last_error(handle).decode()
    if status in (RGF_ERR_INVALID_ARGUMENT, RGF_ERR_BOUNDARY_CONDITION, RGF_ERR_CHANNEL_INDEX):
        raise ValueError(msg)
    if status == RGF_ERR_OUT_OF_MEMORY:
        raise MemoryError(msg)
    raise RGFError(status, msg)


def handle(device: int = -1) -> int:
    """The process-wide handle of a device (created on first use).  Calls that
    pass it to the library must hold its lock: use ``using(device)``."""
    lib = load_library()
    with _lock:
        if device not in _handles:
            h = C.c_void_p()
            st = lib.rgf_create(C.byref(h), device)
            if st != RGF_OK:
                raise RGFError(st, lib.rgf_last_error(None).decode())
            _handles[device] = h.value
            _handle_locks[h.value] = threading.RLock()
        return _handles[device]


@contextlib.contextmanager
def using(device: int = -1):
    """The device's handle, held exclusively by this thread for the block."""
    h = handle(device)
    with _handle_locks[h]:
        yield h


def trim(device: int = -1) -> None:
    """Release the device workspace cached by the device's handle (rgf_trim)."""
    lib = load_library()
    with using(device) as h:
        check(lib.rgf_trim(h), h)


def dptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def iptr(a: np.ndarray):
    return a.ctypes.data_as(_ip)
