"""rgf_dense_inverse_noalloc (the entry point for callers that cannot
allocate) through the C ABI against rgf_dense_inverse, the entry point whose
kernels it shares: identical results on two calls in a row, the second on a
smaller matrix, in both precisions; a size beyond its workspace is refused."""
import ctypes as C

import numpy as np
import pytest
import torch

import rgf_amd
from rgf_amd import _lib as L

pytestmark = pytest.mark.gpu


def _matrix(rng, n):
    return (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n) + 4 * np.eye(n)


def _noalloc(A, precision, block=0, with_status=True):
    """Invert A through rgf_dense_inverse_noalloc on planar device copies."""
    lib = L.load_library()
    dt = torch.float64 if precision == "fp64" else torch.float32
    n = A.shape[0]
    Ar = torch.tensor(np.ascontiguousarray(A.real), dtype=dt, device="cuda:0")
    Ai = torch.tensor(np.ascontiguousarray(A.imag), dtype=dt, device="cuda:0")
    st = torch.full((2,), 7, dtype=torch.int32, device="cuda:0")  # cleared by the call
    prec = L.RGF_PRECISION_FP64 if precision == "fp64" else L.RGF_PRECISION_FP32_REFINED
    torch.cuda.synchronize()
    with L.using(0) as h:
        rc = lib.rgf_dense_inverse_noalloc(h, prec, n, C.c_void_p(Ar.data_ptr()), C.c_void_p(Ai.data_ptr()), n,
                                           block, C.c_void_p(st.data_ptr()) if with_status else None, None)
        L.check(rc, h)
        torch.cuda.synchronize()
    words = st.cpu().numpy()
    return (Ar.double() + 1j * Ai.double()).cpu().numpy(), words


@pytest.mark.parametrize("precision,sizes", [("fp64", (700, 300)), ("fp32_refined", (1000, 200))])
def test_noalloc_matches_dense_inverse_twice(precision, sizes):
    rng = np.random.default_rng(31)
    for n in sizes:
        A = _matrix(rng, n)
        if precision == "fp32_refined":
            A = A.astype(np.complex64).astype(np.complex128)  # both entries see the same FP32 values
        X, words = _noalloc(A, precision)
        ref = rgf_amd.dense_inverse(A, precision)
        print(f"{precision} n={n}: max|X - X_ref| = {np.abs(X - ref).max():.3e}", flush=True)
        assert np.array_equal(X, ref), (precision, n)
        assert words[0] == 0
        m = np.frombuffer(words[1:2].astype(np.int32).tobytes(), dtype=np.float32)[0]
        assert 0 < m < 1e3
        Xs, _ = _noalloc(A, precision, with_status=False)  # the synchronous form
        assert np.array_equal(Xs, ref)


def test_noalloc_refuses_sizes_beyond_its_workspace():
    lib = L.load_library()
    with L.using(0) as h:
        rc = lib.rgf_dense_inverse_noalloc(h, L.RGF_PRECISION_FP64, 20000, C.c_void_p(16), C.c_void_p(16), 20000, 0,
                                           None, None)
        assert rc == L.RGF_ERR_INVALID_ARGUMENT
        assert "no-allocation entry point" in lib.rgf_last_error(h).decode()
