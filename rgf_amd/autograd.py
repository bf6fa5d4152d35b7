"""torch.autograd.Function for the transmission T = sum |t_ij|^2 and its adjoint.

When eps requires a gradient, the forward pass and the gradient come from one
call of ``rgf_transmission_gradient`` in librgf (the hand-written adjoint of
the slice recursion); otherwise the forward pass is one S-matrix solve.
CUDA tensors are handed to the library as device pointers and the library
runs on torch's current stream.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib as L
from .api import _options, _parse_selection, _problem_from_tensor, _query, _Sel, _solve


class TransmissionFunction(torch.autograd.Function):
    """T(eps) = sum_ij |t_ij|^2; backward returns dT/d eps (real eps) or
    dT/dRe + i dT/dIm (complex eps, PyTorch's convention for real losses)."""

    @staticmethod
    def forward(ctx, eps, epsilon_L, epsilon_R, k0dx, bc, out, in_, three_d, checkpoint_every, precision):
        lib = L.load_library()
        prob = _problem_from_tensor(eps, epsilon_L, epsilon_R, k0dx, bc, three_d)
        nprop = _query(prob)
        sel = _Sel(_parse_selection(list(out), "out", nprop), _parse_selection(list(in_), "in", nprop))
        shape = tuple(eps.shape)
        n = int(np.prod(shape))
        dev = eps.device.index if eps.is_cuda else -1
        if not ctx.needs_input_grad[0]:
            # no gradient wanted: T from the S block alone (one forward solve)
            S, _ = _solve(prob, list(out), list(in_), precision, dev if dev is not None else -1, None)
            ctx.save_for_backward(torch.zeros(0, device=eps.device))
            return torch.tensor(float(np.sum(np.abs(S) ** 2)), dtype=torch.float64, device=eps.device)
        gre = np.zeros(n)
        gim = np.zeros(n)
        T = C.c_double()
        with L.using(dev if dev is not None else -1) as h:
            if eps.is_cuda:
                L.check(lib.rgf_set_stream(h, C.c_void_p(torch.cuda.current_stream(eps.device).cuda_stream)), h)
            try:
                st = lib.rgf_transmission_gradient(h, C.byref(prob.p), C.byref(sel.s), C.byref(_options(precision)),
                                                   int(checkpoint_every), C.byref(T), L.dptr(gre), L.dptr(gim), None)
            finally:
                if eps.is_cuda:
                    lib.rgf_set_stream(h, None)
            L.check(st, h)
        g = gre.reshape(shape, order="F")
        if eps.is_complex():
            g = g + 1j * gim.reshape(shape, order="F")
        ctx.save_for_backward(torch.as_tensor(g, dtype=eps.dtype if eps.is_complex() else torch.float64,
                                              device=eps.device))
        return torch.tensor(T.value, dtype=torch.float64, device=eps.device)

    @staticmethod
    def backward(ctx, grad_out):
        (g,) = ctx.saved_tensors
        return (grad_out * g).to(g.dtype), None, None, None, None, None, None, None, None, None


def transmission(eps: torch.Tensor, epsilon_L, epsilon_R, k0dx, bc, out=("R",), in_=("L",), three_d: bool = False,
                 checkpoint_every: int = 32, precision: str = "fp64") -> torch.Tensor:
    """Differentiable T = sum |t_ij|^2 of the slab eps (MATLAB index order);
    precision 'fp64' or 'fp32_refined' (as transmission_gradient)."""
    return TransmissionFunction.apply(eps, epsilon_L, epsilon_R, k0dx, bc, tuple(out), tuple(in_), three_d,
                                      checkpoint_every, precision)
