"""Expectation values and adjoint gradients on the device (include/clifft_svm.h,
clifft_svm_adjoint_*), and the torch.autograd.Function around them.

No counterpart in the reference, which evaluates EXP_VAL probes
(src/clifft/svm/svm_kernels.inl:2031-2173) but has no gradients.  For a
noiseless, measurement-free program compiled with keep_rotations=True, each
R_Z(theta_i / pi) is one ARRAY_ROT / EXPAND_ROT instruction
(program.param_instr[i]) whose phase is exp(i * param_sign[i] * theta_i).

    adj = Adjoint(program)
    exp_vals, grads = adj.run(thetas, coeffs)            # numpy in, numpy out
    F = expectation(adj, thetas_tensor, coeffs_tensor)    # torch, differentiable

F[b] = sum_j coeffs[j] * exp_vals[b, j]; dF/dtheta comes from one forward and
one backward (adjoint) sweep per theta vector in a hand-written HIP kernel.
"""

from __future__ import annotations

import ctypes
import weakref

import numpy as np

from . import _lib
from .svm import device_program


class Adjoint:
    def __init__(self, program, param_instr=None, param_sign=None):
        self.dp = device_program(program)
        prog = self.dp.program
        pi = prog.param_instr if param_instr is None else param_instr
        ps = prog.param_sign if param_sign is None else param_sign
        if pi is None:
            raise ValueError("program has no parameter slots (compile with keep_rotations=True)")
        self.param_instr = np.ascontiguousarray(pi, dtype=np.uint32)
        self.param_sign = np.ascontiguousarray(ps, dtype=np.int8)
        if np.any(self.param_instr == np.uint32(0xFFFFFFFF)):
            raise ValueError("a parameter was compiled away (fused or dropped); compile with keep_rotations=True")
        self.num_params = len(self.param_instr)
        self.num_exp_vals = prog.num_exp_vals
        h = ctypes.c_void_p()
        L = _lib.lib()
        _lib.check(L.clifft_svm_adjoint_create(self.dp.handle, self.param_instr.ctypes.data,
                                               self.param_sign.ctypes.data, self.num_params, ctypes.byref(h)),
                   "adjoint_create")
        self.handle = h
        self._fin = weakref.finalize(self, L.clifft_svm_adjoint_destroy, h)

    def run(self, thetas, coeffs=None, grads: bool = True):
        """Host arrays: thetas [B, P] -> (exp_vals [B, n_ev], grads [B, P] or None)."""
        th = np.ascontiguousarray(np.asarray(thetas, dtype=np.float64).reshape(-1, self.num_params))
        B = th.shape[0]
        ev = np.zeros((B, self.num_exp_vals), np.float64)
        g = np.zeros((B, self.num_params), np.float64) if grads else None
        c = None
        if grads:
            c = np.ascontiguousarray(np.asarray(coeffs, dtype=np.float64).reshape(-1))
            if c.size != self.num_exp_vals:
                raise ValueError(f"expected {self.num_exp_vals} coefficients, got {c.size}")
        _lib.check(_lib.lib().clifft_svm_adjoint_run(self.handle, th.ctypes.data if th.size else None, B,
                                                     None if c is None else c.ctypes.data,
                                                     ev.ctypes.data if ev.size else None,
                                                     None if g is None else g.ctypes.data, 0, None),
                   "adjoint_run")
        return ev, g

    def run_device(self, thetas, batch: int, coeffs, exp_vals, grads, stream=None):
        """Device pointers (ints or None); asynchronous on `stream`."""
        _lib.check(_lib.lib().clifft_svm_adjoint_run(self.handle, thetas, int(batch), coeffs, exp_vals, grads,
                                                     _lib.REC_DEVICE, stream), "adjoint_run")


def _run_torch(adj: Adjoint, thetas, coeffs, want_grad: bool):
    """Launch on a side stream ordered after, and before, torch's current
    stream (which may be the null stream, a handle the C ABI reads as "the
    library's own stream")."""
    import torch
    cur = torch.cuda.current_stream(thetas.device)
    side = torch.cuda.Stream(device=thetas.device)
    side.wait_stream(cur)
    if thetas.dim() != 2 or thetas.shape[1] != adj.num_params:
        raise ValueError(f"thetas must have shape [batch, {adj.num_params}], got {tuple(thetas.shape)}")
    if coeffs.numel() != adj.num_exp_vals:
        raise ValueError(f"expected {adj.num_exp_vals} coefficients, got {coeffs.numel()}")
    with torch.cuda.stream(side):
        th = thetas.detach().to(torch.float64).contiguous()
        B = th.shape[0]
        ev = torch.empty((B, adj.num_exp_vals), dtype=torch.float64, device=th.device)
        g = torch.empty((B, adj.num_params), dtype=torch.float64, device=th.device) if want_grad else None
        c = coeffs.detach().to(device=th.device, dtype=torch.float64).reshape(-1).contiguous()
        _lib.check(_lib.lib().clifft_svm_adjoint_run(adj.handle, th.data_ptr() if th.numel() else None, B,
                                                     c.data_ptr() if c.numel() else None,
                                                     ev.data_ptr() if ev.numel() else None,
                                                     None if g is None else g.data_ptr(), _lib.REC_DEVICE,
                                                     side.cuda_stream),
                   "adjoint_run")
    cur.wait_stream(side)
    for t in (th, ev, c) + ((g,) if g is not None else ()):
        t.record_stream(cur)
    return ev, g, c


def _function():
    import torch

    class ExpectationFunction(torch.autograd.Function):
        """F[b] = sum_j coeffs[j] <P_j>(thetas[b]).  dF/dthetas is the
        adjoint gradient computed in the same kernel; dF/dcoeffs[j] is
        <P_j>(thetas[b]) itself.  backward multiplies both by the incoming
        grad."""

        @staticmethod
        def forward(ctx, thetas, adj, coeffs):
            want = thetas.requires_grad
            ev, g, c = _run_torch(adj, thetas, coeffs, want)
            ctx.save_for_backward(g if want else None, ev if coeffs.requires_grad else None)
            ctx.coeffs_shape = coeffs.shape
            return ev @ c

        @staticmethod
        def backward(ctx, grad_out):
            g, ev = ctx.saved_tensors
            d_th = grad_out.unsqueeze(1) * g if ctx.needs_input_grad[0] else None
            d_c = (grad_out @ ev).reshape(ctx.coeffs_shape) if ctx.needs_input_grad[2] else None
            return d_th, None, d_c

    return ExpectationFunction


_FN = None


def expectation(adj, thetas, coeffs):
    """Differentiable F(thetas) for a CUDA float64 tensor thetas [B, P]."""
    global _FN
    if _FN is None:
        _FN = _function()
    if not isinstance(adj, Adjoint):
        adj = Adjoint(adj)
    return _FN.apply(thetas, adj, coeffs)


def exp_vals_torch(adj: Adjoint, thetas):
    """Forward only: exp_vals [B, n_ev] for a CUDA float64 tensor."""
    import torch
    ev, _, _ = _run_torch(adj, thetas, torch.zeros(adj.num_exp_vals, dtype=torch.float64, device=thetas.device),
                          False)
    return ev
