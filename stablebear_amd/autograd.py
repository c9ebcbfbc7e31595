"""Differentiable pdist (p = 2) and l2_kernel for PyTorch-ROCm.

``pdist_l2(values, times, offsets)`` returns the compressed L2 distance vector
(``n(n-1)/2`` entries, index ``j(j-1)/2 + i`` for i < j) and
``l2_kernel(values, times, offsets)`` the compressed Gram vector (``n(n+1)/2``
entries, index ``j(j+1)/2 + i`` for i <= j). Both are ``torch.autograd``
functions whose backward is the hand-written HIP kernel behind
``sbamd_plan_backward``: it returns dL/dvalues and dL/dtimes (breakpoint
times) indexed like the inputs.

Conventions of the gradient (no reference counterpart): the first breakpoint
(t = 0) is fixed (zero gradient); where two PCFs have a breakpoint at the same
time the derivative is one-sided (f's breakpoint sees g's left-limit value,
g's sees f's right-limit value); pairs at distance 0 contribute nothing to the
distance gradient; the last value of a PCF only has a finite gradient when the
integrand of the final unbounded piece is zero.
"""

from __future__ import annotations

import numpy as np
import torch

from . import runtime


def _prepare(values, times, offsets, precision):
    if values.dtype != torch.float64 or times.dtype != torch.float64:
        raise TypeError("autograd pairwise integration needs float64 values and times")
    if values.shape != times.shape or values.dim() != 1:
        raise ValueError("values and times must be 1-D tensors of the same length")
    off = offsets.detach().cpu().numpy() if isinstance(offsets, torch.Tensor) else np.asarray(offsets)
    off = np.ascontiguousarray(off, dtype=np.int64)
    if off[-1] != values.numel():
        raise ValueError("offsets[-1] must equal the number of points")
    csr = runtime.csr_from_arrays(off, times.detach().cpu().numpy(), values.detach().cpu().numpy())
    for k in range(len(off) - 1):
        if off[k + 1] == off[k]:
            raise ValueError("autograd inputs need at least one point per PCF")
    return runtime.DeviceSet(csr, precision)


class _Pairwise(torch.autograd.Function):
    @staticmethod
    def forward(ctx, values, times, offsets, op, precision):
        ds = _prepare(values, times, offsets, precision)
        plan = runtime.Plan(ds, op=op, p=2.0 if op == "lp" else 1.0)
        out_f64 = precision == "fp64"
        out = torch.empty(plan.out_count, dtype=torch.float64 if out_f64 else torch.float32,
                          device=values.device)
        stream = torch.cuda.current_stream(values.device).cuda_stream
        if plan.out_count:
            plan.execute(out.data_ptr(), out_f64, stream=stream)
        ctx.plan = plan
        ctx.dset = ds
        ctx.out_f64 = out_f64
        ctx.save_for_backward(out)
        ctx.npts = values.numel()
        return out

    @staticmethod
    def backward(ctx, grad_out):
        (out,) = ctx.saved_tensors
        dev = out.device
        g = grad_out.to(out.dtype).contiguous()
        gv = torch.empty(ctx.npts, dtype=torch.float64, device=dev)
        gt = torch.empty(ctx.npts, dtype=torch.float64, device=dev)
        stream = torch.cuda.current_stream(dev).cuda_stream
        if ctx.plan.out_count:
            ctx.plan.backward(g.data_ptr(), out.data_ptr(), ctx.out_f64, gv.data_ptr(),
                              gt.data_ptr(), True, stream=stream)
        else:
            gv.zero_()
            gt.zero_()
        return gv, gt, None, None, None


def pdist_l2(values: torch.Tensor, times: torch.Tensor, offsets, precision: str = "fp32c"):
    """Differentiable compressed L2 distances of the PCFs given in CSR form."""
    return _Pairwise.apply(values, times, offsets, "lp", precision)


def l2_kernel(values: torch.Tensor, times: torch.Tensor, offsets, precision: str = "fp32c"):
    """Differentiable compressed L2 Gram matrix of the PCFs given in CSR form."""
    return _Pairwise.apply(values, times, offsets, "dot", precision)


__all__ = ["pdist_l2", "l2_kernel"]
