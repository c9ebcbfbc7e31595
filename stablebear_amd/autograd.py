"""Differentiable pdist (p = 2) and l2_kernel for PyTorch-ROCm.

``pdist_l2(values, times, offsets)`` returns the compressed L2 distance vector
(``n(n-1)/2`` entries, index ``j(j-1)/2 + i`` for i < j) and
``l2_kernel(values, times, offsets)`` the compressed Gram vector (``n(n+1)/2``
entries, index ``j(j+1)/2 + i`` for i <= j). Both are ``torch.autograd``
functions whose backward is the hand-written HIP kernel behind
``sbamd_plan_backward``: it returns dL/dvalues and dL/dtimes (breakpoint
times) indexed like the inputs.

Everything stays on the device: the PCF set is validated and packed on the
GPU from the CSR tensors as they lie, and the set and its tile plan are kept
while the offsets and times are unchanged, so a training step that only moves
the values refreshes them with one gather launch (``sbamd_set_update_values``)
instead of rebuilding the set.

Conventions of the gradient (no reference counterpart): the first breakpoint
(t = 0) is fixed (zero gradient); where breakpoints of two different PCFs
coincide the derivative does not exist, and the backward takes the mean of
the partner's left and right limits there, so the two one-sided results of
the pair add up to the derivative of moving the shared time; pairs at
distance 0 contribute nothing to the distance gradient; the last value of a
PCF only has a finite gradient when the integrand of the final unbounded
piece is zero.
"""

from __future__ import annotations

import collections

import torch

from . import runtime

_CACHE_SIZE = 4


class _Entry:
    """A device set and its plan for one (offsets, times) pair."""

    def __init__(self, off, times, values, op, precision):
        self.off = off.clone()
        self.times = times.detach().clone()
        self.dset = runtime.DeviceSet.from_device(off, times.detach(), values.detach(), precision)
        self.plan = runtime.Plan(self.dset, op=op, p=2.0 if op == "lp" else 1.0)
        self.owner = None  # the forward whose values the set holds

    def matches(self, off, times):
        return (self.off.shape == off.shape and self.times.shape == times.shape
                and bool(torch.equal(self.off, off)) and bool(torch.equal(self.times, times)))


_cache: "collections.OrderedDict[tuple, list[_Entry]]" = collections.OrderedDict()


def _offsets_tensor(offsets, device):
    if isinstance(offsets, torch.Tensor):
        off = offsets.detach()
    else:
        off = torch.as_tensor(offsets)
    return off.to(device=device, dtype=torch.int64).contiguous()


def _entry(values, times, offsets, op, precision):
    if values.dtype != torch.float64 or times.dtype != torch.float64:
        raise TypeError("autograd pairwise integration needs float64 values and times")
    if values.shape != times.shape or values.dim() != 1:
        raise ValueError("values and times must be 1-D tensors of the same length")
    if not values.is_cuda or times.device != values.device:
        raise ValueError("values and times must be on the same GPU")
    off = _offsets_tensor(offsets, values.device)
    if off.dim() != 1 or off.numel() < 1:
        raise ValueError("offsets must be a 1-D tensor of n + 1 entries")
    times = times.detach().contiguous()
    key = (values.device.index, op, precision, off.numel(), times.numel())
    bucket = _cache.get(key, [])
    for e in bucket:
        if e.matches(off, times):
            _cache.move_to_end(key)
            return e, True
    if int(off[-1].item()) != values.numel():
        raise ValueError("offsets[-1] must equal the number of points")
    if off.numel() > 1 and bool((off[1:] == off[:-1]).any().item()):
        raise ValueError("autograd inputs need at least one point per PCF")
    e = _Entry(off, times, values.detach().contiguous(), op, precision)
    bucket = [e] + bucket[: _CACHE_SIZE - 1]
    _cache[key] = bucket
    _cache.move_to_end(key)
    while len(_cache) > _CACHE_SIZE:
        _cache.popitem(last=False)
    return e, False


class _Pairwise(torch.autograd.Function):
    @staticmethod
    def forward(ctx, values, times, offsets, op, precision):
        e, hit = _entry(values, times, offsets, op, precision)
        stream = torch.cuda.current_stream(values.device).cuda_stream
        vals = values.detach().contiguous()
        token = object()
        if hit:
            e.dset.update_values(vals.data_ptr(), stream)
        e.owner = token
        plan = e.plan
        out_f64 = precision == "fp64"
        out = torch.empty(plan.out_count, dtype=torch.float64 if out_f64 else torch.float32,
                          device=values.device)
        if plan.out_count:
            plan.execute(out.data_ptr(), out_f64, stream=stream)
        ctx.entry = e
        ctx.token = token
        ctx.out_f64 = out_f64
        ctx.save_for_backward(out, vals)
        ctx.npts = values.numel()
        return out

    @staticmethod
    def backward(ctx, grad_out):
        out, vals = ctx.saved_tensors
        e = ctx.entry
        dev = out.device
        stream = torch.cuda.current_stream(dev).cuda_stream
        if e.owner is not ctx.token:
            # a later forward on the same offsets and times loaded other values
            e.dset.update_values(vals.data_ptr(), stream)
            e.owner = ctx.token
        g = grad_out.to(out.dtype).contiguous()
        gv = torch.empty(ctx.npts, dtype=torch.float64, device=dev)
        gt = torch.empty(ctx.npts, dtype=torch.float64, device=dev)
        if e.plan.out_count:
            e.plan.backward(g.data_ptr(), out.data_ptr(), ctx.out_f64, gv.data_ptr(),
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


def clear_cache() -> None:
    """Drop the cached device sets and plans."""
    _cache.clear()


__all__ = ["pdist_l2", "l2_kernel", "clear_cache"]
