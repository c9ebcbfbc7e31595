"""Many independent pdist / l2_kernel problems in one launch.

No reference counterpart: the reference runs one ``pdist`` per problem, each
with its own host pipeline. Here every problem's PCFs are staged once on the
device, one tile list spans all problems (longest tiles first), and a single
persistent launch of the tile kernel computes every matrix; the launch can be
captured in a hipGraph and replayed.
"""

from __future__ import annotations

import numpy as np

from . import runtime
from .distance import _DISTANCE_BACKEND_MAP, _resolve
from .distance_matrix import DistanceMatrix
from .symmetric_matrix import SymmetricMatrix
from .typing import pcf32, pcf64


class BatchedPlan:
    """Device-resident batch of problems and their single tile plan.

    ``problems``: list of CSR triples ``(offsets, t, v)`` (or ``runtime.Csr``).
    ``op``: "l1", "lp" (with ``p``) or "dot" (kernel, SymmetricMatrix layout).
    """

    def __init__(self, problems, op="l1", p=1.0, precision="fp32c", shape=None,
                 persistent=True):
        csrs = [c if isinstance(c, runtime.Csr) else runtime.csr_from_arrays(*c) for c in problems]
        if not csrs:
            raise ValueError("need at least one problem")
        self.sets = [runtime.DeviceSet(c, precision) for c in csrs]
        self.sizes = [s.n for s in self.sets]
        self.op = op
        self.plan = runtime.Plan(self.sets, op=op, p=p, shape=shape, batched=True)
        self.plan.set_persistent(persistent)
        self.offsets = [self.plan.out_offset(b) for b in range(len(self.sets))]

    @property
    def out_count(self):
        return self.plan.out_count

    def execute(self, out_ptr: int, out_f64: bool, stream: int | None = None):
        self.plan.execute(out_ptr, out_f64, stream=stream)

    def graph(self, out_ptr: int, out_f64: bool) -> runtime.Graph:
        return self.plan.graph(out_ptr, out_f64)

    def split(self, flat: np.ndarray):
        """Per-problem compressed vectors from the concatenated output."""
        ends = self.offsets[1:] + [self.out_count]
        return [flat[a:b] for a, b in zip(self.offsets, ends)]


def _problems(containers, backend_map):
    out = []
    dtype = None
    for c in containers:
        _, X = _resolve(backend_map, c)
        if len(X.shape) != 1:
            raise ValueError("1d tensor expected.")
        if dtype is not None and X.dtype is not dtype:
            raise TypeError("All problems must have the same dtype.")
        dtype = X.dtype
        out.append(runtime.csr_from_pcfs(X.flat_pcfs(), X.dtype))
    return out, dtype


def pdist_batched(containers, p=1, precision=None) -> list[DistanceMatrix]:
    """``pdist`` of each 1-D PCF container, all computed in one launch."""
    if p < 1:
        raise ValueError("p must be >= 1.")
    csrs, dtype = _problems(containers, _DISTANCE_BACKEND_MAP)
    prec = "fp32c" if dtype is pcf32 else (precision or "fp64")
    op = "l1" if p == 1 else "lp"
    bp = BatchedPlan(csrs, op=op, p=float(p), precision=prec)
    out_f64 = dtype is pcf64
    flat = bp.plan.execute_to_host(out_f64)
    res = []
    for n, part in zip(bp.sizes, bp.split(flat)):
        res.append(DistanceMatrix._from_storage(n, np.array(part)))
    return res


def l2_kernel_batched(containers, precision=None) -> list[SymmetricMatrix]:
    """``l2_kernel`` of each 1-D PCF container, all computed in one launch."""
    csrs, dtype = _problems(containers, _DISTANCE_BACKEND_MAP)
    prec = "fp32c" if dtype is pcf32 else (precision or "fp64")
    bp = BatchedPlan(csrs, op="dot", precision=prec)
    flat = bp.plan.execute_to_host(dtype is pcf64)
    return [SymmetricMatrix._from_storage(n, np.array(part))
            for n, part in zip(bp.sizes, bp.split(flat))]


__all__ = ["BatchedPlan", "pdist_batched", "l2_kernel_batched"]
