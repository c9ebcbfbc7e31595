"""Row reductions of pairwise outputs on the GPU: per-row sums, means, minima,
maxima and the k nearest columns of a distance matrix (``DistanceMatrix``,
compressed, rows over j != i) or of a dense cross-distance matrix.

Inputs already on the device (torch CUDA tensors: 2-D dense, or the 1-D
compressed storage of an n x n distance matrix with ``n=``) are read in place
and results stay on the device; host inputs (``DistanceMatrix``, numpy or
``FloatTensor``) are uploaded once and results come back as numpy arrays.
The kernels (``sbamd_rows_reduce`` / ``sbamd_rows_knn``) read every element
once. The reference has no counterpart: its callers reduce the returned
matrices on the host (``D.mean(axis=1)`` in its cdist benchmark).
"""

from __future__ import annotations

import numpy as np

from . import _lib as L
from . import runtime
from .distance_matrix import DistanceMatrix
from .tensor import FloatTensor


def _device_matrix(D, n=None):
    """(tensor on the device, out kind, rows, cols, came from the host)."""
    import torch

    if isinstance(D, DistanceMatrix):
        return torch.from_numpy(np.ascontiguousarray(D.storage)).cuda(), L.OUT_DISTMAT, D.size, D.size, True
    if isinstance(D, FloatTensor):
        D = D.to_numpy()
    host = isinstance(D, np.ndarray)
    t = torch.from_numpy(np.ascontiguousarray(D)).cuda() if host else D
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError("expected a DistanceMatrix, a numpy / FloatTensor matrix or a CUDA tensor")
    if t.dtype not in (torch.float32, torch.float64):
        raise TypeError("matrix elements must be float32 or float64")
    t = t.contiguous()
    if t.dim() == 2:
        return t, L.OUT_DENSE, t.shape[0], t.shape[1], host
    if t.dim() == 1:
        if n is None:
            raise ValueError("compressed storage needs n (the matrix size)")
        if t.numel() != n * (n - 1) // 2:
            raise ValueError(f"compressed storage of a {n} x {n} distance matrix has "
                             f"{n * (n - 1) // 2} entries, got {t.numel()}")
        return t, L.OUT_DISTMAT, n, n, host
    raise ValueError("expected a 2-D dense matrix or 1-D compressed storage")


def _dtype(t):
    import torch

    return L.F64 if t.dtype == torch.float64 else L.F32


def row_stats(D, n=None):
    """Per-row (sum, mean, min, max) in FP64 over j != i (distance matrices)
    or all columns (dense)."""
    import torch

    t, kind, nr, nc, host = _device_matrix(D, n)
    out = torch.empty((3, nr), dtype=torch.float64, device=t.device)
    stream = torch.cuda.current_stream(t.device).cuda_stream
    L.check(L.lib().sbamd_rows_reduce(runtime.context(t.device.index), kind, nr, nc, t.data_ptr(),
                                      _dtype(t), out[0].data_ptr(), out[1].data_ptr(),
                                      out[2].data_ptr(), stream))
    entries = nc - 1 if kind == L.OUT_DISTMAT else nc
    s, lo, hi = out[0], out[1], out[2]
    mean = s / entries if entries > 0 else torch.full_like(s, float("nan"))
    res = (s, mean, lo, hi)
    if host:
        torch.cuda.synchronize(t.device)
        return tuple(x.cpu().numpy() for x in res)
    return res


def row_means(D, n=None):
    """Mean of every row (over j != i for distance matrices)."""
    return row_stats(D, n)[1]


def row_min(D, n=None):
    return row_stats(D, n)[2]


def row_max(D, n=None):
    return row_stats(D, n)[3]


def knn(D, k: int, n=None):
    """The k nearest columns of every row: (values, indices), each
    n_rows x k, ascending by (value, column); distance matrices exclude the
    diagonal. Rows with fewer than k entries are padded with inf / -1."""
    import torch

    if not 1 <= int(k) <= 32:
        raise ValueError("k must be in [1, 32]")
    t, kind, nr, nc, host = _device_matrix(D, n)
    vals = torch.empty((nr, k), dtype=t.dtype, device=t.device)
    idx = torch.empty((nr, k), dtype=torch.int64, device=t.device)
    stream = torch.cuda.current_stream(t.device).cuda_stream
    L.check(L.lib().sbamd_rows_knn(runtime.context(t.device.index), kind, nr, nc, t.data_ptr(),
                                   _dtype(t), int(k), vals.data_ptr(), idx.data_ptr(), stream))
    if host:
        torch.cuda.synchronize(t.device)
        return vals.cpu().numpy(), idx.cpu().numpy()
    return vals, idx
