"""``l2_kernel`` (``stablebear/inner_product.py:13-50``): the L2 Gram matrix
``K_ij = int f_i f_j dt`` in compressed symmetric storage, computed on the GPU.
Backends ``InnerProduct_f32_f32`` / ``InnerProduct_f64_f64`` mirror
``src/python/functional/py_inner_product.cpp:24-90``."""

from __future__ import annotations

import numpy as np

from . import _lib as L
from . import runtime, system
from .async_task import _run_task
from .distance import _precision_for, _resolve
from .symmetric_matrix import SymmetricMatrix
from .typing import float32, float64, pcf32, pcf64


class _InnerProductBackend:
    pcf_dtype = pcf64
    np_dtype = np.float64

    @classmethod
    def l2(cls, fs, precision=None):
        n = fs.shape[0]
        sm = SymmetricMatrix(n, dtype=float64 if cls.np_dtype == np.float64 else float32)
        total = n * (n + 1) // 2
        if total == 0:
            return runtime.done_task("Computing kernel matrix.", 0), sm
        system._verbose("Integral computation on HIP device")
        csr = runtime.csr_from_pcfs(fs.flat_pcfs(), cls.pcf_dtype)
        task = runtime.submit(csr, None, L.OP_DOT, 1.0, L.OUT_SYMMAT,
                              _precision_for(cls.pcf_dtype, precision), sm._data,
                              "Computing kernel matrix.", total)
        return task, sm


class InnerProduct_f32_f32(_InnerProductBackend):
    pcf_dtype = pcf32
    np_dtype = np.float32


class InnerProduct_f64_f64(_InnerProductBackend):
    pcf_dtype = pcf64
    np_dtype = np.float64


_INNER_PRODUCT_BACKEND_MAP = {pcf32: InnerProduct_f32_f32, pcf64: InnerProduct_f64_f64}


def l2_kernel(fs, verbose=False, precision=None) -> SymmetricMatrix:
    """Pairwise L2 kernel matrix ``K_ij = <f_i, f_j>`` of a 1-D tensor of PCFs."""
    backend, X = _resolve(_INNER_PRODUCT_BACKEND_MAP, fs)
    if len(X.shape) != 1:
        raise ValueError("1d tensor expected.")
    task, sm = backend.l2(X, precision)
    _run_task(lambda: task, verbose=verbose)
    return SymmetricMatrix(sm)
