"""``pdist`` / ``cdist`` / ``lp_distance`` (``stablebear/distance.py:17-154``).

The backend classes ``Distance_f32_f32`` / ``Distance_f64_f64`` mirror the
reference's pybind11 plug-in objects (``src/python/functional/py_distance.cpp:
40-238``): static methods that start an asynchronous job and return
``(task, result)``; the result container is filled when the task finishes.
All integrals run on the GPU through ``libstablebear_amd.so``.
"""

from __future__ import annotations

import numpy as np

from . import _lib as L
from . import runtime, system
from .async_task import _run_task
from .distance_matrix import DistanceMatrix
from .pcf import Pcf, _has_matching_types
from .tensor import FloatTensor, PcfTensor, to_pcf_tensor
from .typing import float32, float64, pcf32, pcf64


def _precision_for(dtype, precision=None):
    if dtype is pcf32:
        return L.PREC_FP32C
    prec = precision or system.get_precision()
    return runtime.PRECISIONS[prec]


class _DistanceBackend:
    pcf_dtype = pcf64
    np_dtype = np.float64

    @classmethod
    def _pdist(cls, fs: PcfTensor, op: int, p: float, precision=None):
        n = fs.shape[0]
        dm = DistanceMatrix(n, dtype=float64 if cls.np_dtype == np.float64 else float32)
        total = n * (n - 1) // 2
        if total == 0:
            return runtime.done_task("Computing distance matrix.", 0), dm
        system._verbose("Integral computation on HIP device")
        csr = runtime.csr_from_pcfs(fs.flat_pcfs(), cls.pcf_dtype)
        task = runtime.submit(csr, None, op, p, L.OUT_DISTMAT,
                              _precision_for(cls.pcf_dtype, precision), dm._data,
                              "Computing distance matrix.", total)
        return task, dm

    @classmethod
    def pdist_l1(cls, fs, precision=None):
        return cls._pdist(fs, L.OP_L1, 1.0, precision)

    @classmethod
    def pdist_lp(cls, fs, p, precision=None):
        return cls._pdist(fs, L.OP_LP, float(p), precision)

    @classmethod
    def _cdist(cls, X: PcfTensor, Y: PcfTensor, op: int, p: float, precision=None):
        shape = tuple(X.shape) + tuple(Y.shape)
        out = np.zeros(shape, dtype=cls.np_dtype)
        nx, ny = X.size, Y.size
        if nx == 0 or ny == 0:
            return runtime.done_task("Computing cross-distances.", 0), out
        system._verbose("Cross-distance computation on HIP device")
        cx = runtime.csr_from_pcfs(X.flat_pcfs(), cls.pcf_dtype)
        cy = runtime.csr_from_pcfs(Y.flat_pcfs(), cls.pcf_dtype)
        flat = out.reshape(-1)
        task = runtime.submit(cx, cy, op, p, L.OUT_DENSE, _precision_for(cls.pcf_dtype, precision),
                              flat, "Computing cross-distances.", nx * ny)
        return task, out

    @classmethod
    def cdist_l1(cls, X, Y, precision=None):
        return cls._cdist(X, Y, L.OP_L1, 1.0, precision)

    @classmethod
    def cdist_lp(cls, X, Y, p, precision=None):
        return cls._cdist(X, Y, L.OP_LP, float(p), precision)

    @classmethod
    def lp_distance_l1(cls, f: Pcf, g: Pcf):
        return runtime.lp_distance(f, g, 1.0)

    @classmethod
    def lp_distance_lp(cls, f: Pcf, g: Pcf, p):
        return runtime.lp_distance(f, g, float(p))


class Distance_f32_f32(_DistanceBackend):
    pcf_dtype = pcf32
    np_dtype = np.float32


class Distance_f64_f64(_DistanceBackend):
    pcf_dtype = pcf64
    np_dtype = np.float64


_DISTANCE_BACKEND_MAP = {pcf32: Distance_f32_f32, pcf64: Distance_f64_f64}


def _resolve(backend_map, *inputs):
    """Coerce inputs to PcfTensors and pick the backend (``base_tensor.py:485-505``)."""
    tensors = []
    for x in inputs:
        t = to_pcf_tensor(x)
        if t.dtype not in backend_map:
            raise TypeError(f"Operation not supported for tensors of this type (dtype {t.dtype})")
        tensors.append(t)
    dtypes = [t.dtype for t in tensors]
    if len(set(id(d) for d in dtypes)) > 1:
        names = ", ".join(d.__name__ for d in dtypes)
        raise TypeError(f"All inputs must have the same dtype (got {names}).")
    return (backend_map[dtypes[0]], *tensors)


def lp_distance(f: Pcf, g: Pcf, p=1) -> float:
    """L_p distance between two PCFs, ``(int |f - g|^p dt)^(1/p)``."""
    if p < 1:
        raise ValueError("p must be >= 1.")
    if not isinstance(f, Pcf) or not isinstance(g, Pcf):
        raise TypeError("Both f and g must be Pcf objects.")
    if not _has_matching_types(f, g):
        raise TypeError("f and g must have the same dtype.")
    backend = _DISTANCE_BACKEND_MAP.get(f.dtype)
    if backend is None:
        raise TypeError(f"lp_distance is not supported for this PCF type (vtype={f.vtype}).")
    if p == 1:
        return float(backend.lp_distance_l1(f, g))
    return float(backend.lp_distance_lp(f, g, float(p)))


def pdist(fs, p=1, verbose=False, precision=None) -> DistanceMatrix:
    """Pairwise L_p distance matrix of a 1-D tensor of PCFs (compressed storage).

    ``precision`` ("fp64" or "fp32c", float64 PCFs only) overrides
    ``system.set_precision``; float32 PCFs always use the compensated FP32 form.
    """
    if p < 1:
        raise ValueError("p must be >= 1.")
    backend, X = _resolve(_DISTANCE_BACKEND_MAP, fs)
    if len(X.shape) != 1:
        raise ValueError("1d tensor expected.")
    if p == 1:
        task, dm = backend.pdist_l1(X, precision)
    else:
        task, dm = backend.pdist_lp(X, float(p), precision)
    _run_task(lambda: task, verbose=verbose)
    return DistanceMatrix(dm)


def cdist(X, Y, p=1, verbose=False, precision=None) -> FloatTensor:
    """Cross distances, shape ``(*X.shape, *Y.shape)``."""
    if p < 1:
        raise ValueError("p must be >= 1.")
    backend, Xt, Yt = _resolve(_DISTANCE_BACKEND_MAP, X, Y)
    if p == 1:
        task, out = backend.cdist_l1(Xt, Yt, precision)
    else:
        task, out = backend.cdist_lp(Xt, Yt, float(p), precision)
    _run_task(lambda: task, verbose=verbose)
    return FloatTensor(out)
