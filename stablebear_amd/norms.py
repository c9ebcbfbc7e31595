"""``lp_norm`` (``stablebear/norms.py:10-53``): the L_p norm of every PCF of a
tensor, ``(int_0^MAX |f(t)|^p dt)^(1/p)``, one GPU lane per PCF. Backends
``Norms_f32_f32`` / ``Norms_f64_f64`` mirror
``src/python/functional/py_norms.cpp:24-68``: static ``lpnorm_l1`` /
``lpnorm_lp`` returning ``(task, out)``; ``out`` has the input's shape and
value type."""

from __future__ import annotations

import numpy as np

from . import runtime, system
from .async_task import _run_task
from .distance import _precision_for, _resolve
from .tensor import FloatTensor
from .typing import pcf32, pcf64


class _NormsBackend:
    pcf_dtype = pcf64
    np_dtype = np.float64

    @classmethod
    def _norms(cls, fs, p: float, precision=None):
        out = np.zeros(tuple(fs.shape), dtype=cls.np_dtype)
        if out.size == 0:
            return runtime.done_task("Computing norms.", 0), out
        system._verbose("Norm computation on HIP device")
        csr = runtime.csr_from_pcfs(fs.flat_pcfs(), cls.pcf_dtype)
        runtime.lp_norms(csr, p, _precision_for(cls.pcf_dtype, precision), out.reshape(-1))
        return runtime.done_task("Computing norms.", out.size), out

    @classmethod
    def lpnorm_l1(cls, fs, precision=None):
        return cls._norms(fs, 1.0, precision)

    @classmethod
    def lpnorm_lp(cls, fs, p, precision=None):
        return cls._norms(fs, float(p), precision)


class Norms_f32_f32(_NormsBackend):
    pcf_dtype = pcf32
    np_dtype = np.float32


class Norms_f64_f64(_NormsBackend):
    pcf_dtype = pcf64
    np_dtype = np.float64


_NORMS_BACKEND_MAP = {pcf32: Norms_f32_f32, pcf64: Norms_f64_f64}


def lp_norm(fs, p=1, verbose=False, precision=None) -> FloatTensor:
    """L_p norm of each PCF in ``fs``; a tensor of the same shape as ``fs``.

    ``precision`` ("fp64" or "fp32c", float64 PCFs only) overrides
    ``system.set_precision``; float32 PCFs always use the compensated FP32 form.
    """
    if p < 1:
        raise ValueError("p must be >= 1.")
    backend, X = _resolve(_NORMS_BACKEND_MAP, fs)
    if p == 1:
        task, out = backend.lpnorm_l1(X, precision)
    else:
        task, out = backend.lpnorm_lp(X, float(p), precision)
    _run_task(lambda: task, verbose=verbose)
    return FloatTensor(out)
