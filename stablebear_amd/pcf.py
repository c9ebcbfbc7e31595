"""Piecewise constant functions (the reference's ``stablebear.Pcf``).

A PCF is a sequence of (time, value) rows ``(t_0, v_0), ..., (t_{n-1},
v_{n-1})`` with ``t_0 = 0`` and non-decreasing times; it takes ``v_k`` on
``[t_k, t_{k+1})`` and ``v_{n-1}`` on ``[t_{n-1}, inf)``. Construction rules
follow ``src/python/pypcf_support.hpp:36-89`` (empty input is the zero PCF
``[(0, 0)]``; an (m, 2) shape is required; unordered times or a first time
other than 0 raise ``ValueError``).
"""

from __future__ import annotations

import numpy as np

from .typing import NP_TO_PCF, NP_TO_SCALAR, PCF_TO_NP, pcf32, pcf32i, pcf64, pcf64i

_ALLOWED = (np.dtype(np.float32), np.dtype(np.float64), np.dtype(np.int32), np.dtype(np.int64))


def _validated(arr: np.ndarray) -> np.ndarray:
    if arr.size == 0:
        return np.zeros((1, 2), dtype=arr.dtype)
    if arr.ndim != 2:
        raise RuntimeError("Input array should have two dimensions (time + value).")
    if arr.shape[1] != 2:
        raise RuntimeError(
            f"Input array should be Nx2 (supplied shape is ({arr.shape[0]}, {arr.shape[1]})).")
    t = arr[:, 0]
    if np.any(t[1:] < t[:-1]):
        raise ValueError("Breakpoints must be given in non-decreasing time order.")
    if t[0] != 0:
        raise ValueError(f"The first breakpoint must have time t=0 (got t={float(t[0]):f}).")
    return arr


class Pcf:
    """A piecewise constant function backed by an (m, 2) numpy array."""

    def __init__(self, arr, dtype=None):
        if isinstance(arr, Pcf):
            self._arr = arr._arr
            return
        if isinstance(arr, list):
            np_dtype = PCF_TO_NP.get(dtype, dtype) if dtype is not None else np.float64
            a = np.array(arr, dtype=np_dtype)
        elif isinstance(arr, np.ndarray):
            a = arr
            if dtype is not None:
                np_dtype = np.dtype(PCF_TO_NP.get(dtype, dtype))
                if a.dtype != np_dtype:
                    a = a.astype(np_dtype)
        else:
            raise ValueError(
                f"Tried to create PCF from unsupported input data of type {type(arr)}.")
        if a.dtype not in _ALLOWED:
            raise ValueError("Unsupported array type (must be np.float32/64 or np.int32/64)")
        a = np.ascontiguousarray(a)
        self._arr = _validated(a).copy()
        self._arr.setflags(write=False)

    # --- dtype information -------------------------------------------------
    @property
    def dtype(self):
        """The PCF dtype token (``pcf32``, ``pcf64``, ``pcf32i`` or ``pcf64i``)."""
        return NP_TO_PCF[self._arr.dtype]

    @property
    def ttype(self):
        return NP_TO_SCALAR[self._arr.dtype]

    @property
    def vtype(self):
        return NP_TO_SCALAR[self._arr.dtype]

    @property
    def size(self) -> int:
        """Number of breakpoints."""
        return int(self._arr.shape[0])

    def __len__(self):
        return self.size

    # --- conversions -------------------------------------------------------
    def to_numpy(self) -> np.ndarray:
        """The (m, 2) array of (time, value) rows."""
        return self._arr.copy()

    def __array__(self, dtype=None, copy=None):
        a = self._arr.copy()
        return a if dtype is None else a.astype(dtype)

    def astype(self, dtype):
        """Copy cast to ``pcf32``, ``pcf64``, ``pcf32i``, ``pcf64i`` or a numpy dtype."""
        np_dtype = PCF_TO_NP.get(dtype, dtype)
        return Pcf(self._arr.astype(np_dtype))

    @property
    def times(self) -> np.ndarray:
        return self._arr[:, 0]

    @property
    def values(self) -> np.ndarray:
        return self._arr[:, 1]

    def __call__(self, t):
        """Evaluate at time(s) ``t`` (value of the last breakpoint at or before t)."""
        ts = self._arr[:, 0]
        idx = np.searchsorted(ts, np.asarray(t), side="right") - 1
        idx = np.clip(idx, 0, len(ts) - 1)
        out = self._arr[:, 1][idx]
        return out.item() if np.ndim(out) == 0 else out

    def __eq__(self, other):
        if not isinstance(other, Pcf):
            return NotImplemented
        return self._arr.dtype == other._arr.dtype and np.array_equal(self._arr, other._arr)

    def __hash__(self):
        return id(self)

    def __repr__(self):
        return f"Pcf(size={self.size}, dtype={self.dtype})"

    def __reduce__(self):
        return (Pcf, (self._arr.copy(),))


def _has_matching_types(f: Pcf, g: Pcf) -> bool:
    return f._arr.dtype == g._arr.dtype


class Rectangle:
    """One rectangle of the walk: [left, right) with values (f_value, g_value)."""

    __slots__ = ("left", "right", "f_value", "g_value")

    def __init__(self, left, right, f_value, g_value):
        self.left = float(left)
        self.right = float(right)
        self.f_value = float(f_value)
        self.g_value = float(g_value)

    def __repr__(self):
        return (f"Rectangle(l={self.left}, r={self.right}, fv={self.f_value}, "
                f"gv={self.g_value})")


def iterate_rectangles(f: Pcf, g: Pcf, a=0.0, b=float("inf")):
    """Rectangles of the merge walk of f and g over [a, b), in time order.

    Mirrors ``stablebear.iterate_rectangles`` (``stablebear/functional/pcf.py:347-374``;
    walk: ``include/sbear/algorithms/functional/iterate_rectangles.hpp:37-104``).
    Runs on the GPU.
    """
    if not isinstance(f, Pcf) or not isinstance(g, Pcf):
        raise TypeError("Both f and g must be Pcf objects.")
    if not _has_matching_types(f, g):
        raise TypeError("f and g must have the same dtype.")
    if f.dtype not in (pcf32, pcf64):
        raise TypeError("iterate_rectangles is not supported for this PCF type.")
    from . import runtime

    return [Rectangle(*r) for r in runtime.rectangles(f, g, a, b)]


__all__ = ["Pcf", "Rectangle", "iterate_rectangles", "pcf32", "pcf64", "pcf32i", "pcf64i"]
