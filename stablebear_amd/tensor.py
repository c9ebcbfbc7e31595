"""Tensors of PCFs and of floats (the part of the reference's tensor library
that the pairwise path consumes: shaped containers, strided / reversed /
multi-dimensional views, element assignment)."""

from __future__ import annotations

import numpy as np

from .pcf import Pcf
from .typing import (NP_TO_SCALAR, PCF_TO_NP, SCALAR_TO_NP, float32, float64, pcf32, pcf32i,
                     pcf64, pcf64i)

_ZERO = {dt: Pcf(np.zeros((1, 2), dtype=PCF_TO_NP[dt])) for dt in (pcf32, pcf64, pcf32i, pcf64i)}


class Tensor:
    """Common base of the tensor wrappers."""

    @property
    def shape(self):
        return tuple(self._a.shape)

    @property
    def ndim(self):
        return self._a.ndim

    def __len__(self):
        return self._a.shape[0]

    @property
    def size(self):
        return int(self._a.size)


class PcfTensor(Tensor):
    """N-d tensor of PCFs with a single PCF dtype; indexing returns a Pcf or a view."""

    def __init__(self, data=None, dtype=None):
        if isinstance(data, PcfTensor):
            self._a = data._a
            self.dtype = data.dtype
            return
        if isinstance(data, np.ndarray) and data.dtype == object:
            arr = data
        else:
            items = [] if data is None else data
            arr = np.empty(len(items), dtype=object) if not _is_nested(items) else None
            if arr is None:
                arr = _nested_to_object(items)
            else:
                for k, x in enumerate(items):
                    arr[k] = x if isinstance(x, Pcf) else Pcf(x)
        flat = [x for x in arr.ravel()]
        if dtype is None:
            dtype = flat[0].dtype if flat else pcf64
        for x in flat:
            if not isinstance(x, Pcf):
                raise TypeError("PcfTensor elements must be Pcf objects")
            if x.dtype is not dtype:
                raise TypeError(f"All PCFs must have the same dtype (got {x.dtype} and {dtype}).")
        self._a = arr
        self.dtype = dtype

    @classmethod
    def _wrap(cls, arr, dtype):
        t = cls.__new__(cls)
        t._a = arr
        t.dtype = dtype
        return t

    def __getitem__(self, idx):
        r = self._a[idx]
        if isinstance(r, np.ndarray):
            return PcfTensor._wrap(r, self.dtype)
        return r

    def __setitem__(self, idx, value):
        if isinstance(value, Pcf):
            if value.dtype is not self.dtype:
                raise TypeError(f"Cannot assign a {value.dtype} PCF into a {self.dtype} tensor")
            view = self._a[idx]
            if isinstance(view, np.ndarray):
                if not np.shares_memory(view, self._a) and view.size:
                    raise IndexError("PCF assignment needs basic (view) indexing")
                for j in np.ndindex(*view.shape):
                    view[j] = value
            else:
                self._a[idx] = value
        elif isinstance(value, PcfTensor):
            if value.dtype is not self.dtype:
                raise TypeError("dtype mismatch in PcfTensor assignment")
            self._a[idx] = value._a
        else:
            raise TypeError("PcfTensor elements must be Pcf objects")

    def flat_pcfs(self):
        """PCFs in row-major (C) order of this view."""
        return list(self._a.ravel(order="C"))

    def __iter__(self):
        for k in range(len(self)):
            yield self[k]

    def __repr__(self):
        return f"PcfTensor(shape={self.shape}, dtype={self.dtype})"


class IntPcfTensor(PcfTensor):
    """Tensor of integer PCFs (accepted as input, rejected by the pairwise ops)."""


class FloatTensor(Tensor):
    """N-d float tensor (results of cdist)."""

    def __init__(self, data, dtype=None):
        if isinstance(data, FloatTensor):
            data = data._a
        a = np.asarray(data)
        if dtype is not None:
            a = a.astype(SCALAR_TO_NP.get(dtype, dtype))
        elif a.dtype not in (np.float32, np.float64):
            a = a.astype(np.float64)
        self._a = a

    @property
    def dtype(self):
        return NP_TO_SCALAR[self._a.dtype]

    def __getitem__(self, idx):
        r = self._a[idx]
        if isinstance(r, np.ndarray):
            return FloatTensor(r)
        return float(r)

    def __setitem__(self, idx, value):
        self._a[idx] = value

    def __array__(self, dtype=None, copy=None):
        return self._a if dtype is None else self._a.astype(dtype)

    def to_numpy(self):
        return np.array(self._a)

    def __repr__(self):
        return f"FloatTensor(shape={self.shape}, dtype={self.dtype})"


def _is_nested(items):
    return bool(items) and isinstance(items[0], (list, tuple)) and items[0] and isinstance(
        items[0][0], Pcf)


def _nested_to_object(items):
    shape = []
    x = items
    while isinstance(x, (list, tuple)):
        shape.append(len(x))
        x = x[0] if x else None
    arr = np.empty(tuple(shape), dtype=object)
    for idx in np.ndindex(*shape):
        v = items
        for k in idx:
            v = v[k]
        arr[idx] = v
    return arr


def _is_scalar_index(idx, ndim):
    if isinstance(idx, (int, np.integer)):
        return ndim == 1
    if isinstance(idx, tuple):
        return len(idx) == ndim and all(isinstance(k, (int, np.integer)) for k in idx)
    return False


def zeros(shape, dtype=pcf32):
    """Tensor of zero PCFs (``[(0, 0)]``); float dtypes give a FloatTensor.

    The default dtype is pcf32, as in the reference (``tensor_create.py:82``)."""
    if isinstance(shape, int):
        shape = (shape,)
    shape = tuple(shape)
    if dtype in (float32, float64):
        return FloatTensor(np.zeros(shape, dtype=SCALAR_TO_NP[dtype]))
    if dtype not in _ZERO:
        raise TypeError(f"Unsupported dtype {dtype}")
    arr = np.empty(shape, dtype=object)
    z = _ZERO[dtype]
    for idx in np.ndindex(*shape):
        arr[idx] = z
    cls = IntPcfTensor if dtype in (pcf32i, pcf64i) else PcfTensor
    return cls._wrap(arr, dtype)


PcfContainerLike = PcfTensor | list | Pcf


def to_pcf_tensor(fs) -> PcfTensor:
    """Coerce a PcfTensor, a list of Pcf (or nested lists) or a single Pcf."""
    if isinstance(fs, PcfTensor):
        return fs
    if isinstance(fs, Pcf):
        return PcfTensor([fs])
    if isinstance(fs, (list, tuple)):
        if not fs:
            return PcfTensor._wrap(np.empty((0,), dtype=object), pcf64)
        first = fs[0] if not isinstance(fs[0], (list, tuple)) else fs[0][0]
        if isinstance(first, Pcf):
            t = PcfTensor(list(fs))
            if t.dtype in (pcf32i, pcf64i):
                return IntPcfTensor._wrap(t._a, t.dtype)
            return t
    raise TypeError("Input should be convertible to a PcfTensor.")
