"""Compressed symmetric matrix (``stablebear.SymmetricMatrix``).

Storage and index map follow ``include/sbear/symmetric_matrix.hpp:25-107``:
``n(n+1)/2`` entries of the lower triangle including the diagonal, entry
``(i, j)`` at ``max*(max+1)/2 + min``. No sign rule. Python behaviour follows
``stablebear/symmetric_matrix.py``.
"""

from __future__ import annotations

import numpy as np

from .typing import float32, float64

_NP = {float32: np.float32, float64: np.float64}


def storage_size(n: int) -> int:
    return n * (n + 1) // 2


def compressed_index(i: int, j: int) -> int:
    r, c = (i, j) if i >= j else (j, i)
    return r * (r + 1) // 2 + c


class SymmetricMatrix:
    """Symmetric n x n matrix in compressed lower-triangle storage."""

    def __init__(self, n_or_data, dtype=None):
        if isinstance(n_or_data, SymmetricMatrix):
            self._n = n_or_data._n
            self._data = n_or_data._data
        elif isinstance(n_or_data, (int, np.integer)):
            n = int(n_or_data)
            if n < 0:
                raise ValueError("matrix size must be non-negative")
            if dtype is None:
                dtype = float64
            if dtype not in _NP:
                raise TypeError(f"Unsupported dtype {dtype}; use float32 or float64")
            self._n = n
            self._data = np.zeros(storage_size(n), dtype=_NP[dtype])
        else:
            raise TypeError(f"Expected int or SymmetricMatrix; got {type(n_or_data)}")

    @classmethod
    def _from_storage(cls, n: int, data: np.ndarray) -> "SymmetricMatrix":
        m = cls.__new__(cls)
        m._n = int(n)
        if data.shape != (storage_size(n),):
            raise ValueError("storage has the wrong length")
        m._data = data
        return m

    @property
    def dtype(self):
        return float32 if self._data.dtype == np.float32 else float64

    @property
    def size(self) -> int:
        return self._n

    @property
    def storage_count(self) -> int:
        return storage_size(self._n)

    @property
    def storage(self) -> np.ndarray:
        return self._data

    def _resolve_ij(self, ij):
        i, j = ij
        n = self._n
        if i < 0:
            i += n
        if j < 0:
            j += n
        if not (0 <= i < n and 0 <= j < n):
            raise IndexError(f"index ({ij[0]}, {ij[1]}) is out of bounds for a {n}x{n} matrix")
        return int(i), int(j)

    def __getitem__(self, ij):
        i, j = self._resolve_ij(ij)
        return float(self._data[compressed_index(i, j)])

    def __setitem__(self, ij, value):
        i, j = self._resolve_ij(ij)
        self._data[compressed_index(i, j)] = value

    def to_dense(self) -> np.ndarray:
        n = self._n
        out = np.zeros((n, n), dtype=self._data.dtype)
        if n > 0:
            r, c = np.tril_indices(n, 0)
            out[r, c] = self._data
            out[c, r] = self._data
        return out

    def to_numpy(self) -> np.ndarray:
        return self.to_dense()

    def __array__(self, dtype=None, copy=None):
        arr = self.to_dense()
        return arr if dtype is None else arr.astype(dtype, copy=False)

    @classmethod
    def from_dense(cls, array):
        array = np.asarray(array)
        if array.dtype not in (np.float32, np.float64):
            raise TypeError(f"Unsupported dtype {array.dtype}")
        if array.ndim != 2:
            raise ValueError("Expected a 2-D array")
        n = array.shape[0]
        if array.shape[1] != n:
            raise ValueError("Expected a square array")
        r, c = np.tril_indices(n, 0)
        if np.any(array[r, c] != array[c, r]):
            raise ValueError("Matrix must be symmetric")
        return cls._from_storage(n, np.ascontiguousarray(array[r, c]))

    def allclose(self, other, atol=1e-8, rtol=1e-5) -> bool:
        other = other if isinstance(other, SymmetricMatrix) else SymmetricMatrix(other)
        if other._n != self._n:
            return False
        return bool(np.allclose(self._data, other._data, atol=atol, rtol=rtol))

    def __eq__(self, other):
        if not isinstance(other, SymmetricMatrix):
            return NotImplemented
        return self._n == other._n and np.array_equal(self._data, other._data)

    __hash__ = None

    def __reduce__(self):
        return (SymmetricMatrix._from_storage, (self._n, self._data.copy()))

    def __repr__(self):
        suffix = "_f32" if self._data.dtype == np.float32 else "_f64"
        return f"SymmetricMatrix{suffix}(size={self._n})"
