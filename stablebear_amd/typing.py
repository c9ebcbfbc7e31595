"""Element dtypes (the reference's ``stablebear/typing.py`` tokens that the
pairwise path uses)."""

import numpy as np


class dtype:
    """Element type token; each dtype is a singleton (e.g. ``pcf64``)."""

    def __init__(self, name: str, doc: str = ""):
        self._name = name
        self.__doc__ = doc

    @property
    def name(self) -> str:
        return self._name

    def __repr__(self):
        return f"stablebear_amd.{self._name}"

    def __str__(self):
        return self._name

    @property
    def __name__(self):
        return self._name

    def __hash__(self):
        return id(self)

    def __reduce__(self):
        return self._name


float32 = dtype("float32", "32-bit floating-point scalar dtype.")
float64 = dtype("float64", "64-bit floating-point scalar dtype.")
int32 = dtype("int32", "32-bit integer scalar dtype.")
int64 = dtype("int64", "64-bit integer scalar dtype.")

pcf32 = dtype("pcf32", "32-bit piecewise constant function dtype.")
pcf64 = dtype("pcf64", "64-bit piecewise constant function dtype.")
pcf32i = dtype("pcf32i", "32-bit integer piecewise constant function dtype.")
pcf64i = dtype("pcf64i", "64-bit integer piecewise constant function dtype.")

PCF_TO_NP = {pcf32: np.float32, pcf64: np.float64, pcf32i: np.int32, pcf64i: np.int64}
NP_TO_PCF = {np.dtype(np.float32): pcf32, np.dtype(np.float64): pcf64,
             np.dtype(np.int32): pcf32i, np.dtype(np.int64): pcf64i}
NP_TO_SCALAR = {np.dtype(np.float32): float32, np.dtype(np.float64): float64,
                np.dtype(np.int32): int32, np.dtype(np.int64): int64}
SCALAR_TO_NP = {float32: np.float32, float64: np.float64, int32: np.int32, int64: np.int64}
