"""Saving and loading in the reference's binary format (``stablebear.save`` /
``stablebear.load``, stablebear/io.py; format: include/sbear/io.hpp).

Files written here are read by the reference and the other way round: a
header (``"\\1MPCF"``, endianness, format version 2, version / date /
platform strings), one tensor or one object, elements in row-major order. The
byte-level parsing and writing is native (``sbamd_io_*`` in the C-ABI);
``load_device_set`` goes from a file of PCFs straight to a device set,
copying the file's payload to the GPU once and packing it there.

Supported: tensors of pcf32 / pcf64 / pcf32i / pcf64i (``PcfTensor``,
``IntPcfTensor``), float32 / float64 (``FloatTensor``), int32 / int64 /
uint32 / uint64 / bool (numpy arrays), symmetric and distance matrices (as
numpy object arrays of ``SymmetricMatrix`` / ``DistanceMatrix``); objects
``Pcf``, ``DistanceMatrix`` and ``SymmetricMatrix``. Point clouds and
barcodes (persistence, outside the pairwise path) raise ``TypeError``.
"""

from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _lib as L
from .distance_matrix import DistanceMatrix
from .pcf import Pcf
from .symmetric_matrix import SymmetricMatrix
from .tensor import FloatTensor, IntPcfTensor, PcfTensor
from .typing import pcf32, pcf32i, pcf64, pcf64i

# element type codes (include/sbear/io/tensor_io.hpp:92-121)
_PCF_CODE = {pcf32: (100, 32), pcf64: (100, 64), pcf32i: (101, 32), pcf64i: (101, 64)}
_PCF_FROM_CODE = {v: k for k, v in _PCF_CODE.items()}
_PCF_NP = {pcf32: np.float32, pcf64: np.float64, pcf32i: np.int32, pcf64i: np.int64}
_NUM_CODE = {np.dtype(np.float32): (1, 32), np.dtype(np.float64): (1, 64),
             np.dtype(np.int32): (2, 32), np.dtype(np.int64): (2, 64),
             np.dtype(np.uint32): (3, 32), np.dtype(np.uint64): (3, 64),
             np.dtype(np.bool_): (4, 8)}
_NUM_FROM_CODE = {v: k for k, v in _NUM_CODE.items()}
_MATRIX = {1120: DistanceMatrix, 1100: SymmetricMatrix}
_UNSUPPORTED = {1000: "point clouds", 10000: "barcodes"}


def _header(kind, base, sub, shape=()) -> bytes:
    shp = (ctypes.c_uint64 * max(1, len(shape)))(*shape)
    size = ctypes.c_uint64()
    L.check(L.lib().sbamd_io_header(kind, base, sub, len(shape), shp, None, 0, ctypes.byref(size)))
    buf = ctypes.create_string_buffer(size.value)
    L.check(L.lib().sbamd_io_header(kind, base, sub, len(shape), shp, buf, size.value,
                                    ctypes.byref(size)))
    return buf.raw


def _pcf_elements(pcfs, np_dt) -> bytes:
    """[u64 points][(t, v) x points] per PCF, in order."""
    if np_dt in (np.float32, np.float64):
        n = len(pcfs)
        lens = np.fromiter((p._arr.shape[0] for p in pcfs), dtype=np.int64, count=n)
        off = np.zeros(n + 1, dtype=np.int64)
        np.cumsum(lens, out=off[1:])
        pts = (np.concatenate([p._arr for p in pcfs], axis=0).astype(np_dt, copy=False)
               if n else np.zeros((0, 2), dtype=np_dt))
        t = np.ascontiguousarray(pts[:, 0])
        v = np.ascontiguousarray(pts[:, 1])
        dt = L.F64 if np_dt == np.float64 else L.F32
        size = ctypes.c_uint64()
        args = (n, off.ctypes.data_as(L.c_i64p), t.ctypes.data if t.size else None,
                v.ctypes.data if v.size else None, dt)
        L.check(L.lib().sbamd_io_write_pcfs(*args, None, 0, ctypes.byref(size)))
        out = np.empty(size.value, dtype=np.uint8)
        L.check(L.lib().sbamd_io_write_pcfs(*args, out.ctypes.data if out.size else None, size.value,
                                            ctypes.byref(size)))
        return out.tobytes()
    parts = []
    for p in pcfs:
        a = np.ascontiguousarray(p._arr, dtype=np_dt)
        parts.append(np.uint64(a.shape[0]).tobytes())
        parts.append(a.tobytes())
    return b"".join(parts)


def _matrix_element(m) -> bytes:
    return np.uint64(m.size).tobytes() + np.ascontiguousarray(m.storage).tobytes()


def _matrix_code(m):
    base = 1120 if isinstance(m, DistanceMatrix) else 1100
    return base, 64 if m.storage.dtype == np.float64 else 32


def dumps(item) -> bytes:
    """The file image of a tensor or object, as ``save`` writes it."""
    if isinstance(item, Pcf):
        base, sub = _PCF_CODE[item.dtype]
        return (_header(L.IO_OBJECT, base, sub)
                + _pcf_elements([item], _PCF_NP[item.dtype]))
    if isinstance(item, (DistanceMatrix, SymmetricMatrix)):
        base, sub = _matrix_code(item)
        return _header(L.IO_OBJECT, base, sub) + _matrix_element(item)
    if isinstance(item, PcfTensor):
        base, sub = _PCF_CODE[item.dtype]
        flat = list(item._a.ravel(order="C"))
        return (_header(L.IO_TENSOR, base, sub, item.shape)
                + _pcf_elements(flat, _PCF_NP[item.dtype]))
    if isinstance(item, FloatTensor):
        item = item.to_numpy()
    if isinstance(item, np.ndarray) and item.dtype == object:
        flat = list(item.ravel(order="C"))
        if flat and all(isinstance(m, (DistanceMatrix, SymmetricMatrix)) for m in flat):
            codes = {_matrix_code(m) for m in flat}
            if len(codes) != 1:
                raise TypeError("all matrices of a tensor must be of one type and dtype")
            base, sub = codes.pop()
            return (_header(L.IO_TENSOR, base, sub, item.shape)
                    + b"".join(_matrix_element(m) for m in flat))
        raise TypeError("object arrays must hold DistanceMatrix / SymmetricMatrix elements")
    if isinstance(item, np.ndarray):
        code = _NUM_CODE.get(item.dtype)
        if code is None:
            raise TypeError(f"Unsupported tensor dtype {item.dtype}")
        return _header(L.IO_TENSOR, *code, item.shape) + np.ascontiguousarray(item).tobytes()
    raise TypeError(f"Unsupported item type {type(item).__name__}")


def dumps_csr(offsets, t, v) -> np.ndarray:
    """The file image (uint8 array) of a 1-D pcf32 / pcf64 tensor given in CSR
    form (offsets n+1, times, values), written natively without Pcf objects."""
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    t = np.ascontiguousarray(t)
    np_dt = np.float64 if t.dtype == np.float64 else np.float32
    t = t.astype(np_dt, copy=False)
    v = np.ascontiguousarray(v, dtype=np_dt)
    n = len(offsets) - 1
    head = np.frombuffer(_header(L.IO_TENSOR, 100, 64 if np_dt == np.float64 else 32, (n,)),
                         dtype=np.uint8)
    dt = L.F64 if np_dt == np.float64 else L.F32
    size = ctypes.c_uint64()
    args = (n, offsets.ctypes.data_as(L.c_i64p), t.ctypes.data if t.size else None,
            v.ctypes.data if v.size else None, dt)
    L.check(L.lib().sbamd_io_write_pcfs(*args, None, 0, ctypes.byref(size)))
    out = np.empty(head.size + size.value, dtype=np.uint8)
    out[: head.size] = head
    body = out[head.size:]
    L.check(L.lib().sbamd_io_write_pcfs(*args, body.ctypes.data if body.size else None,
                                        size.value, ctypes.byref(size)))
    return out


def load_csr(file):
    """(offsets, t, v) numpy arrays of a file holding a tensor or object of
    pcf32 / pcf64 (elements in row-major order), without Pcf objects."""
    img = _image(file)
    info = _info(img)
    if info.base_format != 100:
        raise TypeError("the file does not hold PCFs with floating-point times and values")
    counts, starts = _index(img, info)
    np_dt = np.float64 if info.sub_format == 64 else np.float32
    offsets = np.zeros(len(counts) + 1, dtype=np.int64)
    np.cumsum(counts.astype(np.int64), out=offsets[1:])
    pb = 2 * np.dtype(np_dt).itemsize
    if offsets[-1]:
        pts = np.concatenate([img[int(s): int(s) + int(c) * pb] for c, s in zip(counts, starts)])
        pts = pts.view(np_dt).reshape(-1, 2)
    else:
        pts = np.zeros((0, 2), dtype=np_dt)
    return offsets, np.ascontiguousarray(pts[:, 0]), np.ascontiguousarray(pts[:, 1])


def save(item, file) -> None:
    """Save a tensor or object (``Pcf``, ``DistanceMatrix``, ``SymmetricMatrix``)
    in the reference's binary format. ``file`` is a path or a binary file."""
    data = dumps(item)
    if isinstance(file, (str, os.PathLike)):
        with open(file, "wb") as fh:
            fh.write(data)
    else:
        file.write(data)


def _image(file) -> np.ndarray:
    """The file's bytes as a uint8 array (memory-mapped for paths; images
    already in memory are used as they are)."""
    if isinstance(file, np.ndarray):
        return file.view(np.uint8).ravel()
    if isinstance(file, (bytes, bytearray, memoryview)):
        return np.frombuffer(file, dtype=np.uint8)
    if isinstance(file, (str, os.PathLike)):
        if os.path.getsize(file) == 0:
            return np.zeros(0, dtype=np.uint8)
        return np.memmap(file, dtype=np.uint8, mode="r")
    return np.frombuffer(file.read(), dtype=np.uint8)


def _info(img) -> L.IoInfo:
    info = L.IoInfo()
    L.check(L.lib().sbamd_io_parse(img.ctypes.data if img.size else None, img.size,
                                   ctypes.byref(info)))
    return info


def _index(img, info):
    n = info.elements
    counts = np.zeros(n, dtype=np.uint64)
    starts = np.zeros(n, dtype=np.uint64)
    end = ctypes.c_uint64()
    L.check(L.lib().sbamd_io_index(img.ctypes.data if img.size else None, img.size,
                                   ctypes.byref(info), counts.ctypes.data_as(L.c_u64p),
                                   starts.ctypes.data_as(L.c_u64p), ctypes.byref(end)))
    return counts, starts


def loads(img):
    """The tensor or object of a file image (bytes or a uint8 array)."""
    if not isinstance(img, np.ndarray):
        img = np.frombuffer(bytes(img), dtype=np.uint8)
    info = _info(img)
    base, sub = info.base_format, info.sub_format
    if base in _UNSUPPORTED:
        raise TypeError(f"{_UNSUPPORTED[base]} are not supported by this build")
    shape = tuple(info.shape[k] for k in range(info.ndim))
    if info.kind == L.IO_TENSOR and base in (1, 2, 3, 4):
        dt = _NUM_FROM_CODE[(base, sub)]
        a = np.frombuffer(img, dtype=dt, count=info.elements, offset=info.payload).copy()
        a = a.reshape(shape) if info.elements else np.zeros(shape if info.ndim else (0,), dt)
        return FloatTensor(a) if base == 1 else a
    counts, starts = _index(img, info)
    if base in (100, 101):
        pdt = _PCF_FROM_CODE[(base, sub)]
        npdt = _PCF_NP[pdt]
        pcfs = [Pcf(np.frombuffer(img, dtype=npdt, count=2 * int(c), offset=int(s)).reshape(-1, 2)
                    .copy(), dtype=pdt)
                for c, s in zip(counts, starts)]
        if info.kind == L.IO_OBJECT:
            return pcfs[0]
        arr = np.empty(len(pcfs), dtype=object)
        for k, p in enumerate(pcfs):
            arr[k] = p
        arr = arr.reshape(shape) if info.ndim else arr
        cls = IntPcfTensor if base == 101 else PcfTensor
        return cls._wrap(arr, pdt)
    mcls = _MATRIX[base]
    vdt = np.float64 if sub == 64 else np.float32
    mats = []
    for c, s in zip(counts, starts):
        n = int(c)
        cnt = n * (n - 1) // 2 if base == 1120 else n * (n + 1) // 2
        data = np.frombuffer(img, dtype=vdt, count=cnt, offset=int(s)).copy()
        mats.append(mcls._from_storage(n, data))
    if info.kind == L.IO_OBJECT:
        return mats[0]
    arr = np.empty(len(mats), dtype=object)
    for k, m in enumerate(mats):
        arr[k] = m
    return arr.reshape(shape) if info.ndim else arr


def load(file):
    """Load a tensor or object saved in the reference's binary format; the
    result has the saved type, dtype and shape. ``file`` is a path or a
    binary file."""
    return loads(_image(file))


def load_device_set(file, precision: str = "fp64", device: int | None = None):
    """A device set (``runtime.DeviceSet``) straight from a file (path, file
    object, or an image in memory) holding a tensor (any shape, row-major
    order) or an object of pcf32 / pcf64: the payload is copied to the GPU
    once as it lies and validated and packed there, with no per-PCF host
    objects."""
    from . import runtime

    img = _image(file)
    if precision not in runtime.PRECISIONS:
        raise ValueError(f"precision must be one of {sorted(runtime.PRECISIONS)}")
    ctx = runtime.context(device)
    h = ctypes.c_void_p()
    L.check(L.lib().sbamd_set_create_image(ctx, img.ctypes.data if img.size else None, img.size,
                                           runtime.PRECISIONS[precision], ctypes.byref(h)))
    return runtime.DeviceSet.from_handle(h, precision, device)
