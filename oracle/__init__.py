"""CPU oracle for the pairwise PCF integration path.

TEST INFRASTRUCTURE ONLY. Imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` as the checker; the product
package ``stablebear_amd`` never imports it.

Three independent checkers live here:

* ``liboracle.so`` -- a plain-C restatement of the reference walk
  (``include/sbear/algorithms/functional/iterate_rectangles.hpp:23-104``),
  ``integrate()`` (``matrix_integrate.hpp:18-29``), the row-parallel
  ``CpuPairwiseIntegrationTask`` (``matrix_integrate.hpp:35-95``) and
  ``CdistTask`` (``src/python/functional/py_distance.cpp:101-136``), plus an
  analytic gradient.
* ``_ref/libsbref.so`` (optional) -- the reference's own header-only walk,
  compiled from the reference sources by ``oracle/Makefile``.
* ``exact_*`` -- exact rational arithmetic with ``fractions.Fraction`` on small
  inputs (order-independent, so it pins the accumulation to within one
  rounding of the exact value).
"""

from __future__ import annotations

import ctypes
import os
from fractions import Fraction

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))

OP_L1, OP_LP, OP_DOT = 0, 1, 2
KIND_PDIST, KIND_KERNEL, KIND_CDIST = 0, 1, 2

_f64p = ctypes.POINTER(ctypes.c_double)
_f32p = ctypes.POINTER(ctypes.c_float)
_i64p = ctypes.POINTER(ctypes.c_int64)


def _ptr(a, typ):
    if a is None:
        return None
    return a.ctypes.data_as(typ)


def _declare(lib, prefix):
    L = ctypes.c_long
    for suf, fp, ft in (("f64", _f64p, ctypes.c_double), ("f32", _f32p, ctypes.c_float)):
        fn = getattr(lib, f"{prefix}_rects_{suf}")
        fn.restype = L
        fn.argtypes = [fp, fp, L, fp, fp, L, ft, ft, fp, L]
        fn = getattr(lib, f"{prefix}_pairwise_{suf}")
        fn.restype = ctypes.c_int
        fn.argtypes = [ctypes.c_int, ft, ctypes.c_int, L, _i64p, fp, fp, L, _i64p, fp, fp, fp,
                       ctypes.c_int]
    fn = getattr(lib, f"{prefix}_pairwise_rows_f64")
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, L, _i64p, _f64p, _f64p, L, _i64p,
                   _f64p, _f64p, L, L, _f64p, ctypes.c_int]


def _load_oracle():
    path = os.path.join(_HERE, "liboracle.so")
    if not os.path.exists(path):
        raise RuntimeError(f"oracle library missing: {path} (run `make -C oracle`)")
    lib = ctypes.CDLL(path)
    _declare(lib, "orc")
    for suf, fp, ft in (("f64", _f64p, ctypes.c_double), ("f32", _f32p, ctypes.c_float)):
        fn = getattr(lib, f"orc_integrate_{suf}")
        fn.restype = ft
        fn.argtypes = [ctypes.c_int, ft, fp, fp, ctypes.c_long, fp, fp, ctypes.c_long, ft, ft]
        fn = getattr(lib, f"orc_finalize_{suf}")
        fn.restype = ft
        fn.argtypes = [ctypes.c_int, ft, ft]
    lib.orc_grad_f64.restype = ctypes.c_int
    lib.orc_grad_f64.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_long, _i64p, _f64p, _f64p,
                                 _f64p, _f64p, _f64p, _f64p, ctypes.c_int]
    return lib


_lib = _load_oracle()
_ref_lib = None


def ref_available() -> bool:
    """True when ``oracle/_ref/libsbref.so`` (the compiled reference walk) exists."""
    return os.path.exists(os.path.join(_HERE, "_ref", "libsbref.so"))


def _load_ref():
    global _ref_lib
    if _ref_lib is None:
        path = os.path.join(_HERE, "_ref", "libsbref.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/_ref/libsbref.so is not built")
        lib = ctypes.CDLL(path)
        _declare(lib, "ref")
        _ref_lib = lib
    return _ref_lib


def default_threads() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        return os.cpu_count() or 1


def _split(pcf, dtype):
    a = np.ascontiguousarray(np.asarray(pcf, dtype=dtype).reshape(-1, 2))
    return np.ascontiguousarray(a[:, 0]), np.ascontiguousarray(a[:, 1])


def _suffix(dtype):
    return "f64" if np.dtype(dtype) == np.float64 else "f32"


def rects(f, g, a=0.0, b=None, dtype=np.float64, use_ref=False):
    """Rectangles (left, right, fv, gv) of the reference walk over [a, b)."""
    dtype = np.dtype(dtype)
    if b is None:
        b = np.finfo(dtype).max
    ft, fv = _split(f, dtype)
    gt, gv = _split(g, dtype)
    lib = _load_ref() if use_ref else _lib
    prefix = "ref" if use_ref else "orc"
    suf = _suffix(dtype)
    fp = _f64p if suf == "f64" else _f32p
    fn = getattr(lib, f"{prefix}_rects_{suf}")
    cap = len(ft) + len(gt) + 4
    out = np.zeros((cap, 4), dtype=dtype)
    cnt = fn(_ptr(ft, fp), _ptr(fv, fp), len(ft), _ptr(gt, fp), _ptr(gv, fp), len(gt), a, b,
             _ptr(out, fp), cap)
    assert cnt <= cap
    return out[:cnt].copy()


def integrate(op, p, f, g, a=0.0, b=None, dtype=np.float64):
    """Raw integral sum (r - l) * op(fv, gv) before the finalizer."""
    dtype = np.dtype(dtype)
    if b is None:
        b = np.finfo(dtype).max
    ft, fv = _split(f, dtype)
    gt, gv = _split(g, dtype)
    suf = _suffix(dtype)
    fp = _f64p if suf == "f64" else _f32p
    fn = getattr(_lib, f"orc_integrate_{suf}")
    return fn(op, p, _ptr(ft, fp), _ptr(fv, fp), len(ft), _ptr(gt, fp), _ptr(gv, fp), len(gt), a, b)


def finalize(op, p, x, dtype=np.float64):
    fn = getattr(_lib, f"orc_finalize_{_suffix(dtype)}")
    return fn(op, p, x)


def distance(f, g, p=1.0, dtype=np.float64):
    """lp_distance(f, g, p) as the reference computes it (lp_distance.hpp:10-21)."""
    op = OP_L1 if p == 1 else OP_LP
    return finalize(op, p, integrate(op, p, f, g, dtype=dtype), dtype=dtype)


def _csr(off, t, v, dtype):
    off = np.ascontiguousarray(off, dtype=np.int64)
    t = np.ascontiguousarray(t, dtype=dtype)
    v = np.ascontiguousarray(v, dtype=dtype)
    return off, t, v


def out_count(kind, n, n2=0):
    if kind == KIND_PDIST:
        return n * (n - 1) // 2
    if kind == KIND_KERNEL:
        return n * (n + 1) // 2
    return n * n2


def pairwise(op, p, kind, off, t, v, off2=None, t2=None, v2=None, dtype=np.float64,
             nthreads=None, use_ref=False, rows=None):
    """Whole pdist / kernel / cdist matrix in the reference's compressed layout.

    ``rows=(r0, r1)`` computes only rows r0..r1-1 (other entries stay 0)."""
    dtype = np.dtype(dtype)
    off, t, v = _csr(off, t, v, dtype)
    n = len(off) - 1
    n2 = 0
    if kind == KIND_CDIST:
        off2, t2, v2 = _csr(off2, t2, v2, dtype)
        n2 = len(off2) - 1
    out = np.zeros(out_count(kind, n, n2), dtype=dtype)
    if nthreads is None:
        nthreads = default_threads()
    lib = _load_ref() if use_ref else _lib
    prefix = "ref" if use_ref else "orc"
    suf = _suffix(dtype)
    fp = _f64p if suf == "f64" else _f32p
    if rows is not None:
        if suf != "f64":
            raise ValueError("row sampling is implemented for float64 only")
        fn = getattr(lib, f"{prefix}_pairwise_rows_f64")
        rc = fn(op, p, kind, n, _ptr(off, _i64p), _ptr(t, fp), _ptr(v, fp), n2,
                _ptr(off2, _i64p), _ptr(t2, fp), _ptr(v2, fp), int(rows[0]), int(rows[1]),
                _ptr(out, fp), int(nthreads))
    else:
        fn = getattr(lib, f"{prefix}_pairwise_{suf}")
        rc = fn(op, p, kind, n, _ptr(off, _i64p), _ptr(t, fp), _ptr(v, fp), n2,
                _ptr(off2, _i64p), _ptr(t2, fp), _ptr(v2, fp), _ptr(out, fp), int(nthreads))
    if rc != 0:
        raise RuntimeError(f"oracle pairwise failed ({rc})")
    return out


def lp_norms(p, off, t, v, dtype=np.float64, use_ref=False):
    """L_p norm of each PCF of a CSR set over [0, MAX) (apply_functional.hpp:63-100)."""
    dtype = np.dtype(dtype)
    off, t, v = _csr(off, t, v, dtype)
    n = len(off) - 1
    out = np.zeros(n, dtype=dtype)
    lib = _load_ref() if use_ref else _lib
    suf = _suffix(dtype)
    fp = _f64p if suf == "f64" else _f32p
    ft = ctypes.c_double if suf == "f64" else ctypes.c_float
    fn = getattr(lib, f"{'ref' if use_ref else 'orc'}_lp_norms_{suf}")
    fn.restype = None
    fn.argtypes = [ft, ctypes.c_long, _i64p, fp, fp, fp]
    fn(p, n, _ptr(off, _i64p), _ptr(t, fp), _ptr(v, fp), _ptr(out, fp))
    return out


def grad(op, kind, off, t, v, G, D=None, nthreads=None):
    """Analytic (dL/dv, dL/dt) of L = sum_pairs G * X (X = L2 distance or inner product)."""
    off, t, v = _csr(off, t, v, np.float64)
    n = len(off) - 1
    G = np.ascontiguousarray(G, dtype=np.float64)
    Dp = None
    if D is not None:
        D = np.ascontiguousarray(D, dtype=np.float64)
        Dp = _ptr(D, _f64p)
    gv = np.zeros(len(v), dtype=np.float64)
    gt = np.zeros(len(t), dtype=np.float64)
    rc = _lib.orc_grad_f64(op, kind, n, _ptr(off, _i64p), _ptr(t, _f64p), _ptr(v, _f64p),
                           _ptr(G, _f64p), Dp, _ptr(gv, _f64p), _ptr(gt, _f64p),
                           int(nthreads or default_threads()))
    if rc != 0:
        raise RuntimeError("oracle gradient failed")
    return gv, gt


# --------------------------------------------------------------------------
# Exact rational checker (independent of the C code; small inputs only).
# --------------------------------------------------------------------------

def _frac_points(f):
    return [(Fraction(float(t)), Fraction(float(v))) for t, v in np.asarray(f, dtype=np.float64)]


def exact_integral(op, f, g):
    """Exact sum over union segments of len * op(fv, gv) for op in {L1, L2sq, DOT}.

    ``op`` is "l1", "l2sq" or "dot". Evaluates both PCFs on the sorted union of
    breakpoints (a brute-force restatement of the integral, independent of the
    merge walk); the last segment is taken to be [t_last, inf) and must have a
    zero integrand, otherwise ``inf`` is returned.
    """
    F = _frac_points(f)
    Gp = _frac_points(g)

    def value_at(P, x):
        val = P[0][1]
        for tt, vv in P:
            if tt <= x:
                val = vv
            else:
                break
        return val

    ts = sorted(set([p[0] for p in F] + [p[0] for p in Gp]))
    total = Fraction(0)
    for k in range(len(ts)):
        x = ts[k]
        fv = value_at(F, x)
        gv = value_at(Gp, x)
        if op == "l1":
            w = abs(fv - gv)
        elif op == "l2sq":
            w = (fv - gv) ** 2
        else:
            w = fv * gv
        if k + 1 < len(ts):
            total += (ts[k + 1] - x) * w
        elif w != 0:
            return float("inf")
    return total
