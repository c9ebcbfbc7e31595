"""Host runtime around the C-ABI: the default device context, CSR packing of
PCF containers, asynchronous tasks, and device-resident sets / plans.

Everything here drives ``libstablebear_amd.so``; there is no CPU compute path.
"""

from __future__ import annotations

import ctypes
import threading
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from .pcf import Pcf
from .typing import pcf32, pcf64

_ctx_lock = threading.Lock()
_contexts: dict[int, ctypes.c_void_p] = {}
_default_device = 0


def set_default_device(device: int) -> None:
    global _default_device
    _default_device = int(device)


def context(device: int | None = None) -> ctypes.c_void_p:
    """The process-wide context of a HIP device (created on first use)."""
    dev = _default_device if device is None else int(device)
    with _ctx_lock:
        ctx = _contexts.get(dev)
        if ctx is None:
            h = ctypes.c_void_p()
            L.check(L.lib().sbamd_ctx_create(dev, ctypes.byref(h)))
            ctx = h
            _contexts[dev] = ctx
        return ctx


# ---------------------------------------------------------------------------
# CSR packing
# ---------------------------------------------------------------------------

@dataclass
class Csr:
    """PCF set in CSR form; keeps the numpy arrays alive for the C structs."""

    offsets: np.ndarray  # int64, n + 1
    t: np.ndarray
    v: np.ndarray
    dtype: int  # L.F32 / L.F64

    @property
    def n(self) -> int:
        return len(self.offsets) - 1

    @property
    def strided(self) -> bool:
        """True when t and v are strided views (of AoS points)."""
        return not (self.t.flags["C_CONTIGUOUS"] and self.v.flags["C_CONTIGUOUS"])

    def struct(self) -> L.PcfSet:
        """The C struct of contiguous arrays (strided t / v are copied once)."""
        if self.strided:
            self.t = np.ascontiguousarray(self.t)
            self.v = np.ascontiguousarray(self.v)
        s = L.PcfSet()
        s.n = self.n
        s.offsets = self.offsets.ctypes.data_as(L.c_i64p)
        s.t = self.t.ctypes.data if self.t.size else None
        s.v = self.v.ctypes.data if self.v.size else None
        s.dtype = self.dtype
        return s


def csr_view(csr: Csr) -> L.PcfView:
    """Host view of a CSR set for the GPU packer."""
    view = L.PcfView()
    view.n = csr.n
    view.offsets = csr.offsets.ctypes.data
    view.t = csr.t.ctypes.data if csr.t.size else None
    view.v = csr.v.ctypes.data if csr.v.size else None
    view.t_stride = csr.t.strides[0] if csr.t.ndim else csr.t.itemsize
    view.v_stride = csr.v.strides[0] if csr.v.ndim else csr.v.itemsize
    view.dtype = csr.dtype
    view.location = L.HOST
    return view


def csr_from_arrays(offsets, t, v) -> Csr:
    """CSR set from offsets (n+1), times and values (float32 or float64)."""
    t = np.asarray(t)
    dt = np.float64 if t.dtype == np.float64 else np.float32
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    t = np.ascontiguousarray(t, dtype=dt)
    v = np.ascontiguousarray(v, dtype=dt)
    return Csr(offsets, t, v, L.F64 if dt == np.float64 else L.F32)


def csr_from_pcfs(pcfs, dtype) -> Csr:
    """CSR set from a list of Pcf of one dtype (pcf32 or pcf64). The points
    are concatenated once; t and v are strided views of them (AoS), which the
    jobs take as they are and other callers copy when they need flat arrays."""
    np_dt = np.float64 if dtype is pcf64 else np.float32
    n = len(pcfs)
    lens = np.fromiter((p._arr.shape[0] for p in pcfs), dtype=np.int64, count=n)
    offsets = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(lens, out=offsets[1:])
    if n:
        pts = np.ascontiguousarray(np.concatenate([p._arr for p in pcfs], axis=0), dtype=np_dt)
        t = pts[:, 0]
        v = pts[:, 1]
    else:
        t = np.zeros(0, dtype=np_dt)
        v = np.zeros(0, dtype=np_dt)
    return Csr(offsets, t, v, L.F64 if np_dt == np.float64 else L.F32)


# ---------------------------------------------------------------------------
# Tasks (the reference's StoppableTask protocol)
# ---------------------------------------------------------------------------

class Task:
    """Handle of an asynchronous pairwise job: wait_for / progress / stop.

    Mirrors the methods ``stablebear/async_task.py`` polls
    (``include/sbear/task.hpp:15-193``). Holds the host buffers the job reads
    and writes until it is destroyed.
    """

    def __init__(self, job: ctypes.c_void_p | None, keepalive, desc: str, total: int):
        self._job = job
        self._keep = keepalive
        self._desc = desc
        self._total = int(total)
        self._done = job is None
        self._error = None

    def wait_for(self, timeout_ms: int) -> bool:
        if self._done:
            if self._error:
                err, self._error = self._error, None
                raise err
            return True
        rc = L.lib().sbamd_job_wait(self._job, int(timeout_ms))
        if rc == L.TIMEOUT:
            return False
        self._done = True
        try:
            L.check(rc)
        finally:
            self._release()
        return True

    def wait(self) -> None:
        while not self.wait_for(1000):
            pass

    def request_stop(self) -> None:
        if self._job is not None and not self._done:
            L.lib().sbamd_job_cancel(self._job)

    def work_total(self) -> int:
        return self._total

    def work_completed(self) -> int:
        if self._job is None:
            return self._total
        done = ctypes.c_uint64()
        tot = ctypes.c_uint64()
        L.lib().sbamd_job_progress(self._job, ctypes.byref(done), ctypes.byref(tot))
        return int(done.value)

    def work_step(self) -> int:
        return 0

    def work_step_desc(self) -> str:
        return self._desc

    def work_step_unit(self) -> str:
        return "integral"

    def _release(self):
        if self._job is not None:
            L.lib().sbamd_job_destroy(self._job)
            self._job = None
        self._keep = None

    def __del__(self):
        try:
            if self._job is not None:
                L.lib().sbamd_job_cancel(self._job)
                self._release()
        except Exception:
            pass


def done_task(desc="", total=0) -> Task:
    return Task(None, None, desc, total)


def set_output_budget(nbytes: int, device: int | None = None) -> None:
    """Device bytes a host-to-host job may give its output at once (0: half
    the free memory); larger outputs run in bands of rows."""
    L.check(L.lib().sbamd_ctx_set_output_budget(context(device), int(nbytes)))


def set_job_tile_shape(rows=0, cols=0, cap=0, lane_pairs=0, device: int | None = None) -> None:
    """Tile shape of the host-to-host jobs' plans (0: defaults)."""
    sh = L.TileShape(int(rows), int(cols), int(cap), int(lane_pairs))
    L.check(L.lib().sbamd_ctx_set_tile_shape(context(device), ctypes.byref(sh)))


def autotune_cache_stats():
    """(hits, misses) of the autotuner's per-class shape cache in this process."""
    h, m = ctypes.c_uint64(), ctypes.c_uint64()
    L.check(L.lib().sbamd_autotune_cache_stats(ctypes.byref(h), ctypes.byref(m)))
    return h.value, m.value


def submit(rows: Csr, cols: Csr | None, op: int, p: float, out_kind: int, precision: int,
           out: np.ndarray, desc: str, total: int) -> Task:
    """Start a host-to-host pairwise job writing into ``out`` (kept alive by
    the task). Sets of strided (AoS) points go in as views, unsplit; their
    points are checked on the device during the job."""
    job = ctypes.c_void_p()
    out_dtype = L.F64 if out.dtype == np.float64 else L.F32
    out_ptr = out.ctypes.data if out.size else None
    if rows.strided or (cols is not None and cols.strided):
        rs = csr_view(rows)
        cs = csr_view(cols) if cols is not None else None
        rc = L.lib().sbamd_pairwise_submit_view(context(), ctypes.byref(rs),
                                                ctypes.byref(cs) if cs is not None else None, op,
                                                float(p), out_kind, precision, out_ptr, out_dtype,
                                                ctypes.byref(job))
    else:
        rs = rows.struct()
        cs = cols.struct() if cols is not None else None
        rc = L.lib().sbamd_pairwise_submit(context(), ctypes.byref(rs),
                                           ctypes.byref(cs) if cs is not None else None, op,
                                           float(p), out_kind, precision, out_ptr, out_dtype,
                                           ctypes.byref(job))
    L.check(rc)
    return Task(job, (rows, cols, rs, cs, out), desc, total)


# ---------------------------------------------------------------------------
# Single-pair helpers
# ---------------------------------------------------------------------------

def _one(p: Pcf) -> Csr:
    return csr_from_pcfs([p], p.dtype)


def lp_distance(f: Pcf, g: Pcf, p: float) -> float:
    a, b = _one(f), _one(g)
    sa, sb_ = a.struct(), b.struct()
    out = ctypes.c_double()
    L.check(L.lib().sbamd_lp_distance(context(), ctypes.byref(sa), ctypes.byref(sb_), float(p),
                                      ctypes.byref(out)))
    return float(out.value)


def lp_norms(csr: Csr, p: float, precision: int, out: np.ndarray) -> None:
    """Per-PCF L_p norms of a CSR set into ``out`` (n values, float32 or float64)."""
    s = csr.struct()
    out_dtype = L.F64 if out.dtype == np.float64 else L.F32
    L.check(L.lib().sbamd_lp_norm(context(), ctypes.byref(s), float(p), precision,
                                  out.ctypes.data if out.size else None, out_dtype))


def rectangles(f: Pcf, g: Pcf, a: float, b: float) -> np.ndarray:
    A, B = _one(f), _one(g)
    sa, sb_ = A.struct(), B.struct()
    np_dt = np.float64 if f.dtype is pcf64 else np.float32
    cap = f.size + g.size + 4
    out = np.zeros((cap, 4), dtype=np_dt)
    cnt = ctypes.c_int64()
    L.check(L.lib().sbamd_iterate_rectangles(context(), ctypes.byref(sa), ctypes.byref(sb_),
                                             float(a), float(b), out.ctypes.data, cap,
                                             ctypes.byref(cnt)))
    if cnt.value > cap:
        raise RuntimeError("rectangle count exceeds the walk bound")
    return out[: cnt.value]


def count_segments(rows: Csr, cols: Csr | None = None, include_diagonal=False) -> int:
    """Exact union-segment count W of a pdist / kernel / cdist (host only)."""
    rs = rows.struct()
    cs = cols.struct() if cols is not None else None
    w = ctypes.c_uint64()
    L.check(L.lib().sbamd_count_segments(ctypes.byref(rs), ctypes.byref(cs) if cs else None,
                                         1 if include_diagonal else 0, ctypes.byref(w)))
    return int(w.value)


# ---------------------------------------------------------------------------
# Device-resident sets and plans (zero-copy execution on device buffers)
# ---------------------------------------------------------------------------

PRECISIONS = {"fp64": L.PREC_FP64, "fp32c": L.PREC_FP32C}


class DeviceSet:
    """A PCF set validated, sorted and uploaded once for a given precision.

    ``pack="host"`` packs the records on the host and uploads them;
    ``pack="gpu"`` copies the CSR arrays to the device as they are and
    validates and packs there (the same records, bit for bit)."""

    def __init__(self, csr: Csr | None, precision: str = "fp64", device: int | None = None,
                 pack: str = "host", view: L.PcfView | None = None, keep=None):
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(PRECISIONS)}")
        if pack not in ("host", "gpu"):
            raise ValueError("pack must be 'host' or 'gpu'")
        self.ctx = context(device)
        self.csr = csr
        self.precision = precision
        self._keep = keep
        h = ctypes.c_void_p()
        if view is not None:
            L.check(L.lib().sbamd_set_create_gpu(self.ctx, ctypes.byref(view),
                                                 PRECISIONS[precision], ctypes.byref(h)))
        elif pack == "gpu":
            view = csr_view(csr)
            L.check(L.lib().sbamd_set_create_gpu(self.ctx, ctypes.byref(view),
                                                 PRECISIONS[precision], ctypes.byref(h)))
        else:
            s = csr.struct()
            L.check(L.lib().sbamd_set_create(self.ctx, ctypes.byref(s), PRECISIONS[precision],
                                             ctypes.byref(h)))
        self.handle = h
        n, rec, pts = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        L.check(L.lib().sbamd_set_info(h, ctypes.byref(n), ctypes.byref(rec), ctypes.byref(pts)))
        self.n, self.records, self.points = n.value, rec.value, pts.value

    @classmethod
    def from_handle(cls, handle, precision: str, device: int | None = None, keep=None):
        """Wrap a set handle created by the C-ABI (the object owns it)."""
        self = cls.__new__(cls)
        self.ctx = context(device)
        self.csr = None
        self.precision = precision
        self._keep = keep
        self.handle = handle
        n, rec, pts = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        L.check(L.lib().sbamd_set_info(handle, ctypes.byref(n), ctypes.byref(rec),
                                       ctypes.byref(pts)))
        self.n, self.records, self.points = n.value, rec.value, pts.value
        return self

    @classmethod
    def from_device(cls, offsets, t, v, precision: str = "fp64", device: int | None = None):
        """Set from CSR arrays already in device memory (torch tensors on the
        GPU: offsets int64 (n+1), t and v float32/float64 of one dtype, any
        element strides); validated and packed on the GPU, read in place."""
        import torch

        if offsets.dtype != torch.int64 or offsets.dim() != 1 or not offsets.is_contiguous():
            raise ValueError("offsets must be a contiguous 1-D int64 tensor")
        if t.dtype != v.dtype or t.dtype not in (torch.float32, torch.float64):
            raise TypeError("t and v must both be float32 or both float64")
        if t.dim() != 1 or v.dim() != 1 or t.numel() != v.numel():
            raise ValueError("t and v must be 1-D tensors of equal length")
        es = t.element_size()
        view = L.PcfView()
        view.n = offsets.numel() - 1
        view.offsets = offsets.data_ptr()
        view.t = t.data_ptr() if t.numel() else None
        view.v = v.data_ptr() if v.numel() else None
        view.t_stride = t.stride(0) * es
        view.v_stride = v.stride(0) * es
        view.dtype = L.F64 if t.dtype == torch.float64 else L.F32
        view.location = L.DEVICE
        return cls(None, precision, device, view=view, keep=(offsets, t, v))

    def export(self):
        """The packed device layout as numpy arrays: (records as raw bytes,
        sorted record offsets, permutation, source points)."""
        rb = ctypes.c_int32()
        L.check(L.lib().sbamd_set_export(self.handle, None, None, None, None, ctypes.byref(rb)))
        rec = np.zeros(self.records * rb.value, dtype=np.uint8)
        roff = np.zeros(self.n + 1, dtype=np.int64)
        perm = np.zeros(self.n, dtype=np.int64)
        src = np.zeros(self.records, dtype=np.int64)
        L.check(L.lib().sbamd_set_export(
            self.handle, rec.ctypes.data if rec.size else None, roff.ctypes.data_as(L.c_i64p),
            perm.ctypes.data_as(L.c_i64p) if perm.size else None,
            src.ctypes.data_as(L.c_i64p) if src.size else None, ctypes.byref(rb)))
        return rec, roff, perm, src

    def update_values(self, v_device_ptr: int, stream: int | None = None) -> None:
        L.check(L.lib().sbamd_set_update_values(self.handle, v_device_ptr, stream))

    def lp_norm(self, p: float, out_ptr: int, out_f64: bool, stream: int | None = None) -> None:
        """Per-PCF L_p norms into device memory (n values, original order)."""
        L.check(L.lib().sbamd_set_lp_norm(self.handle, float(p), out_ptr,
                                          L.F64 if out_f64 else L.F32, stream))

    def close(self):
        if getattr(self, "handle", None):
            L.lib().sbamd_set_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


OPS = {"l1": L.OP_L1, "lp": L.OP_LP, "dot": L.OP_DOT}


class Plan:
    """Tile plan of one pdist / l2_kernel / cdist (or a batch of pdists)."""

    def __init__(self, rows, cols=None, op="l1", p=1.0, out_kind=None, shape=None,
                 batched=False):
        self._sets = list(rows) if batched else [rows] + ([cols] if cols is not None else [])
        first = self._sets[0]
        self.ctx = first.ctx
        self.op = op
        self.p = float(p)
        if out_kind is None:
            out_kind = L.OUT_DENSE if cols is not None else (
                L.OUT_SYMMAT if op == "dot" else L.OUT_DISTMAT)
        self.out_kind = out_kind
        sh = L.TileShape()
        if shape:
            sh.rows, sh.cols, sh.cap = shape[:3]
            if len(shape) > 3:
                sh.lane_pairs = shape[3]
        h = ctypes.c_void_p()
        if batched:
            arr = (ctypes.c_void_p * len(self._sets))(*[s.handle.value for s in self._sets])
            L.check(L.lib().sbamd_plan_create_batched(self.ctx, len(self._sets), arr, OPS[op],
                                                      self.p, out_kind, ctypes.byref(sh),
                                                      ctypes.byref(h)))
        else:
            L.check(L.lib().sbamd_plan_create(self.ctx, first.handle,
                                              cols.handle if cols is not None else None,
                                              OPS[op], self.p, out_kind, ctypes.byref(sh),
                                              ctypes.byref(h)))
        self.handle = h
        self.refresh_stats()

    def refresh_stats(self):
        st = L.PlanStats()
        L.check(L.lib().sbamd_plan_info(self.handle, ctypes.byref(st)))
        self.pairs = st.pairs
        self.segments = st.segments
        self.out_count = st.out_count
        self.tiles = st.tiles
        self.input_bytes = st.input_bytes
        self.shape = (st.rows, st.cols, st.cap)
        self.lane_pairs = st.lane_pairs
        self.problems = st.problems
        self.long_pairs = bool(st.long_pairs)

    def set_long_pairs(self, mode):
        """Long-pair kernel: False / "off", True / "on", or "auto" (the default)."""
        code = {False: 0, "off": 0, True: 1, "on": 1, "auto": 2}[mode]
        L.check(L.lib().sbamd_plan_set_long_pairs(self.handle, code))
        self.refresh_stats()

    def out_offset(self, problem: int) -> int:
        off = ctypes.c_uint64()
        L.check(L.lib().sbamd_plan_out_offset(self.handle, problem, ctypes.byref(off)))
        return int(off.value)

    def set_fold_threshold(self, segments: int):
        L.check(L.lib().sbamd_plan_set_fold_threshold(self.handle, int(segments)))

    def set_persistent(self, on: bool):
        L.check(L.lib().sbamd_plan_set_persistent(self.handle, 1 if on else 0))

    def set_shape(self, rows, cols, cap, lane_pairs=0):
        sh = L.TileShape(rows, cols, cap, lane_pairs)
        L.check(L.lib().sbamd_plan_set_shape(self.handle, ctypes.byref(sh)))
        self.refresh_stats()

    def execute(self, out_ptr: int, out_f64: bool, row_sum_ptr: int | None = None,
                stream: int | None = None):
        L.check(L.lib().sbamd_plan_execute(self.handle, out_ptr, L.F64 if out_f64 else L.F32,
                                           row_sum_ptr, stream))

    def execute_to_host(self, out_f64: bool) -> np.ndarray:
        out = np.empty(self.out_count, dtype=np.float64 if out_f64 else np.float32)
        L.check(L.lib().sbamd_plan_execute_to_host(self.handle, out.ctypes.data if out.size else None,
                                                   L.F64 if out_f64 else L.F32))
        return out

    def backward(self, grad_out_ptr, fwd_out_ptr, out_f64, grad_v_ptr, grad_t_ptr, grad_f64,
                 stream=None):
        L.check(L.lib().sbamd_plan_backward(self.handle, grad_out_ptr, fwd_out_ptr,
                                            L.F64 if out_f64 else L.F32, grad_v_ptr, grad_t_ptr,
                                            L.F64 if grad_f64 else L.F32, stream))

    def backward_noalloc(self, grad_out_ptr, fwd_out_ptr, out_f64, grad_v_ptr, grad_t_ptr,
                         grad_f64, stream=None):
        """``backward`` with no device allocation (a static working array)."""
        L.check(L.lib().sbamd_plan_backward_noalloc(self.handle, grad_out_ptr, fwd_out_ptr,
                                                    L.F64 if out_f64 else L.F32, grad_v_ptr,
                                                    grad_t_ptr, L.F64 if grad_f64 else L.F32, stream))

    def autotune(self, out_ptr: int, out_f64: bool, reps: int = 3):
        sh = L.TileShape()
        ms = ctypes.c_double()
        L.check(L.lib().sbamd_autotune(self.handle, out_ptr, L.F64 if out_f64 else L.F32,
                                       int(reps), ctypes.byref(sh), ctypes.byref(ms)))
        self.refresh_stats()
        return (sh.rows, sh.cols, sh.cap), ms.value

    def set_class_shapes(self, shapes):
        """One (rows, cols, cap, lane_pairs) per length class (4 of them)."""
        arr = (L.TileShape * 4)(*[L.TileShape(*sh) for sh in shapes])
        L.check(L.lib().sbamd_plan_set_class_shapes(self.handle, arr))
        self.refresh_stats()

    def class_shapes(self):
        """[(rows, cols, cap, lane_pairs, tiles)] of the four length classes."""
        sh = (L.TileShape * 4)()
        tiles = np.zeros(4, dtype=np.int64)
        L.check(L.lib().sbamd_plan_class_shapes(self.handle, sh, tiles.ctypes.data_as(L.c_i64p)))
        return [(s.rows, s.cols, s.cap, s.lane_pairs, int(t)) for s, t in zip(sh, tiles)]

    def graph(self, out_ptr: int, out_f64: bool) -> "Graph":
        return Graph(self, out_ptr, out_f64)

    def close(self):
        if getattr(self, "handle", None):
            L.lib().sbamd_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Graph:
    """hipGraph holding one captured forward launch of a plan."""

    def __init__(self, plan: Plan, out_ptr: int, out_f64: bool):
        self.plan = plan
        h = ctypes.c_void_p()
        L.check(L.lib().sbamd_graph_create(plan.handle, out_ptr, L.F64 if out_f64 else L.F32,
                                           ctypes.byref(h)))
        self.handle = h

    def launch(self, stream: int | None = None):
        L.check(L.lib().sbamd_graph_launch(self.handle, stream))

    def close(self):
        if getattr(self, "handle", None):
            L.lib().sbamd_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PlainKernel:
    """The plain FP64 GPU kernel (one lane per pair over raw breakpoints in
    global memory): the measured GPU baseline of the tiled kernels."""

    def __init__(self, csr: Csr, device: int | None = None):
        self.ctx = context(device)
        self.csr = csr
        self.n = csr.n
        s = csr.struct()
        h = ctypes.c_void_p()
        L.check(L.lib().sbamd_plain_create(self.ctx, ctypes.byref(s), ctypes.byref(h)))
        self.handle = h

    def execute(self, op: str, p: float, out_kind: int, out_ptr: int, stream: int | None = None):
        L.check(L.lib().sbamd_plain_execute(self.handle, OPS[op], float(p), out_kind, out_ptr,
                                            stream))

    def close(self):
        if getattr(self, "handle", None):
            L.lib().sbamd_plain_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def synchronize(device: int | None = None):
    L.check(L.lib().sbamd_ctx_synchronize(context(device)))
