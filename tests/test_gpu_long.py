"""GPU tests of the long-pair kernel (one wavefront per pair, the pair's union
segments split across the lanes): forced on against the oracle for every
output kind and form, the automatic choice, the job path, lp_distance, and
the autotuner's choice between the two kernels."""

import numpy as np
import pytest
import torch

import oracle
from stablebear_amd import _lib as L
from stablebear_amd import runtime, workloads

pytestmark = pytest.mark.gpu


def _long_set(n, m_lo, m_hi, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    m = rng.integers(m_lo, m_hi + 1, size=n).astype(np.int64)
    off, t = workloads._times(rng, m, scale)
    return off, t, workloads._normal_values(rng, off)


def _edge():
    ref = np.load(__import__("os").path.join(__import__("os").path.dirname(__file__), "golden",
                                             "ref_vectors.npz"))
    return ref["edge_off"], ref["edge_t"], ref["edge_v"]


def _run(plan, f64, row_sum=None):
    out = torch.empty(plan.out_count, dtype=torch.float64 if f64 else torch.float32, device="cuda")
    plan.execute(out.data_ptr(), f64, None if row_sum is None else row_sum.data_ptr(),
                 torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return out.cpu().numpy().astype(np.float64)


def _check(got, exp, prec, op):
    """FP64: 1e-12 (64 partial sums instead of one running sum); FP32C: 1e-6.
    Relative per entry, normwise for the signed inner product."""
    tol = 1e-12 if prec == "fp64" else 1e-6
    same = got == exp  # includes the infinite entries of PCFs that end non-zero
    np.testing.assert_array_equal(np.isinf(got), np.isinf(exp))
    fin = ~same & np.isfinite(exp)
    err = np.abs(got[fin] - exp[fin])
    if op == "dot":
        scale = np.max(np.abs(exp[np.isfinite(exp)]))
        assert err.size == 0 or np.max(err) <= tol * scale
    else:
        assert np.all(exp[fin] != 0)
        assert err.size == 0 or np.max(err / np.abs(exp[fin])) <= tol


_OPS = {"l1": (oracle.OP_L1, 1.0), "lp": (oracle.OP_LP, 2.0), "dot": (oracle.OP_DOT, 1.0)}


@pytest.mark.parametrize("prec", ["fp32c", "fp64"])
@pytest.mark.parametrize("op", ["l1", "lp", "dot"])
@pytest.mark.parametrize("which", ["c2", "few_long", "edge"])
def test_forced_long_matches_oracle(gpu, prec, op, which):
    if which == "c2":
        off, t, v = workloads.c2(40, seed=21)
    elif which == "few_long":
        off, t, v = _long_set(7, 3000, 9000, seed=22)
    else:
        off, t, v = _edge()
    ds = runtime.DeviceSet(runtime.csr_from_arrays(off, t, v), prec)
    plan = runtime.Plan(ds, op=op, p=_OPS[op][1])
    plan.set_long_pairs("on")
    assert plan.long_pairs
    kind = oracle.KIND_KERNEL if op == "dot" else oracle.KIND_PDIST
    exp = oracle.pairwise(_OPS[op][0], _OPS[op][1], kind, off, t, v)
    _check(_run(plan, prec == "fp64"), exp, prec, op)


@pytest.mark.parametrize("prec", ["fp32c", "fp64"])
def test_forced_long_cdist_with_row_sums(gpu, prec):
    X = _long_set(5, 2000, 6000, seed=23, scale=10.0)
    Y = _long_set(9, 500, 5000, seed=24, scale=10.0)
    sx = runtime.DeviceSet(runtime.csr_from_arrays(*X), prec)
    sy = runtime.DeviceSet(runtime.csr_from_arrays(*Y), prec)
    plan = runtime.Plan(sx, sy, op="l1")
    plan.set_long_pairs("on")
    rs = torch.full((5,), 7.0, dtype=torch.float64, device="cuda")
    got = _run(plan, prec == "fp64", rs)
    exp = oracle.pairwise(oracle.OP_L1, 1.0, oracle.KIND_CDIST, *X, *Y)
    _check(got, exp, prec, "l1")
    np.testing.assert_allclose(rs.cpu().numpy(), exp.reshape(5, 9).sum(axis=1), rtol=1e-6)


def test_auto_uses_long_for_few_long_fp32c_pairs(gpu):
    off, t, v = _long_set(8, 10000, 20000, seed=25)
    csr = runtime.csr_from_arrays(off, t, v)
    p32 = runtime.Plan(runtime.DeviceSet(csr, "fp32c"), op="l1")
    p64 = runtime.Plan(runtime.DeviceSet(csr, "fp64"), op="l1")
    assert p32.long_pairs and not p64.long_pairs
    exp = oracle.pairwise(oracle.OP_L1, 1.0, oracle.KIND_PDIST, off, t, v)
    _check(_run(p32, False), exp, "fp32c", "l1")
    np.testing.assert_array_equal(_run(p64, True), exp)  # tile kernel: bitwise
    many = runtime.Plan(runtime.DeviceSet(runtime.csr_from_arrays(*workloads.c2(200)), "fp32c"),
                        op="l1")
    assert not many.long_pairs


def test_job_path_and_lp_distance_on_long_pcfs(gpu):
    off, t, v = _long_set(6, 8000, 12000, seed=26)
    csr = runtime.csr_from_arrays(off, t, v)
    out = np.zeros(15, dtype=np.float32)
    task = runtime.submit(csr, None, L.OP_L1, 1.0, L.OUT_DISTMAT, L.PREC_FP32C, out, "d", 15)
    task.wait()
    exp = oracle.pairwise(oracle.OP_L1, 1.0, oracle.KIND_PDIST, off, t, v)
    _check(out.astype(np.float64), exp, "fp32c", "l1")
    # lp_distance of two long pcf32 PCFs (FP32C, long-pair kernel) against the
    # FP64 integral of the same float32 inputs
    from stablebear_amd.pcf import Pcf
    pts = np.stack([t, v], axis=1).astype(np.float32)
    f, g = Pcf(pts[off[0]:off[1]]), Pcf(pts[off[1]:off[2]])
    got = runtime.lp_distance(f, g, 1.0)
    t32, v32 = t.astype(np.float32).astype(np.float64), v.astype(np.float32).astype(np.float64)
    fx = (np.array([0, off[1]]), t32[:off[1]], v32[:off[1]])
    gy = (np.array([0, off[2] - off[1]]), t32[off[1]:off[2]], v32[off[1]:off[2]])
    exp1 = oracle.pairwise(oracle.OP_L1, 1.0, oracle.KIND_CDIST, *fx, *gy)
    assert got == pytest.approx(float(exp1[0]), rel=1e-6)


def test_autotune_weighs_long_against_tiles(gpu):
    off, t, v = _long_set(12, 5000, 15000, seed=27)
    ds = runtime.DeviceSet(runtime.csr_from_arrays(off, t, v), "fp32c")
    plan = runtime.Plan(ds, op="l1")
    out = torch.empty(plan.out_count, dtype=torch.float32, device="cuda")
    shape, ms = plan.autotune(out.data_ptr(), False, reps=1)
    assert ms > 0
    exp = oracle.pairwise(oracle.OP_L1, 1.0, oracle.KIND_PDIST, off, t, v)
    _check(_run(plan, False), exp, "fp32c", "l1")
