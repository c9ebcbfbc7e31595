"""GPU tests of the per-PCF L_p norm (``lp_norm``): the reference's own known
answers (test/python/test_norms.py, test_norms_extended.py) through the public
API, and the device kernel against the oracle through the C-ABI: FP64 L1
bit-exact, FP64 p != 1 within 1e-13 (the device ``pow`` may differ from the
host's by an ulp), the FP32 forms within 1e-6 of the FP64 oracle."""

import json
import os

import numpy as np
import pytest
import torch

import oracle
import stablebear_amd as sb
from stablebear_amd import runtime, workloads

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(__file__)
with open(os.path.join(HERE, "golden", "reference_cases.json")) as fh:
    CASES = json.load(fh)
REF = np.load(os.path.join(HERE, "golden", "ref_vectors.npz"))


def _device_norms(off, t, v, p, prec, out_f64=True):
    csr = runtime.csr_from_arrays(off, t, v)
    ds = runtime.DeviceSet(csr, prec)
    out = torch.full((len(off) - 1,), float("nan"),
                     dtype=torch.float64 if out_f64 else torch.float32, device="cuda")
    ds.lp_norm(p, out.data_ptr(), out_f64, torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return out.cpu().numpy().astype(np.float64)


def _check(got, exp, rtol):
    np.testing.assert_array_equal(np.isinf(got), np.isinf(exp))
    assert not np.any(np.isnan(got))
    fin = np.isfinite(exp) & (got != exp)
    if fin.any():
        assert np.max(np.abs(got[fin] - exp[fin]) / np.abs(exp[fin])) <= rtol


@pytest.mark.parametrize("case", CASES["lp_norm"], ids=lambda c: c["src"])
def test_known_answers(gpu, case, pcf_dtype):
    np_dt = np.float32 if pcf_dtype is sb.pcf32 else np.float64
    X = sb.zeros((1,), dtype=pcf_dtype)
    X[0] = sb.Pcf(np.array(case["pcf"], dtype=np_dt), dtype=pcf_dtype)
    got = float(sb.lp_norm(X, p=case["p"])[0])
    if "abs" in case:
        assert got == pytest.approx(case["expected"], abs=case["abs"])
    else:
        assert got == pytest.approx(case["expected"], rel=case.get("rel", 1e-7), abs=1e-300)


def test_reference_table_shape_and_values(gpu):
    """test/python/test_norms.py:26-58: a 2 x 3 tensor keeps its shape."""
    X = sb.zeros((2, 3), dtype=sb.pcf64)
    cases = CASES["lp_norm"][:6]
    for k, c in enumerate(cases):
        X[k // 3, k % 3] = sb.Pcf(np.array(c["pcf"], dtype=np.float64))
    got = sb.lp_norm(X, p=1)
    assert isinstance(got, sb.FloatTensor)
    assert got.shape == (2, 3)
    exp = np.array([c["expected"] for c in cases], dtype=np.float64).reshape(2, 3)
    np.testing.assert_array_equal(np.asarray(got), exp)


def test_negative_values_and_float32_result(gpu):
    """|f| enters the norm; pcf32 inputs give float32 norms."""
    X = sb.zeros((2,), dtype=sb.pcf32)
    X[0] = sb.Pcf(np.array([[0.0, 3.0], [1.0, 0.0]], dtype=np.float32), dtype=sb.pcf32)
    X[1] = sb.Pcf(np.array([[0.0, -3.0], [1.0, 0.0]], dtype=np.float32), dtype=sb.pcf32)
    for p in (1, 2):
        got = np.asarray(sb.lp_norm(X, p=p))
        assert got.dtype == np.float32
        assert got[0] == got[1] == pytest.approx(3.0, rel=1e-6)


def test_empty_tensor(gpu):
    got = sb.lp_norm(sb.zeros((0,), dtype=sb.pcf64))
    assert got.shape == (0,)


@pytest.mark.parametrize("p", [1.0, 2.0, 3.5])
@pytest.mark.parametrize("which", ["c2", "edge", "dups"])
def test_fp64_matches_oracle(gpu, p, which):
    if which == "c2":
        off, t, v = workloads.c2(3000, seed=71)
    elif which == "edge":
        off, t, v = REF["edge_off"], REF["edge_t"], REF["edge_v"]
    else:
        off, t, v = workloads.c2(500, seed=72)
        t = t.copy()
        t[off[:-1] + 3] = t[off[:-1] + 2]  # duplicate times: zero-length pieces
    got = _device_norms(off, t, v, p, "fp64")
    exp = oracle.lp_norms(p, off, t, v)
    if p == 1.0:
        np.testing.assert_array_equal(got, exp)
    else:
        _check(got, exp, 1e-13)


@pytest.mark.parametrize("p", [1.0, 2.0, 3.5])
@pytest.mark.parametrize("which", ["c2", "edge"])
def test_fp32c_matches_fp64_oracle(gpu, p, which):
    if which == "c2":
        off, t, v = workloads.c2(3000, seed=73)
    else:
        off, t, v = REF["edge_off"], REF["edge_t"], REF["edge_v"]
    got = _device_norms(off, t, v, p, "fp32c")
    exp = oracle.lp_norms(p, off, t, v)
    _check(got, exp, 1e-6)


@pytest.mark.parametrize("p", [1.0, 2.0])
def test_float32_inputs_match_fp64_oracle(gpu, p):
    """pcf32 sets (FP32 times and values): the compensated FP32 sum stays within
    1e-6 of the FP64 norm of the same float32 data (PCFs end at 0, so the
    domain end MAX does not enter)."""
    off, t, v = workloads.c2(3000, seed=74)
    t32, v32 = t.astype(np.float32), v.astype(np.float32)
    got = _device_norms(off, t32, v32, p, "fp32c", out_f64=False)
    exp = oracle.lp_norms(p, off, t32.astype(np.float64), v32.astype(np.float64))
    _check(got, exp, 1e-6)


def test_large_set_scaling_property(gpu):
    """At 10^6 PCFs: ||2f||_1 == 2 ||f||_1 bit for bit (doubling is exact),
    and a sample of rows equals the oracle bit for bit."""
    rng = np.random.default_rng(75)
    m = rng.integers(2, 41, size=1_000_000).astype(np.int64)
    off, t = workloads._times(rng, m, 1.0)
    v = workloads._normal_values(rng, off)
    a = _device_norms(off, t, v, 1.0, "fp64")
    b = _device_norms(off, t, 2.0 * v, 1.0, "fp64")
    np.testing.assert_array_equal(b, 2.0 * a)
    rows = np.random.default_rng(5).choice(len(off) - 1, 2000, replace=False)
    for i in rows[:50]:
        sl = slice(off[i], off[i + 1])
        exp = oracle.lp_norms(1.0, np.array([0, off[i + 1] - off[i]]), t[sl], v[sl])
        assert a[i] == exp[0]


def test_rejects_p_below_one_on_device_set(gpu):
    off, t, v = workloads.c1(10, seed=76)
    with pytest.raises(ValueError, match="p must be >= 1"):
        _device_norms(off, t, v, 0.5, "fp64")
