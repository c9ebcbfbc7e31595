"""GPU parity: the HIP path against the oracle, the reference's known answers
and the reference's compiled-walk vectors, through the public API (which
calls the C-ABI). FP64 L1 / inner products are bit-exact with the reference
CPU integral; FP64 L_p uses d*d / sqrt at p = 2 and pow otherwise (tolerance
1e-13 relative); FP32C is held to 1e-6 max relative error on non-negative
integrands and 1e-6 normwise (|K - K_ref| / int |f||g|) for the signed kernel.
"""

import json
import os

import numpy as np
import pytest

import oracle
import stablebear_amd as sb
from stablebear_amd import workloads

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
CASES = json.load(open(os.path.join(GOLD, "reference_cases.json")))
REF = np.load(os.path.join(GOLD, "ref_vectors.npz"))
DBL_MAX = np.finfo(np.float64).max


def _num(x):
    return {"DBL_MAX": DBL_MAX, "inf": float("inf")}.get(x, x) if isinstance(x, str) else float(x)


def _pcfs(off, t, v, dtype=np.float64):
    return sb.PcfTensor(workloads.to_pcfs(off, t, v, dtype))


def _dense_pd(c, n):
    d = np.zeros((n, n))
    r, cc = np.tril_indices(n, -1)
    d[r, cc] = c
    d[cc, r] = c
    return d


def test_device_visible(gpu):
    assert sb._lib.device_count() >= 1


# ------------------------------------------------- known answers ----------

@pytest.mark.parametrize("case", CASES["rects"], ids=lambda c: c["src"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_iterate_rectangles_known_answers(gpu, case, dtype):
    f = sb.Pcf(np.array(case["f"], dtype=dtype))
    g = sb.Pcf(np.array(case["g"], dtype=dtype))
    b = _num(case["b"])
    if dtype == np.float32 and b == DBL_MAX:
        b = float(np.finfo(np.float32).max)
    rects = sb.iterate_rectangles(f, g, _num(case["a"]), b)
    exp = [[_num(x) for x in r] for r in case["expected"]]
    assert len(rects) == len(exp)
    for r, e in zip(rects, exp):
        e_r = e[1] if not (dtype == np.float32 and e[1] == DBL_MAX) else b
        assert (r.left, r.right, r.f_value, r.g_value) == pytest.approx((e[0], e_r, e[2], e[3]))


def test_iterate_rectangles_properties(gpu):
    f = sb.Pcf(np.array([[0, 5], [1, 3], [3, 0]], dtype=np.float64))
    assert all(r.f_value == r.g_value for r in sb.iterate_rectangles(f, f))
    f = sb.Pcf(np.array([[0, 4], [2, 0]], dtype=np.float64))
    g = sb.Pcf(np.array([[0, 0], [5, 3], [7, 0]], dtype=np.float64))
    assert all(r.f_value == 0 or r.g_value == 0 for r in sb.iterate_rectangles(f, g, b=10.0))


@pytest.mark.parametrize("case", CASES["lp_distance"], ids=lambda c: c["src"])
def test_lp_distance_known_answers(gpu, pcf_dtype, case):
    dt = np.float32 if pcf_dtype is sb.pcf32 else np.float64
    f = sb.Pcf(np.array(case["f"], dtype=dt))
    g = sb.Pcf(np.array(case["g"], dtype=dt))
    got = sb.lp_distance(f, g, p=case["p"])
    assert isinstance(got, float)
    assert got == pytest.approx(case["expected"], abs=case.get("abs", 1e-5))
    assert sb.lp_distance(g, f, p=case["p"]) == pytest.approx(got, rel=1e-12, abs=1e-12)


@pytest.mark.parametrize("case", CASES["pdist"], ids=lambda c: c["src"])
def test_pdist_known_answers(gpu, pcf_dtype, case):
    dt = np.float32 if pcf_dtype is sb.pcf32 else np.float64
    X = sb.PcfTensor([sb.Pcf(np.array(p, dtype=dt)) for p in case["pcfs"]])
    D = sb.pdist(X, p=case["p"])
    assert isinstance(D, sb.DistanceMatrix)
    np.testing.assert_allclose(D.to_dense(), case["expected_dense"], rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("case", CASES["l2_kernel"], ids=lambda c: c["src"])
def test_kernel_known_answers(gpu, pcf_dtype, case):
    dt = np.float32 if pcf_dtype is sb.pcf32 else np.float64
    X = sb.PcfTensor([sb.Pcf(np.array(p, dtype=dt)) for p in case["pcfs"]])
    K = sb.l2_kernel(X)
    assert isinstance(K, sb.SymmetricMatrix)
    np.testing.assert_allclose(K.to_dense(), case["expected_dense"], rtol=1e-5)


# ------------------------------------------ reference-walk vectors --------

_OPS = {"l1": 1, "l2": 2, "l3": 3}


@pytest.mark.parametrize("name", ["c1", "c2", "edge", "c4"])
@pytest.mark.parametrize("oname", ["l1", "l2", "l3"])
def test_pdist_fp64_matches_reference_vectors(gpu, name, oname):
    X = _pcfs(REF[f"{name}_off"], REF[f"{name}_t"], REF[f"{name}_v"])
    got = sb.pdist(X, p=_OPS[oname]).storage
    exp = REF[f"{name}_{oname}_f64"]
    if oname == "l1":
        np.testing.assert_array_equal(got, exp)
    else:
        np.testing.assert_allclose(got, exp, rtol=1e-13, atol=0)


@pytest.mark.parametrize("name", ["c1", "c2", "edge", "c4"])
def test_kernel_fp64_matches_reference_vectors(gpu, name):
    X = _pcfs(REF[f"{name}_off"], REF[f"{name}_t"], REF[f"{name}_v"])
    np.testing.assert_array_equal(sb.l2_kernel(X).storage, REF[f"{name}_dot_f64"])


@pytest.mark.parametrize("name", ["c1", "c2", "edge", "c4"])
@pytest.mark.parametrize("oname", ["l1", "l2", "dot"])
def test_pcf32_matches_reference_float_path(gpu, name, oname):
    """pcf32 inputs: compensated FP32 vs the reference's naive-float CPU result
    (the reference test tolerance, atol 1e-5 + rtol 1e-5 of the magnitude) and
    vs the FP64 integral of the same float32 inputs (1e-6)."""
    off, t, v = REF[f"{name}_off"], REF[f"{name}_t"], REF[f"{name}_v"]
    X = _pcfs(off, t, v, np.float32)
    if oname == "dot":
        got = sb.l2_kernel(X).storage
        kind, op, p = oracle.KIND_KERNEL, oracle.OP_DOT, 1.0
    else:
        got = sb.pdist(X, p=_OPS[oname]).storage
        kind, op, p = oracle.KIND_PDIST, (oracle.OP_L1 if oname == "l1" else oracle.OP_LP), _OPS[oname]
    ref32 = REF[f"{name}_{oname}_f32"].astype(np.float64)
    np.testing.assert_allclose(got, ref32, rtol=1e-5, atol=1e-5 * max(1.0, np.abs(ref32).max()))
    t32, v32 = t.astype(np.float32), v.astype(np.float32)
    exact = oracle.pairwise(op, p, kind, off, t32, v32)
    if oname == "dot":
        scale = oracle.pairwise(op, p, kind, off, t32, np.abs(v32))
        _assert_fp32c(got, exact, signed=True, scale=np.maximum(scale, 1e-30))
    else:
        _assert_fp32c(got, exact)


def _assert_fp32c(got, exp, signed=False, scale=None, tol=1e-6):
    got = np.asarray(got, dtype=np.float64)
    exp = np.asarray(exp, dtype=np.float64)
    same = got == exp  # covers equal infinities (PCFs ending at different values)
    with np.errstate(invalid="ignore"):
        if signed:
            s = scale if scale is not None else np.maximum(np.abs(exp), 1e-30)
            err = np.where(same, 0.0, np.abs(got - exp) / s)
        else:
            err = np.where(same, 0.0, np.abs(got - exp) / np.where(exp == 0, 1.0, np.abs(exp)))
    assert not np.any(np.isnan(err)), "non-finite mismatch"
    assert err.max() <= tol, f"max rel err {err.max():.3e}"
    return float(err.max())


@pytest.mark.parametrize("oname", ["l1", "l2"])
def test_cdist_matches_reference_vectors(gpu, oname):
    X = _pcfs(REF["cx_off"], REF["cx_t"], REF["cx_v"])
    Y = _pcfs(REF["cy_off"], REF["cy_t"], REF["cy_v"])
    D = np.asarray(sb.cdist(X, Y, p=_OPS[oname])).reshape(-1)
    exp = REF[f"cdist_{oname}_f64"]
    if oname == "l1":
        np.testing.assert_array_equal(D, exp)
    else:
        np.testing.assert_allclose(D, exp, rtol=1e-13)


# ----------------------------------------------- oracle at scale ----------

def test_c1_full_fp64_bitwise_and_fp32c(gpu):
    off, t, v = workloads.c1()
    X = _pcfs(off, t, v)
    exp = oracle.pairwise(oracle.OP_L1, 1.0, oracle.KIND_PDIST, off, t, v)
    np.testing.assert_array_equal(sb.pdist(X).storage, exp)
    _assert_fp32c(sb.pdist(X, precision="fp32c").storage, exp)


@pytest.mark.parametrize("n", [15, 16, 17, 31, 33, 257])
def test_tile_boundary_sizes(gpu, n):
    off, t, v = workloads.c2(n, seed=n)
    X = _pcfs(off, t, v)
    np.testing.assert_array_equal(sb.pdist(X).storage,
                                  oracle.pairwise(oracle.OP_L1, 1.0, 0, off, t, v))
    np.testing.assert_array_equal(sb.l2_kernel(X).storage,
                                  oracle.pairwise(oracle.OP_DOT, 1.0, 1, off, t, v))


def test_c2_subset_all_forms(gpu):
    off, t, v = workloads.c2(1200)
    X = _pcfs(off, t, v)
    for p, op in ((1, oracle.OP_L1), (2, oracle.OP_LP)):
        exp = oracle.pairwise(op, float(p), oracle.KIND_PDIST, off, t, v)
        got = sb.pdist(X, p=p).storage
        if p == 1:
            np.testing.assert_array_equal(got, exp)
        else:
            np.testing.assert_allclose(got, exp, rtol=1e-13)
        _assert_fp32c(sb.pdist(X, p=p, precision="fp32c").storage, exp)
    K = oracle.pairwise(oracle.OP_DOT, 1.0, oracle.KIND_KERNEL, off, t, v)
    np.testing.assert_array_equal(sb.l2_kernel(X).storage, K)
    absX = _pcfs(off, t, np.abs(v))
    Kabs = oracle.pairwise(oracle.OP_DOT, 1.0, oracle.KIND_KERNEL, off, t, np.abs(v))
    _assert_fp32c(sb.l2_kernel(X, precision="fp32c").storage, K, signed=True, scale=Kabs)
    del absX


def test_long_pcfs_many_phases_and_mixed_lengths(gpu):
    """m from 2 to 6000 in one set: many LDS phases, long/short pairs in a tile."""
    rng = np.random.default_rng(77)
    ms = np.concatenate([[2, 3, 6000, 4500], rng.integers(2, 3000, size=40)])
    pcfs = []
    for m in ms:
        tt = np.concatenate([[0.0], np.cumsum(rng.exponential(1.0, size=m - 1))])
        vv = rng.standard_normal(m)
        vv[-1] = 0.0
        pcfs.append(np.stack([tt, vv], 1))
    off = np.concatenate([[0], np.cumsum(ms)]).astype(np.int64)
    t = np.concatenate([p[:, 0] for p in pcfs])
    v = np.concatenate([p[:, 1] for p in pcfs])
    X = _pcfs(off, t, v)
    exp = oracle.pairwise(oracle.OP_L1, 1.0, 0, off, t, v)
    np.testing.assert_array_equal(sb.pdist(X).storage, exp)
    _assert_fp32c(sb.pdist(X, precision="fp32c").storage, exp)


def test_cdist_long_pairs_fp32c_with_fp64_fold(gpu):
    X, Y = workloads.c5(40, m_lo=500, m_hi=5000)
    exp = oracle.pairwise(oracle.OP_L1, 1.0, oracle.KIND_CDIST, *X, *Y)
    D = np.asarray(sb.cdist(_pcfs(*X), _pcfs(*Y))).reshape(-1)
    np.testing.assert_array_equal(D, exp)
    D32 = np.asarray(sb.cdist(_pcfs(*X), _pcfs(*Y), precision="fp32c")).reshape(-1)
    _assert_fp32c(D32, exp)


def test_duplicate_times_ties_and_zero_pcfs(gpu):
    X = _pcfs(REF["edge_off"], REF["edge_t"], REF["edge_v"])
    # empty PCFs are the zero PCF
    Z = sb.PcfTensor([sb.Pcf(np.zeros((0, 2))), sb.Pcf(np.array([[0.0, 2.0], [3.0, 0.0]]))])
    assert sb.pdist(Z)[0, 1] == pytest.approx(6.0)
    np.testing.assert_array_equal(sb.pdist(X).storage, REF["edge_l1_f64"])
    # identical PCFs give exact zeros
    same = sb.PcfTensor([X[0], X[0], X[0]])
    assert np.all(sb.pdist(same).storage == 0.0)


# ------------------------------------------------- drop-in API ------------

def test_pdist_api_semantics(gpu):
    X = sb.zeros((2,), dtype=sb.pcf64)
    X[0] = sb.Pcf(np.array([[0.0, 10.0], [2.0, 5.0], [3.0, 0.0]]))
    X[1] = sb.Pcf(np.array([[0.0, 5.0], [6.0, 0.0]]))
    D = sb.pdist(X)
    assert D.size == 2 and D[0, 0] == 0.0 and D[1, 1] == 0.0
    assert D[0, 1] == pytest.approx(25.0) and D[1, 0] == D[0, 1]
    dense = D.to_dense()
    assert dense.shape == (2, 2) and dense[0, 1] == pytest.approx(25.0)
    Xz = sb.zeros((10, 20))
    sb.pdist(Xz[:, 2])
    sb.l2_kernel(Xz[:, 2])
    assert isinstance(sb.pdist(sb.zeros((3,)), p=3), sb.DistanceMatrix)
    f = sb.Pcf(np.array([[0.0, 10.0], [2.0, 5.0], [3.0, 0.0]]))
    g = sb.Pcf(np.array([[0.0, 5.0], [6.0, 0.0]]))
    assert sb.pdist([f, g])[0, 1] == pytest.approx(25.0)
    C = sb.cdist([sb.Pcf(np.array([[0.0, 1.0], [1.0, 0.0]]))],
                 [sb.Pcf(np.array([[0.0, 2.0], [1.0, 0.0]]))])
    assert C.shape == (1, 1) and float(np.asarray(C).flat[0]) == pytest.approx(1.0)


def _assert_cdist_matches_lp(D, X, Y, p=1):
    import itertools

    for xi in itertools.product(*[range(s) for s in X.shape]):
        for yi in itertools.product(*[range(s) for s in Y.shape]):
            assert D[xi + yi] == pytest.approx(sb.lp_distance(X[xi], Y[yi], p=p))


def test_cdist_multidim_sliced_and_lp(gpu, pcf_dtype):
    dt = np.float32 if pcf_dtype is sb.pcf32 else np.float64
    X_full = sb.zeros((4, 6), dtype=pcf_dtype)
    Y_full = sb.zeros((3, 8), dtype=pcf_dtype)
    for i in range(4):
        for j in range(6):
            X_full[i, j] = sb.Pcf(np.array([[0.0, float(i * 6 + j + 1)], [1.0, 0.0]], dtype=dt))
    for i in range(3):
        for j in range(8):
            Y_full[i, j] = sb.Pcf(np.array([[0.0, float(i * 8 + j)], [1.0, 0.0]], dtype=dt))
    X, Y = X_full[::2, 1::2], Y_full[:, ::3]
    D = sb.cdist(X, Y)
    assert D.shape == (2, 3, 3, 3)
    _assert_cdist_matches_lp(D, X, Y)
    D3 = sb.cdist(X, Y, p=3)
    _assert_cdist_matches_lp(D3, X, Y, p=3)
    Xs = X_full[:, 0]
    Dp = sb.pdist(Xs)
    Dc = sb.cdist(Xs, Xs)
    for i in range(4):
        for j in range(4):
            assert Dc[i, j] == pytest.approx(Dp[i, j], abs=1e-5)


def test_reversed_and_strided_views(gpu, pcf_dtype):
    dt = np.float64 if pcf_dtype is sb.pcf64 else np.float32
    X = sb.zeros((5,), dtype=pcf_dtype)
    for i in range(5):
        X[i] = sb.Pcf(np.array([[0.0, float(i + 1)], [2.0, 0.0]], dtype=dt))
    K = sb.l2_kernel(X[:3][::-1]).to_dense()
    assert np.allclose(K, [[18, 12, 6], [12, 8, 4], [6, 4, 2]])
    D = sb.pdist(X[:3][::-1]).to_dense()
    assert np.allclose(D, [[0, 2, 4], [2, 0, 2], [4, 2, 0]])
    v = np.array([1.0, 3.0, 5.0])
    assert np.allclose(sb.l2_kernel(X[::2]).to_dense(), 2.0 * np.outer(v, v))


def test_verbose_progress_and_precision_switch(gpu, capsys):
    off, t, v = workloads.c2(300)
    X = _pcfs(off, t, v)
    sb.system.set_device_verbose(True)
    try:
        D = sb.pdist(X, verbose=True)
    finally:
        sb.system.set_device_verbose(False)
    assert "HIP" in capsys.readouterr().out
    sb.system.set_precision("fp32c")
    try:
        D32 = sb.pdist(X)
    finally:
        sb.system.set_precision("fp64")
    _assert_fp32c(D32.storage, D.storage)
