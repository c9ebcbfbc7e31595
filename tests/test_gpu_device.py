"""GPU tests of the device-level C-ABI paths: device-resident sets and plans,
tile shapes and the autotuner, persistent launches, hipGraph replays, the
batched launch, the plain FP64 baseline kernel, cdist row sums, and the
asynchronous job protocol (progress and cancellation)."""

import numpy as np
import pytest
import torch

import oracle
import stablebear_amd as sb
from stablebear_amd import _lib as L
from stablebear_amd import runtime, workloads
from stablebear_amd.batched import BatchedPlan

pytestmark = pytest.mark.gpu


def _plan(off, t, v, prec="fp64", op="l1", p=1.0, shape=None):
    ds = runtime.DeviceSet(runtime.csr_from_arrays(off, t, v), prec)
    return ds, runtime.Plan(ds, op=op, p=p, shape=shape)


def _run(plan, f64):
    out = torch.empty(plan.out_count, dtype=torch.float64 if f64 else torch.float32, device="cuda")
    plan.execute(out.data_ptr(), f64, stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("shape", [(16, 16, 24, 1), (16, 16, 48, 1), (8, 8, 32, 1), (16, 16, 2, 1),
                                   (4, 4, 7, 1), (32, 32, 48, 2), (32, 32, 64, 2), (16, 16, 24, 2),
                                   (8, 8, 32, 2), (16, 16, 2, 2), (4, 4, 7, 2), (5, 5, 9, 2),
                                   (32, 16, 64, 1), (16, 32, 48, 1), (8, 32, 16, 1), (32, 8, 7, 1),
                                   (24, 16, 9, 1), (32, 16, 24, 2), (16, 8, 5, 2), (5, 9, 3, 1)])
def test_tile_shapes_all_bit_exact(gpu, shape):
    """Square and rectangular tiles of the pdist triangle (rectangular tiles
    that reach the diagonal skip the pairs on or above it)."""
    off, t, v = workloads.c2(90, seed=3)
    exp = oracle.pairwise(oracle.OP_L1, 1.0, oracle.KIND_PDIST, off, t, v)
    ds, pl = _plan(off, t, v, shape=shape)
    assert pl.shape == shape[:3] and pl.lane_pairs == shape[3]
    np.testing.assert_array_equal(_run(pl, True), exp)


@pytest.mark.parametrize("shape", [(32, 16, 64, 1), (16, 32, 48, 1), (7, 24, 5, 1), (16, 8, 9, 2)])
def test_rectangular_tiles_of_the_kernel_matrix_bit_exact(gpu, shape):
    """The kernel matrix (diagonal included) with rectangular tiles, FP64
    against the oracle bit for bit, and FP32C within 1e-6 normwise."""
    off, t, v = workloads.c2(77, seed=13)
    exp = oracle.pairwise(oracle.OP_DOT, 1.0, oracle.KIND_KERNEL, off, t, v)
    ds, pl = _plan(off, t, v, op="dot", shape=shape)
    np.testing.assert_array_equal(_run(pl, True), exp)
    ds32, pl32 = _plan(off, t, v, prec="fp32c", op="dot", shape=shape)
    got = _run(pl32, False).astype(np.float64)
    assert np.linalg.norm(got - exp) <= 1e-6 * np.linalg.norm(exp)


def test_persistent_launch_matches(gpu):
    off, t, v = workloads.c2(300, seed=4)
    exp = oracle.pairwise(oracle.OP_L1, 1.0, oracle.KIND_PDIST, off, t, v)
    ds, pl = _plan(off, t, v)
    pl.set_persistent(True)
    np.testing.assert_array_equal(_run(pl, True), exp)
    pl.set_persistent(False)
    np.testing.assert_array_equal(_run(pl, True), exp)


def test_execute_to_host_and_fold_threshold(gpu):
    off, t, v = workloads.c2(120, seed=5)
    exp = oracle.pairwise(oracle.OP_LP, 2.0, oracle.KIND_PDIST, off, t, v)
    ds, pl = _plan(off, t, v, prec="fp32c", op="lp", p=2.0)
    got = pl.execute_to_host(False).astype(np.float64)
    assert np.max(np.abs(got - exp) / exp) <= 1e-6
    pl.set_fold_threshold(2)  # fold every pair's partial sums into FP64 each phase
    got2 = pl.execute_to_host(False).astype(np.float64)
    assert np.max(np.abs(got2 - exp) / exp) <= 1e-6


def test_autotune_picks_a_valid_shape(gpu):
    off, t, v = workloads.c2(400, seed=6)
    exp = oracle.pairwise(oracle.OP_L1, 1.0, oracle.KIND_PDIST, off, t, v)
    ds, pl = _plan(off, t, v, prec="fp32c")
    out = torch.empty(pl.out_count, dtype=torch.float32, device="cuda")
    shape, ms = pl.autotune(out.data_ptr(), False, reps=1)
    assert ms > 0 and pl.shape == shape
    got = _run(pl, False).astype(np.float64)
    assert np.max(np.abs(got - exp) / exp) <= 1e-6


def test_per_class_autotune_bit_exact_and_cached(gpu):
    """PCFs of several length classes (5..3000 records): the autotuner picks a
    shape per class, each class's tiles run in a launch of their own, FP64
    results stay bit-exact, and a second plan of the same shape of data takes
    every class's shape from the cache."""
    rng = np.random.default_rng(51)
    m = np.clip(np.round(rng.lognormal(4.5, 1.2, 600)), 5, 3000).astype(np.int64)
    off, t = workloads._times(rng, m, 1.0)
    v = rng.integers(-50, 51, size=int(off[-1])).astype(np.float64)
    v[off[1:] - 1] = 0.0
    exp = oracle.pairwise(oracle.OP_L1, 1.0, oracle.KIND_PDIST, off, t, v)
    ds, pl = _plan(off, t, v)
    # every class its own shape and launch
    pl.set_class_shapes([(32, 32, 64, 2), (16, 16, 48, 1), (16, 16, 96, 1), (8, 8, 32, 2)])
    cls = pl.class_shapes()
    assert [c[:4] for c in cls] == [(32, 32, 64, 2), (16, 16, 48, 1), (16, 16, 96, 1), (8, 8, 32, 2)]
    used = [c for c in cls if c[4] > 0]
    assert len(used) >= 3 and sum(c[4] for c in cls) == pl.tiles
    np.testing.assert_array_equal(_run(pl, True), exp)
    # the autotuner times shapes per class (or keeps one launch of one shape
    # when that is faster); a second plan of the same data hits the cache
    out = torch.empty(pl.out_count, dtype=torch.float64, device="cuda")
    h0, m0 = runtime.autotune_cache_stats()
    pl.autotune(out.data_ptr(), True, reps=1)
    h1, m1 = runtime.autotune_cache_stats()
    assert m1 - m0 == len(used) and h1 == h0
    np.testing.assert_array_equal(_run(pl, True), exp)
    ds2, pl2 = _plan(off, t, v)
    pl2.autotune(out.data_ptr(), True, reps=1)
    h2, m2 = runtime.autotune_cache_stats()
    assert h2 - h1 == len(used) and m2 == m1
    np.testing.assert_array_equal(_run(pl2, True), exp)
    pl2.set_shape(16, 16, 48, 1)  # back to one shape for every class
    assert all(c[4] == 0 for c in pl2.class_shapes()[1:])
    np.testing.assert_array_equal(_run(pl2, True), exp)


def test_graph_replays_match_eager(gpu):
    off, t, v = workloads.c2(250, seed=8)
    ds, pl = _plan(off, t, v, prec="fp32c")
    eager = _run(pl, False)
    out = torch.zeros(pl.out_count, dtype=torch.float32, device="cuda")
    g = pl.graph(out.data_ptr(), False)
    for k in range(4):
        out.fill_(-1.0)
        torch.cuda.synchronize()
        g.launch(torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(out.cpu().numpy(), eager)
        _ = torch.randn(1000, device="cuda").sum().item()  # eager work between replays


def test_stale_graph_is_refused(gpu):
    """A graph binds its plan's tile list and launch parameters: after the
    plan is reshaped, retuned or closed, launching it raises instead of
    reading freed buffers; a new capture works."""
    off, t, v = workloads.c2(120, seed=18)
    exp = oracle.pairwise(oracle.OP_L1, 1.0, oracle.KIND_PDIST, off, t, v)
    ds, pl = _plan(off, t, v)
    out = torch.zeros(pl.out_count, dtype=torch.float64, device="cuda")
    g = pl.graph(out.data_ptr(), True)
    pl.set_shape(8, 8, 16, 1)
    with pytest.raises(ValueError, match="stale"):
        g.launch(torch.cuda.current_stream().cuda_stream)
    g2 = pl.graph(out.data_ptr(), True)
    g2.launch(torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(out.cpu().numpy(), exp)
    pl.set_persistent(True)
    with pytest.raises(ValueError, match="stale"):
        g2.launch(torch.cuda.current_stream().cuda_stream)
    g3 = pl.graph(out.data_ptr(), True)
    pl.close()
    with pytest.raises(ValueError, match="stale"):
        g3.launch(torch.cuda.current_stream().cuda_stream)


def test_batched_plan_sees_value_updates(gpu):
    """Batched plans copy their members' records; value updates of a member
    set reach both eager executes and replays of an earlier capture."""
    probs = [workloads.c3_problem(b, n=int(n)) for b, n in enumerate([40, 75, 9])]
    bp = BatchedPlan(probs, op="l1", precision="fp64")
    out = torch.zeros(bp.out_count, dtype=torch.float64, device="cuda")
    g = bp.graph(out.data_ptr(), True)
    off, t, v = probs[1]
    v2 = v * 0.5 + 3.0
    v2[off[1:] - 1] = 0.0
    d_v2 = torch.from_numpy(v2).cuda()
    bp.sets[1].update_values(d_v2.data_ptr(), torch.cuda.current_stream().cuda_stream)
    exps = [oracle.pairwise(oracle.OP_L1, 1.0, oracle.KIND_PDIST, *pr) for pr in probs]
    exps[1] = oracle.pairwise(oracle.OP_L1, 1.0, oracle.KIND_PDIST, off, t, v2)
    for launch in (lambda: g.launch(torch.cuda.current_stream().cuda_stream),
                   lambda: bp.execute(out.data_ptr(), True, torch.cuda.current_stream().cuda_stream)):
        out.fill_(-1.0)
        launch()
        torch.cuda.synchronize()
        for got, exp in zip(bp.split(out.cpu().numpy()), exps):
            np.testing.assert_array_equal(got, exp)


def test_batched_persistent_graph_matches_oracle(gpu):
    probs = [workloads.c3_problem(b, n=int(n)) for b, n in enumerate([70, 5, 133, 1, 64, 200, 2])]
    bp = BatchedPlan(probs, op="l1", precision="fp32c")
    assert bp.plan.problems == len(probs)
    out = torch.zeros(bp.out_count, dtype=torch.float32, device="cuda")
    g = bp.graph(out.data_ptr(), False)
    for _ in range(3):
        out.fill_(-7.0)
        g.launch(torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        parts = bp.split(out.cpu().numpy())
        for (off, t, v), got in zip(probs, parts):
            exp = oracle.pairwise(oracle.OP_L1, 1.0, oracle.KIND_PDIST, off, t, v)
            assert got.shape == exp.shape
            if exp.size:
                nz = exp != 0
                assert np.max(np.abs(got[nz] - exp[nz]) / exp[nz]) <= 1e-6


def test_batched_python_api_fp64_bit_exact(gpu):
    probs = [workloads.c3_problem(b, n=n) for b, n in enumerate([40, 3, 77])]
    Xs = [sb.PcfTensor(workloads.to_pcfs(*pr)) for pr in probs]
    Ds = sb.pdist_batched(Xs)
    Ks = sb.l2_kernel_batched(Xs)
    for (off, t, v), D, K in zip(probs, Ds, Ks):
        np.testing.assert_array_equal(D.storage, oracle.pairwise(oracle.OP_L1, 1.0, 0, off, t, v))
        np.testing.assert_array_equal(K.storage, oracle.pairwise(oracle.OP_DOT, 1.0, 1, off, t, v))
    D2 = sb.pdist_batched(Xs, p=2)
    for (off, t, v), D in zip(probs, D2):
        np.testing.assert_allclose(D.storage, oracle.pairwise(oracle.OP_LP, 2.0, 0, off, t, v),
                                   rtol=1e-13)


@pytest.mark.parametrize("op,kind,out_kind", [("l1", 0, L.OUT_DISTMAT), ("dot", 1, L.OUT_SYMMAT)])
def test_plain_fp64_kernel_bit_exact(gpu, op, kind, out_kind):
    off, t, v = workloads.c2(150, seed=9)
    pk = runtime.PlainKernel(runtime.csr_from_arrays(off, t, v))
    n = len(off) - 1
    cnt = oracle.out_count(kind, n)
    out = torch.empty(cnt, dtype=torch.float64, device="cuda")
    pk.execute(op, 1.0, out_kind, out.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    exp = oracle.pairwise(oracle.OP_L1 if op == "l1" else oracle.OP_DOT, 1.0, kind, off, t, v)
    np.testing.assert_array_equal(out.cpu().numpy(), exp)


def test_plain_kernel_on_duplicates_and_ties(gpu):
    ref = np.load(__import__("os").path.join(__import__("os").path.dirname(__file__), "golden",
                                             "ref_vectors.npz"))
    off, t, v = ref["edge_off"], ref["edge_t"], ref["edge_v"]
    pk = runtime.PlainKernel(runtime.csr_from_arrays(off, t, v))
    out = torch.empty(len(ref["edge_l1_f64"]), dtype=torch.float64, device="cuda")
    pk.execute("l1", 1.0, L.OUT_DISTMAT, out.data_ptr(), None)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(out.cpu().numpy(), ref["edge_l1_f64"])


def test_cdist_row_sums(gpu):
    X, Y = workloads.c5(60, m_lo=50, m_hi=400)
    exp = oracle.pairwise(oracle.OP_L1, 1.0, oracle.KIND_CDIST, *X, *Y).reshape(60, 60)
    sx = runtime.DeviceSet(runtime.csr_from_arrays(*X), "fp32c")
    sy = runtime.DeviceSet(runtime.csr_from_arrays(*Y), "fp32c")
    pl = runtime.Plan(sx, sy, op="l1")
    out = torch.empty(pl.out_count, dtype=torch.float32, device="cuda")
    rs = torch.full((60,), 123.0, dtype=torch.float64, device="cuda")
    pl.execute(out.data_ptr(), False, rs.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    D = out.cpu().numpy().reshape(60, 60).astype(np.float64)
    assert np.max(np.abs(D - exp) / exp) <= 1e-6
    np.testing.assert_allclose(rs.cpu().numpy() / 60, exp.mean(axis=1), rtol=1e-6)


def test_job_progress_and_cancel(gpu):
    off, t, v = workloads.c2(3000, seed=10)
    csr = runtime.csr_from_arrays(off, t, v)
    n = 3000
    out = np.zeros(n * (n - 1) // 2)
    task = runtime.submit(csr, None, L.OP_L1, 1.0, L.OUT_DISTMAT, L.PREC_FP64, out, "d", len(out))
    task.wait()
    assert task.work_total() == len(out)
    np.testing.assert_array_equal(out, oracle.pairwise(oracle.OP_L1, 1.0, 0, off, t, v))
    out2 = np.zeros_like(out)
    task2 = runtime.submit(csr, None, L.OP_L1, 1.0, L.OUT_DISTMAT, L.PREC_FP64, out2, "d", len(out))
    task2.request_stop()
    with pytest.raises(L.CancelledError):
        task2.wait()


@pytest.mark.parametrize("kind", ["pdist", "kernel", "cdist"])
def test_job_bands_under_small_output_budget(gpu, kind):
    """An output far above the output budget runs in bands of rows (each a
    cross plan against the earlier rows plus the band's own pairs, downloads
    overlapped); FP64 results stay bit-exact with the oracle, FP32C within
    1e-6, and progress reaches the total."""
    off, t, v = workloads.c2(1500, seed=41)
    csr = runtime.csr_from_arrays(off, t, v)
    n = 1500
    if kind == "cdist":
        yo, yt, yv = workloads.c2(700, seed=42)
        cols = runtime.csr_from_arrays(yo, yt, yv)
        total = n * 700
        op, okind, out_kind = L.OP_L1, oracle.KIND_CDIST, L.OUT_DENSE
        exp = oracle.pairwise(oracle.OP_L1, 1.0, okind, off, t, v, yo, yt, yv)
    elif kind == "kernel":
        cols, total = None, n * (n + 1) // 2
        op, okind, out_kind = L.OP_DOT, oracle.KIND_KERNEL, L.OUT_SYMMAT
        exp = oracle.pairwise(oracle.OP_DOT, 1.0, okind, off, t, v)
    else:
        cols, total = None, n * (n - 1) // 2
        op, okind, out_kind = L.OP_L1, oracle.KIND_PDIST, L.OUT_DISTMAT
        exp = oracle.pairwise(oracle.OP_L1, 1.0, okind, off, t, v)
    try:
        runtime.set_output_budget(total * 8 // 9)  # three buffers: bands of ~1/27 of the output
        out = np.full(total, np.nan)
        task = runtime.submit(csr, cols, op, 1.0, out_kind, L.PREC_FP64, out, "d", total)
        task.wait()
        assert task.work_completed() == total
        np.testing.assert_array_equal(out, exp)
        out32 = np.full(total, np.nan, dtype=np.float32)
        task = runtime.submit(csr, cols, op, 1.0, out_kind, L.PREC_FP32C, out32, "d", total)
        task.wait()
        if kind == "kernel":
            scale = oracle.pairwise(oracle.OP_DOT, 1.0, okind, off, t, np.abs(v))
            assert np.max(np.abs(out32 - exp) / scale) <= 1e-6
        else:
            assert np.max(np.abs(out32 - exp) / exp) <= 1e-6
    finally:
        runtime.set_output_budget(0)


def _aos(off, t, v):
    """A CSR set whose times and values are strided views of (m, 2) points,
    as a list of Pcf arrays concatenates to."""
    pts = np.ascontiguousarray(np.stack([t, v], axis=1))
    return runtime.Csr(np.asarray(off, dtype=np.int64), pts[:, 0], pts[:, 1], L.F64)


@pytest.mark.parametrize("kind,budget", [("pdist", 0), ("kernel", 0), ("cdist", 0), ("pdist", 9), ("cdist", 9)])
def test_job_from_strided_points_matches(gpu, kind, budget):
    """Jobs handed AoS points as strided views (sbamd_pairwise_submit_view)
    give the results of the contiguous-array jobs, bit for bit in FP64, with
    one plan or in bands (a small output budget)."""
    n = 700
    off, t, v = workloads.c2(n, seed=44)
    rows = _aos(off, t, v)
    assert rows.strided
    cols = None
    if kind == "cdist":
        yo, yt, yv = workloads.c2(300, seed=45)
        cols = _aos(yo, yt, yv)
        total, op, out_kind = n * 300, L.OP_L1, L.OUT_DENSE
        exp = oracle.pairwise(oracle.OP_L1, 1.0, oracle.KIND_CDIST, off, t, v, yo, yt, yv)
    elif kind == "kernel":
        total, op, out_kind = n * (n + 1) // 2, L.OP_DOT, L.OUT_SYMMAT
        exp = oracle.pairwise(oracle.OP_DOT, 1.0, oracle.KIND_KERNEL, off, t, v)
    else:
        total, op, out_kind = n * (n - 1) // 2, L.OP_L1, L.OUT_DISTMAT
        exp = oracle.pairwise(oracle.OP_L1, 1.0, oracle.KIND_PDIST, off, t, v)
    try:
        runtime.set_output_budget(total * 8 // budget if budget else 0)
        out = np.full(total, np.nan)
        task = runtime.submit(rows, cols, op, 1.0, out_kind, L.PREC_FP64, out, "d", total)
        task.wait()
        np.testing.assert_array_equal(out, exp)
    finally:
        runtime.set_output_budget(0)


@pytest.mark.parametrize("bad,msg", [("order", "non-decreasing"), ("start", "t=0"), ("inf", "finite")])
def test_job_from_strided_points_reports_bad_points(gpu, bad, msg):
    """A strided job checks the points on the device: the wait reports the
    first offending PCF with the host rules' message."""
    off, t, v = workloads.c2(400, seed=46)
    t = t.copy()
    i = 123
    if bad == "order":
        t[off[i] + 3] = t[off[i] + 1]
    elif bad == "start":
        t[off[i]] = 1e-9
    else:
        t[off[i] + 2] = np.inf
    rows = _aos(off, t, v)
    total = 400 * 399 // 2
    out = np.zeros(total)
    task = runtime.submit(rows, None, L.OP_L1, 1.0, L.OUT_DISTMAT, L.PREC_FP64, out, "d", total)
    with pytest.raises(ValueError, match=msg):
        task.wait()


def test_device_set_value_update(gpu):
    off, t, v = workloads.c2(80, seed=12)
    ds, pl = _plan(off, t, v)
    v2 = v * 1.5 - 0.25
    v2[off[1:] - 1] = 0.0
    vd = torch.from_numpy(v2).cuda()
    ds.update_values(vd.data_ptr(), torch.cuda.current_stream().cuda_stream)
    np.testing.assert_array_equal(_run(pl, True), oracle.pairwise(oracle.OP_L1, 1.0, 0, off, t, v2))


def test_bad_inputs_raise(gpu):
    with pytest.raises(ValueError, match="non-decreasing"):
        runtime.DeviceSet(runtime.csr_from_arrays([0, 3], [0.0, 2.0, 1.0], [1.0, 1.0, 0.0]))
    with pytest.raises(ValueError, match="p must be"):
        off, t, v = workloads.c1(5)
        _plan(off, t, v, op="lp", p=0.5)
